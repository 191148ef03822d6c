"""The mixed-precision triangle classifier never calls a ray that the FP64
Moller-Trumbore test hits a certain miss, nor one that it misses a certain
hit (CPU, randomized adversarial cases near edges, vertices, the grazing band
and both segment ends).  The same check with the margins forced to zero
finds violations of both kinds, so the cases exercise the bounds."""
import ctypes as C
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
P = C.POINTER(C.c_int64)


@pytest.fixture(scope="module")
def fcheck(tmp_path_factory):
    out = tmp_path_factory.mktemp("filter") / "libfcheck.so"
    src = os.path.join(HERE, "support", "filter_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", src, "-o", str(out)],
                   check=True)
    lib = C.CDLL(str(out))
    lib.filter_check.argtypes = [C.c_uint64, C.c_int64, C.c_int, P, P, P]
    lib.filter_check.restype = C.c_int64
    return lib


def _run(lib, mode):
    hits, sure, false_sure = C.c_int64(), C.c_int64(), C.c_int64()
    bad = lib.filter_check(0x5EED00F1, 2_000_000, mode, C.byref(hits), C.byref(sure), C.byref(false_sure))
    return bad, hits.value, sure.value, false_sure.value


def test_classifier_is_conservative(fcheck):
    bad, hits, sure, false_sure = _run(fcheck, 0)
    assert hits > 100_000
    assert bad == 0          # no FP64 hit called a certain miss
    assert false_sure == 0   # no FP64 miss called a certain hit
    assert sure > hits // 4  # most clear hits need no FP64 test


def test_zero_margins_are_caught(fcheck):
    bad, hits, sure, false_sure = _run(fcheck, 1)
    assert bad > 0 and false_sure > 0


@pytest.fixture(scope="module")
def fadv(fcheck):
    fcheck.filter_check_adv.argtypes = [C.c_uint64, C.c_int64, C.c_double, P, P, P]
    fcheck.filter_check_adv.restype = C.c_int64
    fcheck.filter_check_sray.argtypes = [C.c_uint64, C.c_int64, P, P, P]
    fcheck.filter_check_sray.restype = C.c_int64
    return fcheck


def _adv(lib, scale, n=300_000):
    fs, near, hits = C.c_int64(), C.c_int64(), C.c_int64()
    fm = lib.filter_check_adv(0x5EED00F2, n, scale, C.byref(fs), C.byref(near), C.byref(hits))
    return fm, fs.value, near.value, hits.value


def test_boundary_cases_with_inputs_at_their_bounds(fadv):
    # FP64 decisions a tiny distance from each threshold (U = 0, V = 0,
    # U + V = a, T = 0, T = L a, |a| ~ 1e-10), FP32 origin, direction and
    # length moved to the full extent of their error bounds against the
    # decision: no violation with the margins as built, nor with half of
    # them (margins equal to the first-order bounds)
    for scale in (1.0, 0.5):
        fm, fs, near, hits = _adv(fadv, scale)
        assert near > 150_000 and hits > 50_000
        assert fm == 0 and fs == 0, (scale, fm, fs)


def test_boundary_search_finds_violations_below_the_bounds(fadv):
    # the same search with the margins at a quarter of the built ones (half
    # the first-order bounds) finds FP64 hits called certain misses, and at a
    # tenth certain hits the FP64 test misses too: the cases reach the bounds
    fm, fs, _, _ = _adv(fadv, 0.25)
    assert fm > 0
    fm, fs, _, _ = _adv(fadv, 0.1)
    assert fm > 0 and fs > 0


def test_rays_formed_as_the_kernel_forms_them(fadv):
    # anchored FP32 offsets, fused r2, reciprocal square root 1 ulp off either
    # way: the classifier against the FP64 ray from the FP64 points
    fs, sure, hits = C.c_int64(), C.c_int64(), C.c_int64()
    fm = fadv.filter_check_sray(0x5EED00F3, 1_000_000, C.byref(fs), C.byref(sure), C.byref(hits))
    assert hits.value > 100_000 and sure.value > 1_000
    assert fm == 0 and fs.value == 0
