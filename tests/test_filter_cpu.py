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
