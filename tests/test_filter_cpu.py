"""The mixed-precision triangle filter never rejects a ray that the FP64
Moller-Trumbore test hits (CPU, randomized adversarial cases near edges,
vertices, the grazing band and both segment ends).  The same check with the
margins forced to zero finds violations, so the cases exercise the bounds."""
import ctypes as C
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fcheck(tmp_path_factory):
    out = tmp_path_factory.mktemp("filter") / "libfcheck.so"
    src = os.path.join(HERE, "support", "filter_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", src, "-o", str(out)],
                   check=True)
    lib = C.CDLL(str(out))
    lib.filter_check.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.POINTER(C.c_int64)]
    lib.filter_check.restype = C.c_int64
    return lib


def test_filter_keeps_every_fp64_hit(fcheck):
    hits = C.c_int64()
    bad = fcheck.filter_check(0x5EED00F1, 2_000_000, 0, C.byref(hits))
    assert hits.value > 100_000
    assert bad == 0


def test_zero_margins_are_caught(fcheck):
    hits = C.c_int64()
    assert fcheck.filter_check(0x5EED00F1, 2_000_000, 1, C.byref(hits)) > 0
