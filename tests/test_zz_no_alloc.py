"""The no-allocation entry point (rvf_compute_view_factors_noalloc) against
rvf_compute_view_factors, through the C ABI.

Both run the same kernels in the same order, so every result is expected bit
for bit equal: F (written to a caller-owned device buffer), A, F_group,
A_group, group tags and the traversal counters.  Each case makes two calls in a
row on the shared static scratch array, the second with a smaller input, so a
value left over from the first call would show in the second.
"""
import ctypes as C

import numpy as np
import pytest

from rvf_amd import _lib, api
from rvf_amd import mesh as M

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20  # RVF_NOALLOC_SCRATCH_BYTES in include/rvf.h


def _opts(nquad, method="quad", precision="fp64", self_vf=False, flags=0):
    o = _lib.Opts()
    o.nquad = nquad
    o.method = _lib.METHODS[method]
    o.precision = _lib.PRECISIONS[precision]
    o.n_samples = 64
    o.seed = 0x5EED
    o.flags = flags | (_lib.RVF_SELF_VF if self_vf else 0)
    o.variant = 0
    return o


def _device(mesh, obs, parity):
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, parity) if len(obs) else None
    return ctx, dm, do


def _outputs(G):
    return (np.empty(0), np.empty((G, G), order="F"), np.empty(G), np.empty(G, np.int32), _lib.Stats())


def solve_alloc(mesh, nquad, obs=(), parity="geometric", **kw):
    ctx, dm, do = _device(mesh, obs, parity)
    N, G = mesh.n_elems, len(dm.tags)
    F = np.empty((N, N), order="F")
    A = np.empty(N)
    _, Fg, Ag, tags, st = _outputs(G)
    o = _opts(nquad, **kw)
    _lib.check(_lib.lib().rvf_compute_view_factors(ctx.h, dm.h, do.h if do else None, C.byref(o), _lib.ptr(F),
                                                   _lib.ptr(A), _lib.ptr(Fg), _lib.ptr(Ag), _lib.ptr(tags),
                                                   C.byref(st)))
    dm.close()
    if do:
        do.close()
    return dict(F=F, A=A, Fg=Fg, Ag=Ag, tags=tags, st=st.as_dict())


def _hip():
    # the caller's device buffer for F comes from the HIP runtime the library
    # itself links (libamdhip64.so.7 is already loaded with it)
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    for f in (h.hipMalloc, h.hipFree, h.hipMemset, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


def solve_noalloc(mesh, nquad, obs=(), parity="geometric", **kw):
    ctx, dm, do = _device(mesh, obs, parity)
    N, G = mesh.n_elems, len(dm.tags)
    hip = _hip()
    F_dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(F_dev), 8 * N * N) == 0
    # all-ones bytes are NaN doubles: an entry the solve does not write shows
    assert hip.hipMemset(F_dev, 0xFF, 8 * N * N) == 0 and hip.hipDeviceSynchronize() == 0
    A = np.full(N, np.nan)
    _, Fg, Ag, tags, st = _outputs(G)
    o = _opts(nquad, **kw)
    print(f"noalloc call: N={N} nquad={nquad} obs={tuple(obs)} {kw}", flush=True)
    _lib.check(_lib.lib().rvf_compute_view_factors_noalloc(
        ctx.h, dm.h, do.h if do else None, C.byref(o), F_dev, _lib.ptr(A), _lib.ptr(Fg), _lib.ptr(Ag),
        _lib.ptr(tags), C.byref(st)))
    F = np.empty((N, N), order="F")
    assert hip.hipMemcpy(_lib.ptr(F), F_dev, 8 * N * N, 2) == 0  # hipMemcpyDeviceToHost
    assert hip.hipFree(F_dev) == 0
    dm.close()
    if do:
        do.close()
    return dict(F=F, A=A, Fg=Fg, Ag=Ag, tags=tags, st=st.as_dict())


def same(a, b):
    assert np.array_equal(a["F"], b["F"]), f"F differs: max {np.nanmax(np.abs(a['F'] - b['F']))}"
    assert np.array_equal(a["A"], b["A"])
    assert np.array_equal(a["Fg"], b["Fg"])
    assert np.array_equal(a["Ag"], b["Ag"])
    assert np.array_equal(a["tags"], b["tags"])
    for k in ("rays", "retests", "culled_rays", "sure_hits", "singular_pairs", "tiles", "unique_pairs"):
        assert a["st"][k] == b["st"][k], k


CASES = {
    # occluded plates + seeded sphere, mixed precision (split traversal, hint, culling)
    "occluded_mixed": (lambda: M.occluded_plates(n=14, subdiv=2), 3, (3,), dict(precision="mixed")),
    "occluded_fp64": (lambda: M.occluded_plates(n=10, subdiv=2), 2, (3,), dict(precision="fp64")),
    # curved Quad8 side + Tri6 caps with Duffy singular pairs
    "duffy": (lambda: M.cylinder(n_around=16, n_along=6, cap_tris=40), 3, (), dict(method="duffy")),
    "mc": (lambda: M.parallel_plates(6), 2, (), dict(method="mc")),
    "cube_self_mixed": (lambda: M.unit_cube(6), 3, (), dict(precision="mixed", self_vf=True)),
    "cube_fp64": (lambda: M.unit_cube(3), 4, (), dict(precision="fp64")),
}

SEQUENCES = [("occluded_mixed", "cube_fp64"), ("duffy", "mc"), ("occluded_mixed", "occluded_fp64"),
             ("cube_self_mixed", "duffy")]


@pytest.mark.parametrize("first,second", SEQUENCES)
def test_noalloc_equals_allocating_entry_two_calls(first, second):
    m1, nq1, obs1, kw1 = CASES[first]
    m2, nq2, obs2, kw2 = CASES[second]
    mesh1, mesh2 = m1(), m2()
    assert mesh2.n_elems < mesh1.n_elems  # the second call is the smaller one
    ref1 = solve_alloc(mesh1, nq1, obs1, **kw1)
    ref2 = solve_alloc(mesh2, nq2, obs2, **kw2)
    got1 = solve_noalloc(mesh1, nq1, obs1, **kw1)
    got2 = solve_noalloc(mesh2, nq2, obs2, **kw2)
    same(got1, ref1)
    same(got2, ref2)


def test_noalloc_refuses_oversize_and_allocating_flags():
    ctx = api.context(0)
    # 90x90 plates at nquad 3: the partial sums alone pass 64 MiB
    big = M.parallel_plates(90)
    dm = api.DeviceMesh(ctx, big)
    G = len(dm.tags)
    _, Fg, Ag, tags, st = _outputs(G)
    A = np.empty(big.n_elems)
    L = _lib.lib()
    o = _opts(3)
    rc = L.rvf_compute_view_factors_noalloc(ctx.h, dm.h, None, C.byref(o), None, _lib.ptr(A), _lib.ptr(Fg),
                                            _lib.ptr(Ag), _lib.ptr(tags), C.byref(st))
    msg = L.rvf_last_error().decode()
    assert rc == -3 and "bytes of scratch" in msg and str(SCRATCH_BYTES) in msg, (rc, msg)
    for flag in (_lib.RVF_AUDIT, _lib.RVF_GRAPH, _lib.RVF_TIMING):
        o = _opts(2, flags=flag)
        rc = L.rvf_compute_view_factors_noalloc(ctx.h, dm.h, None, C.byref(o), None, _lib.ptr(A), _lib.ptr(Fg),
                                                _lib.ptr(Ag), _lib.ptr(tags), C.byref(st))
        assert rc == -1 and "need device allocations" in L.rvf_last_error().decode()
    dm.close()
    # the refusals leave the entry point usable
    small = M.parallel_plates(4)
    same(solve_noalloc(small, 2), solve_alloc(small, 2))
