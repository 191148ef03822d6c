"""ctypes wrapper of the C oracle — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` (to check the GPU result) and
``bench.py`` (``cpu_baseline`` leg).  Never imported by the product package.

Also holds numpy restatements of the reference's host-side reductions
(src/Results.jl:56-127) used to check the library's device reductions.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librvf_oracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_gl1d.argtypes = [C.c_int, P, P]
        L.orc_quad_rule.argtypes = [C.c_int, P, P]
        L.orc_tri_rule.argtypes = [C.c_int, P, P]
        L.orc_shape.argtypes = [C.c_int, C.c_double, C.c_double, P, P, P]
        L.orc_eval_point.argtypes = [C.c_int, P, P, C.c_double, C.c_double, P, P, P]
        L.orc_vf_kernel.argtypes = [P, P, P, P]
        L.orc_vf_kernel.restype = C.c_double
        L.orc_vf_kernel_2d.argtypes = [P, P, P, P]
        L.orc_vf_kernel_2d.restype = C.c_double
        L.orc_seg_hits_brute.argtypes = [P, C.c_int64, P, P]
        L.orc_pair.argtypes = [P, P, P, C.c_int64, C.c_int64, C.c_int, P]
        L.orc_pair.restype = C.c_double
        L.orc_pair_duffy.argtypes = [P, P, P, C.c_int64, C.c_int64, C.c_int]
        L.orc_pair_duffy.restype = C.c_double
        L.orc_set_norm.argtypes = [C.c_int]
        L.orc_mc_expect.argtypes = [P, P, P, C.c_int64, C.c_int64, C.c_int, C.c_int]
        L.orc_mc_expect.restype = C.c_double
        L.orc_bvh_export.argtypes = [P, C.c_int64, C.c_int, C.c_int, P, P, P]
        L.orc_ray_hits.argtypes = [P, C.c_int64, C.c_int, P, P, C.c_double]
        L.orc_ray_hits_brute.argtypes = [P, C.c_int64, C.c_int, P, P, C.c_double]
        L.orc_visible_batch.argtypes = [P, C.c_int64, C.c_int, P, P, C.c_int64, P]
        L.orc_xorshift_stream.argtypes = [C.c_uint64, C.c_int64, C.c_int, P]
        L.orc_init_rng.argtypes = [C.c_uint64, C.c_int64]
        L.orc_init_rng.restype = C.c_uint64
        L.orc_singularity.argtypes = [P, C.c_int, P, C.c_int, P, P]
        L.orc_duffy_rule.argtypes = [C.c_int, C.c_int, P, P, P, P, P]
        L.orc_duffy_rule.restype = C.c_int64
        L.orc_duffy_map.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, P]
        L.orc_compute.argtypes = [P, C.c_int64, C.c_int64, P, P, P, C.c_int, P, P, C.c_int64, P, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int64,
                                  C.c_int64, P, P, P, P, C.c_int, P, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def soup_mode(parity_mode: str, soup_layout: str) -> int:
    """0 = reference reading (Q1: AABB reads [xyz,vertex], Moller-Trumbore
    reads [vertex,xyz]); 1/2 = geometric reading of an [xyz,vertex] /
    [vertex,xyz] soup."""
    if parity_mode == "reference":
        return 0
    return 1 if soup_layout == "xyz_vertex" else 2


def gl1d(n):
    p = np.zeros(8); w = np.zeros(8)
    k = lib().orc_gl1d(n, _p(p), _p(w))
    return p[:k].copy(), w[:k].copy()


def quad_rule(n):
    p = np.zeros(2 * 64); w = np.zeros(64)
    k = lib().orc_quad_rule(n, _p(p), _p(w))
    return p[:2 * k].reshape(k, 2).copy(), w[:k].copy()


def tri_rule(n):
    p = np.zeros(2 * 13); w = np.zeros(13)
    k = lib().orc_tri_rule(n, _p(p), _p(w))
    return p[:2 * k].reshape(k, 2).copy(), w[:k].copy()


def shape(fam, x, e):
    N = np.zeros(8); dx = np.zeros(8); de = np.zeros(8)
    k = lib().orc_shape(fam, x, e, _p(N), _p(dx), _p(de))
    return N[:k].copy(), dx[:k].copy(), de[:k].copy()


def eval_point(fam, coords, nodes, x, e):
    coords = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
    nodes = np.ascontiguousarray(nodes, np.int32)
    X = np.zeros(3); n = np.zeros(3); dA = np.zeros(1)
    lib().orc_eval_point(fam, _p(coords), _p(nodes), x, e, _p(X), _p(n), _p(dA))
    return X, n, float(dA[0])


def vf_kernel(xi, ni, xj, nj):
    a = [np.ascontiguousarray(v, np.float64) for v in (xi, ni, xj, nj)]
    return lib().orc_vf_kernel(*[_p(v) for v in a])


def pair(mesh, i, j, nquad):
    Li = np.zeros(1)
    r = lib().orc_pair(_p(mesh.coords), _p(mesh.family), _p(mesh.nodes), i, j, nquad, _p(Li))
    return r, float(Li[0])


def pair_duffy(mesh, i, j, nquad):
    return lib().orc_pair_duffy(_p(mesh.coords), _p(mesh.family), _p(mesh.nodes), i, j, nquad)


def mc_expect(mesh, i, j, n_samples, g=3):
    """Expectation of the Monte Carlo pair estimate raw_ij of two quadrilaterals
    (the stream's sample k puts both points in parameter cell k), by Gauss
    quadrature inside each paired cell."""
    return lib().orc_mc_expect(_p(mesh.coords), _p(mesh.family), _p(mesh.nodes), i, j, n_samples, g)


def duffy_rule(kind, nquad, ci, cj):
    """The Duffy point rule for one pair of shared corners (kind 1: common
    vertex, corners ci[0], cj[0]; kind 2: common edge, ci[0:2], cj[0:2]):
    points (u, v, s, t) in [0,1]^4, weights without the kernel and area
    elements, and the region of each point."""
    a = np.ascontiguousarray(list(ci) + [0] * (2 - len(ci)), np.int32)
    b = np.ascontiguousarray(list(cj) + [0] * (2 - len(cj)), np.int32)
    n = lib().orc_duffy_rule(kind, nquad, _p(a), _p(b), None, None, None)
    uvst = np.zeros((n, 4)); w = np.zeros(n); reg = np.zeros(n, np.int32)
    lib().orc_duffy_rule(kind, nquad, _p(a), _p(b), _p(uvst), _p(w), _p(reg))
    return uvst, w, reg


def duffy_map(kind, region, r, e1, e2, e3):
    """One Duffy region map: kind 1 common-vertex offsets (du, dv, ds, dt),
    kind 2 common-edge (u, v, s, t) in the bottom-edge frame."""
    out = np.zeros(4)
    lib().orc_duffy_map(kind, region, r, e1, e2, e3, _p(out))
    return out


SEGMENTS = 3  # soup mode of a curve mesh's segment soup (T, 2, 3)


def _soup(soup, mode):
    return np.ascontiguousarray(soup, np.float64).reshape(-1, 2 if mode == SEGMENTS else 3, 3)


def bvh_export(soup, mode):
    soup = _soup(soup, mode)
    T = soup.shape[0]
    cap = max(4, 2 * T)
    box = np.zeros(6 * cap); meta = np.zeros(5 * cap, np.int32); tri = np.zeros(max(T, 1), np.int32)
    nn = lib().orc_bvh_export(_p(soup), T, mode, cap, _p(box), _p(meta), _p(tri))
    assert nn >= 0
    return box[:6 * nn].reshape(nn, 6), meta[:5 * nn].reshape(nn, 5), tri[:T]


def ray_hits(soup, mode, o, d, tmax, brute=False):
    soup = np.ascontiguousarray(soup, np.float64).reshape(-1, 3, 3)
    o = np.ascontiguousarray(o, np.float64); d = np.ascontiguousarray(d, np.float64)
    f = lib().orc_ray_hits_brute if brute else lib().orc_ray_hits
    return bool(f(_p(soup), soup.shape[0], mode, _p(o), _p(d), tmax))


def visible_batch(soup, mode, xi, xj):
    soup = _soup(soup, mode)
    xi = np.ascontiguousarray(xi, np.float64).reshape(-1, 3)
    xj = np.ascontiguousarray(xj, np.float64).reshape(-1, 3)
    out = np.zeros(xi.shape[0], np.uint8)
    lib().orc_visible_batch(_p(soup), soup.shape[0], mode, _p(xi), _p(xj), xi.shape[0], _p(out))
    return out.astype(bool)


def seg_hits_brute(soup, xi, xj):
    """Any segment of a (T, 2, 3) soup blocks the 2-D path xi -> xj (no tree)."""
    soup = _soup(soup, SEGMENTS)
    xi = np.ascontiguousarray(xi, np.float64); xj = np.ascontiguousarray(xj, np.float64)
    return bool(lib().orc_seg_hits_brute(_p(soup), soup.shape[0], _p(xi), _p(xj)))


def vf_kernel_2d(xi, ni, xj, nj):
    a = [np.ascontiguousarray(v, np.float64) for v in (xi, ni, xj, nj)]
    return lib().orc_vf_kernel_2d(*[_p(v) for v in a])


def xorshift_stream(seed, tid, n):
    out = np.zeros(n)
    lib().orc_xorshift_stream(C.c_uint64(seed), tid, n, _p(out))
    return out


def compute(mesh, nquad=4, obstruction_groups=(), parity_mode="reference", method="quad", n_samples=256,
            seed=0, nthreads=0, schedule="static", rows=None, want_raw=True, audit=False, count=False,
            self_vf=False, norm="plain"):
    """Restatement of ``_compute_cpu`` (src/Assembly.jl:102-225).  Returns a
    dict with ``raw`` (N x N, Fortran-ordered like Julia; rows outside
    ``rows`` stay zero), ``A``, optional audit arrays and test counters.
    ``self_vf`` integrates the diagonal pairs (i, i) too (Assembly.jl:173,193).
    ``norm``: "plain" or "scaled", the two forms StaticArrays' norm takes
    depending on its version (quirk Q13)."""
    lib().orc_set_norm(1 if norm == "scaled" else 0)
    try:
        return _compute(mesh, nquad, obstruction_groups, parity_mode, method, n_samples, seed, nthreads, schedule,
                        rows, want_raw, audit, count, self_vf)
    finally:
        lib().orc_set_norm(0)


def _compute(mesh, nquad, obstruction_groups, parity_mode, method, n_samples, seed, nthreads, schedule, rows,
             want_raw, audit, count, self_vf):
    N = mesh.n_elems
    soup, sg = mesh.merged_soup()
    obs = np.ascontiguousarray(list(obstruction_groups), np.int32)
    mode = soup_mode(parity_mode, mesh.soup_layout)
    raw = np.zeros((N, N), order="F") if want_raw else None
    A = np.zeros(N)
    ac = np.zeros((N, N), np.uint32, order="F") if audit else None
    ah = np.zeros((N, N), np.uint32, order="F") if audit else None
    cnt = np.zeros(3, np.int64)
    r0, r1 = (0, N) if rows is None else rows
    meth = {"quad": 0, "mc": 1, "duffy": 2}[method]
    rc = lib().orc_compute(_p(mesh.coords), mesh.n_nodes, N, _p(mesh.family), _p(mesh.nodes), _p(mesh.group),
                           nquad, _p(soup), _p(sg), soup.shape[0], _p(obs), len(obs), mode, meth, n_samples,
                           C.c_uint64(seed), nthreads, 0 if schedule == "static" else 1, r0, r1, _p(raw), _p(A),
                           _p(ac), _p(ah), 1 if count else 0, _p(cnt), 1 if self_vf else 0)
    assert rc == 0
    return {"raw": raw, "A": A, "audit_count": ac, "audit_hash": ah,
            "rays": int(cnt[0]), "aabb_tests": int(cnt[1]), "tri_tests": int(cnt[2])}


def aggregate(mesh, F, A):
    """``_aggregate`` (src/Results.jl:56-82): sorted tags; A_g F_gh =
    sum_{i in g} A_i sum_{j in h} F_ij, sums in element order."""
    tags = sorted(mesh.group_tags)
    ge = mesh.group_elems
    G = len(tags)
    Ag = np.zeros(G)
    for k, t in enumerate(tags):
        for e in ge[t]:
            Ag[k] += A[e]
    Fg = np.zeros((G, G))
    for gi, ti in enumerate(tags):
        for e in ge[ti]:
            for gj, tj in enumerate(tags):
                s = 0.0
                for f in ge[tj]:
                    s += F[e, f]
                Fg[gi, gj] += A[e] * s
        Fg[gi, :] /= Ag[gi]
    return tags, [mesh.group_tags[t] for t in tags], Fg, Ag


def reciprocity_error(F, A):
    """max relative reciprocity error as check_reciprocity (Results.jl:105-114)."""
    AF = A[:, None] * F
    iu = np.triu_indices(F.shape[0], 1)
    num = np.abs(AF[iu] - AF.T[iu])
    den = np.maximum(AF[iu], 1e-30)
    return float((num / den).max()) if num.size else 0.0
