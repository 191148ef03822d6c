"""Curve meshes (mesh_dim 1: Line3 / Line2 elements, per unit depth) in the
oracle.  The reference computes them on the CPU only (src/Assembly.jl:83-86)
and tests only the Line2 shape functions (test/runtests.jl:241-247); the rest
is pinned by the mathematics:

  * Line2 shape functions: partition of unity and nodal values
    (runtests.jl:241-247); Line3 the same, plus the nodal delta at its
    midpoint; element geometry: length |x2 - x1|, normal = tangent turned 90
    degrees counter-clockwise (Geometry.jl:219-221).
  * the 2-D kernel cos cos / (2 r): 1 / (2 d) for facing points, zero behind
    either point and for coincident points.
  * Hottel's crossed strings, the exact 2-D view factor between straight
    strips: parallel strips sqrt(1 + (h/w)^2) - h/w; perpendicular strips
    sharing an edge (1 + H/W - sqrt(1 + (H/W)^2)) / 2; parallel strips with a
    half-infinite baffle (a string wraps round its end); arcs of a circle
    F_0k = (chord((k+1)a) + chord((k-1)a) - 2 chord(k a)) / (2 arc), and an arc
    with itself 1 - chord/arc.  Quadrature converges to them (O(h) where
    visibility or a shared corner cuts the integrand).
  * segment occlusion: the reference's tree traversal decides as a brute-force
    any-hit over all segments (the path's box test only prunes).
  * Monte Carlo on curves (the library's own stream layout, stratum pairing of
    the reference's CPU sampler): converges to the exact value as h^2.
"""
import math

import numpy as np
import pytest

from rvf_amd import mesh as M


def _fg(orc, m, **kw):
    r = orc.compute(m, **kw)
    F = r["raw"] / r["A"][:, None]
    tags, names, Fg, Ag = orc.aggregate(m, F, r["A"])
    return r, F, Fg


def test_line_shapes(orc):
    for x in np.linspace(-1, 1, 7):
        N, dN, _ = orc.shape(5, x, 0.0)
        assert abs(N.sum() - 1) < 1e-14 and abs(dN.sum()) < 1e-14
        N3, dN3, _ = orc.shape(4, x, 0.0)
        assert abs(N3.sum() - 1) < 1e-14 and abs(dN3.sum()) < 1e-14
    assert np.allclose(orc.shape(5, -1.0, 0.0)[0], [1, 0], atol=1e-14)
    assert np.allclose(orc.shape(5, 1.0, 0.0)[0], [0, 1], atol=1e-14)
    for x, want in ((-1.0, [1, 0, 0]), (1.0, [0, 1, 0]), (0.0, [0, 0, 1])):
        assert np.allclose(orc.shape(4, x, 0.0)[0], want, atol=1e-14)


def test_line_geometry(orc):
    coords = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0], [1.5, 2.0, 0.0]])
    for fam, nodes in ((5, [1, 2]), (4, [1, 2, 3])):
        x, n, dL = orc.eval_point(fam, coords, np.array(nodes + [0] * (8 - len(nodes)), np.int32), 0.3, 0.0)
        assert np.allclose(x, [0.5 * 1.3 * 3, 0.5 * 1.3 * 4, 0.0])
        assert np.allclose(n, [-0.8, 0.6, 0.0]) and abs(dL - 2.5) < 1e-14  # length 5 over [-1, 1]
    m = M.parallel_lines(4, 1.0, 2.0, order=2)
    r = orc.compute(m, nquad=3)
    assert np.allclose(r["A"], 0.5, atol=1e-14)


def test_kernel_2d(orc):
    z = np.zeros(3)
    k = orc.vf_kernel_2d([0, 0, 0], [0, 1, 0], [0, 2.0, 0], [0, -1, 0])
    assert abs(k - 0.25) < 1e-15
    assert orc.vf_kernel_2d([0, 0, 0], [0, -1, 0], [0, 2.0, 0], [0, -1, 0]) == 0.0
    assert orc.vf_kernel_2d([0, 0, 0], [0, 1, 0], [0, 2.0, 0], [0, 1, 0]) == 0.0
    assert orc.vf_kernel_2d(z, [0, 1, 0], z, [0, -1, 0]) == 0.0
    # oblique: cos_i = cos_j = 1/sqrt2, r = sqrt2 -> (1/2) / (2 sqrt2)
    k = orc.vf_kernel_2d([0, 0, 0], [0, 1, 0], [1.0, 1.0, 0], [0, -1, 0])
    assert abs(k - 0.5 / (2 * math.sqrt(2))) < 1e-15


@pytest.mark.parametrize("order", [1, 2])
def test_parallel_strips_crossed_strings(orc, order):
    for h, w in ((1.0, 1.0), (0.5, 2.0)):
        m = M.parallel_lines(8, h, w, order)
        r, F, Fg = _fg(orc, m, nquad=4)
        exact = math.sqrt(1 + (h / w) ** 2) - h / w
        assert abs(Fg[0, 1] - exact) < 1e-7, (h, w, Fg[0, 1], exact)
        assert abs(Fg[1, 0] - exact) < 1e-7
        assert np.allclose(r["A"].sum(), 2 * w)


def test_perpendicular_strips_converge(orc):
    exact = (2 - math.sqrt(2)) / 2
    err = []
    for n in (8, 32):
        r, F, Fg = _fg(orc, M.square_enclosure_2d(n, 1, baffle=False), nquad=4)
        err.append(abs(Fg[0, 3] - exact))
        assert abs(Fg[0, 2] - (math.sqrt(2) - 1)) < 1e-12          # bottom -> top: parallel, exact
    assert err[1] < 1e-3 and 3.0 < err[0] / err[1] < 5.0            # first order at the shared corner


def test_baffle_crossed_strings(orc):
    # strips y = 0 and y = 1 on [0, 1]; baffle y = 0.5 from x = 0.6 to 2: the
    # string (1,0)-(1,1) wraps round (0.6, 0.5)
    exact = (2 * math.sqrt(2) - 1 - 2 * math.sqrt(0.16 + 0.25)) / 2
    errs = []
    for n in (16, 64):
        m = M.parallel_lines(n, 1.0, 1.0, 1)
        m.group_tri_soup = {9: np.array([[[0.6, 0.5, 0.0], [2.0, 0.5, 0.0]]])}
        r, F, Fg = _fg(orc, m, nquad=4, obstruction_groups=[9])
        errs.append(abs(Fg[0, 1] - exact))
    assert errs[1] < 3e-4 and errs[1] < errs[0]
    # a baffle across the whole view blocks everything
    m = M.parallel_lines(8, 1.0, 1.0, 2)
    m.group_tri_soup = {9: np.array([[[-5.0, 0.5, 0.0], [5.0, 0.5, 0.0]]])}
    r, F, Fg = _fg(orc, m, nquad=3, obstruction_groups=[9])
    assert Fg[0, 1] == 0.0


def test_circle_arcs(orc):
    n = 16
    a = 2 * math.pi / n
    r, F, Fg = _fg(orc, M.circle_interior(n, 1.0, 2), nquad=6, self_vf=True)
    chord = lambda t: 2 * math.sin(abs(t) / 2)
    for k in (1, 2, 5, 8):
        exact = (2 * chord(k * a) - chord((k + 1) * a) - chord((k - 1) * a)) / (2 * a)
        assert abs(F[0, k] - exact) < 2e-4 * exact + 1e-7, (k, F[0, k], exact)
    assert np.all(np.abs(F.sum(1) - 1) < 5e-4)
    # An arc with itself: on a circle the kernel is r / (8 R^2), so the Gauss
    # sum over the exact arc is S = sum_pq w_p w_q (a/2)^2 chord(t_p - t_q) / 8 / a.
    # The Line3 arc differs from it by its O(h^2) geometry error; S itself
    # approaches 1 - chord/arc from below as nquad grows (the integrand has a
    # kink on the diagonal).
    dev = []
    for n_, nq in ((16, 6), (64, 6)):
        a_ = 2 * math.pi / n_
        Fq = _fg(orc, M.circle_interior(n_, 1.0, 2), nquad=nq, self_vf=True)[1]
        p, w = orc.gl1d(nq)
        t = a_ * (1 + p) / 2
        S = sum(w[i] * w[j] * (a_ / 2) ** 2 * chord(t[i] - t[j]) / 8 for i in range(nq) for j in range(nq)) / a_
        dev.append(abs(Fq[0, 0] / S - 1))
        assert S < 1 - chord(a_) / a_
    assert dev[0] < 5e-3 and 10 < dev[0] / dev[1] < 20
    selfs = [_fg(orc, M.circle_interior(n, 1.0, 2), nquad=nq, self_vf=True)[1][0, 0] for nq in (4, 6, 8)]
    assert selfs[0] < selfs[1] < selfs[2] < 1 - chord(a) / a


def test_segment_tree_matches_brute_force(orc):
    rng = np.random.default_rng(7)
    T = 200
    p0 = rng.uniform(0, 1, (T, 3)); p0[:, 2] = 0
    d = rng.normal(0, 0.08, (T, 3)); d[:, 2] = 0
    soup = np.ascontiguousarray(np.stack([p0, p0 + d], axis=1))
    xi = rng.uniform(0, 1, (500, 3)); xj = rng.uniform(0, 1, (500, 3))
    xi[:, 2] = xj[:, 2] = 0
    vis = orc.visible_batch(soup, orc.SEGMENTS, xi, xj)
    brute = np.array([not orc.seg_hits_brute(soup, xi[k], xj[k]) for k in range(500)])
    assert np.array_equal(vis, brute)
    assert 0 < vis.sum() < 500
    box, meta, idx = orc.bvh_export(soup, orc.SEGMENTS)
    assert sorted(idx) == list(range(1, T + 1))
    leaves = meta[:, 0] == 0
    assert np.all(meta[leaves, 3] <= 8)


def test_monte_carlo_curves_converge(orc):
    # sample k of both elements lies in stratum k (the pairing of the
    # reference's _stratified_line_points, MCKernel.jl:107-115, 236-262), so
    # the estimate weighs pairs of nearby strata: its bias falls as h^2
    exact = math.sqrt(2) - 1
    bias = {}
    for n in (4, 8, 16):
        m = M.parallel_lines(n, 1.0, 1.0, 1)
        est = []
        for seed in range(4):
            r, F, Fg = _fg(orc, m, method="mc", n_samples=64, seed=seed)
            est.append(Fg[0, 1])
            assert np.allclose(r["A"], 1.0 / n)  # straight Line2: dL is constant, the MC length exact
        bias[n] = abs(np.mean(est) - exact)
    assert bias[16] < 5e-4
    assert 3.0 < bias[4] / bias[8] < 5.0 and 3.0 < bias[8] / bias[16] < 5.0
