"""Pins the oracle's Duffy / Sauter-Schwab path (src/DuffyKernel.jl:95-487),
which no reference test covers and whose maps are not a valid change of
variables (the reference's results are not the integral).  What fixes it:

  * common vertex: the eight region maps split [0,1]^4 (corner-shifted
    coordinates du, dv on element i, ds, dt on element j) by which coordinate
    is largest; weighted by rho^3 instead of their Jacobian rho^3 eta2, the
    rule integrates g * W with W = max(m_i, m_j) / min(m_i, m_j),
    m_i = max(du, dv), m_j = max(ds, dt).  For monomials g this integral has a
    closed form (derived below from the distribution of the maximum of two
    uniforms), and Gauss-Legendre is exact on each region while the degree in
    rho stays <= 2 nquad - 1: the rule must reproduce it to rounding, for every
    pair of shared corners.  A constant integrates to 2.
  * common edge: in the bottom-edge frame each region's points lie on the
    hyperplanes that make its Jacobian vanish (region 1: s = u - t + v;
    2: s = u - v; 3: s = u - t; 4: t = v; 5: s = u and t = v); regions 2, 4
    and 5 never leave the square, so their weights sum to int rho^3 = 1/4;
    regions 1 and 3 lose points outside it.  Any other shared edge is the
    bottom-edge rule carried by the rotation of the square that takes the
    bottom edge (corners 1, 2) onto it.
  * the pair value is the kernel summed over that rule with the Quad8 area
    elements of the [0,1]^2 parametrisation (4 dA).
  * every region map, vertex and edge, bit for bit against vectors of the
    reference's own map expressions (tests/golden/duffy_map_vectors.json).
"""
import itertools
import json
import os

import numpy as np
import pytest

from rvf_amd import mesh as M

CORNER_UV = {1: (0.0, 0.0), 2: (1.0, 0.0), 3: (1.0, 1.0), 4: (0.0, 1.0)}


def vertex_moment(a, b, c, d):
    """int over [0,1]^4 of du^a dv^b ds^c dt^d * max(m_i, m_j) / min(m_i, m_j).

    m = max(du, dv) has density 2 m; given m, one coordinate equals m and the
    other is uniform on [0, m], each way with probability 1/2, so
    E[du^a dv^b | m] = m^(a+b) (1/(a+1) + 1/(b+1)) / 2.  What remains is
    4 int int x^(p+1) y^(q+1) max(x,y)/min(x,y) dx dy
      = 4 (1/(q+1) + 1/(p+1)) / (p+q+4)."""
    ca = (1 / (a + 1) + 1 / (b + 1)) / 2
    cb = (1 / (c + 1) + 1 / (d + 1)) / 2
    p, q = a + b, c + d
    return 4 * ca * cb * (1 / (q + 1) + 1 / (p + 1)) / (p + q + 4)


def shifted(uvst, ci, cj):
    u0, v0 = CORNER_UV[ci]
    s0, t0 = CORNER_UV[cj]
    return np.abs(uvst - np.array([u0, v0, s0, t0]))


@pytest.mark.parametrize("nq", [2, 3, 4, 5])
def test_vertex_rule_closed_form(orc, nq):
    for ci, cj in itertools.product(range(1, 5), repeat=2):
        uvst, w, reg = orc.duffy_rule(1, nq, [ci], [cj])
        assert len(w) == 8 * nq ** 4  # the shifted maps never leave the square
        assert np.all((uvst >= 0) & (uvst <= 1))
        assert abs(w.sum() - 2.0) < 1e-14
        d = shifted(uvst, ci, cj)
        top = 2 * nq - 4  # degree in rho: a + b + c + d + 3 <= 2 nq - 1
        for a, b, c, e in itertools.product(range(top + 1), repeat=4):
            if a + b + c + e > top:
                continue
            got = float(np.sum(w * d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c * d[:, 3] ** e))
            want = vertex_moment(a, b, c, e)
            assert abs(got - want) < 1e-13, (ci, cj, (a, b, c, e), got, want)


def test_vertex_rule_is_not_the_integral(orc):
    # with the true Jacobian the constant would integrate to 1 and du to 1/2
    uvst, w, _ = orc.duffy_rule(1, 4, [1], [1])
    assert abs(w.sum() - 1.0) > 0.9
    assert abs(np.sum(w * uvst[:, 0]) - vertex_moment(1, 0, 0, 0)) < 1e-14
    assert abs(vertex_moment(1, 0, 0, 0) - 0.5) > 0.1


@pytest.mark.parametrize("nq", [2, 3, 5])
def test_edge_rule_bottom_frame(orc, nq):
    uvst, w, reg = orc.duffy_rule(2, nq, [1, 2], [1, 2])
    assert np.all((uvst >= 0) & (uvst <= 1))
    u, v, s, t = uvst.T
    rel = {1: [s - (u - t + v)], 2: [s - (u - v)], 3: [s - (u - t)], 4: [t - v], 5: [s - u, t - v]}
    for r, exprs in rel.items():
        k = reg == r
        assert k.any()
        for e in exprs:
            assert np.abs(e[k]).max() < 4e-16, r
    for r in (2, 4, 5):
        assert np.sum(reg == r) == nq ** 4 and abs(w[reg == r].sum() - 0.25) < 1e-15
    for r in (1, 3):  # (the 2-point rule happens to keep all its points)
        kept = np.sum(reg == r)
        assert kept <= nq ** 4 and (kept < nq ** 4 and w[reg == r].sum() < 0.25 or nq == 2)


def square_rotation(c1, c2):
    """Quarter turns k (counter-clockwise about the square's centre) that take
    the bottom edge's corners {(0,0), (1,0)} onto {uv(c1), uv(c2)}."""
    want = {CORNER_UV[c1], CORNER_UV[c2]}
    for k in range(4):
        got = {tuple(rot(np.array([p]), k)[0]) for p in ((0.0, 0.0), (1.0, 0.0))}
        if got == want:
            return k
    raise AssertionError("not an edge")


def rot(p, k):
    c, s = [(1, 0), (0, 1), (-1, 0), (0, -1)][k]
    q = p - 0.5
    return np.stack([c * q[:, 0] - s * q[:, 1], s * q[:, 0] + c * q[:, 1]], 1) + 0.5


def test_edge_rule_rotations(orc):
    base, wb, rb = orc.duffy_rule(2, 3, [1, 2], [1, 2])
    edges = [(1, 2), (2, 3), (3, 4), (4, 1), (2, 1), (3, 2), (4, 3), (1, 4)]
    for ei, ej in itertools.product(edges, repeat=2):
        uvst, w, reg = orc.duffy_rule(2, 3, list(ei), list(ej))
        ki, kj = square_rotation(*ei), square_rotation(*ej)
        assert np.array_equal(w, wb) and np.array_equal(reg, rb)
        assert np.abs(uvst[:, :2] - rot(base[:, :2], ki)).max() < 1e-15, (ei, ki)
        assert np.abs(uvst[:, 2:] - rot(base[:, 2:], kj)).max() < 1e-15, (ej, kj)


def shared_corners(m, i, j):
    ci, cj = [], []
    for a in range(4):
        for b in range(4):
            if m.nodes[i, a] == m.nodes[j, b]:
                ci.append(a + 1); cj.append(b + 1)
    return ci, cj


def test_pair_is_kernel_over_rule(orc):
    m = M.cylinder(n_around=12, n_along=4, cap_tris=30)
    side = np.nonzero(m.family == 0)[0]
    seen = {1: 0, 2: 0}
    for i in side[:6]:
        for j in side:
            ci, cj = shared_corners(m, i, j)
            if i == j or len(ci) not in (1, 2):
                continue
            kind = len(ci)
            for nq in (2, 3):
                uvst, w, _ = orc.duffy_rule(kind, nq, ci[:kind], cj[:kind])
                tot = 0.0
                for (u, v, s, t), wk in zip(uvst, w):
                    xi, ni, ai = orc.eval_point(0, m.coords, m.nodes[i], 2 * u - 1, 2 * v - 1)
                    xj, nj, aj = orc.eval_point(0, m.coords, m.nodes[j], 2 * s - 1, 2 * t - 1)
                    tot += wk * orc.vf_kernel(xi, ni, xj, nj) * (4 * ai) * (4 * aj)
                got = orc.pair_duffy(m, int(i), int(j), nq)
                assert abs(got - tot) <= 1e-13 * abs(tot), (i, j, kind, got, tot)
                assert got != orc.pair(m, int(i), int(j), nq)[0]
            seen[kind] += 1
    assert seen[1] > 0 and seen[2] > 0


def test_region_maps_match_reference_vectors(orc):
    # vectors of the reference's own map expressions (tests/golden/
    # make_duffy_map_vectors.py), compared bit for bit
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "duffy_map_vectors.json")) as f:
        data = json.load(f)
    pts = data["inputs"]
    for kind, name in ((1, "vertex"), (2, "edge")):
        regions = data["maps"][name]
        assert sorted(map(int, regions)) == list(range(1, 9 if kind == 1 else 6))
        for r, outs in regions.items():
            for p, want in zip(pts, outs):
                got = orc.duffy_map(kind, int(r), *p)
                assert np.array_equal(got, np.array(want)), (name, r, p, got, want)
