"""Self view factors in the oracle (``self_vf``: the loop j >= i,
src/Assembly.jl:173,193; the reference has no test of it).  Pinned by:

  * every pair (i, j), i != j, and its blocked-ray audit are bit-identical
    with and without self_vf;
  * the diagonal is the pair integral of the element with itself, bit for bit
    (orc_pair / orc_pair_duffy, pinned against the reference's own cases
    elsewhere), for quadrature, Duffy and the occluded solve (no ray blocked);
  * flat elements see none of themselves: F_ii = 0 exactly (the kernel's two
    cosines are exact negatives of each other);
  * inside a sphere the kernel is 1 / (4 pi R^2) for every point pair, so the
    quadrature gives F_ii = A_i / (4 pi R^2) (1 - c_i) with
    c_i = sum_p (w_p dA_p)^2 / A_i^2 (the coincident points p = q contribute
    0); what remains is the Quad8 geometry error, which falls as h^2;
  * Monte Carlo: pairs unchanged, the curved side sees itself, flat caps not.
"""
import math

import numpy as np
import pytest

from rvf_amd import mesh as M

from support.concave import sphere_groups


def _offdiag(a):
    return a[~np.eye(a.shape[0], dtype=bool)]


@pytest.mark.parametrize("method", ["quad", "duffy", "mc"])
def test_pairs_unchanged_diagonal_is_self_pair(orc, method):
    m = M.cylinder(n_around=8, n_along=3, cap_tris=20)
    kw = dict(nquad=3, method=method, n_samples=49, seed=11)
    r0 = orc.compute(m, **kw)
    r1 = orc.compute(m, self_vf=True, **kw)
    assert np.array_equal(r0["A"], r1["A"])
    assert np.array_equal(_offdiag(r0["raw"]), _offdiag(r1["raw"]))
    assert not np.any(np.diag(r0["raw"]))
    d = np.diag(r1["raw"])
    for i in range(m.n_elems):
        if method == "quad":
            want = orc.pair(m, i, i, 3)[0]
        elif method == "duffy":
            want = orc.pair_duffy(m, i, i, 3)
        else:
            continue
        assert d[i] == want, (i, d[i], want)
    side = m.family == 0
    if method != "duffy":  # (Duffy: a Quad8 with itself is COMMON_EDGE on corners 1, 2)
        assert np.all(d[side] > 0)          # the curved side sees itself
    # flat Tri6 caps do not (Monte Carlo samples carry rounding off the plane)
    assert np.all(np.abs(d[~side]) < 1e-20) if method == "mc" else not np.any(d[~side])


def test_occluded_pairs_unchanged(orc):
    m = sphere_groups(4, occluders=True)
    kw = dict(nquad=3, obstruction_groups=[9], parity_mode="geometric", audit=True)
    r0 = orc.compute(m, **kw)
    r1 = orc.compute(m, self_vf=True, **kw)
    for k in ("raw", "audit_count", "audit_hash"):
        assert np.array_equal(_offdiag(r0[k]), _offdiag(r1[k])), k
    blocked = np.diag(r1["audit_count"])
    assert list(np.nonzero(blocked)[0]) == [10]       # the box cuts element 10's own chords
    free = orc.compute(M.sphere_interior(4, 1.0, 2), nquad=3, self_vf=True)
    d, d_free = np.diag(r1["raw"]), np.diag(free["raw"])
    assert np.array_equal(np.delete(d, 10), np.delete(d_free, 10))
    assert 0 < d[10] < d_free[10]


def test_flat_elements_zero(orc):
    m = M.unit_cube(3)
    r = orc.compute(m, nquad=4, self_vf=True)
    assert not np.any(np.diag(r["raw"]))
    assert np.array_equal(_offdiag(r["raw"]), _offdiag(orc.compute(m, nquad=4)["raw"]))


def _missing_diagonal(orc, m, nq):
    """c_i = sum_p (w_p dA_p)^2 / A_i^2 over the element's Gauss points."""
    p, w = orc.gl1d(nq)
    c = np.empty(m.n_elems)
    for e in range(m.n_elems):
        wd = np.array([w[k] * w[l] * orc.eval_point(0, m.coords, m.nodes[e], p[k], p[l])[2]
                       for k in range(nq) for l in range(nq)])
        c[e] = (wd @ wd) / wd.sum() ** 2
    return c


def test_sphere_closed_form_and_convergence(orc):
    dev = {}
    for n in (3, 6):
        m = M.sphere_interior(n, 1.0, 2)
        r = orc.compute(m, nquad=3, self_vf=True)
        A = r["A"]
        F = r["raw"] / A[:, None]
        c = _missing_diagonal(orc, m, 3)
        ratio = np.diag(F) * 4 * math.pi / A
        dev[n] = ratio / (1 - c) - 1
        # off-diagonal pairs: F_ij = A_j / (4 pi R^2) to the same geometry error
        off = _offdiag(F * 4 * math.pi / A[None, :])
        assert np.all(np.abs(off - 1) < (0.02 if n == 3 else 0.005)), (n, off.min(), off.max())
        # closure: row sums reach 1 - sum_p (w_p dA_p)^2 / (4 pi A_i) only with the diagonal
        rs0 = F.sum(1) - np.diag(F)
        assert np.all(F.sum(1) > rs0)
        assert np.all(np.abs(F.sum(1) - (1 - c * A / (4 * math.pi))) < 2e-3), n
    assert np.abs(dev[6]).max() < 0.06
    # second order: halving h divides the geometry error by about 4
    q = np.abs(dev[3]).max() / np.abs(dev[6]).max()
    assert 3.0 < q < 5.0, q
