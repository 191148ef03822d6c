"""GPU parity for Monte Carlo (same seeded xorshift64 streams as the oracle:
bit-identical) and Duffy / Sauter-Schwab singular pairs (FP64, terms are
non-negative, so any summation order is within 1e-12 relative)."""
import math

import numpy as np
import pytest

from rvf_amd import api
from rvf_amd import mesh as M

pytestmark = pytest.mark.gpu


def gpu(mesh, nquad, method, obs=(), precision="fp64", parity="reference", n_samples=64, seed=0):
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, parity) if len(obs) else None
    plan = api.Plan(ctx, dm, do, nquad, method, precision, n_samples=n_samples, seed=seed, want_F=True)
    plan.run(); plan.sync()
    F, A, rs, Fg, Ag, tags, _, _ = plan.fetch()
    st = plan.stats()
    plan.close(); dm.close()
    if do is not None:
        do.close()
    return F, A, rs, Fg, Ag, st


@pytest.mark.parametrize("n_samples", [1, 10, 64, 100])
def test_mc_bitexact_plates(orc, n_samples):
    m = M.parallel_plates(5)
    F, A, rs, Fg, Ag, _ = gpu(m, 3, "mc", n_samples=n_samples, seed=0xC0FFEE)
    r = orc.compute(m, method="mc", n_samples=n_samples, seed=0xC0FFEE)
    assert np.array_equal(A, r["A"])
    assert np.array_equal(F, r["raw"] / r["A"][:, None])


def test_mc_bitexact_curved_mixed_families(orc):
    m = M.cylinder(n_around=12, n_along=3, cap_tris=30)
    F, A, *_ = gpu(m, 3, "mc", n_samples=36, seed=77)
    r = orc.compute(m, method="mc", n_samples=36, seed=77)
    assert np.array_equal(A, r["A"])
    assert np.array_equal(F, r["raw"] / r["A"][:, None])


def test_mc_bitexact_tri3_first_ragged(orc):
    # element 0 has 3 nodes and the last strip is partly empty: lanes past the
    # last element must sample element 0 with its own (Tri3) shape
    q = M.parallel_plates(3)
    tris = []
    for e in range(q.n_elems):
        a, b, c, d = q.nodes[e, :4]
        tris += [[a, b, c, 0, 0, 0, 0, 0], [a, c, d, 0, 0, 0, 0, 0]]
    m = M.MeshData(q.coords, np.full(len(tris), 3, np.int8), np.array(tris, np.int32), np.repeat(q.group, 2),
                   q.group_tags)
    F, A, *_ = gpu(m, 3, "mc", n_samples=16, seed=21)
    r = orc.compute(m, method="mc", n_samples=16, seed=21)
    assert np.array_equal(A, r["A"])
    assert np.array_equal(F, r["raw"] / r["A"][:, None])


def test_mc_occluded_bitexact(orc):
    m = M.occluded_plates(n=6, subdiv=2)
    F, A, *_ = gpu(m, 3, "mc", obs=[3], n_samples=25, seed=5)
    r = orc.compute(m, method="mc", obstruction_groups=[3], n_samples=25, seed=5)
    assert np.array_equal(F, r["raw"] / r["A"][:, None])


def test_mc_group_level_within_3sigma():
    # sigma from R = 8 seeds (the per-pair stratum pairing of the reference
    # biases single pairs but not the group level)
    m = M.parallel_plates(12)
    Fq = gpu(m, 5, "quad")[3][0, 1]
    est = np.array([gpu(m, 3, "mc", n_samples=256, seed=s)[3][0, 1] for s in range(8)])
    sd = est.std(ddof=1) / math.sqrt(len(est))
    assert abs(est.mean() - Fq) < 3 * sd + 1e-12


@pytest.mark.parametrize("nquad", [2, 3])
def test_duffy_matches_oracle(orc, nquad):
    m = M.cylinder(n_around=12, n_along=4, cap_tris=30)
    F, A, rs, Fg, Ag, st = gpu(m, nquad, "duffy")
    r = orc.compute(m, nquad=nquad, method="duffy")
    Fo = r["raw"] / r["A"][:, None]
    assert np.array_equal(A, r["A"])
    assert st["singular_pairs"] > 0
    rel = np.abs(F - Fo) / np.maximum(np.abs(Fo), 1e-300)
    assert float(rel.max()) <= 1e-12
    assert np.allclose(rs, Fo.sum(1), rtol=1e-12, atol=1e-15)
    tags, names, Fg_o, Ag_o = orc.aggregate(m, Fo, r["A"])
    assert np.allclose(Fg, Fg_o, rtol=1e-12, atol=1e-15)


def test_duffy_only_touches_singular_pairs(orc):
    m = M.cylinder(n_around=12, n_along=4, cap_tris=30)
    Fd = gpu(m, 3, "duffy")[0]
    Fq = gpu(m, 3, "quad")[0]
    diff = Fd != Fq
    # every differing entry is a Quad8 pair sharing a corner node
    for i, j in zip(*np.nonzero(diff)):
        assert m.family[i] == 0 and m.family[j] == 0
        assert set(m.nodes[i, :4]) & set(m.nodes[j, :4])
