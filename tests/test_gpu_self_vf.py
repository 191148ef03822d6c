"""GPU parity of self view factors (RVF_SELF_VF: the pairs (i, i) too,
src/Assembly.jl:173,193) against the oracle run with self_vf.

Tolerances as for the other solves:
  FP64       F and A bit-identical, the diagonal included; row sums and
             F_group <= 1e-12 relative; blocked-ray counts and hashes of
             every pair (diagonal included) identical.
  mixed      max |dF| <= 1e-6 and closure <= 1e-6 against FP64 with
             self_vf, no flipped visibility decision.
  Monte Carlo  bit-identical (the pair (i, i) draws from its own stream).
The sphere mesh puts element 0 alone in a column chunk and ends another chunk
at slot 64, so two pair tiles hold nothing but an element with itself.
"""
import numpy as np
import pytest

from rvf_amd import api
from rvf_amd import mesh as M

from support.concave import sphere_groups

pytestmark = pytest.mark.gpu


def gpu(mesh, nquad, obs=(), precision="fp64", parity="reference", method="quad", variant=0, self_vf=True,
        n_samples=64, seed=0, graph=False, runs=1):
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, parity) if len(obs) else None
    plan = api.Plan(ctx, dm, do, nquad, method, precision, n_samples=n_samples, seed=seed, want_F=True,
                    audit=do is not None, variant=variant, self_vf=self_vf, graph=graph)
    out = []
    for _ in range(runs):
        plan.run(); plan.sync()
        F, A, rs, Fg, Ag, tags, ac, ah = plan.fetch()
        out.append(dict(F=F, A=A, rs=rs, Fg=Fg, Ag=Ag, tags=tags, ac=ac, ah=ah, stats=plan.stats()))
    plan.close(); dm.close()
    if do is not None:
        do.close()
    return out[0] if runs == 1 else out


def check(orc, mesh, nquad, obs=(), parity="reference", method="quad", variant=0, **kw):
    g = gpu(mesh, nquad, obs, "fp64", parity, method, variant, **kw)
    r = orc.compute(mesh, nquad=nquad, obstruction_groups=obs, parity_mode=parity, method=method,
                    audit=len(obs) > 0, self_vf=True, n_samples=kw.get("n_samples", 64), seed=kw.get("seed", 0))
    F = r["raw"] / r["A"][:, None]
    assert np.array_equal(g["A"], r["A"]), "areas differ"
    if method == "duffy":  # singular pairs: wave-summed non-negative terms
        rel = np.abs(g["F"] - F) / np.maximum(np.abs(F), 1e-300)
        assert float(rel.max()) <= 1e-12
    else:
        assert np.array_equal(np.diag(g["F"]), np.diag(F)), f"diagonal differs: {np.abs(np.diag(g['F'] - F)).max()}"
        assert np.array_equal(g["F"], F), f"F differs: max {np.abs(g['F'] - F).max()}"
    tags, names, Fg, Ag = orc.aggregate(mesh, F, r["A"])
    assert g["tags"] == tags
    assert np.allclose(g["Fg"], Fg, rtol=1e-12, atol=1e-15)
    assert np.allclose(g["rs"], F.sum(1), rtol=1e-12, atol=1e-15)
    if len(obs):  # the device audits each pair once, at [i + N j] for i <= j
        iu = np.triu_indices(mesh.n_elems, 0)
        assert np.array_equal(g["ac"][iu], r["audit_count"][iu]), "blocked-ray counts differ"
        assert np.array_equal(g["ah"][iu], r["audit_hash"][iu]), "blocked-ray masks differ"
    return g, F


def test_sphere_fp64_bitexact(orc):
    g, F = check(orc, sphere_groups(4), 3)
    assert np.all(np.diag(F) > 0)
    # without self_vf the same solve has a zero diagonal and the same pairs
    g0 = gpu(sphere_groups(4), 3, self_vf=False)
    off = ~np.eye(F.shape[0], dtype=bool)
    assert np.array_equal(g0["F"][off], g["F"][off]) and not np.any(np.diag(g0["F"]))
    assert np.allclose(g["rs"] - g0["rs"], np.diag(F), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("nquad", [2, 3, 5])
def test_curved_mixed_families_fp64(orc, nquad):
    check(orc, M.cylinder(n_around=12, n_along=4, cap_tris=40), nquad)


@pytest.mark.parametrize("variant", api.FP64_VARIANTS)
@pytest.mark.parametrize("parity", ["geometric", "reference"])
def test_occluded_fp64_variants(orc, parity, variant):
    g, _ = check(orc, sphere_groups(4, occluders=True), 3, obs=[9], parity=parity, variant=variant)
    if parity == "geometric":
        assert list(np.nonzero(np.diag(g["ac"]))[0]) == [10]


def test_duffy_fp64(orc):
    g, F = check(orc, M.cylinder(n_around=12, n_along=4, cap_tris=30), 3, method="duffy")
    assert g["stats"]["singular_pairs"] > 0


def test_monte_carlo_bitexact(orc):
    check(orc, M.cylinder(n_around=8, n_along=3, cap_tris=20), 3, method="mc", n_samples=49, seed=11)


@pytest.mark.parametrize("variant", [0] + list(api.MIXED_VARIANTS))
def test_mixed_occluded_tolerance(variant):
    m = sphere_groups(4, occluders=True)
    a = gpu(m, 3, [9], "fp64", "geometric")
    b = gpu(m, 3, [9], "mixed", "geometric", variant=variant)
    assert float(np.abs(a["F"] - b["F"]).max()) <= 1e-6
    assert float(np.abs(np.diag(a["F"]) - np.diag(b["F"])).max()) <= 1e-6
    assert float(np.abs(a["rs"] - b["rs"]).max()) <= 1e-6
    ca, cb = a["ac"].astype(np.int64), b["ac"].astype(np.int64)
    flips = int(np.abs(ca - cb).sum() + 2 * np.sum((ca == cb) & (a["ah"] != b["ah"])))
    assert flips == 0, flips


def test_mixed_unoccluded_tolerance():
    m = M.cylinder(n_around=16, n_along=4, cap_tris=40)
    a, b = gpu(m, 4, precision="fp64"), gpu(m, 4, precision="mixed")
    assert float(np.abs(a["F"] - b["F"]).max()) <= 1e-6
    assert float(np.abs(a["rs"] - b["rs"]).max()) <= 1e-6


def test_graph_replays_match_eager():
    m = sphere_groups(4, occluders=True)
    ref = gpu(m, 3, [9], "mixed", "geometric")
    for got in gpu(m, 3, [9], "mixed", "geometric", graph=True, runs=3):
        assert np.array_equal(got["F"], ref["F"]) and np.array_equal(got["rs"], ref["rs"])
        assert np.array_equal(got["Fg"], ref["Fg"])


def test_server_matches_standalone():
    ctx = api.context(0)
    reqs = [(sphere_groups(4, occluders=True), [9], "mixed"), (M.cylinder(n_around=12, n_along=4, cap_tris=30), [],
                                                               "fp64")]
    ref = [gpu(m, 3, obs, prec, "geometric") for m, obs, prec in reqs]
    for direct in (-1, 1):
        srv = api.Server(ctx, max_active=2, round_tiles=8, direct_tiles=direct)
        tickets = [srv.submit(m, 3, obs, "geometric", prec, want_F=True, self_vf=True) for m, obs, prec in reqs]
        for t, r in zip(tickets, ref):
            assert srv.wait(t, 120000) == api.REQ_DONE
            F, A, rs, Fg, Ag, tags, st = srv.result(t)
            assert np.array_equal(F, r["F"]) and np.array_equal(rs, r["rs"]) and np.array_equal(Fg, r["Fg"])
        srv.close()
