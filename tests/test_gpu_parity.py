"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
inputs.

Tolerances (stated per test):
  FP64 mode   F, A bit-identical (the kernel sums in the reference's order);
              F_group / row sums <= 1e-12 relative; visibility: per-pair
              blocked-ray count and order-free hash identical.
  mixed mode  max |dF| <= 1e-6, closure |rowsum - rowsum_fp64| <= 1e-6,
              flipped visibility decisions reported (0 expected).
"""
import numpy as np
import pytest

from rvf_amd import api
from rvf_amd import mesh as M

pytestmark = pytest.mark.gpu


def gpu(mesh, nquad, obs=(), precision="fp64", parity="reference", audit=False, want_F=True, variant=0):
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, parity) if len(obs) else None
    plan = api.Plan(ctx, dm, do, nquad, "quad", precision, want_F=want_F, audit=audit, variant=variant)
    plan.run(); plan.sync()
    F, A, rs, Fg, Ag, tags, ac, ah = plan.fetch()
    st = plan.stats()
    plan.close(); dm.close()
    if do is not None:
        do.close()
    return dict(F=F, A=A, rs=rs, Fg=Fg, Ag=Ag, tags=tags, ac=ac, ah=ah, stats=st)


def oracle(orc, mesh, nquad, obs=(), parity="reference", audit=False, rows=None):
    r = orc.compute(mesh, nquad=nquad, obstruction_groups=obs, parity_mode=parity, audit=audit, rows=rows)
    F = r["raw"] / r["A"][:, None]
    return r, F


def check_fp64(orc, mesh, nquad, obs=(), parity="reference", variant=0):
    g = gpu(mesh, nquad, obs, "fp64", parity, audit=len(obs) > 0, variant=variant)
    r, F = oracle(orc, mesh, nquad, obs, parity, audit=len(obs) > 0)
    assert np.array_equal(g["A"], r["A"]), "areas differ"
    assert np.array_equal(g["F"], F), f"F differs: max {np.abs(g['F'] - F).max()}"
    tags, names, Fg, Ag = orc.aggregate(mesh, F, r["A"])
    assert g["tags"] == tags
    assert np.allclose(g["Fg"], Fg, rtol=1e-12, atol=1e-15)
    assert np.allclose(g["Ag"], Ag, rtol=1e-13, atol=0)
    assert np.allclose(g["rs"], F.sum(1), rtol=1e-12, atol=1e-15)
    if len(obs):
        iu = np.triu_indices(mesh.n_elems, 1)
        assert np.array_equal(g["ac"][iu], r["audit_count"][iu]), "blocked-ray counts differ"
        assert np.array_equal(g["ah"][iu], r["audit_hash"][iu]), "blocked-ray masks differ"
    return g, r, F


def test_c1_plates_fp64_bitexact(orc):
    m = M.parallel_plates(16)
    g, r, F = check_fp64(orc, m, 4)
    assert abs(g["Fg"][0, 1] - 0.1998248957) < 1e-6


def test_cube_fp64_bitexact(orc):
    check_fp64(orc, M.unit_cube(8), 3)


@pytest.mark.parametrize("nquad", [1, 2, 3, 4, 5, 6, 8])
def test_nquad_sweep_ragged(orc, nquad):
    # 2 x 5 x 5 = 50 elements: not a multiple of the 64-lane strip
    check_fp64(orc, M.parallel_plates(5), nquad)


@pytest.mark.parametrize("nquad", [2, 3, 4, 5])
def test_curved_mixed_families(orc, nquad):
    # Quad8 side + Tri6 caps (no Duffy): 16*4 + 2*~40 elements
    check_fp64(orc, M.cylinder(n_around=16, n_along=4, cap_tris=40), nquad)


def test_random_boxes_fp64(orc):
    rng = np.random.default_rng(0x5EED0005)
    for _ in range(3):
        mesh, obs, nq = M.random_box_enclosure(rng, int(rng.integers(60, 300)), max_occluders=0)
        check_fp64(orc, mesh, nq)


@pytest.mark.parametrize("parity", ["reference", "geometric"])
def test_occluded_fp64_visibility_bitexact(orc, parity):
    m = M.occluded_plates(n=8, subdiv=2)
    g, r, F = check_fp64(orc, m, 3, obs=[3], parity=parity)
    assert g["ac"].sum() > 0  # the sphere blocks something


@pytest.mark.parametrize("variant", api.FP64_VARIANTS)
@pytest.mark.parametrize("parity", ["reference", "geometric"])
def test_occluded_fp64_variants_bitexact(orc, parity, variant):
    # every traversal variant gives the oracle's F and visibility bit for bit
    m = M.occluded_plates(n=8, subdiv=2)
    g, r, F = check_fp64(orc, m, 3, obs=[3], parity=parity, variant=variant)
    assert g["ac"].sum() > 0


@pytest.mark.parametrize("variant", [api.VARIANT_QUEUE(2), api.VARIANT_LANE])
def test_occluded_fp64_random_boxes_nquad5(orc, variant):
    # Quad8/Tri3/Quad4 mixes with box occluders; nquad 5 makes the (p, q)
    # combinations of a Quad8 pair (625) span several queue batches
    rng = np.random.default_rng(0x5EED0007)
    done = 0
    while done < 2:
        mesh, obs, nq = M.random_box_enclosure(rng, int(rng.integers(80, 160)), max_occluders=2,
                                               max_occluder_tris=300)
        if not obs:
            continue
        check_fp64(orc, mesh, 5, obs=obs, parity="geometric", variant=variant)
        done += 1


def test_obstruction_group_that_radiates(orc):
    # obstruction groups that are also radiating: one tree per active set
    # obstruction_groups \ {g_i, g_j} (src/Assembly.jl:119-143)
    m = M.unit_cube(4)
    m.group_tri_soup = M.build_group_obs_soups(m)
    V, T = M.icosphere(1, 0.2, (0.5, 0.5, 0.5), seed=3)
    m.group_tri_soup[9] = V[T]
    g, r, F = check_fp64(orc, m, 2, obs=[1, 2, 9], parity="geometric")


def test_unsorted_interleaved_groups(orc):
    m = M.parallel_plates(6)
    perm = np.random.default_rng(1).permutation(m.n_elems)
    m2 = M.MeshData(m.coords, m.family[perm], m.nodes[perm], m.group[perm] * 7 - 3,
                    {t * 7 - 3: n for t, n in m.group_tags.items()})
    check_fp64(orc, m2, 3)


def test_tiny_meshes(orc):
    m = M.parallel_plates(1)          # N = 2
    check_fp64(orc, m, 3)
    one = M.MeshData(m.coords, m.family[:1], m.nodes[:1], m.group[:1], {1: "bottom"})
    g = gpu(one, 3)
    assert g["F"].shape == (1, 1) and g["F"][0, 0] == 0.0 and g["A"][0] > 0


def test_degenerate_element_is_an_error():
    m = M.parallel_plates(2)
    m.nodes[0, :4] = m.nodes[0, 0]  # collapse the element to a point: dA = 0 everywhere
    with pytest.raises(api.RVFError):
        gpu(m, 2)


def test_invalid_inputs_are_errors():
    m = M.parallel_plates(2)
    bad = M.MeshData(m.coords.copy(), m.family, m.nodes.copy(), m.group, m.group_tags)
    bad.nodes[0, 0] = 10 ** 6
    with pytest.raises(api.RVFError):
        api.DeviceMesh(api.context(0), bad)
    nanm = M.MeshData(m.coords.copy(), m.family, m.nodes, m.group, m.group_tags)
    nanm.coords[0, 0] = np.nan
    with pytest.raises(api.RVFError):
        api.DeviceMesh(api.context(0), nanm)
    with pytest.raises(api.RVFError):
        gpu(m, 9)


# ------------------------------------------------------------------------------
# mixed precision
# ------------------------------------------------------------------------------
def mixed_vs_fp64(mesh, nquad, obs=(), parity="reference"):
    a = gpu(mesh, nquad, obs, "fp64", parity, audit=len(obs) > 0)
    b = gpu(mesh, nquad, obs, "mixed", parity, audit=len(obs) > 0)
    dF = float(np.abs(a["F"] - b["F"]).max())
    dclose = float(np.abs(a["rs"] - b["rs"]).max())
    flips = 0
    if len(obs):
        iu = np.triu_indices(mesh.n_elems, 1)
        ca, cb = a["ac"][iu].astype(np.int64), b["ac"][iu].astype(np.int64)
        flips = int(np.abs(ca - cb).sum() + 2 * np.sum((ca == cb) & (a["ah"][iu] != b["ah"][iu])))
    return dF, dclose, flips, a, b


def test_mixed_cube_tolerance():
    dF, dclose, _, a, b = mixed_vs_fp64(M.unit_cube(12), 3)
    assert dF <= 1e-6 and dclose <= 1e-6, (dF, dclose)
    assert np.allclose(a["Fg"], b["Fg"], atol=1e-6)


@pytest.mark.parametrize("nquad", [2, 4, 5])
def test_mixed_curved_tolerance(nquad):
    dF, dclose, _, a, b = mixed_vs_fp64(M.cylinder(n_around=16, n_along=4, cap_tris=40), nquad)
    assert dF <= 1e-6 and dclose <= 1e-6, (dF, dclose)


@pytest.mark.parametrize("parity", ["reference", "geometric"])
def test_mixed_occluded_flips(parity):
    dF, dclose, flips, a, b = mixed_vs_fp64(M.occluded_plates(n=12, subdiv=3), 3, obs=[3], parity=parity)
    assert dF <= 1e-6 and dclose <= 1e-6, (dF, dclose)
    assert flips == 0, flips
    assert b["stats"]["rays"] > 0


def test_mixed_random_boxes_with_occluders():
    rng = np.random.default_rng(0x5EED0006)
    for _ in range(3):
        mesh, obs, nq = M.random_box_enclosure(rng, int(rng.integers(100, 400)), max_occluders=3,
                                               max_occluder_tris=600)
        if not obs:
            continue
        dF, dclose, flips, a, b = mixed_vs_fp64(mesh, nq, obs=obs, parity="geometric")
        assert dF <= 1e-6 and dclose <= 1e-6 and flips == 0, (dF, dclose, flips)


# ------------------------------------------------------------------------------
# full-size workload: 13,824-element cube (size-independent properties + rows)
# ------------------------------------------------------------------------------
def test_cube48_full_size(orc):
    m = M.unit_cube(48)
    a = gpu(m, 3, (), "fp64")
    N = m.n_elems
    # oracle rows: first strip, a strip in the middle and the last rows
    for r0, r1 in ((0, 64), (N // 2, N // 2 + 32), (N - 40, N)):
        r, _ = oracle(orc, m, 3, rows=(r0, r1))
        for i in range(r0, r1):
            ref = r["raw"][i, i + 1:] / r["A"][i]
            assert np.array_equal(a["F"][i, i + 1:], ref)
            assert np.array_equal(a["F"][i + 1:, i], r["raw"][i, i + 1:] / r["A"][i + 1:])
        assert np.array_equal(a["A"], r["A"])
    # properties: reciprocity, closure, symmetry of A F
    AF = a["A"][:, None] * a["F"]
    assert np.abs(AF - AF.T).max() <= 1e-15
    assert np.allclose(a["rs"], a["F"].sum(1), rtol=1e-12)
    assert np.allclose(a["Fg"].sum(1), 1.0, atol=2e-2)
    b = gpu(m, 3, (), "mixed")
    assert float(np.abs(a["F"] - b["F"]).max()) <= 1e-6
    assert float(np.abs(a["rs"] - b["rs"]).max()) <= 1e-6


@pytest.mark.parametrize("variant", api.EXACT_COUNT_VARIANTS)
def test_fp64_traversal_counts_match_oracle(orc, variant):
    # the FP64 kernels walk the canonical stackless order (GPUBVH.jl:202-285):
    # rays, box tests and triangle tests equal the oracle's counts exactly
    m = M.occluded_plates(n=6, subdiv=2)
    g = gpu(m, 2, [3], "fp64", "geometric", want_F=False, variant=variant)
    r = orc.compute(m, nquad=2, obstruction_groups=[3], parity_mode="geometric", count=True, want_raw=False)
    assert (g["stats"]["rays"], g["stats"]["aabb_tests"], g["stats"]["tri_tests"]) == \
        (r["rays"], r["aabb_tests"], r["tri_tests"])


@pytest.mark.parametrize("variant", api.MIXED_VARIANTS)
def test_mixed_variants_agree(variant):
    # mixed-precision decisions are the exact FP64 test on the FP32 traversal's
    # candidates, so every variant reports the same blocked rays
    m = M.unit_cube(8)
    V, T = M.icosphere(3, 0.3, (0.5, 0.5, 0.5), seed=0x5EED0003)
    m.group_tri_soup = {7: V[T]}
    ref = gpu(m, 3, [7], "fp64", "geometric", audit=True)
    got = gpu(m, 3, [7], "mixed", "geometric", audit=True, variant=variant)
    iu = np.triu_indices(m.n_elems, 1)
    assert np.array_equal(got["ac"][iu], ref["ac"][iu]) and np.array_equal(got["ah"][iu], ref["ah"][iu])
    assert float(np.abs(got["F"] - ref["F"]).max()) <= 1e-6


def test_mixed_cube_with_sphere_flips():
    # the benchmark geometry at reduced size: cube interior + sphere occluder
    m = M.unit_cube(12)
    V, T = M.icosphere(3, 0.3, (0.5, 0.5, 0.5), seed=0x5EED0003)
    m.group_tri_soup = {7: V[T]}
    for parity in ("geometric", "reference"):
        dF, dclose, flips, a, b = mixed_vs_fp64(m, 3, obs=[7], parity=parity)
        assert a["ac"].sum() > 0
        assert flips == 0 and dF <= 1e-6 and dclose <= 1e-6, (parity, flips, dF, dclose)


def test_graph_replay_matches_eager():
    # the captured solve is replayed several times, with eager solves of
    # another plan in between; every replay equals the eager result
    ctx = api.context(0)
    m = M.occluded_plates(n=8, subdiv=2)
    dm = api.DeviceMesh(ctx, m)
    do = api.DeviceObstruction(ctx, dm, [3], "geometric")
    eager = api.Plan(ctx, dm, do, 3, "quad", "mixed")
    eager.run(); eager.sync()
    ref = eager.fetch()
    other = api.Plan(ctx, dm, None, 2, "quad", "fp64")
    g = api.Plan(ctx, dm, do, 3, "quad", "mixed", graph=True)
    for _ in range(3):
        g.run(); g.sync()
        got = g.fetch()
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])
        other.run(); other.sync()
    for pl in (eager, other, g):
        pl.close()
    do.close(); dm.close()


@pytest.mark.parametrize("variant", api.EXACT_COUNT_VARIANTS)
def test_culled_waves_counts_and_values(orc, variant):
    # a small occluder near one corner of the cube: most waves' pairs are all
    # culled and skip the trace; their rays are still counted as the oracle
    # counts them, and F, audits and counts stay exact
    m = M.unit_cube(6)
    V, T = M.icosphere(1, 0.05, (0.12, 0.1, 0.15), seed=3)
    m.group_tri_soup = {7: V[T]}
    g = gpu(m, 2, [7], "fp64", "geometric", audit=True, variant=variant)
    r = orc.compute(m, nquad=2, obstruction_groups=[7], parity_mode="geometric", audit=True, count=True)
    assert (g["stats"]["rays"], g["stats"]["aabb_tests"], g["stats"]["tri_tests"]) == \
        (r["rays"], r["aabb_tests"], r["tri_tests"])
    if variant != api.VARIANT_LANE:  # (the per-lane kernel traces every ray; only the wave kernels cull)
        assert g["stats"]["culled_rays"] > g["stats"]["rays"] // 2
        assert g["stats"]["culled_waves"] > 0  # some waves took the branch that skips the trace
    assert np.array_equal(g["F"], r["raw"] / r["A"][:, None])
    iu = np.triu_indices(m.n_elems, 1)
    assert np.array_equal(g["ac"][iu], r["audit_count"][iu]) and int(g["ac"][iu].sum()) > 0
    mx = gpu(m, 2, [7], "mixed", "geometric", audit=True)
    assert mx["stats"]["culled_rays"] > mx["stats"]["rays"] // 2 and mx["stats"]["culled_waves"] > 0
    assert np.array_equal(mx["ac"][iu], g["ac"][iu]) and np.array_equal(mx["ah"][iu], g["ah"][iu])
    assert float(np.abs(mx["F"] - g["F"]).max()) <= 1e-6


def test_split_variants_refused_past_task_index_width(orc):
    # a split-variant task word holds a triangle index in 19 bits: with 2^19 or
    # more obstruction triangles over all trees the split variants are refused
    # (explicitly or by the autotuner) and the mixed default is the queue form;
    # results stay exact there
    rng = np.random.default_rng(0x5EED0019)
    T = (1 << 19) + 1500
    cen = rng.uniform((0.15, 0.15, 0.35), (0.85, 0.85, 0.65), (T, 1, 3))
    soup = np.ascontiguousarray(cen + rng.uniform(-0.004, 0.004, (T, 3, 3)))
    m = M.parallel_plates(6)
    m.group_tri_soup = {3: soup}
    m.soup_layout = "xyz_vertex"
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, m)
    do = api.DeviceObstruction(ctx, dm, [3], "geometric")
    assert do.n_tris == T
    for prec, v in (("mixed", api.VARIANT_SPLIT(3)), ("mixed", api.VARIANT_SPLIT(1)), ("fp64", api.VARIANT_SPLIT(2))):
        with pytest.raises(api.RVFError, match="2\\^19"):
            api.Plan(ctx, dm, do, 2, "quad", prec, variant=v)
    pl = api.Plan(ctx, dm, do, 2, "quad", "mixed")
    assert pl.kernel_name() == "k_pair<mixed_occ_queue3>"
    with pytest.raises(api.RVFError, match="2\\^19"):
        pl.tune(api.VARIANT_SPLIT(3), 16)
    v, r, meas = pl.autotune(sample_tiles=64)
    assert meas and all(e["variant"] < 20 for e in meas) and v < 20
    pl.close(); do.close(); dm.close()
    g = gpu(m, 2, [3], "fp64", "geometric", audit=True)
    ro = orc.compute(m, nquad=2, obstruction_groups=[3], parity_mode="geometric", audit=True)
    iu = np.triu_indices(m.n_elems, 1)
    assert np.array_equal(g["F"], ro["raw"] / ro["A"][:, None])
    assert np.array_equal(g["ac"][iu], ro["audit_count"][iu]) and int(g["ac"][iu].sum()) > 0
    mx = gpu(m, 2, [3], "mixed", "geometric", audit=True)
    assert np.array_equal(mx["ac"][iu], g["ac"][iu]) and np.array_equal(mx["ah"][iu], g["ah"][iu])
    assert float(np.abs(mx["F"] - g["F"]).max()) <= 1e-6


def test_fetch_into_pinned_buffer():
    # F copied into a reused page-locked buffer equals the pageable copy, bit
    # for bit, over two solves (the buffer keeps no stale values); wrong shapes
    # are refused
    m = M.parallel_plates(10)
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, m)
    plan = api.Plan(ctx, dm, None, 2, "quad", "mixed", want_F=True)
    N = m.n_elems
    buf = api.pinned_empty((N, N))
    buf[:] = np.nan
    for _ in range(2):
        plan.run(); plan.sync()
        F, A, rs, *_ = plan.fetch()
        Fp, Ap, rsp, *_ = plan.fetch(F_out=buf)
        assert Fp is buf and np.array_equal(Fp, F) and np.array_equal(Ap, A) and np.array_equal(rsp, rs)
        buf[:] = -1.0
    with pytest.raises(ValueError):
        plan.fetch(F_out=np.empty((N, N)))  # C order
    with pytest.raises(ValueError):
        plan.fetch(F_out=api.pinned_empty((N, N - 1)))
    del buf
    plan.close(); dm.close()


def test_plan_size_guards_refuse_before_allocating():
    # 2 x 2048^2 = 8,388,608 elements at nquad 8 (64 points per element): the
    # mixed split kernels index the row points (8 x 64 x N FP32 values = 2^32)
    # with 32-bit offsets, so they are refused; the default mixed plan (the
    # queue form) then needs ~1.4e11 pair tiles, past the 2^31 tile indices,
    # and is refused before its tile list is built
    m = M.parallel_plates(2048)
    assert m.n_elems == 1 << 23
    m.group_tri_soup = {3: np.array([[[0.4, 0.4, 0.5], [0.6, 0.4, 0.5], [0.5, 0.6, 0.5]]])}
    m.soup_layout = "xyz_vertex"
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, m)
    do = api.DeviceObstruction(ctx, dm, [3], "geometric")
    with pytest.raises(api.RVFError, match="2\\^32"):
        api.Plan(ctx, dm, do, 8, "quad", "mixed", want_F=False, variant=api.VARIANT_SPLIT(3))
    with pytest.raises(api.RVFError, match="mesh too large for one plan"):
        api.Plan(ctx, dm, do, 8, "quad", "mixed", want_F=False)
    do.close(); dm.close()
