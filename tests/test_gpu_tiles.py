"""Tile width (rvf_plan_tune_tiles): the columns per pair tile change which
wave sums which pairs and, for the wave-queue kernels, how many tiles the
persistent launch claims, but not the pair sums themselves.

Tolerances:
  F, A, per-pair blocked-ray counts and hashes, the diagonal with self_vf:
      bit-identical across widths, and FP64 F bit-identical to the oracle;
  row sums, F_group: the same terms added in another grouping,
      <= 1e-13 relative across widths (<= 1e-12 against the oracle).
"""
import numpy as np
import pytest

from rvf_amd import _lib, api
from rvf_amd import mesh as M

pytestmark = pytest.mark.gpu

WIDTHS = (1, 2, 5, 16)


def _scene():
    m = M.occluded_plates(n=12, subdiv=2)
    return m, [3]


def _solve_widths(mesh, obs, nquad, precision, variant=0, self_vf=False, widths=WIDTHS):
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, "geometric") if len(obs) else None
    plan = api.Plan(ctx, dm, do, nquad, "quad", precision, want_F=True, audit=do is not None, variant=variant,
                    self_vf=self_vf)
    out = {}
    for w in (0,) + tuple(widths):
        plan.tune_tiles(w)
        plan.run(); plan.sync()
        F, A, rs, Fg, Ag, tags, ac, ah = plan.fetch()
        st = plan.stats()
        out[w] = dict(F=F, A=A, rs=rs, Fg=Fg, Ag=Ag, ac=ac, ah=ah, cols=plan.columns, tiles=st["tiles"],
                      rays=st["rays"], culled=st["culled_rays"])
    plan.close(); dm.close()
    if do is not None:
        do.close()
    return out


def _same(a, b, audited):
    assert np.array_equal(a["F"], b["F"]), np.nanmax(np.abs(a["F"] - b["F"]))
    assert np.array_equal(a["A"], b["A"])
    assert np.allclose(a["rs"], b["rs"], rtol=1e-13, atol=1e-15)
    assert np.allclose(a["Fg"], b["Fg"], rtol=1e-13, atol=1e-15)
    assert np.array_equal(a["Ag"], b["Ag"])
    if audited:
        assert np.array_equal(a["ac"], b["ac"]) and np.array_equal(a["ah"], b["ah"])
        assert a["rays"] == b["rays"] and a["culled"] == b["culled"]


@pytest.mark.parametrize("precision,variant", [("mixed", 0), ("fp64", 0), ("fp64", api.VARIANT_LANE),
                                               ("mixed", api.VARIANT_QUEUE(3))])
def test_occluded_results_do_not_depend_on_tile_width(orc, precision, variant):
    m, obs = _scene()
    out = _solve_widths(m, obs, 3, precision, variant)
    lane = variant == api.VARIANT_LANE
    # the kernel's default width: mixed split (the mixed default) 2, FP64
    # queue (the FP64 default) 1, the other wave-queue variants 4, lane 16
    default = 16 if lane else (2 if precision == "mixed" else 1) if variant == 0 else 4
    assert out[0]["cols"] == default
    for w in WIDTHS:
        assert out[w]["cols"] == w
        _same(out[w], out[0], audited=True)
    # narrower tiles, more of them
    assert out[1]["tiles"] > out[2]["tiles"] > out[5]["tiles"] > out[16]["tiles"]
    if precision == "fp64":
        r = orc.compute(m, nquad=3, obstruction_groups=obs, parity_mode="geometric", audit=True)
        F = r["raw"] / r["A"][:, None]
        iu = np.triu_indices(m.n_elems, 1)
        for w in (1, 5):
            assert np.array_equal(out[w]["F"], F)
            assert np.allclose(out[w]["rs"], F.sum(1), rtol=1e-12, atol=1e-15)
            assert np.array_equal(out[w]["ac"][iu], r["audit_count"][iu])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_unoccluded_and_self_vf_tile_widths(precision):
    # unoccluded kernels (static launch, default 16 columns) and the self-pair
    # tiles (a tile that holds only an element with itself) at every width
    m = M.unit_cube(5)
    out = _solve_widths(m, (), 3, precision, self_vf=True, widths=(1, 3, 16))
    assert out[0]["cols"] == 16
    for w in (1, 3):
        _same(out[w], out[0], audited=False)


def test_tile_width_out_of_range_is_refused():
    ctx = api.context(0)
    m, obs = _scene()
    dm = api.DeviceMesh(ctx, m)
    do = api.DeviceObstruction(ctx, dm, obs, "geometric")
    plan = api.Plan(ctx, dm, do, 2, "quad", "mixed")
    for bad in (-1, 17, 1 << 20):
        with pytest.raises(_lib.RVFError, match="columns must be in 0..16"):
            plan.tune_tiles(bad)
    assert plan.columns == 2  # the mixed split default
    plan.run(); plan.sync()  # the plan is unchanged and still runs
    plan.close(); do.close(); dm.close()
