"""Autotuner (rvf_plan_autotune, api.TuneTable): it measures every traversal
variant valid for the plan's precision, keeps the library default unless
another configuration beats it by more than its run-to-run spread, and the solve that
follows gives the same result as the default configuration (bit for bit in
FP64 and in mixed precision, whose blocked rays come from exact FP64 tests);
a persisted table entry is applied without measuring."""
import numpy as np
import pytest

from rvf_amd import api
from rvf_amd import mesh as M

pytestmark = pytest.mark.gpu


def _scene():
    m = M.unit_cube(10)
    V, T = M.icosphere(3, 0.3, (0.5, 0.5, 0.5), seed=0x5EED0003)
    m.group_tri_soup = {7: V[T]}
    return m


def _solve(ctx, dm, do, prec, tune=None, **kw):
    plan = api.Plan(ctx, dm, do, 3, "quad", prec, want_F=True, audit=True)
    out = None
    if tune is not None:
        out = plan.autotune(**tune)
    plan.run(); plan.sync()
    res = plan.fetch()
    name = plan.kernel_name()
    cols = plan.columns
    plan.close()
    return res, out, (name, cols)


def _mixed_split(variant, prec):
    return prec == "mixed" and variant in (api.VARIANT_SPLIT(1), api.VARIANT_SPLIT(2), api.VARIANT_SPLIT(3))


def _default_columns(variant, prec):
    # the mixed split variants run on tiles of 2 columns, the FP64 queue
    # variants 1, the other wave-queue variants 4, the lane variant 16
    if variant == api.VARIANT_LANE:
        return 16
    if _mixed_split(variant, prec):
        return 2
    if prec == "fp64" and variant in (api.VARIANT_QUEUE(1), api.VARIANT_QUEUE(2), api.VARIANT_QUEUE(3)):
        return 1
    return 4


WIDTHS = {1, 2, 4, 8, 16}


def _default_refill(variant, prec):
    # lanes waiting before a refill: 24 for the mixed split variants, 16 for
    # the other wave-queue variants (the lane variant has none)
    if variant == api.VARIANT_LANE:
        return 0
    split = variant in (api.VARIANT_SPLIT(1), api.VARIANT_SPLIT(2), api.VARIANT_SPLIT(3))
    return 24 if (prec == "mixed" and split) else 16


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_autotune_measures_all_variants_and_keeps_results(prec, tmp_path):
    ctx = api.context(0)
    m = _scene()
    dm = api.DeviceMesh(ctx, m)
    do = api.DeviceObstruction(ctx, dm, [7], "geometric")
    ref, _, (_, ref_cols) = _solve(ctx, dm, do, prec)
    table = api.TuneTable(str(tmp_path / "tune.json"))
    got, (v, r, meas), (name, cols) = _solve(ctx, dm, do, prec, tune=dict(sample_tiles=512, table=table))
    variants = api.FP64_VARIANTS if prec == "fp64" else api.MIXED_VARIANTS
    measured = {e["variant"] for e in meas}
    assert set(variants) <= measured, (sorted(measured), variants)
    assert all(e["ms"] > 0 for e in meas)
    # the first len(variants) entries rank the sample, each variant at its
    # default tile width and refill threshold; then the library default
    # (split 3 mixed, queue 1 FP64) is timed twice on the larger set, then the
    # three fastest others, then the choice's other refill thresholds of 8, 16,
    # 24 and 32 (wave-queue variants), then its other tile widths of 4, 8 and 16
    # columns, then (mixed split kernels) their three-wave build.  Another
    # configuration replaces the current choice only if it is faster than the
    # best so far by more than twice the default's spread (at least 1 %).
    assert all(e["columns"] == _default_columns(e["variant"], prec) for e in meas[:len(variants)])
    assert all(e["refill_min"] == _default_refill(e["variant"], prec) for e in meas[:len(variants)])
    rd = lambda v: _default_refill(v, prec)  # noqa: E731
    vdef = api.VARIANT_SPLIT(3) if prec == "mixed" else api.VARIANT_QUEUE(1)
    ranked = [e["variant"] for e in sorted(meas[:len(variants)], key=lambda e: e["ms"]) if e["variant"] != vdef]
    again = meas[len(variants):]
    d1, d2 = again[0], again[1]
    assert (d1["variant"], d1["refill_min"]) == (vdef, rd(vdef)) == (d2["variant"], d2["refill_min"])
    assert [e["variant"] for e in again[2:5]] == ranked[:3]
    best_ms = min(d1["ms"], d2["ms"])
    margin = max(0.01, 2 * abs(d1["ms"] - d2["ms"]) / best_ms)
    choice = (vdef, rd(vdef), 0)
    for e in again[2:5]:
        assert e["refill_min"] == rd(e["variant"])
        if e["ms"] < best_ms * (1 - margin):
            best_ms, choice = e["ms"], (e["variant"], rd(e["variant"]), 0)
    k = 5
    if choice[0] != api.VARIANT_LANE:
        others = sorted({8, 16, 24, 32} - {rd(choice[0])})
        assert sorted(e["refill_min"] for e in again[k:k + len(others)]) == others
        for e in again[k:k + len(others)]:
            assert e["variant"] == choice[0]
            assert e["columns"] == _default_columns(choice[0], prec)
            if e["ms"] < best_ms * (1 - margin):
                best_ms, choice = e["ms"], (e["variant"], e["refill_min"], 0)
        k += len(others)
    nw = len(WIDTHS - {_default_columns(choice[0], prec)})
    widths = [e["columns"] for e in again[k:k + nw]]
    assert sorted(widths) == sorted(WIDTHS - {_default_columns(choice[0], prec)})
    for e in again[k:k + nw]:
        assert e["variant"] == choice[0] and e["refill_min"] == choice[1] and e["waves"] == 0
        if e["ms"] < best_ms * (1 - margin):
            best_ms, choice = e["ms"], (choice[0], choice[1], e["columns"])
    k += nw
    waves = 0
    if _mixed_split(choice[0], prec):
        assert len(again) == k + 1
        e = again[k]
        assert (e["variant"], e["refill_min"], e["waves"]) == (choice[0], choice[1], 3)
        assert e["columns"] == (choice[2] or _default_columns(choice[0], prec))
        if e["ms"] < best_ms * (1 - margin):
            waves = 3
    else:
        assert len(again) == k
    assert (v, r) == choice[:2] and v in variants
    assert cols == (choice[2] or _default_columns(v, prec))
    assert name.startswith("k_pair_w3<" if waves == 3 else "k_pair<"), name
    names = ("F", "A", "rowsum", "F_group", "A_group", "tags", "audit_count", "audit_hash")
    for nm, a, b in zip(names, got, ref):
        if nm in ("rowsum", "F_group") and cols != ref_cols:
            # sums over another tile width: the same terms added in another order
            assert np.allclose(a, b, rtol=1e-13, atol=1e-15), nm
        elif isinstance(a, np.ndarray):
            assert np.array_equal(a, b), nm
        else:
            assert a == b, nm
    # the stored choice is applied by a new plan without measuring
    table2 = api.TuneTable(str(tmp_path / "tune.json"))
    got2, (v2, r2, meas2), name2 = _solve(ctx, dm, do, prec, tune=dict(table=table2))
    assert meas2 == [] and (v2, r2) == (v, r) and name2 == (name, cols)
    assert np.array_equal(got2[0], ref[0])
    do.close(); dm.close()


def test_autotune_unoccluded_plan_is_a_no_op():
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, M.unit_cube(6))
    plan = api.Plan(ctx, dm, None, 3, "quad", "fp64")
    name = plan.kernel_name()
    v, r, meas = plan.autotune()
    assert meas == [] and plan.kernel_name() == name and plan.columns == 16
    plan.close(); dm.close()


def test_three_wave_build_gives_identical_results():
    """rvf_plan_tune_waves: the mixed split kernels' three-wave build runs the
    same tiles with the same arithmetic, so F, the sums and every decision are
    bit-identical; other settings are refused."""
    ctx = api.context(0)
    m = _scene()
    dm = api.DeviceMesh(ctx, m)
    do = api.DeviceObstruction(ctx, dm, [7], "geometric")
    out = {}
    for w in (0, 3, 4):
        plan = api.Plan(ctx, dm, do, 3, "quad", "mixed", want_F=True, audit=True)
        plan.tune_waves(w)
        assert plan.waves == (3 if w == 3 else 0)
        plan.run(); plan.sync()
        out[w] = (plan.fetch(), plan.kernel_name())
        for bad in (1, 2, 5, -1):
            with pytest.raises(RuntimeError):
                plan.tune_waves(bad)
        plan.tune(api.VARIANT_QUEUE(2))  # another kernel: back to its default build
        assert plan.waves == 0
        with pytest.raises(RuntimeError):
            plan.tune_waves(3)
        plan.close()
    assert out[0][1] == out[4][1] == "k_pair<mixed_occ_split3>" and out[3][1] == "k_pair_w3<mixed_occ_split3>"
    for a, b in zip(out[0][0], out[3][0]):
        if isinstance(a, np.ndarray):
            assert np.array_equal(a, b)
        else:
            assert a == b
    do.close(); dm.close()
