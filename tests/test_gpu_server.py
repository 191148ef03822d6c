"""Request server (continuous batching): every result equals the standalone
solve of the same request bit for bit, whatever else is in the mix; admission
follows the device-memory budget; over-budget and cancelled requests end with
the right state."""
import threading

import numpy as np
import pytest

from rvf_amd import api
from rvf_amd import mesh as M
from rvf_amd._lib import RVFError

pytestmark = pytest.mark.gpu


def _requests(seed, n, lo=60, hi=400):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        mesh, obs, nq = M.random_box_enclosure(rng, int(rng.integers(lo, hi)), max_occluders=2,
                                               max_occluder_tris=400)
        out.append((mesh, obs, int(nq), "mixed" if k % 2 else "fp64"))
    return out


def _alone(ctx, mesh, obs, nq, prec):
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, "geometric") if obs else None
    plan = api.Plan(ctx, dm, do, nq, "quad", prec, want_F=True)
    plan.run(); plan.sync()
    F, A, rs, Fg, Ag, tags, _, _ = plan.fetch()
    plan.close(); dm.close()
    if do is not None:
        do.close()
    return F, A, rs, Fg, Ag, tags


@pytest.mark.parametrize("direct_tiles", [-1, 1, 0])
def test_results_identical_alone_and_in_mix(direct_tiles):
    # -1: every request through the persistent batched kernel; 1: every request
    # through direct tile-range launches on the slot streams; 0: the default split
    ctx = api.context(0)
    reqs = _requests(0x5EED0C01, 10)
    ref = [_alone(ctx, *r) for r in reqs]
    # small rounds and a small slot count force interleaving and admission waves
    srv = api.Server(ctx, max_active=4, round_tiles=8, direct_tiles=direct_tiles)
    tickets = [srv.submit(m, nq, obs, "geometric", prec, want_F=True) for m, obs, nq, prec in reqs]
    for t in tickets:
        assert srv.wait(t, 120000) == api.REQ_DONE
    for t, r in zip(tickets, ref):
        F, A, rs, Fg, Ag, tags, st = srv.result(t)
        assert np.array_equal(F, r[0]) and np.array_equal(A, r[1]) and np.array_equal(rs, r[2])
        assert np.array_equal(Fg, r[3]) and np.array_equal(Ag, r[4]) and tags == r[5]
        assert st["t_submit_ms"] <= st["t_admit_ms"] <= st["t_start_ms"] <= st["t_done_ms"]
    c = srv.stats()
    assert c["completed"] == len(reqs) and c["failed"] == 0
    if direct_tiles == -1:
        assert c["rounds"] > len(reqs)  # round_tiles=8 splits every request over rounds
    elif direct_tiles == 1:
        assert c["rounds"] == 0         # whole tile ranges on the slot streams, no rounds
    srv.close()


def test_requests_overlap_on_the_device():
    # a long request and short ones submitted after it: the short ones start
    # while the long one is still running (continuous batching, not a queue)
    ctx = api.context(0)
    big = M.unit_cube(24)
    V, T = M.icosphere(4, 0.3, (0.5, 0.5, 0.5), seed=5)
    big.group_tri_soup = {7: V[T]}
    big.soup_layout = "xyz_vertex"
    srv = api.Server(ctx, round_tiles=256)
    tb = srv.submit(big, 3, [7], "geometric", "mixed")
    small = [srv.submit(m, nq, obs, "geometric", prec) for m, obs, nq, prec in _requests(0x5EED0C02, 4)]
    for t in small + [tb]:
        assert srv.wait(t, 120000) == api.REQ_DONE
    sb = srv.request_stats(tb)
    for t in small:
        s = srv.request_stats(t)
        assert s["t_start_ms"] < sb["t_done_ms"] and s["t_done_ms"] < sb["t_done_ms"]
    srv.close()


def test_admission_respects_the_budget():
    ctx = api.context(0)
    reqs = [(m, obs, 3, prec) for m, obs, _, prec in _requests(0x5EED0C03, 4, 300, 400)]
    srv = api.Server(ctx)
    tickets = [srv.submit(m, nq, obs, "geometric", prec) for m, obs, nq, prec in reqs]
    for t in tickets:
        assert srv.wait(t, 120000) == api.REQ_DONE
    fp = [srv.request_stats(t)["footprint_bytes"] for t in tickets]
    srv.close()
    # a budget that holds any one of them but no two serialises the requests
    budget = 2 * min(fp) - 1
    assert max(fp) <= budget
    srv = api.Server(ctx, hbm_budget=budget, round_tiles=64)
    tickets = [srv.submit(m, nq, obs, "geometric", prec) for m, obs, nq, prec in reqs]
    for t in tickets:
        assert srv.wait(t, 120000) == api.REQ_DONE
    st = [srv.request_stats(t) for t in tickets]
    for a, b in zip(st, st[1:]):
        assert b["t_admit_ms"] >= a["t_done_ms"]
    assert srv.stats()["peak_bytes"] <= budget
    srv.close()


def test_over_budget_request_fails_and_others_run():
    ctx = api.context(0)
    srv = api.Server(ctx, hbm_budget=2 << 20)
    big = srv.submit(M.unit_cube(12), 3, want_F=True)
    small = srv.submit(M.parallel_plates(2), 2)
    assert srv.wait(big, 120000) == api.REQ_FAILED
    with pytest.raises(RVFError, match="budget"):
        srv.result(big)
    assert srv.wait(small, 120000) == api.REQ_DONE
    srv.close()


def test_cancel_pending_request():
    ctx = api.context(0)
    srv = api.Server(ctx, max_active=1, round_tiles=16)
    first = srv.submit(M.unit_cube(16), 3)
    second = srv.submit(M.unit_cube(8), 3)
    srv.cancel(second)
    assert srv.wait(second, 120000) == api.REQ_CANCELLED
    assert srv.wait(first, 120000) == api.REQ_DONE
    with pytest.raises(RVFError):
        srv.result(second)
    srv.close()


def test_concurrent_clients():
    ctx = api.context(0)
    srv = api.Server(ctx, round_tiles=32)
    reqs = _requests(0x5EED0C04, 8)
    ref = [_alone(ctx, *r) for r in reqs]
    results = [None] * len(reqs)

    def client(k):
        m, obs, nq, prec = reqs[k]
        t = srv.submit(m, nq, obs, "geometric", prec)
        assert srv.wait(t, 120000) == api.REQ_DONE
        results[k] = srv.result(t)

    th = [threading.Thread(target=client, args=(k,)) for k in range(len(reqs))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for got, r in zip(results, ref):
        assert np.array_equal(got[2], r[2]) and np.array_equal(got[3], r[3])
    srv.close()
