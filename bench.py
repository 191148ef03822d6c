"""Benchmark: element-pair view factors on one MI355X.

Default workload: the occluded 13,824-element enclosure — the unit-cube
interior (6 faces x 48 x 48 Quad4, inward normals, 6 groups) with the
20,480-triangle perturbed icosphere (radius 0.3, centre (0.5,0.5,0.5)) as an
obstruction-only group 7, nquad = 3, geometric soup reading.  A step is one
full solve (precompute, pair tiles with occlusion, F fill, row/group
reductions) with mesh and BVH resident in HBM; the F matrix stays on the
device.  Reported: FP32+refine ("mixed") pairs/s (N^2 ordered pairs) next to
the plain FP64 path, max |dF| / closure / flipped visibility decisions of
mixed vs FP64, p50/p99 solve latency, the pair kernel's roofline position and
the oracle (C restatement of the reference CPU path) timed on a row sample.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_FP32 = 157.3e12   # MI355X FP32 vector (AMD datasheet)
VALU_PEAK_FP64 = 78.6e12    # MI355X FP64 vector
TRAFFIC_COPY_BYTES = 256 << 20  # calibration copy: reads and writes this many bytes once


def build_workload(name):
    import __graft_entry__ as G
    G.load_package()
    from rvf_amd import mesh as M
    if name == "cube_occluded":
        m = M.unit_cube(48)
        V, T = M.icosphere(5, 0.3, (0.5, 0.5, 0.5), seed=0x5EED0003)
        m.group_tri_soup = {7: V[T]}
        m.soup_layout = "xyz_vertex"
        desc = ("unit-cube enclosure 6x48x48 Quad4 (13,824 elements) + 20,480-triangle icosphere occluder "
                "(obstruction-only group 7), nquad=3")
        return m, [7], 3, "geometric", desc
    if name == "cube":
        return M.unit_cube(48), [], 3, "reference", "unit-cube enclosure 6x48x48 Quad4 (13,824 elements), nquad=3"
    if name == "plates_occluded":
        m = M.occluded_plates(96, 5)
        return m, [3], 3, "geometric", ("occluded parallel plates 2x96x96 Quad4 (18,432 elements) + 20,480-"
                                        "triangle icosphere, nquad=3")
    if name == "plates":
        return M.parallel_plates(16), [], 4, "reference", "parallel plates 2x16x16 Quad4 (512 elements), nquad=4"
    raise SystemExit(f"unknown workload {name}")


def flop_model(N, npts, rays=0, aabb=0, tri=0):
    """Algorithmic flops per solve: per unique pair npi*npj*28 + npi*3, plus
    15 per ray, 24 per AABB test and 52 per triangle test."""
    pairs = N * (N - 1) // 2
    return pairs * (npts * npts * 28 + npts * 3) + 15 * rays + 24 * aabb + 52 * tri


def run_mode(api, ctx, dm, do, nquad, precision, steps, warmup, audit=False, tune=None, self_vf=False, lat_min=0,
             method="quad", n_samples=256, seed=0):
    """Times `steps` solves (HIP events on the context stream around them, the
    pair kernel's own events per solve); then, untimed, further solves until
    `lat_min` latency samples exist.  Returns (plan, ms per solve, latencies,
    pair-kernel ms of the timed solves)."""
    plan = api.Plan(ctx, dm, do, nquad, method, precision, want_F=True, audit=audit, timing=True, self_vf=self_vf,
                    n_samples=n_samples, seed=seed)
    if tune is not None and do is not None:
        # autotuner (setup, outside the timed region): time each traversal
        # configuration over a tile sample and keep the fastest
        t0 = time.perf_counter()
        v, r, meas = plan.autotune(sample_tiles=tune["sample_tiles"])
        tune.update(variant=v, refill_min=r, columns=plan.columns, waves=plan.waves, kernel=plan.kernel_name(),
                    seconds=time.perf_counter() - t0, measurements=meas)
    for _ in range(warmup):
        plan.run(); plan.sync()
    lat, pair_ms = [], []
    ctx.sync()
    ctx.record(0)
    for _ in range(steps):
        t0 = time.perf_counter()
        plan.run(); plan.sync()
        lat.append((time.perf_counter() - t0) * 1e3)
        pair_ms.append(plan.stats()["pair_ms"])
    ctx.record(1)
    total_ms = ctx.elapsed(0, 1)
    while len(lat) < lat_min:  # latency samples past the timed region
        t0 = time.perf_counter()
        plan.run(); plan.sync()
        lat.append((time.perf_counter() - t0) * 1e3)
    return plan, total_ms / steps, np.array(lat), np.array(pair_ms)


def end_to_end(plan, reps=3, F_out=None):
    """Solve plus the D2H copy of F (N x N doubles), areas, row sums and the
    group sums, as the reference's hook returns them (GPUAssembly.jl:83-104):
    into new pageable arrays each time, or (F_out) into one reused buffer."""
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.run(); plan.sync()
        plan.fetch(F_out=F_out)
        ms.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ms))


def graph_steady_state(api, ctx, dm, do, nquad, precision, tune, reps=10):
    """Steady-state solves replayed from a captured hipGraph (RVF_GRAPH) with
    the tuned traversal configuration: median solve time."""
    plan = api.Plan(ctx, dm, do, nquad, "quad", precision, want_F=True, timing=False, graph=True)
    if tune is not None and do is not None and "variant" in tune:
        plan.tune(tune["variant"], tune["refill_min"])
        plan.tune_tiles(tune["columns"])
        plan.tune_waves(tune["waves"])
    plan.run(); plan.sync()  # capture
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.run(); plan.sync()
        ms.append((time.perf_counter() - t0) * 1e3)
    name = plan.kernel_name()
    plan.close()
    return {"ms_p50": float(np.median(ms)), "replays": reps, "kernel": name}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of the pair kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary, calibrated against the copy kernel
    (profiles/traffic_<workload>.json); None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("kernel") != kernel:
        return None
    return t["bytes_per_launch"]


def cpu_baseline(mesh, nquad, obs, parity, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.lib()
    cores = os.cpu_count() or 1
    N = mesh.n_elems
    # rows spread over the matrix; grow the sample until the budget is used
    rng = np.random.default_rng(1)
    pairs = 0
    t0 = time.perf_counter()
    nrows = 0
    while time.perf_counter() - t0 < budget_s and nrows < N:
        r0 = int(rng.integers(0, N - 8))
        out = orc.compute(mesh, nquad=nquad, obstruction_groups=obs, parity_mode=parity, rows=(r0, r0 + 8),
                          want_raw=False, nthreads=cores)
        pairs += sum(N - 1 - i for i in range(r0, r0 + 8))
        nrows += 8
    dt = time.perf_counter() - t0
    # areas pass (serial) is included in the first call; negligible vs the rows
    return {"value": pairs / dt, "unit": "unique pairs/s", "cores": cores, "kind": "port",
            "sample": f"{nrows} rows x (N-1-i) upper-triangle pairs of the same workload, {dt:.1f} s, "
                      f"C restatement of _compute_cpu with OpenMP (contiguous row chunks), "
                      f"host {platform.processor() or platform.machine()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cube_occluded")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-secondary", dest="secondary", action="store_false")
    ap.add_argument("--serving-requests", type=int, default=24)
    ap.add_argument("--serving-rate", type=float, default=8.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--no-serving", dest="serving", action="store_false")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default traversal configuration instead of autotuning it")
    ap.add_argument("--tune-tiles", type=int, default=8192, help="tiles the autotuner times per configuration")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one copy of known size, then one solve (for rocprofv3 --pmc runs)")
    ap.add_argument("--probe-precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--section", default=None,
                    choices=["occluded_plates", "curved_cylinder", "unoccluded_cube", "self_vf_sphere", "curves",
                             "serving"],
                    help="run only this secondary section and print its JSON (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; sharding pair tiles over several GPUs is designed but not built",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as G
    G.load_package()
    from rvf_amd import api
    if args.section:
        ctx = api.context(0)
        fn = {"occluded_plates": lambda: secondary_plates(api, ctx, args.steps),
              "curved_cylinder": lambda: secondary_cylinder(api, ctx),
              "unoccluded_cube": lambda: secondary_cube(api, ctx, args.steps),
              "self_vf_sphere": lambda: secondary_self_vf(api, ctx, args.steps),
              "curves": lambda: secondary_curves(api, ctx, args.steps, 0.0),
              "serving": lambda: serving_stream(api, ctx, args.serving_requests, args.serving_rate)}[args.section]
        print(json.dumps({args.section: fn()}))
        return
    mesh, obs, nquad, parity, desc = build_workload(args.workload)
    N = mesh.n_elems
    ctx = api.context(0)
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, parity) if obs else None
    npts = nquad * nquad

    if args.traffic_probe:
        ctx.copy_probe(TRAFFIC_COPY_BYTES, 1)
        plan = api.Plan(ctx, dm, do, nquad, "quad", args.probe_precision, want_F=True, timing=True)
        plan.run(); plan.sync()
        print(json.dumps({"copy_bytes": TRAFFIC_COPY_BYTES, "kernel": plan.kernel_name(),
                          "pair_ms": plan.stats()["pair_ms"]}))
        return

    tune_m = {"sample_tiles": args.tune_tiles} if args.autotune else None
    tune_d = {"sample_tiles": args.tune_tiles} if args.autotune else None
    plan_m, ms_m, lat_m, pair_m = run_mode(api, ctx, dm, do, nquad, "mixed", args.steps, args.warmup, tune=tune_m,
                                           lat_min=30)
    st_m = plan_m.stats()
    plan_d, ms_d, lat_d, pair_d = run_mode(api, ctx, dm, do, nquad, "fp64", max(2, args.steps // 2),
                                           min(args.warmup, 1), tune=tune_d, lat_min=10)
    st_d = plan_d.stats()
    e2e_m = end_to_end(plan_m)
    Fpin = api.pinned_empty((N, N))
    e2e_pin = end_to_end(plan_m, F_out=Fpin)
    del Fpin
    graph_m = graph_steady_state(api, ctx, dm, do, nquad, "mixed", tune_m)

    # verification of mixed against FP64 on the same inputs (outside the timed region)
    ver = {}
    if not args.no_verify:
        Fm, Am, rsm, Fgm, *_ = plan_m.fetch()
        Fd, Ad, rsd, Fgd, *_ = plan_d.fetch()
        ver["max_abs_dF"] = float(np.abs(Fm - Fd).max())
        ver["closure_err"] = float(np.abs(rsm - rsd).max())
        ver["max_abs_dF_group"] = float(np.abs(Fgm - Fgd).max())
        ver["rowsum_minus_1_max_fp64"] = float(np.abs(rsd - 1.0).max())
        del Fm, Fd
        if do is not None:
            # the audited plans run the timed (tuned) traversal configurations
            pa = api.Plan(ctx, dm, do, nquad, "quad", "fp64", want_F=False, audit=True, timing=False)
            if tune_d and "variant" in tune_d:
                pa.tune(tune_d["variant"], tune_d["refill_min"])
                pa.tune_tiles(tune_d["columns"])
                pa.tune_waves(tune_d["waves"])
            pa.run(); pa.sync()
            *_, ca, ha = pa.fetch(F=False)
            ver["audited_kernels"] = {"fp64": pa.kernel_name()}
            pa.close()
            pb = api.Plan(ctx, dm, do, nquad, "quad", "mixed", want_F=False, audit=True, timing=False)
            if tune_m and "variant" in tune_m:
                pb.tune(tune_m["variant"], tune_m["refill_min"])
                pb.tune_tiles(tune_m["columns"])
                pb.tune_waves(tune_m["waves"])
            pb.run(); pb.sync()
            *_, cb, hb = pb.fetch(F=False)
            ver["audited_kernels"]["mixed"] = pb.kernel_name()
            pb.close()
            iu = np.triu_indices(N, 1)
            ca, cb, ha, hb = ca[iu].astype(np.int64), cb[iu].astype(np.int64), ha[iu], hb[iu]
            ver["flipped_visibility"] = int(np.abs(ca - cb).sum() + 2 * np.sum((ca == cb) & (ha != hb)))
            ver["blocked_rays_fp64"] = int(ca.sum())
            del ca, cb, ha, hb

    ordered = float(N) * N
    value = ordered / (ms_m * 1e-3)
    # roofline for the dominant kernel (mixed pair kernel)
    flops = flop_model(N, npts, st_d.get("rays", 0), st_d.get("aabb_tests", 0), st_d.get("tri_tests", 0))
    pair_avg_s = float(pair_m.mean()) * 1e-3
    achieved = flops / pair_avg_s
    roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / VALU_PEAK_FP32, "traffic": measured_traffic(args.workload, plan_m.kernel_name()),
            "kernel": plan_m.kernel_name(),
            "flops_per_launch": flops, "avg_launch_ms": float(pair_m.mean())}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, nquad, obs, parity, args.cpu_seconds)
    out = {
        "metric": "view-factor pairs/s (occluded, FP32+refine vs FP64) at |dF|<=1e-6; p50/p99",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_m, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 bulk + fp64 accumulate/refine", "data": "synthetic (deterministic seeded meshes)",
        "config": {"workload": desc, "elements": N, "ordered_pairs": int(ordered),
                   "unique_pairs": N * (N - 1) // 2, "precision": "mixed", "soup_reading": parity},
        "fp64": {"pairs_per_s": ordered / (ms_d * 1e-3), "ms_per_step": ms_d,
                 "pair_kernel_ms": float(pair_d.mean())},
        "speedup_mixed_vs_fp64": ms_d / ms_m,
        "latency_ms": {"p50": float(np.percentile(lat_m, 50)), "p99": float(np.percentile(lat_m, 99)),
                       "solves": int(len(lat_m)), "fp64_p50": float(np.percentile(lat_d, 50)),
                       "fp64_p99": float(np.percentile(lat_d, 99)), "fp64_solves": int(len(lat_d))},
        "end_to_end": {"ms": e2e_m, "pairs_per_s": ordered / (e2e_m * 1e-3),
                       "includes": "solve + D2H of F (N^2 x 8 B) + areas, row sums, group sums (PCIe-inclusive)",
                       "ms_pinned_F": e2e_pin,
                       "pinned_F": "the same into one reused page-locked F buffer (api.pinned_empty)"},
        "graph_replay": graph_m,
        "kernel_ms": {"mixed": {k: st_m[k] for k in ("precompute_ms", "pair_ms", "fill_ms", "reduce_ms")},
                      "fp64": {k: st_d[k] for k in ("precompute_ms", "pair_ms", "fill_ms", "reduce_ms")}},
        "rays": st_m.get("rays"), "fp64_retests": st_m.get("retests"),
        "traversal": {"mixed_kernel": plan_m.kernel_name(), "fp64_kernel": plan_d.kernel_name(),
                      "fp64_box_tests": st_d.get("aabb_tests"), "fp64_triangle_tests": st_d.get("tri_tests"),
                      "mixed_slot_steps": st_m.get("slot_steps"), "mixed_task_batches": st_m.get("task_batches"),
                      "mixed_sure_hits": st_m.get("sure_hits"), "mixed_exact_batches": st_m.get("exact_batches"),
                      "mixed_culled_rays": st_m.get("culled_rays"), "fp64_culled_rays": st_d.get("culled_rays")},
        "verification": ver, "roofline": roof, "cpu_baseline": cpu,
    }
    if args.autotune:
        out["autotune"] = {"mixed": tune_m, "fp64": tune_d}
    if cpu:
        out["gpu_vs_cpu_unique_pairs"] = (N * (N - 1) / 2) / (ms_m * 1e-3) / cpu["value"]
    plan_m.close(); plan_d.close()
    if args.secondary and args.workload != "plates_occluded":
        out["occluded_plates"] = secondary_plates(api, ctx, args.steps)
    if args.secondary:
        out["curved_cylinder"] = secondary_cylinder(api, ctx)
    if args.secondary and args.workload != "cube":
        out["unoccluded_cube"] = secondary_cube(api, ctx, args.steps)
    if args.secondary:
        out["self_vf_sphere"] = secondary_self_vf(api, ctx, args.steps)
        out["curves"] = secondary_curves(api, ctx, args.steps, 0.0 if args.no_cpu else 5.0)
    if args.serving and args.serving_requests > 0:
        out["serving"] = serving_stream(api, ctx, args.serving_requests, args.serving_rate)
    print(json.dumps(out))


def serving_stream(api, ctx, nreq, rate, seed=0x5EED0005):
    """The serving stream: seeded Poisson arrivals of random closed box
    enclosures (sides U[0.5, 2], 500-20,000 elements log-uniform, per-face
    Quad4/Tri3/Quad8, 0-3 box occluders of <= 8k Tri3 each, nquad U{2..5}),
    submitted to the request server at their arrival times, after one untimed
    warm-up request.  Per precision:
    throughput (ordered element pairs of all requests / makespan) and latency
    percentiles (submission to result on the server's clock); mixed against
    FP64 per request: max |row-sum difference| and max |F_group difference|."""
    import threading
    sys.path.insert(0, ROOT)
    from rvf_amd import mesh as M
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(nreq):
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(20000)))))
        mesh, obs, nq = M.random_box_enclosure(rng, n, max_occluders=3, max_occluder_tris=8000)
        reqs.append((mesh, obs, int(nq)))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    res = {"requests": nreq, "arrival_rate_per_s": rate, "seed": seed,
           "elements": [int(m.n_elems) for m, _, _ in reqs], "nquad": [nq for _, _, nq in reqs],
           "occluders": [len(o) for _, o, _ in reqs]}
    wmesh, wobs, wnq = M.random_box_enclosure(np.random.default_rng(seed + 1), 600, max_occluders=1,
                                              max_occluder_tris=500)
    outs = {}
    # the Poisson stream in each precision, the same requests with every tile
    # batched, and the same requests submitted at once ("burst": the makespan
    # is then the device's, not the arrivals')
    for prec, direct, burst in (("mixed", 0, False), ("fp64", 0, False), ("mixed_batched", -1, False),
                                ("mixed_burst", 0, True), ("fp64_burst", 0, True)):
        srv = api.Server(ctx, direct_tiles=direct)
        prec_name, prec = prec, prec.split("_")[0]
        # warm-up request (untimed): first use of the server's streams and of
        # the batched and direct kernels, allocator growth
        w = srv.submit(wmesh, wnq, wobs, "geometric", prec)
        if srv.wait(w, 600000) != api.REQ_DONE:
            raise RuntimeError("serving warm-up request failed")
        srv.release(w)
        c0 = srv.stats()
        tickets = [None] * nreq
        t0 = time.perf_counter()

        def submitter():
            for k, (mesh, obs, nq) in enumerate(reqs):
                dt = (0.0 if burst else arrivals[k]) - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                tickets[k] = srv.submit(mesh, nq, obs, "geometric", prec)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for t in tickets:
            if srv.wait(t, 600000) != api.REQ_DONE:
                raise RuntimeError("serving request failed")
        got = [srv.result(t) for t in tickets]
        st = [g[6] for g in got]
        lat = np.array([s["t_done_ms"] - s["t_submit_ms"] for s in st])
        span_s = (max(s["t_done_ms"] for s in st) - min(s["t_submit_ms"] for s in st)) * 1e-3
        pairs = float(sum(float(m.n_elems) ** 2 for m, _, _ in reqs))
        c = srv.stats()
        for t in tickets:
            srv.release(t)
        srv.close()
        outs[prec_name] = got
        tiles = c["tiles"] - c0["tiles"]
        res[prec_name] = {"pairs_per_s": pairs / span_s, "makespan_s": span_s,
                          "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                         "max": float(lat.max())},
                          "rounds": c["rounds"] - c0["rounds"], "peak_device_bytes": c["peak_bytes"],
                          "tiles": tiles, "batched_tiles_share": (c["batched_tiles"] - c0["batched_tiles"]) / max(1, tiles),
                          "direct_tiles_option": "default (512)" if direct == 0 else "never (every tile batched)",
                          "arrivals": "all at once" if burst else f"Poisson {rate}/s",
                          "worker_ms": {k: c[k] - c0[k] for k in ("launch_ms", "admit_ms", "wait_ms", "retire_ms")}}
    res["speedup_mixed_vs_fp64"] = res["fp64"]["makespan_s"] / res["mixed"]["makespan_s"]
    res["speedup_burst_mixed_vs_fp64"] = res["fp64_burst"]["makespan_s"] / res["mixed_burst"]["makespan_s"]
    res["closure_err"] = float(max(np.abs(a[2] - b[2]).max() for a, b in zip(outs["mixed"], outs["fp64"])))
    res["max_abs_dF_group"] = float(max(np.abs(a[3] - b[3]).max() for a, b in zip(outs["mixed"], outs["fp64"])))
    # per request, untimed: standalone solves of both precisions with audits;
    # element-level max |dF| and flipped decisions of mixed against FP64, and
    # the server's results (row sums, F_group) against the standalone ones
    per = []
    identical = 0
    for k, (mesh, obs, nq) in enumerate(reqs):
        dm = api.DeviceMesh(ctx, mesh)
        do = api.DeviceObstruction(ctx, dm, obs, "geometric") if obs else None
        alone = {}
        for prec in ("mixed", "fp64"):
            pl = api.Plan(ctx, dm, do, nq, "quad", prec, want_F=True, audit=do is not None, timing=False)
            pl.run(); pl.sync()
            alone[prec] = pl.fetch()
            pl.close()
        Fm, Fd = alone["mixed"][0], alone["fp64"][0]
        flips = 0
        if do is not None:
            iu = np.triu_indices(mesh.n_elems, 1)
            ca, cb = alone["fp64"][6][iu].astype(np.int64), alone["mixed"][6][iu].astype(np.int64)
            ha, hb = alone["fp64"][7][iu], alone["mixed"][7][iu]
            flips = int(np.abs(ca - cb).sum() + 2 * np.sum((ca == cb) & (ha != hb)))
            do.close()
        dm.close()
        same = all(np.array_equal(outs[p][k][2], alone[p.split("_")[0]][2]) and
                   np.array_equal(outs[p][k][3], alone[p.split("_")[0]][3]) for p in outs)
        identical += same
        per.append({"elements": int(mesh.n_elems), "max_abs_dF": float(np.abs(Fm - Fd).max()),
                    "closure_err": float(np.abs(alone["mixed"][2] - alone["fp64"][2]).max()), "flips": flips})
        del alone, Fm, Fd
    res["per_request"] = per
    res["max_abs_dF"] = max(r["max_abs_dF"] for r in per)
    res["flipped_visibility"] = sum(r["flips"] for r in per)
    res["server_equals_standalone"] = f"{identical} of {len(reqs)} requests (row sums and F_group bit-identical)"
    return res


def audit_flips(api, ctx, dm, do, nquad, N):
    """Blocked-ray counts and order-free hashes of the default FP64 and mixed
    traversals: flipped visibility decisions of mixed against FP64."""
    aud = {}
    for prec in ("fp64", "mixed"):
        pl = api.Plan(ctx, dm, do, nquad, "quad", prec, want_F=False, audit=True, timing=False)
        pl.run(); pl.sync()
        *_, c, h = pl.fetch(F=False)
        aud[prec] = (c, h, pl.kernel_name())
        pl.close()
    iu = np.triu_indices(N, 1)
    ca, ha = aud["fp64"][0][iu].astype(np.int64), aud["fp64"][1][iu]
    cb, hb = aud["mixed"][0][iu].astype(np.int64), aud["mixed"][1][iu]
    flips = int(np.abs(ca - cb).sum() + 2 * np.sum((ca == cb) & (ha != hb)))
    return flips, int(ca.sum()), {"fp64": aud["fp64"][2], "mixed": aud["mixed"][2]}


def secondary_plates(api, ctx, steps):
    """Occluded parallel plates: 2 x 96 x 96 Quad4 (18,432 elements) with the
    20,480-triangle perturbed sphere between them, nquad 3, geometric soup
    reading, default traversal configurations.  Mixed and FP64 solve times,
    mixed against FP64 on the whole matrix (max |dF|, closure, flipped
    decisions from audits) and the mixed pair kernel's roofline."""
    mesh, obs, nquad, parity, desc = build_workload("plates_occluded")
    N = mesh.n_elems
    dm = api.DeviceMesh(ctx, mesh)
    do = api.DeviceObstruction(ctx, dm, obs, parity)
    res = {"workload": desc}
    Fs, st = {}, {}
    for prec, k in (("mixed", max(3, steps // 4)), ("fp64", 2)):
        plan, ms, lat, pair = run_mode(api, ctx, dm, do, nquad, prec, k, 1)
        st[prec] = plan.stats()
        F, A, rs, *_ = plan.fetch()
        Fs[prec] = (F, rs)
        res[prec] = {"ms_per_step": ms, "pairs_per_s": float(N) * N / (ms * 1e-3), "pair_kernel_ms": float(pair.mean()),
                     "kernel": plan.kernel_name(), "steps": k}
        plan.close()
    res["speedup_mixed_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    res["max_abs_dF"] = float(np.abs(Fs["mixed"][0] - Fs["fp64"][0]).max())
    res["closure_err"] = float(np.abs(Fs["mixed"][1] - Fs["fp64"][1]).max())
    del Fs
    res["flipped_visibility"], res["blocked_rays_fp64"], res["audited_kernels"] = audit_flips(api, ctx, dm, do, nquad, N)
    d = st["fp64"]
    flops = flop_model(N, nquad * nquad, d.get("rays", 0), d.get("aabb_tests", 0), d.get("tri_tests", 0))
    ach = flops / (res["mixed"]["pair_kernel_ms"] * 1e-3)
    res["roofline_mixed"] = {"bound": "valu", "achieved": ach / 1e12, "peak": VALU_PEAK_FP32 / 1e12,
                             "unit": "TFLOP/s", "frac": ach / VALU_PEAK_FP32, "flops_per_launch": flops,
                             "rays": d.get("rays"), "box_tests": d.get("aabb_tests"), "triangle_tests": d.get("tri_tests")}
    do.close(); dm.close()
    return res


# Flops per evaluation of a curved element (point, normal, area element) and
# per Duffy map, for the Duffy and Monte Carlo rooflines.
EVAL_FLOPS = 90
MAP_FLOPS = 20


def secondary_cylinder(api, ctx, steps=3, mc_steps=2, seeds=4):
    """Second-order curved geometry: the interior of the closed cylinder,
    128 x 64 Quad8 side and 2 x 3,000 Tri6 Delaunay caps (14,192 elements).
    Duffy (Sauter-Schwab) on the Quad8 pairs sharing a corner with nquad 5
    elsewhere: solve time, the k_duffy and pair kernels' durations and
    rooflines; stratified Monte Carlo at 256 samples per pair: solve time and
    roofline, and F_group over `seeds` seeds against the deterministic solve
    (the reference's stream pairs cell k of one element with cell k of the
    other, which biases each pair's estimate)."""
    from rvf_amd import mesh as M
    mesh = M.cylinder()
    N = mesh.n_elems
    dm = api.DeviceMesh(ctx, mesh)
    res = {"workload": f"cylinder interior r=0.5 h=2: 128x64 Quad8 side + 2x3,000 Tri6 caps ({N:,} elements)"}
    # Duffy, nquad 5
    plan, ms, lat, pair = run_mode(api, ctx, dm, None, 5, "fp64", steps, 1, method="duffy")
    st = plan.stats()
    Fd, Ad, rsd, Fgd, Agd, tags, *_ = plan.fetch()
    nquad_pts = {0: 25, 1: 13}
    fam = mesh.family
    nq, nt = int((fam == 0).sum()), int((fam == 1).sum())
    quad_flops = (nq * (nq - 1) // 2 * (25 * 25 * 28 + 25 * 3) + nq * nt * (25 * 13 * 28 + 25 * 3)
                  + nt * (nt - 1) // 2 * (13 * 13 * 28 + 13 * 3))
    # singular pairs by shared corners: 1 corner -> vertex rule (8 regions), 2 -> edge rule (5)
    by = {}
    for e in np.nonzero(fam == 0)[0]:
        for c in mesh.nodes[e, :4]:
            by.setdefault(int(c), set()).add(int(e))
    shared = {}
    for es in by.values():
        for x in es:
            for y in es:
                if x < y:
                    shared[x, y] = shared.get((x, y), 0) + 1
    nv = sum(1 for v in shared.values() if v == 1)
    ne = sum(1 for v in shared.values() if v >= 2)
    duffy_evals = (8 * nv + 5 * ne) * 5 ** 4
    duffy_flops = duffy_evals * (28 + 2 * EVAL_FLOPS + MAP_FLOPS)
    res["duffy_nquad5"] = {
        "ms_per_step": ms, "pair_kernel_ms": float(pair.mean()), "duffy_kernel_ms": st["duffy_ms"],
        "singular_pairs": st["singular_pairs"], "vertex_pairs": nv, "edge_pairs": ne, "kernel": plan.kernel_name(),
        "pairs_per_s": float(N) * N / (ms * 1e-3),
        "roofline_pair": {"achieved": quad_flops / (float(pair.mean()) * 1e-3) / 1e12, "peak": VALU_PEAK_FP64 / 1e12,
                          "unit": "TFLOP/s", "flops_per_launch": quad_flops},
        "roofline_duffy": {"achieved": duffy_flops / (st["duffy_ms"] * 1e-3) / 1e12, "peak": VALU_PEAK_FP64 / 1e12,
                           "unit": "TFLOP/s", "flops_per_launch": duffy_flops,
                           "model": f"(8 vertex + 5 edge regions) x nquad^4 evaluations x (28 + 2 x {EVAL_FLOPS} + {MAP_FLOPS})"},
        "F_group": Fgd.tolist(), "group_tags": tags}
    for k in ("roofline_pair", "roofline_duffy"):
        r = res["duffy_nquad5"][k]
        r["frac"] = r["achieved"] / r["peak"]
    plan.close()
    del Fd
    # Monte Carlo, 256 samples per pair
    mc_flops = N * (N - 1) // 2 * 256 * (2 * EVAL_FLOPS + 28)
    fg = []
    for k in range(seeds):
        plan, ms, lat, pair = run_mode(api, ctx, dm, None, 3, "fp64", mc_steps if k == 0 else 1, 0 if k else 1,
                                       method="mc", n_samples=256, seed=0x5EED0004 + k)
        if k == 0:
            res["monte_carlo_256"] = {"ms_per_step": ms, "pair_kernel_ms": float(pair.mean()), "kernel": plan.kernel_name(),
                                      "pairs_per_s": float(N) * N / (ms * 1e-3),
                                      "roofline": {"achieved": mc_flops / (float(pair.mean()) * 1e-3) / 1e12,
                                                   "peak": VALU_PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                                                   "flops_per_launch": mc_flops,
                                                   "model": f"256 x (2 x {EVAL_FLOPS} + 28) per unique pair"}}
            r = res["monte_carlo_256"]["roofline"]
            r["frac"] = r["achieved"] / r["peak"]
        fg.append(plan.fetch(F=False)[3])
        plan.close()
    fg = np.array(fg)
    res["monte_carlo_256"]["F_group_mean"] = fg.mean(0).tolist()
    res["monte_carlo_256"]["F_group_sigma_of_mean"] = (fg.std(0, ddof=1) / np.sqrt(seeds)).tolist()
    res["monte_carlo_256"]["F_group_minus_deterministic"] = (fg.mean(0) - Fgd).tolist()
    dm.close()
    return res


def secondary_cube(api, ctx, steps):
    """The unoccluded 13,824-element cube (nquad 3): mixed and FP64 solve
    times, the pair kernels' durations and mixed-vs-FP64 differences."""
    mesh, obs, nquad, parity, desc = build_workload("cube")
    N = mesh.n_elems
    dm = api.DeviceMesh(ctx, mesh)
    res = {"workload": desc}
    Fs = {}
    for prec in ("mixed", "fp64"):
        plan, ms, lat, pair = run_mode(api, ctx, dm, None, nquad, prec, steps, 1)
        st = plan.stats()
        F, A, rs, Fg, *_ = plan.fetch()
        Fs[prec] = (F, rs)
        res[prec] = {"ms_per_step": ms, "pairs_per_s": float(N) * N / (ms * 1e-3),
                     "pair_kernel_ms": float(pair.mean()), "fill_ms": st["fill_ms"], "reduce_ms": st["reduce_ms"],
                     "precompute_ms": st["precompute_ms"],
                     "pair_kernel_tflops": flop_model(N, nquad * nquad) / (float(pair.mean()) * 1e-3) / 1e12}
        plan.close()
    res["speedup_mixed_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    res["max_abs_dF"] = float(np.abs(Fs["mixed"][0] - Fs["fp64"][0]).max())
    res["closure_err"] = float(np.abs(Fs["mixed"][1] - Fs["fp64"][1]).max())
    res["rowsum_minus_1_max_fp64"] = float(np.abs(Fs["fp64"][1] - 1.0).max())
    dm.close()
    return res


def secondary_self_vf(api, ctx, steps, n=48):
    """Self view factors on a concave surface: the interior of a unit sphere
    as 6 n^2 curved Quad8 elements (13,824 at n = 48), nquad 3, unoccluded.
    Mixed and FP64 solve times with and without the pairs (i, i), and the
    closure the diagonal restores (inside a sphere every F_ij, the diagonal
    included, is A_j / 4 pi up to the quadrature's missing p = q terms)."""
    from rvf_amd import mesh as M
    mesh = M.sphere_interior(n, 1.0, 2)
    N = mesh.n_elems
    dm = api.DeviceMesh(ctx, mesh)
    res = {"workload": f"sphere interior, cubed sphere 6x{n}x{n} Quad8 ({N:,} elements), nquad=3"}
    out = {}
    for prec in ("mixed", "fp64"):
        for self_vf in (False, True):
            plan, ms, lat, pair = run_mode(api, ctx, dm, None, 3, prec, steps, 1, self_vf=self_vf)
            F, A, rs, *_ = plan.fetch()
            out[prec, self_vf] = (F if self_vf else None, rs)
            res[f"{prec}{'_self_vf' if self_vf else ''}"] = {
                "ms_per_step": ms, "pair_kernel_ms": float(pair.mean()),
                "pairs_per_s": (N * (N + 1) / 2 if self_vf else N * (N - 1) / 2) / (ms * 1e-3)}
            plan.close()
    for prec in ("mixed", "fp64"):
        res[f"{prec}_self_vf_overhead"] = res[f"{prec}_self_vf"]["ms_per_step"] / res[prec]["ms_per_step"] - 1
    Fm, rsm = out["mixed", True]
    Fd, rsd = out["fp64", True]
    res["max_abs_dF_mixed_vs_fp64"] = float(np.abs(Fm - Fd).max())
    res["max_abs_dFii_mixed_vs_fp64"] = float(np.abs(np.diag(Fm) - np.diag(Fd)).max())
    res["closure_err_mixed_vs_fp64"] = float(np.abs(rsm - rsd).max())
    res["rowsum_minus_1_max_fp64"] = {"self_vf": float(np.abs(rsd - 1).max()),
                                      "without": float(np.abs(out["fp64", False][1] - 1).max())}
    dm.close()
    return res


def secondary_curves(api, ctx, steps, cpu_seconds):
    """Curve meshes (per unit depth; CPU only in the reference): the interior
    of a circle as 8,192 Line3 arcs with self view factors, and the unit
    square's interior as 4 x 2,048 Line2 elements around an 8-segment baffle
    (obstruction-only), nquad 4; mixed and FP64 solve times, the mixed-vs-FP64
    differences, and the oracle's rate on a sample of the circle's rows."""
    from rvf_amd import mesh as M
    res = {}
    for name, mesh, obs, self_vf in (
            ("circle", M.circle_interior(8192, 1.0, 2), [], True),
            ("square_baffle", M.square_enclosure_2d(2048, 1, baffle=True), [5], False)):
        N = mesh.n_elems
        dm = api.DeviceMesh(ctx, mesh)
        do = api.DeviceObstruction(ctx, dm, obs, "geometric") if obs else None
        r = {"workload": f"{name}: {N:,} {'Line3' if mesh.family[0] == 4 else 'Line2'} elements, nquad=4"
                         + (", self_vf" if self_vf else "") + (", 8 occluding segments" if obs else "")}
        Fs = {}
        for prec in ("mixed", "fp64"):
            plan, ms, lat, pair = run_mode(api, ctx, dm, do, 4, prec, steps, 1, self_vf=self_vf)
            st = plan.stats()
            F, A, rs, *_ = plan.fetch()
            Fs[prec] = (F, rs)
            pairs = N * (N + 1) / 2 if self_vf else N * (N - 1) / 2
            r[prec] = {"ms_per_step": ms, "pair_kernel_ms": float(pair.mean()), "unique_pairs_per_s": pairs / (ms * 1e-3),
                       "kernel": plan.kernel_name(), "rays": st.get("rays")}
            plan.close()
        r["max_abs_dF"] = float(np.abs(Fs["mixed"][0] - Fs["fp64"][0]).max())
        r["closure_err"] = float(np.abs(Fs["mixed"][1] - Fs["fp64"][1]).max())
        r["rowsum_minus_1_max_fp64"] = float(np.abs(Fs["fp64"][1] - 1.0).max())
        if do is not None:
            do.close()
        dm.close()
        if cpu_seconds > 0 and name == "circle":
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as orc
            rng = np.random.default_rng(2)
            t0, pairs, rows = time.perf_counter(), 0, 0
            while time.perf_counter() - t0 < cpu_seconds and rows < N:
                r0 = int(rng.integers(0, N - 64))
                orc.compute(mesh, nquad=4, rows=(r0, r0 + 64), want_raw=False, nthreads=os.cpu_count() or 1,
                            self_vf=True)
                pairs += sum(N - i for i in range(r0, r0 + 64))
                rows += 64
            cpu = pairs / (time.perf_counter() - t0)
            r["cpu_oracle_unique_pairs_per_s"] = cpu
            r["gpu_fp64_vs_cpu"] = r["fp64"]["unique_pairs_per_s"] / cpu
        res[name] = r
    return res


if __name__ == "__main__":
    main()
