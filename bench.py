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


def run_mode(api, ctx, dm, do, nquad, precision, steps, warmup, audit=False, tune=None, self_vf=False):
    plan = api.Plan(ctx, dm, do, nquad, "quad", precision, want_F=True, audit=audit, timing=True, self_vf=self_vf)
    if tune is not None and do is not None:
        # autotuner (setup, outside the timed region): time each traversal
        # configuration over a tile sample and keep the fastest
        t0 = time.perf_counter()
        v, r, meas = plan.autotune(sample_tiles=tune["sample_tiles"])
        tune.update(variant=v, refill_min=r, kernel=plan.kernel_name(), seconds=time.perf_counter() - t0,
                    measurements=meas)
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
    return plan, total_ms / steps, np.array(lat), np.array(pair_ms)


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; sharding pair tiles over several GPUs is designed but not built",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as G
    G.load_package()
    from rvf_amd import api
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
    plan_m, ms_m, lat_m, pair_m = run_mode(api, ctx, dm, do, nquad, "mixed", args.steps, args.warmup, tune=tune_m)
    st_m = plan_m.stats()
    plan_d, ms_d, lat_d, pair_d = run_mode(api, ctx, dm, do, nquad, "fp64", max(2, args.steps // 2),
                                           min(args.warmup, 1), tune=tune_d)
    st_d = plan_d.stats()

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
            pa = api.Plan(ctx, dm, do, nquad, "quad", "fp64", want_F=False, audit=True, timing=False)
            pa.run(); pa.sync()
            *_, ca, ha = pa.fetch(F=False)
            pa.close()
            pb = api.Plan(ctx, dm, do, nquad, "quad", "mixed", want_F=False, audit=True, timing=False)
            pb.run(); pb.sync()
            *_, cb, hb = pb.fetch(F=False)
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
                       "fp64_p50": float(np.percentile(lat_d, 50))},
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
    for prec in ("mixed", "fp64"):
        srv = api.Server(ctx)
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
                dt = arrivals[k] - (time.perf_counter() - t0)
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
        srv.close()
        outs[prec] = got
        res[prec] = {"pairs_per_s": pairs / span_s, "makespan_s": span_s,
                     "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                    "max": float(lat.max())},
                     "rounds": c["rounds"] - c0["rounds"], "peak_device_bytes": c["peak_bytes"],
                     "worker_ms": {k: c[k] - c0[k] for k in ("launch_ms", "admit_ms", "wait_ms", "retire_ms")}}
    res["speedup_mixed_vs_fp64"] = res["fp64"]["makespan_s"] / res["mixed"]["makespan_s"]
    res["closure_err"] = float(max(np.abs(a[2] - b[2]).max() for a, b in zip(outs["mixed"], outs["fp64"])))
    res["max_abs_dF_group"] = float(max(np.abs(a[3] - b[3]).max() for a, b in zip(outs["mixed"], outs["fp64"])))
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
