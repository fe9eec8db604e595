"""Benchmark of the octree-FEM path on one MI355X.

Headline (value): octree build + 2:1 corner balance, octants/s, on the
lognormal-4M workload: 4,194,304 points with coordinates exp(N(-1,0.5))
clipped to [0,1), seed 1, maxDepth 30, maxNumPts 8.  One step =
points2Octree + balanceOctree over the point cloud already resident in HBM.

Also reported on the same line:
  roofline      the FP64 -lap+I matvec on that octree's DA (100 launches
                back to back, one hipEvent pair on the library stream around
                them), algorithmic bytes 34 E + 16 Nn per launch against 8 TB/s;
  solve         Poisson -lap u + u = f, homogeneous Dirichlet, MG(5)-V(2,2)
                Jacobi(0.8)-preconditioned CG to 1e-8, all-FP64 against the
                FP32 V-cycle inside FP64 CG: solves/s, p50/p99 latency per
                solve, error of the mixed solve against the FP64 one; on the
                lognormal-4M octree and on gauss4-16M (16,777,216 points,
                4 Gaussians sigma 0.02, seed 2, maxNumPts 4);
  serve         a served stream of Poisson problems: the first --serve-requests
                (default 400) of a 2,000-request stream (point counts
                log-uniform in [2^12, 2^22], three distributions, maxNumPts
                1/4/8, Poisson arrivals at 200/s, 200 GB HBM budget, problems
                kept resident): solves/s and p50/p99 latency, FP64 and FP32-MG;
  cpu_baseline  the oracle (CPU restatement, one thread) timed on the headline
                workload (all its points by default, about 15 s; --cpu-sample N
                takes the first N).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "matvec_f64_lognormal-4M_traffic.txt")


class HipEvents:
    """hipEvent timing on the library's stream (libamdhip64)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def read_traffic():
    """HBM bytes per matvec launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel of known size), if that summary exists."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solve_section(dgx, ob, nlevels, solves, warmup, seed, graph=True):
    """FP64 and FP32-V-cycle solves of the Dirichlet problem on octree ob."""
    # the hierarchy is built twice: the first build in a process also pays
    # the one-time start-up of rocSOLVER (coarse Cholesky)
    setups = []
    for _ in range(2):
        t0 = time.perf_counter()
        mg = dgx.MG(ob, nlevels)
        dgx.synchronize()
        setups.append(time.perf_counter() - t0)
    t_setup = setups[-1]
    n = mg.n_nodes
    f = dgx.DeviceArray.from_host(np.random.default_rng(seed).standard_normal(n))
    b = dgx.DeviceArray((n,), np.float64)
    mg.rhs(f, b)
    out = {"levels": mg.nlevels, "level_octants": mg.level_octants, "level_nodes": mg.level_nodes,
           "mg_setup_s": t_setup, "mg_first_setup_s": setups[0]}
    sol = {}
    for name, fp32 in (("fp64", False), ("fp32_mg", True)):
        u = dgx.DeviceArray((n,), np.float64)
        for _ in range(warmup):
            mg.solve(b, u, fp32=fp32, graph=graph)
        lat = []
        st = None
        for _ in range(solves):
            t0 = time.perf_counter()
            st = mg.solve(b, u, fp32=fp32, graph=graph)
            dgx.synchronize()
            lat.append(time.perf_counter() - t0)
        sol[name] = u.to_host()
        lat = np.array(lat)
        out[name] = {"solves_per_s": float(len(lat) / lat.sum()), "p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3), "iterations": st["iterations"],
                     "rel_residual": st["rel_residual"]}
    out["speedup_fp32_mg_vs_fp64"] = out["fp32_mg"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["rel_err_fp32_mg_vs_fp64"] = float(np.linalg.norm(sol["fp32_mg"] - sol["fp64"]) /
                                           np.linalg.norm(sol["fp64"]))
    del mg
    return out


def serve_section(dgx, count, slots, budget):
    """Served stream: `count` requests (log-uniform 2^12..2^22 points, three
    distributions, maxNumPts 1/4/8, seed 3) submitted at their Poisson arrival
    times (200/s); latency = completion - arrival, on the server clock."""
    reqs = dgx.serve_workload(2000)[:count]
    t0 = time.perf_counter()
    clouds = [dgx.serve_points(r) for r in reqs]
    t_gen = time.perf_counter() - t0
    out = {"requests": count, "rate_per_s": 200.0, "slots": slots, "hbm_budget_GB": budget / 1e9,
           "points_total": int(sum(r["n"] for r in reqs)), "generate_s": t_gen}
    for name, fp32 in (("fp64", False), ("fp32_mg", True)):
        sv = dgx.Server(slots=slots, hbm_budget=budget, nlevels=5, fp32=fp32, keep_resident=True)
        start = time.perf_counter()
        ids = []
        for k, (r, c) in enumerate(zip(reqs, clouds)):
            dt = r["arrival"] - (time.perf_counter() - start)
            if dt > 0:
                time.sleep(dt)
            ids.append(sv.submit(c, 30, r["max_num_pts"], r["seed"]))
            if (k + 1) % 500 == 0:
                log(f"serve {name}: {k + 1}/{count} submitted")
        res = []
        for k, i in enumerate(ids):
            res.append(sv.wait(i))
            if (k + 1) % 100 == 0:
                log(f"serve {name}: {k + 1}/{count} done at {time.perf_counter() - start:.1f} s")
        elapsed = time.perf_counter() - start
        sv.close()
        lat = np.array([x["t_done"] - x["t_submit"] for x in res])
        run = np.array([x["t_done"] - x["t_start"] for x in res])
        out[name] = {"solves_per_s": len(res) / elapsed, "elapsed_s": elapsed,
                     "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "service_p50_ms": float(np.percentile(run, 50) * 1e3),
                     "service_p99_ms": float(np.percentile(run, 99) * 1e3),
                     "max_rel_residual": float(max(x["rel_residual"] for x in res)),
                     "iterations_max": int(max(x["iterations"] for x in res)),
                     "octants_total": int(sum(x["octants"] for x in res))}
        log(f"serve {name}: {json.dumps(out[name])}")
    out["speedup_fp32_mg_vs_fp64"] = out["fp32_mg"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    del clouds
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--matvecs", type=int, default=100)
    ap.add_argument("--solves", type=int, default=20)
    ap.add_argument("--no-solve", action="store_true", help="skip the Poisson solves")
    ap.add_argument("--no-large", action="store_true", help="skip the gauss4-16M solve")
    ap.add_argument("--no-graph", action="store_true", help="solve without hipGraph replay (for tracing runs)")
    ap.add_argument("--serve-requests", type=int, default=400,
                    help="requests of the served stream, a prefix of the 2,000-request stream (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=0, help="points in the oracle's CPU sample (0: all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--matvec-only", action="store_true",
                    help="build once, run the matvecs and a 256 MiB device copy (for counter runs), print nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import dendro4_amd as dgx
    dgx.init(0)
    stream = dgx.stream()
    ev = HipEvents()

    W = "lognormal-4M"
    MD, MP = 30, dgx.workload_max_num_pts(W)
    pts = dgx.workload_points(W)
    N = len(pts)
    dpts = dgx.DeviceArray.from_host(pts)
    dgx.synchronize()

    def step():
        ou = dgx.points2octree(dpts, MD, MP)
        ob = dgx.balance_octree(ou)
        return ou, ob

    for _ in range(args.warmup):
        step()
    dgx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ou = dgx.points2octree(dpts, MD, MP)
    dgx.synchronize()
    t_p2o = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ou, ob = step()
    dgx.synchronize()
    t_build = (time.perf_counter() - t0) / args.steps
    U, Bn = len(ou), len(ob)
    octants_per_s = Bn / t_build
    build_bytes = 24 * N + 2 * 28 * U + 28 * Bn
    log(f"build+balance {t_build * 1e3:.2f} ms/step ({Bn} octants)")

    # DA + FP64 matvec
    da = dgx.DA(ob)
    E, Nn = da.element_size, da.n_nodes
    x = dgx.DeviceArray.from_host(np.random.default_rng(1).standard_normal(Nn))
    y = dgx.DeviceArray((Nn,), np.float64)
    for _ in range(5):
        da.matvec(x, y)
    dgx.synchronize()
    evs = [(ev.event(), ev.event()) for _ in range(args.matvecs)]
    for a, b in evs:
        ev.record(a, stream)
        da.matvec(x, y)
        ev.record(b, stream)
    dgx.synchronize()
    mv_ms = [ev.elapsed_ms(a, b) for a, b in evs]
    # back-to-back launches: one event pair around the whole run
    s0, s1 = ev.event(), ev.event()
    ev.record(s0, stream)
    for _ in range(args.matvecs):
        da.matvec(x, y)
    ev.record(s1, stream)
    dgx.synchronize()
    # average launch duration = GPU time per matvec when matvecs run back to
    # back (the host enqueues ahead); the per-launch event pairs above also
    # hold the event overhead and any host gap before each launch
    mv_avg_ms = ev.elapsed_ms(s0, s1) / args.matvecs
    mv_bracket_ms = float(np.mean(mv_ms))
    log(f"matvec per-launch events {mv_bracket_ms * 1e3:.1f} us, back-to-back {mv_avg_ms * 1e3:.1f} us")
    mv_bytes = 34 * E + 16 * Nn
    achieved = mv_bytes / (mv_avg_ms * 1e-3) / 1e9
    log(f"matvec {mv_avg_ms * 1e3:.1f} us, {achieved:.0f} GB/s")
    block_info = da.block_info
    del da, x, y
    if args.matvec_only:
        # reference copy kernel of known size for scaling the FETCH_SIZE / WRITE_SIZE counters
        nb = 256 << 20
        src = dgx.DeviceArray((nb // 8,), np.float64)
        dst = dgx.DeviceArray((nb // 8,), np.float64)
        hip = ev.hip
        hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        for _ in range(3):
            assert hip.hipMemcpyAsync(dst.ptr, src.ptr, nb, 3, stream) == 0  # hipMemcpyDeviceToDevice
        dgx.synchronize()
        print(json.dumps({"matvec_avg_ms": mv_avg_ms, "algorithmic_bytes": mv_bytes, "GB_per_s": achieved,
                          "copy_bytes": nb}))
        return

    solve = None
    if not args.no_solve:
        solve = {"lognormal-4M": solve_section(dgx, ob, 5, args.solves, 2, 2, not args.no_graph)}
        log(f"solve lognormal-4M: {json.dumps(solve['lognormal-4M'])}")
        if not args.no_large:
            del ou, ob, dpts
            W3 = "gauss4-16M"
            p3 = dgx.DeviceArray.from_host(dgx.workload_points(W3))
            dgx.synchronize()
            t0 = time.perf_counter()
            o3 = dgx.balance_octree(dgx.points2octree(p3, MD, dgx.workload_max_num_pts(W3)))
            dgx.synchronize()
            t3 = time.perf_counter() - t0
            del p3
            s3 = solve_section(dgx, o3, 5, max(5, args.solves // 2), 2, 2, not args.no_graph)
            s3["build_balance_s"] = t3
            s3["balanced_octants"] = len(o3)
            solve[W3] = s3
            log(f"solve {W3}: {json.dumps(s3)}")
            del o3

    serve = None
    if args.serve_requests > 0:
        serve = serve_section(dgx, args.serve_requests, args.serve_slots, 200e9)

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as O
        if not os.path.exists(O.LIB_PATH):
            import __graft_entry__
            __graft_entry__.build_oracle()
        sub = pts[: args.cpu_sample] if args.cpu_sample > 0 else pts
        t0 = time.perf_counter()
        ru = O.points2octree(sub, MD, MP)
        rb = O.balance(ru)
        tc = time.perf_counter() - t0
        which = "all" if len(sub) == len(pts) else "the first"
        cpu = {"value": len(rb) / tc, "unit": "octants/s", "cores": 1, "kind": "port",
               "sample": f"oracle points2Octree+balanceOctree on {which} {len(sub)} points of the lognormal-4M "
                         f"cloud ({len(rb)} balanced octants, {tc:.1f} s, 1 thread)"}

    out = {
        "metric": "octants/s (build+2:1 balance); Poisson solves/s & p99 latency, FP32-MG vs FP64",
        "value": octants_per_s,
        "unit": "octants/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_build * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "lognormal-4M: points2Octree + balanceOctree(corner) of 4,194,304 points, "
                               "coordinates exp(N(-1,0.5)) clipped to [0,1), seed 1, maxDepth 30, maxNumPts 8; "
                               "100 FP64 -lap+I matvecs on its DA",
                   "points": N, "octants": U, "balanced_octants": Bn, "elements": E, "nodes": Nn},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "kernel": "FP64 -lap+I matvec: k_matvec_pipe<double,false,4,0> + k_matvec_shared<double,false,0>",
                     "algorithmic_bytes": mv_bytes, "avg_launch_ms": mv_avg_ms},
        "matvec": {"da_blocks": block_info, "avg_ms": mv_avg_ms, "bracketed_avg_ms": mv_bracket_ms,
                   "bracketed_min_ms": float(np.min(mv_ms)),
                   "GB_per_s": achieved, "flops": 332 * E, "GFLOP_per_s": 332 * E / (mv_avg_ms * 1e-3) / 1e9},
        "build_balance": {"s_per_step": t_build, "p2o_s": t_p2o, "algorithmic_bytes": build_bytes,
                          "GB_per_s": build_bytes / t_build / 1e9},
        "solve": solve,
        "serve": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
