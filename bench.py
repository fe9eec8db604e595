"""Benchmark of the octree-FEM path on one MI355X.

Headline (value): octree build + 2:1 corner balance, octants/s, on the
lognormal-4M workload: 4,194,304 points with coordinates exp(N(-1,0.5))
clipped to [0,1), seed 1, maxDepth 30, maxNumPts 8.  One step =
points2Octree + balanceOctree over the point cloud already resident in HBM.

Also reported on the same line:
  build_balance  the headline step against the HBM roofline: algorithmic bytes
                 24 N + 2 * 28 U + 28 B per step;
  roofline       the FP64 -lap+I matvec (both passes) on the gauss4-16M DA,
                 whose 628 MB per matvec exceed the 256 MB Infinity Cache: 100
                 launches back to back between one hipEvent pair on the library
                 stream, algorithmic bytes 34 E + 16 Nn per launch against 8 TB/s;
                 `traffic` is the FETCH_SIZE + WRITE_SIZE measurement committed
                 under profiles/;
  matvec         the same on the lognormal-4M DA (cache-resident);
  solve          Poisson -lap u + u = f, homogeneous Dirichlet, MG(5)-V(2,2)
                 Jacobi(0.8)-preconditioned CG to 1e-8, all-FP64 against the
                 FP32 V-cycle inside FP64 CG: solves/s, p50/p99 latency per
                 solve, error of the mixed solve against the FP64 one; on the
                 lognormal-4M octree and on gauss4-16M (16,777,216 points,
                 4 Gaussians sigma 0.02, seed 2, maxNumPts 4);
  large          gauss16-128M (134,217,728 points, 16 Gaussians, seed 4,
                 maxNumPts 4): build + balance octants/s, DA, FP64 matvec GB/s,
                 7-level MG set-up and solves, peak device memory;
  serve          the 2,000-request served stream (point counts log-uniform in
                 [2^12, 2^22], three distributions, maxNumPts 1/4/8, Poisson
                 arrivals at 200/s, 200 GB HBM budget, problems kept resident):
                 solves/s and p50/p99 latency, FP64 and FP32-MG, in a process
                 of its own as a server runs (--serve-in-process: in this one,
                 after the other sections);
  cpu_baseline   the oracle (CPU restatement, OpenMP sorts as the reference's
                 omp_par::merge_sort) on all host cores, timed on the whole
                 headline workload.
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "matvec_f64_gauss4-16M_traffic.txt")
MATVEC_KERNEL = "k_matvec_pipe<double,false,4,0> + k_matvec_shared<double,false,0>"


class HipEvents:
    """hipEvent timing on the library's stream (libamdhip64)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

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


def time_matvecs(dgx, ev, stream, da, count):
    """Back-to-back FP64 matvecs between one event pair: GPU time per matvec."""
    x = dgx.DeviceArray.from_host(np.random.default_rng(1).standard_normal(da.n_nodes))
    y = dgx.DeviceArray((da.n_nodes,), np.float64)
    for _ in range(5):
        da.matvec(x, y)
    dgx.synchronize()
    s0, s1 = ev.event(), ev.event()
    ev.record(s0, stream)
    for _ in range(count):
        da.matvec(x, y)
    ev.record(s1, stream)
    dgx.synchronize()
    ms = ev.elapsed_ms(s0, s1) / count
    E, Nn = da.element_size, da.n_nodes
    nbytes = 34 * E + 16 * Nn
    return {"elements": E, "nodes": Nn, "avg_ms": ms, "algorithmic_bytes": nbytes,
            "GB_per_s": nbytes / (ms * 1e-3) / 1e9, "flops": 332 * E,
            "GFLOP_per_s": 332 * E / (ms * 1e-3) / 1e9, "da_blocks": da.block_info}


def solve_section(dgx, ob, nlevels, solves, warmup, seed, graph=True):
    """FP64 and FP32-V-cycle solves of the Dirichlet problem on octree ob."""
    setups = []
    mg = None
    for _ in range(2):  # the first set-up in a process also pays one-time runtime costs
        mg = None  # the previous hierarchy is freed before the next is built
        t0 = time.perf_counter()
        mg = dgx.MG(ob, nlevels)
        dgx.synchronize()
        setups.append(time.perf_counter() - t0)
    n = mg.n_nodes
    f = dgx.DeviceArray.from_host(np.random.default_rng(seed).standard_normal(n))
    b = dgx.DeviceArray((n,), np.float64)
    mg.rhs(f, b)
    out = {"levels": mg.nlevels, "level_octants": mg.level_octants, "level_nodes": mg.level_nodes,
           "coarse_solver": mg.coarse, "mg_setup_s": setups[-1], "mg_first_setup_s": setups[0]}
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


def build_section(dgx, dpts, md, mnp, steps, warmup):
    """Timed points2Octree + balanceOctree steps on device-resident points."""
    for _ in range(warmup):
        ob = dgx.balance_octree(dgx.points2octree(dpts, md, mnp))
        del ob
    dgx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ou = dgx.points2octree(dpts, md, mnp)
        ob = dgx.balance_octree(ou)
    dgx.synchronize()
    return ou, ob, (time.perf_counter() - t0) / steps


def large_section(dgx, ev, stream, args):
    """gauss16-128M: build + balance, DA, matvec, 7-level MG and solves, peak memory."""
    W = "gauss16-128M"
    dgx.memory_info(reset_peak=True)
    t0 = time.perf_counter()
    pts = dgx.workload_points(W)
    t_gen = time.perf_counter() - t0
    N = len(pts)
    dpts = dgx.DeviceArray.from_host(pts)
    del pts
    dgx.synchronize()
    ou, ob, t_build = build_section(dgx, dpts, 30, dgx.workload_max_num_pts(W), 2, 1)
    U, Bn = len(ou), len(ob)
    del ou, dpts
    mem_build = dgx.memory_info(reset_peak=True)
    log(f"large: build+balance {t_build * 1e3:.1f} ms ({Bn} octants)")
    t0 = time.perf_counter()
    da = dgx.DA(ob)
    dgx.synchronize()
    t_da = time.perf_counter() - t0
    mv = time_matvecs(dgx, ev, stream, da, 20)
    del da
    mem_da = dgx.memory_info(reset_peak=True)
    log(f"large: matvec {mv['avg_ms'] * 1e3:.0f} us, {mv['GB_per_s']:.0f} GB/s")
    s = solve_section(dgx, ob, 7, 3, 1, 4, not args.no_graph)
    mem = dgx.memory_info()
    out = {"workload": "gauss16-128M: 134,217,728 points, 16 Gaussians (sigma log-U[0.005,0.05], centres "
                       "U[0.1,0.9]^3, Dirichlet(1) weights), seed 4, maxDepth 30, maxNumPts 4",
           "points": N, "generate_s": t_gen, "octants": U, "balanced_octants": Bn,
           "build_balance_s": t_build, "octants_per_s": Bn / t_build, "da_build_s": t_da, "matvec": mv,
           "solve": s,
           "peak_device_bytes": {"build_balance": mem_build["library_peak"], "da_and_matvec": mem_da["library_peak"],
                                 "mg_setup_and_solves": mem["library_peak"]},
           "device_used_after_solves": mem["device_used"]}
    return out


def serve_section(dgx, count, slots, budget):
    """Served stream: `count` requests (log-uniform 2^12..2^22 points, three
    distributions, maxNumPts 1/4/8, seed 3) submitted at their Poisson arrival
    times (200/s); latency = completion - arrival, on the server clock."""
    reqs = dgx.serve_workload(2000)[:count]
    t0 = time.perf_counter()
    clouds = [dgx.serve_points(r) for r in reqs]
    t_gen = time.perf_counter() - t0
    log(f"serve: {count} point clouds generated in {t_gen:.1f} s")
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
            if (k + 1) % 250 == 0:
                log(f"serve {name}: {k + 1}/{count} done at {time.perf_counter() - start:.1f} s")
        elapsed = time.perf_counter() - start
        sv.close()
        lat = np.array([x["t_done"] - x["t_submit"] for x in res])
        run = np.array([x["t_done"] - x["t_start"] for x in res])
        out[name] = {"solves_per_s": len(res) / elapsed, "elapsed_s": elapsed,
                     "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "service_p50_ms": float(np.percentile(run, 50) * 1e3),
                     "service_p99_ms": float(np.percentile(run, 99) * 1e3),
                     "octree_s": float(sum(x["s_octree"] for x in res)),
                     "hierarchy_s": float(sum(x["s_hierarchy"] for x in res)),
                     "solve_s": float(sum(x["s_solve"] for x in res)),
                     "max_rel_residual": float(max(x["rel_residual"] for x in res)),
                     "iterations_max": int(max(x["iterations"] for x in res)),
                     "octants_total": int(sum(x["octants"] for x in res))}
        log(f"serve {name}: {json.dumps(out[name])}")
    out["speedup_fp32_mg_vs_fp64"] = out["fp32_mg"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    del clouds
    return out


def serve_in_child(dgx, count, slots):
    """The served stream in a process of its own, as a server runs: this
    process first returns its cached device memory, then waits for the child,
    whose progress lines go to stderr."""
    dgx.trim_cache()
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serve-only", "--serve-requests", str(count),
           "--serve-slots", str(slots)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
    lines = r.stdout.strip().splitlines()
    if r.returncode != 0 or not lines:
        log(f"serve: the served-stream process failed with status {r.returncode}")
        return {"error": f"served-stream process exited with status {r.returncode}", "process": "own process"}
    out = json.loads(lines[-1])
    out["process"] = "own process"
    return out


def cpu_section(pts, md, mnp):
    from oracle import pyoracle as O
    if not os.path.exists(O.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build_oracle()
    threads = O.num_threads()
    t0 = time.perf_counter()
    ru = O.points2octree(pts, md, mnp)
    rb = O.balance(ru)
    tc = time.perf_counter() - t0
    return {"value": len(rb) / tc, "unit": "octants/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(),
            "sample": f"oracle points2Octree+balanceOctree (OpenMP merge sorts) on all {len(pts)} points of the "
                      f"lognormal-4M cloud ({len(rb)} balanced octants, {tc:.2f} s, {threads} threads; "
                      f"host reports {os.cpu_count()} CPUs)"}


def matvec_only(dgx, ev, stream, W):
    """Counter runs: build workload W's DA, run 20 matvecs and three 256 MiB
    device copies (the copy scales FETCH_SIZE / WRITE_SIZE), print one line."""
    dpts = dgx.DeviceArray.from_host(dgx.workload_points(W))
    ob = dgx.balance_octree(dgx.points2octree(dpts, 30, dgx.workload_max_num_pts(W)))
    del dpts
    da = dgx.DA(ob)
    mv = time_matvecs(dgx, ev, stream, da, 20)
    nb = 256 << 20
    src = dgx.DeviceArray((nb // 8,), np.float64)
    dst = dgx.DeviceArray((nb // 8,), np.float64)
    hip = ev.hip
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    for _ in range(3):
        assert hip.hipMemcpyAsync(dst.ptr, src.ptr, nb, 3, stream) == 0  # hipMemcpyDeviceToDevice
    dgx.synchronize()
    print(json.dumps({"workload": W, "matvec_avg_ms": mv["avg_ms"], "algorithmic_bytes": mv["algorithmic_bytes"],
                      "GB_per_s": mv["GB_per_s"], "copy_bytes": nb, "elements": da.element_size,
                      "nodes": da.n_nodes, **da.block_info}))


def da_only(dgx, W):
    """Kernel traces of the set-up: workload W's octree, then 4 DA builds and
    2 MG(5) set-ups one after another, each timed to the end of its work; one
    line."""
    dpts = dgx.DeviceArray.from_host(dgx.workload_points(W))
    ob = dgx.balance_octree(dgx.points2octree(dpts, 30, dgx.workload_max_num_pts(W)))
    del dpts
    dgx.synchronize()
    t_da, t_mg = [], []
    for _ in range(4):
        t0 = time.perf_counter()
        da = dgx.DA(ob)
        dgx.synchronize()
        t_da.append(time.perf_counter() - t0)
        E = da.element_size
        del da
    for _ in range(2):
        t0 = time.perf_counter()
        mg = dgx.MG(ob, 5)
        dgx.synchronize()
        t_mg.append(time.perf_counter() - t0)
        levels = mg.level_octants
        del mg
    print(json.dumps({"workload": W, "elements": E, "da_build_ms": [t * 1e3 for t in t_da],
                      "mg_setup_ms": [t * 1e3 for t in t_mg], "level_octants": levels}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--matvecs", type=int, default=100)
    ap.add_argument("--solves", type=int, default=20)
    ap.add_argument("--no-solve", action="store_true", help="skip the Poisson solves")
    ap.add_argument("--no-large", action="store_true", help="skip gauss16-128M")
    ap.add_argument("--no-graph", action="store_true", help="solve without hipGraph replay (for tracing runs)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the served stream, a prefix of the 2,000-request stream (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-in-process", action="store_true",
                    help="serve in this process, after the other sections, instead of a process of its own")
    ap.add_argument("--serve-only", action="store_true", help="only the served stream (one JSON line)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--matvec-only", metavar="WORKLOAD", default=None,
                    help="build the workload's DA, run matvecs and a 256 MiB device copy (counter runs)")
    ap.add_argument("--da-only", metavar="WORKLOAD", default=None,
                    help="4 DA builds and 2 MG(5) set-ups of the workload's octree (kernel traces)")
    ap.add_argument("--build-only", action="store_true",
                    help="only the timed points2Octree + balanceOctree steps of lognormal-4M (kernel traces)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import dendro4_amd as dgx
    dgx.init(0)
    if args.serve_only:
        print(json.dumps(serve_section(dgx, args.serve_requests, args.serve_slots, 200e9)))
        return
    stream = dgx.stream()
    ev = HipEvents()
    if args.matvec_only:
        matvec_only(dgx, ev, stream, args.matvec_only)
        return
    if args.da_only:
        da_only(dgx, args.da_only)
        return

    W = "lognormal-4M"
    MD, MP = 30, dgx.workload_max_num_pts(W)
    pts = dgx.workload_points(W)
    N = len(pts)
    dpts = dgx.DeviceArray.from_host(pts)
    dgx.synchronize()
    for _ in range(args.warmup):
        dgx.balance_octree(dgx.points2octree(dpts, MD, MP))
    dgx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ou = dgx.points2octree(dpts, MD, MP)
    dgx.synchronize()
    t_p2o = (time.perf_counter() - t0) / args.steps
    ou, ob, t_build = build_section(dgx, dpts, MD, MP, args.steps, 0)
    U, Bn = len(ou), len(ob)
    octants_per_s = Bn / t_build
    if args.build_only:
        print(json.dumps({"workload": W, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_build * 1e3,
                          "p2o_ms": t_p2o * 1e3, "balanced_octants": Bn, "octants_per_s": octants_per_s}))
        return
    build_bytes = 24 * N + 2 * 28 * U + 28 * Bn
    log(f"build+balance {t_build * 1e3:.2f} ms/step ({Bn} octants)")

    da = dgx.DA(ob)
    E, Nn = da.element_size, da.n_nodes
    mv4 = time_matvecs(dgx, ev, stream, da, args.matvecs)
    del da
    log(f"matvec lognormal-4M {mv4['avg_ms'] * 1e3:.1f} us, {mv4['GB_per_s']:.0f} GB/s")

    solve = None
    if not args.no_solve:
        solve = {W: solve_section(dgx, ob, 5, args.solves, 2, 2, not args.no_graph)}
        log(f"solve {W}: {json.dumps(solve[W])}")
    del ou, ob, dpts

    # gauss4-16M: the roofline matvec (628 MB per launch) and the solve of config 3
    W3 = "gauss4-16M"
    p3 = dgx.DeviceArray.from_host(dgx.workload_points(W3))
    dgx.synchronize()
    o3u, o3, t3 = build_section(dgx, p3, MD, dgx.workload_max_num_pts(W3), 2, 1)
    del p3, o3u
    da3 = dgx.DA(o3)
    mv16 = time_matvecs(dgx, ev, stream, da3, args.matvecs)
    del da3
    log(f"matvec {W3} {mv16['avg_ms'] * 1e3:.1f} us, {mv16['GB_per_s']:.0f} GB/s")
    g16 = {"build_balance_s": t3, "balanced_octants": len(o3), "octants_per_s": len(o3) / t3, "matvec": mv16}
    if not args.no_solve:
        g16["solve"] = solve_section(dgx, o3, 5, max(5, args.solves // 2), 2, 2, not args.no_graph)
        log(f"solve {W3}: {json.dumps(g16['solve'])}")
    del o3

    large = None
    if not args.no_large:
        large = large_section(dgx, ev, stream, args)
        log(f"large: {json.dumps(large)}")

    serve = None
    if args.serve_requests > 0:
        if args.serve_in_process:
            serve = serve_section(dgx, args.serve_requests, args.serve_slots, 200e9)
            serve["process"] = "the bench's own, after the other sections"
        else:
            serve = serve_in_child(dgx, args.serve_requests, args.serve_slots)

    cpu = None if args.no_cpu else cpu_section(pts, MD, MP)

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
                               "coordinates exp(N(-1,0.5)) clipped to [0,1), seed 1, maxDepth 30, maxNumPts 8",
                   "points": N, "octants": U, "balanced_octants": Bn, "elements": E, "nodes": Nn},
        "roofline": {"bound": "hbm", "achieved": mv16["GB_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": mv16["GB_per_s"] / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "kernel": f"FP64 -lap+I matvec on the gauss4-16M DA: {MATVEC_KERNEL}",
                     "algorithmic_bytes": mv16["algorithmic_bytes"], "avg_launch_ms": mv16["avg_ms"]},
        "build_balance": {"s_per_step": t_build, "p2o_s": t_p2o, "algorithmic_bytes": build_bytes,
                          "roofline": {"bound": "hbm", "achieved": build_bytes / t_build / 1e9,
                                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": build_bytes / t_build / 1e9 / HBM_PEAK_GBS}},
        "matvec": mv4,
        "solve": solve,
        "gauss4-16M": g16,
        "large": large,
        "serve": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
