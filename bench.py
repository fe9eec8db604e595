"""Benchmark: octree build + 2:1 balance (octants/s) and the FP64 -lap+I matvec
on one MI355X, configuration 2 (4,194,304 log-normal points, maxNumPts=8,
maxDepth=30, corner balance, 100 matvecs).

One step = points2Octree + balanceOctree over the resident point cloud.
Prints one JSON line (see the "metric"/"config"/"roofline"/"cpu_baseline"
fields).  The oracle (CPU restatement) is timed on a bounded sample of the same
workload on the host as the CPU baseline.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--matvecs", type=int, default=100)
    ap.add_argument("--cpu-sample", type=int, default=524288, help="points in the oracle's CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import dendro4_amd as dgx
    dgx.init(0)
    stream = dgx.stream()
    ev = HipEvents()

    N, MP, MD = 4194304, 8, 30
    pts = dgx.config_points(2, N)
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
    mv_avg_ms = float(np.mean(mv_ms))
    mv_bytes = 34 * E + 16 * Nn
    achieved = mv_bytes / (mv_avg_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as O
        if not os.path.exists(O.LIB_PATH):
            import __graft_entry__
            __graft_entry__.build_oracle()
        sub = pts[: args.cpu_sample]
        t0 = time.perf_counter()
        ru = O.points2octree(sub, MD, MP)
        rb = O.balance(ru)
        tc = time.perf_counter() - t0
        cpu = {"value": len(rb) / tc, "unit": "octants/s", "cores": 1, "kind": "port",
               "sample": f"oracle points2Octree+balanceOctree on the first {len(sub)} points of the workload "
                         f"({len(rb)} balanced octants, {tc:.1f} s, 1 thread)"}

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
        "config": {"workload": "config 2: points2Octree + balanceOctree(corner) of 4,194,304 log-normal points "
                               "exp(N(-1,0.5)) clipped to [0,1), seed 1, maxDepth 30, maxNumPts 8; "
                               "100 FP64 -lap+I matvecs on the DA",
                   "points": N, "octants": U, "balanced_octants": Bn, "elements": E, "nodes": Nn},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_matvec<double> (FP64 -lap+I element matvec)",
                     "algorithmic_bytes": mv_bytes, "avg_launch_ms": mv_avg_ms},
        "matvec": {"da_blocks": da.block_info, "avg_ms": mv_avg_ms, "min_ms": float(np.min(mv_ms)), "GB_per_s": achieved,
                   "flops": 332 * E, "GFLOP_per_s": 332 * E / (mv_avg_ms * 1e-3) / 1e9},
        "build_balance": {"s_per_step": t_build, "p2o_s": t_p2o, "algorithmic_bytes": build_bytes,
                          "GB_per_s": build_bytes / t_build / 1e9},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
