"""Benchmark of the fused RK3 time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: hydro module, Kelvin-Helmholtz shear layer on a 3584 x 3584
periodic grid, FP32, fused kick+drift substages (three launches per RK3
step), fixed dt from the CFL limit of the initial state.  Inputs are
resident in HBM before the timed region.  Prints one JSON line with the
whole-job cell-updates/s, the roofline of the fused substage kernel and the
host CPU baseline (the oracle's FP32 instantiation on all host cores, on a
bounded sample of the same workload).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
N = 3584
METRIC = "cell-updates/s per RK3 step + % HBM peak; p50/p99 run latency (batched)"


class HipEvents:
    """hipEvent timing on an explicit stream (the run's own stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                                 ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self, stream):
        assert self.hip.hipStreamSynchronize(ctypes.c_void_p(stream)) == 0


def load_traffic():
    """HBM bytes per launch of the time-step kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload scaled by the
    copy-kernel calibration (None if the summary is absent)."""
    path = os.path.join(ROOT, "profiles", "fused_substage_hydro_3584_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def cpu_baseline(budget_s=15.0):
    """The oracle's FP32 instantiation (OpenMP over rows, all host cores) on
    the same 3584^2 KH workload, for as many steps as fit in budget_s."""
    from oracle.oracle import Oracle
    from fg2_amd import ics
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    x = ics.hydro_kh(N, N, seed=2)
    o = Oracle("hydro", N, N, fp64=False, geometry=dict(l1=1.0, l2=1.0, p1=1, p2=1, c=0.5))
    o.upload(x)
    dt = o.getdt_cfl()
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(steps * dt, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 50:
            break
    rate = N * N * steps / el
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"hydro KH {N}x{N} FP32, {steps} RK3 steps on the host "
                      f"({el:.1f} s, OMP_NUM_THREADS={threads})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--waves", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the time-step path runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import fg2_amd
    from fg2_amd import ics

    ev = HipEvents()
    x = ics.hydro_kh(N, N, seed=2)
    run = fg2_amd.Run("hydro", N, N, geometry=dict(l1=1.0, l2=1.0, p1=1, p2=1, c=0.5))
    run.set_tile(args.tile, args.waves)
    run.upload(x)
    dt = run.getdt_cfl()
    stream = run.stream()
    for k in range(args.warmup):
        run.step_async(k * dt, dt)
    assert run.check() == 0
    e0, e1 = ev.event(), ev.event()
    ev.sync(stream)
    t0 = time.perf_counter()
    ev.record(e0, stream)
    for k in range(args.steps):
        run.step_async(k * dt, dt)
    ev.record(e1, stream)
    ev.sync(stream)
    wall = time.perf_counter() - t0
    gpu_ms = ev.elapsed_ms(e0, e1)
    status = run.check()
    if status != 0:
        print("NaN in the benchmark state", file=sys.stderr)
        sys.exit(1)

    cells = N * N
    value = cells * args.steps / wall
    ms_step = 1e3 * wall / args.steps
    nv = 4
    alg_step = 10 * nv * 4 * cells                      # bytes per RK3 step
    launch_s = gpu_ms * 1e-3 / (3 * args.steps)         # average substage launch
    achieved = (alg_step / 3) / launch_s
    traffic = load_traffic()
    tile_h, waves = run.tile()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"hydro Kelvin-Helmholtz {N}x{N} periodic, FP32, "
                               "fused RK3 substages, fixed CFL dt",
                   "n1": N, "n2": N, "module": "hydro", "dt": dt,
                   "tile_h": tile_h, "waves": waves},
        "ns_per_cell_step": 1e9 * wall / (cells * args.steps),
        "hbm_frac_alg": value * 10 * nv * 4 / HBM_PEAK,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic,
                     "kernel": "fused substage (mean of the 3 stage launches)",
                     "alg_bytes_per_launch": alg_step / 3,
                     "avg_launch_us": launch_s * 1e6},
        "cpu_baseline": None,
    }
    run.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
