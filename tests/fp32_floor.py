"""FP32 floor of the scheme: the oracle's FP32 instantiation against its
FP64 instantiation on the same workload, per field.

    python tests/fp32_floor.py [--which kh|disk|isot|all] [--steps 100] [--out file.json]

A GPU FP32 run cannot be closer to the FP64 oracle than FP32 arithmetic
itself allows.  This script measures that floor on the host with the
oracle's own source evaluated in float (no FMA contraction, the reference's
term order), from the same FP32-rounded initial state and with the same dt
sequence as the GPU parity tests:

  kh    hydro Kelvin-Helmholtz 3584^2 periodic, the CFL dt of the initial
        state, 100 steps (the benchmark workload and its GPU parity test);
  disk  the Keplerian disc of the precision study, 2048^2, 0.9 x the FP64
        CFL dt, 100 steps;
  isot  hydro-isot 1024 x 992 density perturbation (the largest run of the
        batched workload's size range), closed-form dt, 100 steps.

Reported per field: relative L2 (normalised by the FP64 field) of FP32 vs
FP64, and the same for the velocity vector.  The oracle is the reference here;
nothing in the product imports this script.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

DISK_PARAMS = {"es": 1.0e6, "rS": 0.0, "eps": 0.05}


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    num = np.sqrt(np.sum((a - b) ** 2, axis=(0, 1)))
    den = np.sqrt(np.sum(b ** 2, axis=(0, 1)))
    return num / np.where(den > 1e-30, den, 1.0)


def _vel(a, b, sl):
    a = np.asarray(a, np.float64)[..., sl]
    b = np.asarray(b, np.float64)[..., sl]
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def workload(which):
    """(module, n1, n2, params, geometry, FP32-rounded state, dt rule)."""
    from fg2_amd import ics
    if which == "kh":
        x = ics.hydro_kh(3584, 3584, seed=2)
        return "hydro", 3584, 3584, None, None, x, "cfl"
    if which == "disk":
        x = ics.disk_kepler(2048, 2048, seed=4, h_over_r=0.05, eps=DISK_PARAMS["eps"])
        return "disk", 2048, 2048, DISK_PARAMS, None, x, "cfl09"
    if which == "isot":
        x = ics.isot_density(1024, 992, seed=8)
        return "hydro-isot", 1024, 992, None, None, x, "closed"
    raise ValueError(which)


def floor(which, steps=100, progress=True):
    from oracle.oracle import Oracle
    from fg2_amd import FIELDS
    module, n1, n2, params, geo, x, rule = workload(which)
    x = np.asarray(x, np.float32).astype(np.float64)
    o64 = Oracle(module, n1, n2, params=params, geometry=geo)
    o32 = Oracle(module, n1, n2, params=params, geometry=geo, fp64=False)
    o64.upload(x)
    o32.upload(x)
    if rule == "cfl":
        dt = o64.getdt_cfl()
    elif rule == "cfl09":
        dt = 0.9 * o64.getdt_cfl()
    else:
        dt = o64.getdt()
    t0 = time.perf_counter()
    for o, name in ((o64, "fp64"), (o32, "fp32")):
        for k in range(0, steps, 20):
            m = min(20, steps - k)
            assert o.run(k * dt, dt, m) == 0, f"{name} oracle NaN"
            if progress:
                print(f"{which} {name}: {k + m}/{steps} steps, {time.perf_counter() - t0:.0f} s",
                      flush=True)
    z, y = o64.download(), o32.download()
    e = _rel(y, z)
    out = {"workload": which, "module": module, "n1": n1, "n2": n2, "steps": steps, "dt": dt,
           "rel_l2_fp32_vs_fp64": dict(zip(FIELDS[module], e.tolist())),
           "velocity_rel_l2": _vel(y, z, slice(1, 3))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--which", default="all")
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    names = ["kh", "disk", "isot"] if a.which == "all" else a.which.split(",")
    res = {n: floor(n, a.steps) for n in names}
    s = json.dumps(res, indent=1)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s + "\n")


if __name__ == "__main__":
    main()
