"""Precision study: FP32 GPU disk run against the FP64 host oracle.

    python tests/precision_study.py [--n 2048] [--steps 2000] [--out file.json]

Keplerian disc of the disk module (src/disk/eqns.cuh) around a central mass
(GM = 1, Newtonian: rS = 0, Plummer softening length eps = 0.05, a disk
parameter added for this workload) on the module's (log r, theta) grid
(src/disk/config.h:21-34): Sigma ~ R^-1 with 1% seeded density noise,
isothermal sound speed 0.05 v_K (fg2_amd.ics.disk_kepler).  Electron
scattering es = 1e6 keeps the flux-limited radiative diffusion non-stiff.
Both sides start from the same FP32-rounded state and take the same fixed
dt (0.9 x the FP64 CFL step of the initial state).  Reported after the last
step (and the relative L2 after the first 100 steps): per-field relative L2
and Linf error of the FP32 GPU state against the FP64 oracle, the relative
drift of total mass and of angular momentum sum(rho Omega R^2 dV) on each
side, and the wall time of each side.

The oracle is the reference here: this script compares against it and is not
part of the product path.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

PARAMS = {"es": 1.0e6, "rS": 0.0, "eps": 0.05}
FIELDS = ("lnd", "ur", "uz", "Omg", "lne", "lner")


def invariants(x, l1, n1, n2, rin=3.0):
    """FP64 total mass and angular momentum of a disk state (interior)."""
    lnr = l1 * (np.arange(n1) + 0.5) / n1
    th = np.pi * (np.arange(n2) + 0.5) / n2
    sphr = rin * np.exp(lnr)[:, None]
    R = sphr * np.sin(th)[None, :]
    dV = sphr ** 3 * np.sin(th)[None, :]
    rho = np.exp(x[..., 0].astype(np.float64))
    return float(np.sum(rho * dV)), float(np.sum(rho * x[..., 3] * R * R * dV))


def run(n, steps, progress=True):
    import fg2_amd
    from fg2_amd import ics
    from oracle.oracle import Oracle
    x = ics.disk_kepler(n, n, seed=4, h_over_r=0.05, eps=PARAMS["eps"]).astype(np.float32).astype(np.float64)
    o = Oracle("disk", n, n, params=PARAMS)
    o.upload(x)
    dt = 0.9 * o.getdt_cfl()
    l1 = o.geometry()["l1"]
    m0, j0 = invariants(x, l1, n, n)

    g = fg2_amd.Run("disk", n, n, params=PARAMS)
    g.upload(x)
    early = min(100, steps)
    t0 = time.perf_counter()
    for k in range(early):
        g.step_async(k * dt, dt)
    y100 = g.download().astype(np.float64)            # the state after 100 steps
    for k in range(early, steps):
        g.step_async(k * dt, dt)
    gpu_status = g.check()
    gpu_s = time.perf_counter() - t0
    y = g.download().astype(np.float64)
    if progress:
        print(f"gpu: {steps} steps in {gpu_s:.2f} s, status {gpu_status}", flush=True)

    t0 = time.perf_counter()
    done, status = 0, 0
    z100 = None
    while done < steps and status == 0:
        k = min(100, steps - done)
        status = o.run(done * dt, dt, k)
        done += k
        if done == early:
            z100 = o.download()
        if progress:
            print(f"oracle: {done}/{steps} steps, {time.perf_counter() - t0:.1f} s", flush=True)
    cpu_s = time.perf_counter() - t0
    z = o.download()
    l2_100 = np.sqrt(np.sum((y100 - z100) ** 2, axis=(0, 1))) / np.sqrt(np.sum(z100 ** 2, axis=(0, 1)))

    num = np.sqrt(np.sum((y - z) ** 2, axis=(0, 1)))
    den = np.sqrt(np.sum(z ** 2, axis=(0, 1)))
    linf = np.max(np.abs(y - z), axis=(0, 1)) / np.max(np.abs(z), axis=(0, 1))
    mg, jg = invariants(y, l1, n, n)
    mo, jo = invariants(z, l1, n, n)
    return {
        "workload": f"disk Keplerian {n}x{n}, GM=1 (rS=0, softening eps=0.05), es=1e6, c_s=0.05 v_K, "
                    f"{steps} RK3 steps, dt={dt:.4e} (0.9 CFL)",
        "n": n, "steps": steps, "dt": dt,
        "gpu_status": int(gpu_status), "oracle_status": int(status),
        "rel_l2": dict(zip(FIELDS, (num / den).tolist())),
        f"rel_l2_after_{early}_steps": dict(zip(FIELDS, l2_100.tolist())),
        "rel_linf": dict(zip(FIELDS, linf.tolist())),
        "mass_drift": {"gpu_fp32": mg / m0 - 1, "oracle_fp64": mo / m0 - 1},
        "angular_momentum_drift": {"gpu_fp32": jg / j0 - 1, "oracle_fp64": jo / j0 - 1},
        "gpu_s": gpu_s, "oracle_s": cpu_s,
        "oracle_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    r = run(a.n, a.steps)
    s = json.dumps(r)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s + "\n")


if __name__ == "__main__":
    main()
