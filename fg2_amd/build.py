"""Build the gfx950 shared library libfg2_amd.so in-tree with hipcc.

    python -m fg2_amd.build          # or fg2_amd.build.build()

Flags: -O3 for gfx950 only; -ffp-contract=on, so multiply-adds fuse only
within one source expression and every kernel that evaluates the same
expression rounds it the same way; packed-FP32 instruction selection is switched
off (-Xclang -target-feature -Xclang -packed-fp32-ops), so the compiler
emits scalar v_fma_f32 / v_add_f32 for the FP32 stencil arithmetic.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfg2_amd.so")
COPY_EXE = os.path.join(HERE, "hbm_copy")
VALU_EXE = os.path.join(HERE, "valu_peak")
DRIVER_EXE = os.path.join(HERE, "fg2run")
SOURCES = ["kernels.hip", "capi.hip", "ics.hip", "spectra.hip"]
HEADERS = ["fg2_layout.h", "stencil.h", "physics.h", "launch.h", "modules/hydro.h",
           "modules/hydro_isot.h", "modules/disk.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"),
]


def _objects():
    return [os.path.join(CSRC, s.replace(".hip", ".o")) for s in SOURCES]


def build(verbose=False, jobs=2):
    """Compile every source (always) and link libfg2_amd.so."""
    procs = []
    for src, obj in zip(SOURCES, _objects()):
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    while procs:
        _wait(procs.pop(0))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *_objects(), "-o", tmp, "-lhipfft"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the HBM copy calibration executable
    cmd = [HIPCC, *FLAGS, os.path.join(CSRC, "hbm_copy.hip"), "-o", COPY_EXE]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    # the VALU issue ceiling probe
    cmd = [HIPCC, *FLAGS, os.path.join(CSRC, "valu_peak.hip"), "-o", VALU_EXE]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    # the command-line driver, linked against the library (rpath $ORIGIN)
    # (host code only: g++ against the HIP runtime API)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
           os.path.join(CSRC, "driver.cpp"), "-o", DRIVER_EXE,
           "-L", HERE, "-lfg2_amd", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


def _wait(item):
    cmd, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    if out.strip():
        # keep warnings visible but short
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l
                 and "not a recognized feature" not in l]
        if lines:
            sys.stderr.write("\n".join(lines[:40]) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
