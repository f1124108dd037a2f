"""Compile libtorchhip.so for gfx950 from torch_amd/csrc.

Each kernel source is compiled twice: the plain FP64 method (``fv64``,
``-ffp-contract=off`` so products and sums round one at a time like the
reference's NumPy arithmetic) and the FP32 form (``fv32``, contraction and
single-instruction reciprocal/square root). All device code is built with
packed-FP32 instruction selection disabled.

Usage: python -m torch_amd.build [--jobs N]
"""

import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtorchhip.so")
ARCH = "gfx950"

KERNELS = ("fv3d.hip", "fv2d.hip", "fv1d.hip", "fvbatch.hip", "fvmetric.hip")
HOST = ("capi.hip", "batch.hip", "adapt.hip")
COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
PRECISION = {
    "fv64": ["-DFV_REAL=double", "-DFV_NS=fv64", "-ffp-contract=off"],
    # Contraction everywhere except where a pragma turns it off (bookkeeping
    # that must round like the host driver's, and the one per-cell CFL speed
    # formula): plain "fast" would ignore those pragmas.
    "fv32": ["-DFV_REAL=float", "-DFV_NS=fv32", "-DFV_FAST=1", "-ffp-contract=fast-honor-pragmas",
             "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
}


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line and "not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{out}")
    return out


def build(jobs=None, verbose=False):
    cc = hipcc()
    os.makedirs(OBJ, exist_ok=True)
    units = []
    for src in KERNELS:
        for tag, flags in PRECISION.items():
            obj = os.path.join(OBJ, f"{os.path.splitext(src)[0]}_{tag}.o")
            units.append([cc, *COMMON, *flags, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        units.append([cc, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(units), os.cpu_count() or 4, 8)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        for out in pool.map(_run, units):
            if verbose and out.strip():
                print(out)
    objs = [u[-1] for u in units]
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    print(build(args.jobs, args.verbose))


if __name__ == "__main__":
    sys.exit(main())
