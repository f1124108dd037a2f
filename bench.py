"""Benchmark: 3D Cartesian Sod tube on one MI355X, FP32 form and FP64 plain.

Workload (one step = one full Simulation.step on device-resident state):
3D Sod tube along x on a 256^3 uniform grid, gamma 1.4, (rho,u,p) = (1,0,1)
for x <= 0.5 and (0.125,0,0.1) beyond, uniform in y and z, MINBEE limiter,
open x ends, reflective y/z walls, CFL 0.8, synthetic initial data. Each
precision runs W untimed warmup steps and K timed steps (default 200, the
workload's length). The JSON line reports:

- value: whole-job cell-updates/s of the FP32 form (FP32 bulk arithmetic,
  FP64 CFL/balance reductions), with the FP64 plain method beside it;
- roofline: the fused step kernel's algorithmic HBM bytes (40 B per FP32
  cell-update, 80 B per FP64 one: one read and one write of the five
  conserved fields) over its average launch time measured with HIP events
  on the kernel's own stream, against MI355X's 8 TB/s;
- fp32_l1_rel: per-field relative L1 error of the FP32 form against the FP64
  plain method driven with the same dt sequence;
- cpu_baseline: the NumPy oracle (a restatement of the reference's CPU path)
  timed on this host on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "3D cell-updates/s FP32-batched vs FP64 plain; served p50/p99 solve latency"


def sod_primitive(n):
    x = (np.arange(n) + 0.5) / n
    prim = np.zeros((5, n, n, n))
    left = x <= 0.5
    prim[0] = np.where(left, 1.0, 0.125)[None, None, :]
    prim[4] = np.where(left, 1.0, 0.1)[None, None, :]
    return prim


OPTIONS = dict(t_end=10.0, cfl=0.8, limiter="MINBEE", boundaries=["Open", "Reflective", "Reflective"],
               checked=False)


def run_precision(n, precision, steps, warmup, tile=None):
    from torch_amd import Simulation

    sim = Simulation(sod_primitive(n), [1.0, 1.0, 1.0], precision=precision, tile=tile, **OPTIONS)
    for _ in range(warmup):
        sim.step()
    dev = sim._device
    dev.profile(True)
    dts = []
    t0 = time.perf_counter()
    for _ in range(steps):
        info = sim.step()
        dts.append(info.dt)
    wall = time.perf_counter() - t0
    kernel_ms, launches = dev.profile(False)
    fallbacks = sim.fallbacks
    state = sim.state
    sim.close()
    cells = n ** 3
    return dict(
        wall=wall, ms_per_step=1e3 * wall / steps, value=cells * steps / wall,
        kernel_ms=kernel_ms / max(launches, 1), launches=launches, fallbacks=fallbacks, dts=dts,
        state=state,
    )


def roofline(n, bytes_per_cell, kernel_ms):
    achieved = bytes_per_cell * n ** 3 / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}


def fp32_error(n, warmup, steps, dts64):
    """FP32 form driven with the FP64 run's dt sequence, compared per field."""
    from torch_amd import _abi
    from torch_amd.torchSolver3D import config, uprim_to_uconsv_3d
    from torch_amd.torchExecution import DeviceExecution

    state0 = uprim_to_uconsv_3d(sod_primitive(n))
    out = {}
    ex = DeviceExecution()
    for precision in ("fp64", "fp32"):
        cfg = config((n, n, n), None, (1.0, 1.0, 1.0), dim=3, limiter="MINBEE",
                     bcs=("Open", "Reflective", "Reflective"), precision=precision, cfl=0.8)
        dev = _abi.DeviceState(ex.context, cfg)
        dev.upload(state0)
        for dt in dts64:
            dev.hydro(dt)
        out[precision] = dev.download()[0]
        dev.close()
    ex.close()
    a, b = out["fp32"], out["fp64"]
    err = [float(np.sum(np.abs(x - y)) / max(np.sum(np.abs(y)), 1e-300)) for x, y in zip(a, b)]
    return err


def cpu_baseline(budget_s=12.0, n=40):
    """The NumPy oracle on the same workload at n^3, single thread."""
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    from oracle import session

    sim = session.Session(sod_primitive(n), [1.0, 1.0, 1.0], **OPTIONS)
    sim.step()  # warmup
    t0 = time.perf_counter()
    steps = 0
    while steps < 3 or time.perf_counter() - t0 < budget_s:
        sim.step()
        steps += 1
    wall = time.perf_counter() - t0
    return {"value": round(n ** 3 * steps / wall, 1), "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"3D Sod {n}^3 MINBEE CFL 0.8 walls, {steps} steps of the NumPy oracle "
                      f"(restated reference CPU path, 1 thread), {wall:.1f} s"}


def copy_calibration(mib=1024, reps=3):
    """Device-to-device copies of a known size, for scaling rocprof's
    FETCH_SIZE/WRITE_SIZE counters (each copy reads and writes `mib` MiB)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # device to device
    hip.hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_mib": mib, "reps": reps, "GBps_rw": round(2 * n * reps / wall / 1e9, 1)}))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--error-steps", type=int, default=50, help="steps of the FP32-vs-FP64 error pass")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (counter calibration) and exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the workload is the "
              "single-GPU 3D Sod tube; multi-GPU sharding is described, not run).", file=sys.stderr)
        return 2
    n = args.n
    print(f"# 3D Sod {n}^3: FP64 plain, {args.warmup}+{args.steps} steps", file=sys.stderr, flush=True)
    r64 = run_precision(n, "fp64", args.steps, args.warmup)
    print(f"# fp64: {r64['ms_per_step']:.3f} ms/step, kernel {r64['kernel_ms']:.3f} ms", file=sys.stderr,
          flush=True)
    r32 = run_precision(n, "fp32", args.steps, args.warmup)
    print(f"# fp32: {r32['ms_per_step']:.3f} ms/step, kernel {r32['kernel_ms']:.3f} ms", file=sys.stderr,
          flush=True)
    err = fp32_error(n, args.warmup, args.error_steps, r64["dts"][: args.error_steps])
    print(f"# fp32 L1 rel vs fp64 after {args.error_steps} steps: {err}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": METRIC, "value": round(r32["value"], 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D Cartesian Sod tube {n}^3, MINBEE, CFL 0.8, open x / reflective y,z, "
                               "gamma 1.4, full Simulation.step (FP32 form: f32 bulk, f64 reductions)",
                   "cells": n ** 3, "parallelism": "1 GPU"},
        "roofline": roofline(n, 40, r32["kernel_ms"]),
        "fp64_plain": {"value": round(r64["value"], 1), "ms_per_step": round(r64["ms_per_step"], 4),
                       "dtype": "f64", "kernel_ms": round(r64["kernel_ms"], 4),
                       "roofline": roofline(n, 80, r64["kernel_ms"]), "fallbacks": r64["fallbacks"]},
        "fp32_kernel_ms": round(r32["kernel_ms"], 4),
        "fp32_over_fp64": round(r32["value"] / r64["value"], 3),
        "fp32_l1_rel": [float(f"{e:.3e}") for e in err],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
