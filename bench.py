"""Benchmark on one MI355X: the 3D Sod tube (FP32 batched form and FP64 plain
method) and the batched uneven mix.

Headline workload (one step = one full Simulation.step of device-resident
state): 3D Sod tube along x on a 256^3 uniform grid, gamma 1.4,
(rho,u,p) = (1,0,1) for x <= 0.5 and (0.125,0,0.1) beyond, uniform in y and
z, MINBEE limiter, open x ends, reflective y/z walls, CFL 0.8, synthetic
initial data; W untimed warmup steps and K timed steps (default 200, the
workload's length). The JSON line reports:

- value: whole-job cell-updates/s of the FP32 batched form (FP32 bulk
  arithmetic, FP64 CFL/balance reductions; the instance stepped by graph
  replays of the batch engine with the FP32 form's 768-thread 32x24 tile,
  the autotuner's fastest shape for one grid this large);
- fp64_plain: the FP64 plain method through the session step loop;
- fp64_batched: the FP64 plain method as a graph-replayed batch of one;
- roofline: the 3D step kernel's algorithmic HBM bytes (40 B per FP32
  cell-update: one read and one write of the five conserved fields) over
  its average launch time, from HIP events the step graph records around it
  on its stream over the 20 ordinary steps that follow the timed ones
  (fp32_pass_ms: the kernel relaunched alone on the warmed-up state),
  against MI355X's 8 TB/s; traffic from rocprof counter runs when given;
- fp32_error (and fp32_l1_rel): per-field relative L1 and Linf errors of
  the FP32 form against the FP64 plain method over the workload's 200 steps,
  both driven with the FP64 run's dt sequence, and each precision's
  conservation drift per field;
- batched: the uneven mix (256 seeded 1D/2D/3D two-state problems, 100
  steps each, FP32 and FP64 batches) with its throughput and FP32 error;
- served: seeded 3D ignition requests through the continuous-batching
  server (closed loop for capacity, then Poisson arrivals at 80% of it):
  p50/p99 time-to-solution and aggregate cell-updates/s, and six requests
  from the small end (n <= 60; five Arrhenius, one ZND) served in FP32 and
  FP64 to their end time (steps, dt halvings, per-field relative L1 and Linf
  of FP32 against FP64);
- metric_face: the metric/face scheme (curvilinear grid, face boundaries,
  SSPRK2 MUSCL-Rusanov) on a cylindrical-sector blast, FP64 and FP32, with
  its own CPU baseline;
- adaptive: the dyadic AMR leaves (2D and 3D Sod, left half refined twice),
  FP64 and FP32 leaf-updates/s of the device step, the host remesh time, and
  its own CPU baseline;
- legacy: the legacy 1D driver with its full time history, with its own
  CPU baseline;
- checkpoint: save and load rates of a 256^3 FP64 session's checkpoint;
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
WORKLOAD_STEPS = 200  # the 3D Sod workload's length (200 steps at CFL 0.8)
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


KERNEL_STEPS = 20  # steps after the timed ones over which the step kernel is timed


def run_batched_sod(n, steps, warmup, precision="fp32", tile=None):
    """The Sod instance as a batch of one: graph-replayed steps, tuned tile.
    The step kernel's launch time (`kernel_ms`) comes from HIP events that
    the step graph records around it, over the KERNEL_STEPS ordinary steps
    that follow the timed ones (the timed steps themselves run unobserved);
    `pass_ms` is the same kernel relaunched alone on the warmed-up state,
    which reads one unchanging input and so partly from the Infinity Cache."""
    from torch_amd.torchBatch import Batch

    batch = Batch(precision=precision)
    ident = batch.add_problem(sod_primitive(n), [1.0, 1.0, 1.0], t_end=OPTIONS["t_end"], cfl=OPTIONS["cfl"],
                              limiter=OPTIONS["limiter"], boundaries=OPTIONS["boundaries"],
                              max_steps=warmup + steps + KERNEL_STEPS)
    if tile == "tune":
        tile, _ = batch.tune(reps=3)
    else:
        tile = tile or SOD_TILE[precision]
        batch.set_tiles(*tile)
    batch.run(warmup)
    pass_ms = batch.time_pass(reps=10)  # while the instance is still stepping
    t0 = time.perf_counter()
    batch.run(steps)
    wall = time.perf_counter() - t0
    kernel_ms = batch.time_step_kernel(KERNEL_STEPS)
    info = batch.info(ident)
    batch.close()
    if info["steps"] != warmup + steps + KERNEL_STEPS or info["status"] != 0:
        raise RuntimeError(f"batched Sod instance stopped early: {info}")
    return dict(wall=wall, ms_per_step=1e3 * wall / steps, value=n ** 3 * steps / wall, kernel_ms=kernel_ms,
                pass_ms=pass_ms, tile=tile, fallbacks=info["fallbacks"])


def run_uneven(seed, count, steps, precisions=("fp32", "fp64"), tile=None):
    """The batched uneven mix in each precision; FP32 error against FP64."""
    from torch_amd.torchBatch import Batch
    from torch_amd.workloads import uneven_problems

    out, batches, ids, cells = {}, {}, {}, {}
    for precision in precisions:
        t0 = time.perf_counter()
        batch = Batch(precision=precision)
        ids[precision] = []
        by_dim = {1: 0, 2: 0, 3: 0}
        for i, prim, opts in uneven_problems(seed, count):
            ident = batch.add_problem(prim, opts["lengths"], t_end=100.0, cfl=opts["cfl"], gamma=opts["gamma"],
                                      limiter=opts["limiter"], boundaries=opts["boundaries"], max_steps=steps)
            ids[precision].append(ident)
            ncell = int(np.prod(prim.shape[1:]))
            cells[i] = ncell
            by_dim[prim.ndim - 1] += ncell
        if tile:
            batch.set_tiles(*tile)
            used = tile
        else:
            used, _ = batch.tune(reps=2)
        batch.run(0)  # admission: first dt of every instance, graph capture
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        batch.run(steps)
        wall = time.perf_counter() - t0
        done = [batch.info(i) for i in ids[precision]]
        bad = [d for d in done if d["status"] != 0 or d["steps"] != steps]
        total = sum(d["cells"] * d["steps"] for d in done)
        out[precision] = dict(value=total / wall, ms_per_step=1e3 * wall / steps, wall=wall, setup_s=setup,
                              tile=list(used), stopped_early=len(bad),
                              fallback_steps=int(sum(d["fallbacks"] for d in done)),
                              cells_by_dim={str(k): v for k, v in by_dim.items()})
        batches[precision] = batch
        print(f"# uneven {precision}: {out[precision]['ms_per_step']:.3f} ms/step, "
              f"{out[precision]['value']:.4g} cell-updates/s, setup {setup:.1f} s", file=sys.stderr, flush=True)
    err = None
    if "fp32" in batches and "fp64" in batches:
        worst = np.zeros(5)
        per = []
        for a, b in zip(ids["fp32"], ids["fp64"]):
            x, y = batches["fp32"].state(a), batches["fp64"].state(b)
            e = [float(np.sum(np.abs(u - v)) / max(np.sum(np.abs(v)), 1e-300)) for u, v in zip(x, y)]
            per.append(max(e))
            worst[: len(e)] = np.maximum(worst[: len(e)], e)
        err = dict(max_l1_rel_per_field=[float(f"{e:.3e}") for e in worst],
                   median_instance_l1_rel=float(f"{np.median(per):.3e}"))
    for b in batches.values():
        b.close()
    return out, err, sum(cells.values())


def run_served(seed, count, t_run, budget_gib, round_steps, precision="fp32", open_loop=True, policy="backfill"):
    """Served reactive ignition solves: a closed-loop run (all requests at
    t=0) measures capacity, then Poisson arrivals at 80% of it give the
    p50/p99 time-to-solution. Each request's initial arrays arrive with it
    (built before the clock); admission uploads them."""
    from torch_amd.torchServe import Server, ignition_request, summary
    from torch_amd.workloads import served_requests

    specs = served_requests(seed=seed, count=count, t_run=t_run)
    budget = int(budget_gib * (1 << 30))
    out = {}
    server = Server(precision=precision, hbm_budget=budget, round_steps=round_steps, policy=policy)
    closed = summary(server.serve([ignition_request(s, 0.0) for s in specs], prebuild=True))
    phases = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in server.timing.items()}
    phases["graph_instantiated"], phases["graph_updated"] = server.batch._batch.graph_counts()
    server.close()
    capacity = closed["completed"] / closed["makespan_s"]
    print(f"# served closed loop: {closed['makespan_s']:.2f} s for {count} requests, "
          f"{closed['cell_updates_per_s']:.4g} cell-updates/s", file=sys.stderr, flush=True)
    per = lambda k: 1e3 * phases[k + "_s"] / max(phases[k + "_rounds"], 1)
    print(f"# served closed loop phases: admissions {phases['admit_s']:.2f} s, {phases['changed_rounds']} rounds "
          f"after a membership change {phases['changed_s']:.2f} s ({per('changed'):.1f} ms each), "
          f"{phases['steady_rounds']} steady rounds {phases['steady_s']:.2f} s ({per('steady'):.1f} ms each), "
          f"retirements {phases['retire_s']:.2f} s; step graph instantiated {phases['graph_instantiated']} times, "
          f"updated in place {phases['graph_updated']} times", file=sys.stderr, flush=True)
    out["closed_loop"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in closed.items()}
    out["closed_loop"]["host_phases"] = phases
    out["capacity_per_s"] = round(capacity, 3)
    if not open_loop:
        return out
    rng = np.random.default_rng(seed + 1)
    rate = 0.8 * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    arrivals -= arrivals[0]
    server = Server(precision=precision, hbm_budget=budget, round_steps=round_steps, policy=policy)
    opened = summary(server.serve([ignition_request(s, a) for s, a in zip(specs, arrivals)], prebuild=True))
    server.close()
    print(f"# served open loop at {rate:.2f} req/s: p50 {opened['p50_s']:.3f} s, p99 {opened['p99_s']:.3f} s",
          file=sys.stderr, flush=True)
    out["open_loop"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in opened.items()}
    out["arrival_rate_per_s"] = round(rate, 3)
    return out


def served_fp32_error(seed, count, t_run, n_max=60):
    """Requests from the small end of the served workload (the first five
    Arrhenius and the first ZND request with n <= n_max) served to their end
    time in the FP32 form and by the FP64 method: per
    request the steps and dt halvings of each and the FP32 state's relative
    L1 and Linf per field against the FP64 one (the FP64 method matches the
    oracle's reactive session; the tests hold that)."""
    from torch_amd.torchServe import Server, ignition_request
    from torch_amd.workloads import served_requests

    small = [s for s in served_requests(seed=seed, count=count, t_run=t_run) if s["n"] <= n_max]
    specs = [s for s in small if s["model"] == "arrhenius"][:5] + [s for s in small if s["model"] == "znd"][:1]
    got = {}
    for precision in ("fp64", "fp32"):
        server = Server(precision=precision, hbm_budget=0, round_steps=8, keep_results=True)
        try:
            reqs = server.serve([ignition_request(s, 0.0) for s in specs])
        finally:
            server.close()
        got[precision] = {r.ident: r for r in reqs}
    rows = []
    for spec in specs:
        a, b = got["fp32"][spec["index"]], got["fp64"][spec["index"]]
        x, y = a.result[0], b.result[0]
        l1 = [float(np.sum(np.abs(u - v)) / max(np.sum(np.abs(v)), 1e-300)) for u, v in zip(x, y)]
        linf = [float(np.max(np.abs(u - v)) / max(np.max(np.abs(v)), 1e-300)) for u, v in zip(x, y)]
        rows.append(dict(index=spec["index"], n=spec["n"], model=spec["model"], steps=[b.steps, a.steps],
                         halvings=[b.retries, a.retries], l1_rel=[float(f"{e:.3e}") for e in l1],
                         linf_rel=[float(f"{e:.3e}") for e in linf]))
    worst = lambda key: [max(r[key][c] for r in rows) for c in range(5)]
    out = dict(requests=rows, max_l1_rel_per_field=worst("l1_rel"), max_linf_rel_per_field=worst("linf_rel"),
               same_steps_and_halvings=all(r["steps"][0] == r["steps"][1] and r["halvings"][0] == r["halvings"][1]
                                           for r in rows))
    print(f"# served fp32 vs fp64 ({len(rows)} requests, n <= {n_max}): L1 per field "
          f"{out['max_l1_rel_per_field']}, Linf {out['max_linf_rel_per_field']}, "
          f"same steps and halvings: {out['same_steps_and_halvings']}", file=sys.stderr, flush=True)
    return out


def run_reactive(seed, count, t_run, steps, warmup, precision="fp32", tile=None):
    """The served ignition requests admitted together into one batch of fixed
    membership (no arrivals or retirements), stepped by graph replays: the
    reactive kernels' throughput without admission work."""
    from torch_amd.torchServe import Server, ignition_request
    from torch_amd.workloads import served_requests

    specs = served_requests(seed=seed, count=count, t_run=t_run)
    server = Server(precision=precision, hbm_budget=0, round_steps=steps)
    ids = []
    for spec in specs:
        req = ignition_request(spec, 0.0)
        ident = server._admit(req, server.payload(*req.build()))
        if ident is None:
            raise RuntimeError("reactive batch instance not admitted")
        ids.append(ident)
    batch = server.batch
    if tile == "tune":
        tile, _ = batch.tune(reps=2)
    elif tile:
        batch.set_tiles(*tile)
    batch.run(warmup)
    before = [batch.info(i) for i in ids]
    t0 = time.perf_counter()
    batch.run(steps)
    wall = time.perf_counter() - t0
    after = [batch.info(i) for i in ids]
    server.close()
    work = sum(b["cells"] * (b["steps"] - a["steps"]) for a, b in zip(before, after))
    retries = sum(b["retries"] - a["retries"] for a, b in zip(before, after))
    done = sum(b["steps"] - a["steps"] for a, b in zip(before, after))
    return dict(value=round(work / wall, 1), ms_per_replay=round(1e3 * wall / steps, 4), replays=steps,
                steps=done, retries=retries, cells=sum(b["cells"] for b in after), dtype=precision.replace("fp", "f"),
                tile=list(tile) if tile else None)


# HBM bytes per launch of the 3D step kernel on the 256^3 Sod tube with the
# Sod legs' tiles (FP32 32x24, FP64 32x16, z chunks of 24), from rocprofv3
# FETCH_SIZE and WRITE_SIZE counter runs over the 200-step workload (mean per
# dispatch) scaled by a 1 GiB device copy (FETCH_SIZE x 2.000, WRITE_SIZE x
# 1.000); the commands and per-counter values are kept with the profile
# summaries.
TRAFFIC = {("f32", 256, (32, 24, 24)): 1275653237, ("f64", 256, (32, 16, 24)): 2203313824}

# Wave-level VALU instructions per launch of the same kernels (rocprofv3
# --pmc SQ_INSTS_VALU, mean over the dispatches of the same 200-step runs).
VALU_INSTS = {("f32", 256, (32, 24, 24)): 200359171, ("f64", 256, (32, 16, 24)): 298922000}

# VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane
# slots per second (a wave64 instruction holds a SIMD for 4 cycles; FP32
# without packed math and FP64 both issue one lane operation per lane slot).
VALU_PEAK = 256 * 4 * 16 * 2.4e9

# Tiles (BX, BY, z chunk) of the 256^3 Sod legs per precision: FP64 the
# library default; FP32 its 768-thread 32x24 tile, the fastest shape of the
# autotuner's sweep for one grid this large (the library's FP32 default,
# 32x8, is the faster one for mixes of small grids).
SOD_TILE = {"fp32": (32, 24, 24), "fp64": (32, 16, 24)}


def roofline(n, bytes_per_cell, kernel_ms, dtype=None, tile=None):
    """HBM roofline of the step kernel (algorithmic bytes over its launch
    time), with its VALU issue fraction beside it: the kernel issues
    SQ_INSTS_VALU x 64 lane operations per launch, against VALU_PEAK."""
    achieved = bytes_per_cell * n ** 3 / (kernel_ms * 1e-3) / 1e9
    key = (dtype, n, tuple(tile) if tile else None)
    traffic = TRAFFIC.get(key)
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "algorithmic_bytes": bytes_per_cell * n ** 3}
    insts = VALU_INSTS.get(key)
    if insts:
        ops = insts * 64 / (kernel_ms * 1e-3)
        out["valu_issue"] = {"achieved": float(f"{ops:.4g}"), "peak": VALU_PEAK, "unit": "lane-ops/s",
                             "frac": round(ops / VALU_PEAK, 4), "insts_per_launch": insts}
        out["binding"] = ("valu issue and latency: VALU issue at {:.0%} of peak, HBM at {:.0%}"
                          .format(ops / VALU_PEAK, achieved / HBM_PEAK_GBS))
    return out


def fp32_error(n, steps):
    """The FP32 form against the FP64 plain method over `steps` hydro steps
    of the workload (the FP64 run's CFL dt sequence drives both): relative L1
    and Linf errors per field, and each precision's conservation drift per
    field, |final - initial + outflow| over the largest of the three."""
    from torch_amd import _abi
    from torch_amd.torchSolver3D import config, uprim_to_uconsv_3d
    from torch_amd.torchExecution import DeviceExecution

    state0 = uprim_to_uconsv_3d(sod_primitive(n))
    out, drift, dts = {}, {}, []
    ex = DeviceExecution()
    for precision in ("fp64", "fp32"):
        cfg = config((n, n, n), None, (1.0, 1.0, 1.0), dim=3, limiter="MINBEE",
                     bcs=("Open", "Reflective", "Reflective"), precision=precision, cfl=0.8)
        dev = _abi.DeviceState(ex.context, cfg)
        dev.upload(state0)
        initial = dev.totals()
        outward = np.zeros(5)
        for k in range(steps):
            if precision == "fp64":
                dts.append(dev.dt_limit())
            info = dev.hydro(dts[k])
            outward += np.array(info.outward[:5])
        final = dev.totals()
        # Fields whose totals are rounding residues (the y and z momenta of
        # this tube are identically zero) are held to the largest field's scale.
        scale = np.maximum.reduce([np.abs(initial), np.abs(final), np.abs(outward)])
        scale = np.maximum(scale, 1e-12 * np.max(scale))
        drift[precision] = np.abs(final - initial + outward) / scale
        out[precision] = dev.download()[0]
        dev.close()
    ex.close()
    a, b = out["fp32"], out["fp64"]
    l1, linf = [], []
    for x, y in zip(a, b):
        den = np.sum(np.abs(y))
        l1.append(float(np.sum(np.abs(x - y)) / den) if den > 0 else float(np.max(np.abs(x))))
        top = np.max(np.abs(y))
        linf.append(float(np.max(np.abs(x - y)) / top) if top > 0 else float(np.max(np.abs(x))))
    fmt = lambda v: [float(f"{e:.3e}") for e in v]  # noqa: E731
    return {"steps": steps, "l1_rel": fmt(l1), "linf_rel": fmt(linf), "drift_fp32": fmt(drift["fp32"]),
            "drift_fp64": fmt(drift["fp64"]),
            "note": "fields 2 and 3 (y, z momenta) are identically zero in both: their error entries are "
                    "max |FP32|, their drift is relative to 1e-12 of the largest field's scale"}


def host_cpu():
    """The host's CPU model and logical CPU count, for the CPU baselines."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"


def cpu_workers():
    """Worker count of the CPU baseline's pool: the reference's cap of 64
    (torchExecution.py:7-14), held to this host's share of 16 CPUs per GPU."""
    return max(1, min(64, 16, os.cpu_count() or 1))


def cpu_run(n, nz, workers, steps):
    """Seconds per step of the NumPy oracle's Sod session on an n x n x nz
    slab of the 256^3 tube (same cell sizes, the same per-column arithmetic:
    the state is uniform in y and z and the walls mirror it), with the
    reference's 8192-face HLLC batches through its worker pool: one untimed
    step, then `steps` timed ones."""
    from oracle import session

    prim = np.ascontiguousarray(sod_primitive(n)[:, :nz])
    sim = session.Session(prim, [1.0, 1.0, nz / n], workers=workers, **OPTIONS)
    sim.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    wall = time.perf_counter() - t0
    sim.execution.close()
    return wall / steps


def cpu_baseline(n=256, nz=32, steps=2):
    """The NumPy oracle (a restatement of the reference's CPU path, HLLC in
    8192-face batches through the worker pool) on a bounded slab of the
    workload, with one worker (the reference's default) and with a pool."""
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    cells = n * n * nz
    runs = {}
    for workers in sorted({1, cpu_workers()}):
        per = cpu_run(n, nz, workers, steps)
        runs[workers] = {"value": round(cells / per, 1), "s_per_step": round(per, 3), "workers": workers}
        print(f"# cpu baseline, {workers} worker(s): {per:.2f} s per step", file=sys.stderr, flush=True)
    best = max(runs.values(), key=lambda r: r["value"])
    return {"value": best["value"], "unit": "cell-updates/s", "cores": best["workers"], "kind": "port",
            "sample": f"3D Sod tube {n}x{n}x{nz} slab of the {n}^3 workload (same cell sizes and per-column "
                      f"arithmetic), MINBEE CFL 0.8 walls, 1 warmup + {steps} timed steps of the NumPy oracle "
                      "with the reference's 8192-face HLLC batches through its worker pool (OPENBLAS 1 thread), "
                      f"on {host_cpu()}",
            "runs": [runs[w] for w in sorted(runs)]}


def metric_problem(n):
    """A cylindrical-sector blast on the metric/face scheme: (r, theta, z) in
    [0.5, 1.5] x [0, pi/2] x [0, 1], n^3 cells, rho = 1 at rest, p = 10 in a
    torus of radius 0.15 around (r, z) = (1, 0.5), p = 1 elsewhere; walls at
    r = 0.5 and on both theta planes, a pressure outlet (with a backflow
    state) at r = 1.5, periodic in z; Superbee, CFL 0.4."""
    r = 0.5 + (np.arange(n) + 0.5) / n
    z = (np.arange(n) + 0.5) / n
    prim = np.zeros((5, n, n, n))
    prim[0] = 1.0
    hot = ((r[None, None, :] - 1.0) ** 2 + (z[:, None, None] - 0.5) ** 2) < 0.15 ** 2
    prim[4] = np.where(hot, 10.0, 1.0) + 0.0 * prim[0]
    faces = {"x_min": "wall", "x_max": {"type": "pressure_outlet", "pressure": 1.0,
                                        "backflow": [1.0, -0.1, 0.0, 0.0, 1.0]},
             "y_min": "wall", "y_max": "wall", "z_min": "periodic", "z_max": "periodic"}
    opts = dict(geometry="annular", origins=[0.5, 0.0, 0.0], t_end=10.0, cfl=0.4, limiter="Superbee",
                face_boundaries=faces)
    return prim, [1.0, np.pi / 2, 1.0], opts


def run_metric(n, precision, steps, warmup):
    from torch_amd import Simulation

    prim, lengths, opts = metric_problem(n)
    sim = Simulation(prim, lengths, precision=precision, **opts)
    for _ in range(warmup):
        sim.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    wall = time.perf_counter() - t0
    fallbacks = sim.fallbacks
    sim.close()
    return dict(value=round(n ** 3 * steps / wall, 1), ms_per_step=round(1e3 * wall / steps, 4),
                fallbacks=fallbacks, dtype="f64" if precision == "fp64" else "f32")


def cpu_metric_baseline(budget_s=8.0, n=24):
    """The NumPy oracle's metric step on the same workload at n^3, single thread."""
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    from oracle import session

    prim, lengths, opts = metric_problem(n)
    sim = session.Session(prim, lengths, **opts)
    sim.step()
    t0 = time.perf_counter()
    steps = 0
    while steps < 3 or time.perf_counter() - t0 < budget_s:
        sim.step()
        steps += 1
    wall = time.perf_counter() - t0
    return {"value": round(n ** 3 * steps / wall, 1), "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"cylindrical-sector blast {n}^3, {steps} metric steps of the NumPy oracle, {wall:.1f} s"}


def adaptive_problem(nx=512, ny=64, nz=None):
    """A Sod tube on dyadic leaves: nx x ny (x nz) roots over [0, 1] x
    [0, ny/nx] (x [0, nz/nx]; cubic cells), (rho, u, p) = (1, 0, 1) for
    x < 0.5 and (0.125, 0, 0.1) beyond, open x ends, reflective walls
    elsewhere, CFL 0.4, second-order reconstruction; the left half refined
    twice (level 2)."""
    dim = 2 if nz is None else 3
    counts = [nx, ny] + ([nz] if nz else [])
    prim = np.zeros((dim + 2, *counts[::-1]))
    right = (np.arange(nx) + 0.5) / nx >= 0.5
    prim[0] = np.where(right, 0.125, 1.0)
    prim[-1] = np.where(right, 0.1, 1.0)
    opts = dict(boundaries=["Open"] + ["Reflective"] * (dim - 1), t_end=10.0, cfl=0.4)
    return prim, [c / nx for c in counts], opts


def adaptive_session(precision, nx, ny, nz=None, regrid_every=10 ** 9):
    from torch_amd import Simulation

    prim, lengths, opts = adaptive_problem(nx, ny, nz)
    s = Simulation(prim, lengths, precision=precision, **opts).to_adaptive(
        max_level=2, regrid_every=regrid_every, max_cells=2_000_000)
    for _ in range(2):
        s.regrid(refine=s.mesh.centers[:, 0] < 0.5, coarsen=False)
    return s


def run_adaptive(precision, steps, warmup, nx, ny, nz=None):
    s = adaptive_session(precision, nx, ny, nz)
    leaves, faces = len(s.mesh.keys), len(s.mesh.left)
    for _ in range(warmup):
        s.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
    wall = time.perf_counter() - t0
    t1 = time.perf_counter()
    s.regrid()
    regrid_s = time.perf_counter() - t1
    out = dict(value=round(leaves * steps / wall, 1), ms_per_step=round(1e3 * wall / steps, 4), leaves=leaves,
               faces=faces, regrid_s=round(regrid_s, 3), leaves_after_regrid=len(s.mesh.keys),
               dtype="f64" if precision == "fp64" else "f32")
    s.close()
    return out


def cpu_adaptive_baseline(budget_s=8.0, nx=128, ny=16):
    """The NumPy oracle's adaptive cfl + advance on the same workload at
    nx x ny roots (left half at level 2), single thread."""
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    from oracle import adaptive as oa
    from torch_amd import torchAdaptiveGeometry as tag
    from torch_amd.torchAdaptive import remesh
    from torch_amd.torchGeometry import normalize_faces

    prim, lengths, opts = adaptive_problem(nx, ny)
    c = dict(dimension=2, counts_xyz=[nx, ny], lengths_xyz=lengths, boundaries_xyz=opts["boundaries"],
             origins_xyz=[0.0, 0.0], geometry="cartesian", gamma=1.4, cfl=opts["cfl"], reactive=False,
             kinetic_model="arrhenius")
    c["face_boundaries"] = normalize_faces(c, None)
    c["amr"] = tag.settings(max_level=2, max_cells=2_000_000)
    mesh = tag.LeafMesh(c)
    state = prim.reshape(4, -1).copy()
    state[-1] = prim[-1].ravel() / 0.4
    for _ in range(2):
        faces = oa.face_values(mesh, state, [], c, False)[:4]
        mesh, state, _, _, _ = remesh(mesh, state, [], c, faces, mesh.centers[:, 0] < 0.5, False)
    t0 = time.perf_counter()
    steps = 0
    while steps < 2 or time.perf_counter() - t0 < budget_s:
        dt = oa.cfl(mesh, state, [], c)
        state = oa.advance(mesh, state, [], dt, c)[0]
        steps += 1
    wall = time.perf_counter() - t0
    return {"value": round(len(mesh.keys) * steps / wall, 1), "unit": "leaf-updates/s", "cores": 1, "kind": "port",
            "sample": f"2D Sod on {nx}x{ny} roots, left half at level 2 ({len(mesh.keys)} leaves), {steps} "
                      f"cfl+advance steps of the NumPy oracle, {wall:.1f} s"}


def run_legacy(n, t_end):
    """The legacy 1D driver with its full time history (every level
    downloaded as its step ends): Sod on n cells, VanLeer, CFL 0.5."""
    from torch_amd import run_simulation

    t0 = time.perf_counter()
    states, times = run_simulation(1.0, n, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, c_cfl=0.5, t_corte=t_end)
    wall = time.perf_counter() - t0
    steps = states.shape[2] - 1
    return dict(value=round(n * steps / wall, 1), ms_per_step=round(1e3 * wall / steps, 4), steps=steps,
                history_mb=round(states.nbytes / 1e6, 1), dtype="f64")


def cpu_legacy_baseline(n, budget_s=6.0):
    """The NumPy oracle's legacy driver on the same problem, shortened to a
    bounded end time, single thread."""
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    from oracle import legacy

    t_end, wall, steps = 0.002, 0.0, 0
    while wall < budget_s / 4 and t_end < 0.2:
        t0 = time.perf_counter()
        states, _ = legacy.run_simulation(1.0, n, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, c_cfl=0.5, t_corte=t_end)
        wall = time.perf_counter() - t0
        steps = states.shape[2] - 1
        t_end *= 4
    return {"value": round(n * steps / wall, 1), "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"Sod {n} cells to t={t_end / 4:g}, {steps} steps of the NumPy oracle's driver, {wall:.1f} s"}


def run_checkpoint(n):
    """Save and load of a device session's checkpoint: the 3D Sod tube n^3
    in FP64 after two steps, written to a temporary directory, read back
    (with every check of the reader) and compared."""
    import shutil
    import tempfile

    from torch_amd import Simulation

    sim = Simulation(sod_primitive(n), [1.0, 1.0, 1.0], precision="fp64", **OPTIONS)
    sim.step()
    sim.step()
    folder = tempfile.mkdtemp(prefix="torch_ckpt_")
    try:
        path = os.path.join(folder, "sod.npz")
        t0 = time.perf_counter()
        sim.save(path)
        save_s = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        back = Simulation.load(path)
        load_s = time.perf_counter() - t0
        same = bool(np.array_equal(back.state, sim.state)) and back.time == sim.time
        back.close()
    finally:
        shutil.rmtree(folder, ignore_errors=True)
        sim.close()
    if not same:
        raise RuntimeError("checkpoint round trip changed the session")
    return dict(bytes=size, save_s=round(save_s, 3), load_s=round(load_s, 3),
                save_gb_s=round(size / save_s / 1e9, 3), load_gb_s=round(size / load_s / 1e9, 3))


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


LEGS = ("fp64", "fp32", "batched", "error", "uneven", "served", "reactive", "metric", "adaptive", "legacy",
        "checkpoint", "cpu")


def _pair(text):
    a, b = (int(v) for v in text.split(","))
    return a, b


def _tile(text):
    if not text:
        return None
    if text == "tune":
        return text
    parts = tuple(int(v) for v in text.split(","))
    if len(parts) != 3:
        raise argparse.ArgumentTypeError("tile is BX,BY,ZCHUNK")
    return parts


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-slab", type=int, default=32,
                    help="z planes of the CPU baseline's slab of the n^3 tube (n: the whole grid)")
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma list of legs to run (profiling runs pick one): " + ",".join(LEGS))
    ap.add_argument("--tile", type=_tile, default=None,
                    help="BX,BY,ZCHUNK of the batched Sod (default: the library default; 'tune' runs the autotuner)")
    ap.add_argument("--error-steps", type=int, default=200, help="steps of the FP32-vs-FP64 error pass")
    ap.add_argument("--batch-count", type=int, default=256, help="instances of the uneven mix")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--batch-seed", type=int, default=508)
    ap.add_argument("--batch-tile", type=_tile, default=None, help="BX,BY,ZCHUNK of the mix (skips the tuner)")
    ap.add_argument("--serve-count", type=int, default=64, help="served ignition requests")
    ap.add_argument("--serve-seed", type=int, default=355)
    ap.add_argument("--serve-t-run", type=float, default=0.1, help="run length of each served solve")
    ap.add_argument("--serve-budget-gib", type=float, default=4.0, help="HBM budget of the served batch")
    ap.add_argument("--serve-round", type=int, default=4, help="steps between admission checks")
    ap.add_argument("--serve-closed-only", action="store_true", help="skip the open-loop (Poisson) run")
    ap.add_argument("--serve-policy", default="backfill", help="admission order: fifo, backfill or largest")
    ap.add_argument("--reactive-steps", type=int, default=20, help="graph replays of the fixed reactive batch")
    ap.add_argument("--reactive-tile", type=_tile, default=None,
                    help="BX,BY,ZCHUNK of the fixed reactive batch ('tune' runs the autotuner)")
    ap.add_argument("--metric-n", type=int, default=96, help="cells per axis of the metric/face workload")
    ap.add_argument("--metric-steps", type=int, default=20)
    ap.add_argument("--amr-roots", type=_pair, default=(512, 64), help="NX,NY roots of the adaptive workload")
    ap.add_argument("--amr-steps", type=int, default=20)
    ap.add_argument("--amr-roots-3d", type=lambda t: tuple(int(v) for v in t.split(",")), default=(64, 16, 16),
                    help="NX,NY,NZ roots of the 3D adaptive workload")
    ap.add_argument("--legacy-n", type=int, default=4096, help="cells of the legacy 1D history run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (counter calibration) and exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the workloads are "
              "single-GPU; how they would shard is described, not run).", file=sys.stderr)
        return 2
    legs = set(args.legs.split(","))
    if args.no_cpu:
        legs.discard("cpu")
    unknown = legs - set(LEGS)
    if unknown:
        ap.error(f"unknown legs {sorted(unknown)}")
    n = args.n
    say = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    r64 = r64b = r32s = r32 = err = uneven = cpu = None
    if "fp64" in legs:
        say(f"# 3D Sod {n}^3: FP64 plain, {args.warmup}+{args.steps} steps")
        r64 = run_precision(n, "fp64", args.steps, args.warmup, tile=SOD_TILE["fp64"])
        say(f"# fp64 plain: {r64['ms_per_step']:.3f} ms/step, kernel {r64['kernel_ms']:.3f} ms")
    if "fp32" in legs:
        r32s = run_precision(n, "fp32", args.steps, args.warmup, tile=SOD_TILE["fp32"])
        say(f"# fp32 session loop: {r32s['ms_per_step']:.3f} ms/step, kernel {r32s['kernel_ms']:.3f} ms")
    if "batched" in legs:
        r32 = run_batched_sod(n, args.steps, args.warmup, tile=args.tile)
        say(f"# fp32 batched: {r32['ms_per_step']:.3f} ms/step, kernel {r32['kernel_ms']:.3f} ms in steps "
            f"({r32['pass_ms']:.3f} ms relaunched alone), tile {r32['tile']}")
        r64b = run_batched_sod(n, args.steps, args.warmup, precision="fp64")
        say(f"# fp64 batched: {r64b['ms_per_step']:.3f} ms/step, kernel {r64b['kernel_ms']:.3f} ms in steps "
            f"({r64b['pass_ms']:.3f} ms relaunched alone), tile {r64b['tile']}")
    full = None
    if args.steps < WORKLOAD_STEPS and ("batched" in legs or "fp64" in legs):
        # The workload's own length, beside the K-step window (early steps
        # cross more uniform flow, where the step's shortcuts apply).
        full = {"steps": WORKLOAD_STEPS, "warmup": 0}
        if "batched" in legs:
            f32 = run_batched_sod(n, WORKLOAD_STEPS, 0, tile=args.tile)
            t32 = f32["tile"] if f32["tile"] else SOD_TILE["fp32"]
            full["fp32_batched"] = {
                "value": round(f32["value"], 1), "ms_per_step": round(f32["ms_per_step"], 4),
                "pass_ms_start": round(f32["pass_ms"], 4), "kernel_ms_end": round(f32["kernel_ms"], 4),
                "roofline_end": roofline(n, 40, f32["kernel_ms"], "f32", t32)}
        if "fp64" in legs:
            f64 = run_precision(n, "fp64", WORKLOAD_STEPS, 0)
            full["fp64_plain"] = {"value": round(f64["value"], 1), "ms_per_step": round(f64["ms_per_step"], 4),
                                  "kernel_ms": round(f64["kernel_ms"], 4)}
        say(f"# full {WORKLOAD_STEPS}-step run: {full}")
    if "error" in legs:
        err = fp32_error(n, args.error_steps)
        say(f"# fp32 vs fp64 after {args.error_steps} steps: L1 {err['l1_rel']} Linf {err['linf_rel']} "
            f"drift fp32 {err['drift_fp32']}")
    if "uneven" in legs:
        res, uerr, ucells = run_uneven(args.batch_seed, args.batch_count, args.batch_steps, tile=args.batch_tile)
        uneven = {
            "workload": f"{args.batch_count} seeded (rng {args.batch_seed}) two-state problems, dimension "
                        "cycling 1,2,3: 1D N in 256..8192, 2D 64^2..512^2, 3D 32^3..160^3, random limiter, "
                        "gamma 1.4/1.67, open boundaries, CFL 0.4, one graph-replayed batch",
            "instances": args.batch_count, "cells": ucells, "steps": args.batch_steps,
            "unit": "cell-updates/s",
            "fp32": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res["fp32"].items()},
            "fp64": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res["fp64"].items()},
            "fp32_vs_fp64": uerr,
        }
        if r64:
            uneven["fp32_over_fp64_plain_sod"] = round(res["fp32"]["value"] / r64["value"], 3)
    served = None
    if "served" in legs:
        served = run_served(args.serve_seed, args.serve_count, args.serve_t_run, args.serve_budget_gib,
                            args.serve_round, open_loop=not args.serve_closed_only, policy=args.serve_policy)
        served["workload"] = (
            f"{args.serve_count} seeded (rng {args.serve_seed}) 3D ignition solves, n^3 with n in 32..160, "
            f"Arrhenius (ZND for ~10%), Ea/RT0 in [20,30], Q/RT0 in [10,50], Gaussian hot spot T/T0 in "
            f"[1.5,3], run length {args.serve_t_run} (induction ~0.1 of it), VanLeer, CFL 0.4, open "
            f"boundaries; FP32 batch, HBM budget {args.serve_budget_gib} GiB, {args.serve_round} steps/round")
        served["fp32_vs_fp64"] = served_fp32_error(args.serve_seed, args.serve_count, args.serve_t_run)
    if "reactive" in legs:
        served = served or {}
        served["reactive_batch"] = run_reactive(args.serve_seed, args.serve_count, args.serve_t_run,
                                                args.reactive_steps, 3, tile=args.reactive_tile)
        served["reactive_batch"]["workload"] = (
            f"the {args.serve_count} served ignition requests admitted together into one FP32 batch (fixed "
            f"membership), {args.reactive_steps} graph replays after 3")
        say(f"# reactive batch: {served['reactive_batch']['ms_per_replay']} ms/replay, "
            f"{served['reactive_batch']['value']:.4g} cell-updates/s, {served['reactive_batch']['retries']} halvings")
    metric = None
    if "metric" in legs:
        mn = args.metric_n
        metric = {"workload": f"metric/face scheme: cylindrical-sector blast (annular r,theta,z grid) {mn}^3, "
                              "walls, pressure outlet with backflow, periodic z, Superbee, CFL 0.4, SSPRK2 "
                              "MUSCL-Rusanov, full Simulation.step (_metric_step)",
                  "cells": mn ** 3, "steps": args.metric_steps, "unit": "cell-updates/s"}
        for precision in ("fp64", "fp32"):
            metric[precision] = run_metric(mn, precision, args.metric_steps, 2)
            say(f"# metric {precision}: {metric[precision]['ms_per_step']} ms/step")
        if "cpu" in legs:
            metric["cpu_baseline"] = cpu_metric_baseline()
    adaptive = None
    if "adaptive" in legs:
        ax, ay = args.amr_roots
        adaptive = {"workload": f"adaptive (dyadic AMR) leaves: 2D Sod tube on {ax}x{ay} roots, left half "
                                "refined to level 2, open x / reflective y, CFL 0.4, second-order bounded "
                                "least-squares reconstruction, Rusanov faces, SSPRK2; full AdaptiveSimulation.step "
                                "without regrids (regrid_s: one host remesh of the final state)",
                    "steps": args.amr_steps, "unit": "leaf-updates/s"}
        for precision in ("fp64", "fp32"):
            adaptive[precision] = run_adaptive(precision, args.amr_steps, 2, ax, ay)
            say(f"# adaptive {precision}: {adaptive[precision]['ms_per_step']} ms/step, "
                f"{adaptive[precision]['leaves']} leaves")
        bx, by, bz = args.amr_roots_3d
        adaptive["workload_3d"] = (f"the same on {bx}x{by}x{bz} roots in 3D (left half at level 2, reflective "
                                   "y and z walls)")
        for precision in ("fp64", "fp32"):
            adaptive[precision + "_3d"] = run_adaptive(precision, args.amr_steps, 2, bx, by, bz)
            say(f"# adaptive 3D {precision}: {adaptive[precision + '_3d']['ms_per_step']} ms/step, "
                f"{adaptive[precision + '_3d']['leaves']} leaves")
        if "cpu" in legs:
            adaptive["cpu_baseline"] = cpu_adaptive_baseline()
    legacy = None
    if "legacy" in legs:
        ln = args.legacy_n
        legacy = {"workload": f"legacy 1D driver run_simulation with its full time history: Sod on {ln} cells, "
                              "VanLeer, CFL 0.5, open ends, t_corte 0.2, every level downloaded",
                  "unit": "cell-updates/s", **run_legacy(ln, 0.2)}
        say(f"# legacy 1D: {legacy['ms_per_step']} ms/step, {legacy['steps']} levels")
        if "cpu" in legs:
            legacy["cpu_baseline"] = cpu_legacy_baseline(ln)
    checkpoint = None
    if "checkpoint" in legs:
        checkpoint = {"workload": f"checkpoint save + load (every reader check) of the 3D Sod tube {n}^3 FP64 "
                                  "session, NumPy .npz container", **run_checkpoint(n)}
        say(f"# checkpoint: save {checkpoint['save_gb_s']} GB/s, load {checkpoint['load_gb_s']} GB/s")
    if "cpu" in legs:
        cpu = cpu_baseline(n, min(args.cpu_slab, n))
    head = r32 or r32s
    line = {
        "metric": METRIC, "value": round(head["value"], 1) if head else None, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4) if head else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D Cartesian Sod tube {n}^3, MINBEE, CFL 0.8, open x / reflective y,z, "
                               "gamma 1.4, full Simulation.step; FP32 form (f32 bulk, f64 reductions) as a "
                               "graph-replayed batch of one",
                   "cells": n ** 3, "tile": list(r32["tile"]) if r32 else None, "parallelism": "1 GPU"},
        "roofline": roofline(n, 40, head["kernel_ms"], "f32", r32["tile"] if r32 else SOD_TILE["fp32"]) if head else None,
        "fp64_plain": None if not r64 else {
            "value": round(r64["value"], 1), "ms_per_step": round(r64["ms_per_step"], 4), "dtype": "f64",
            "kernel_ms": round(r64["kernel_ms"], 4), "roofline": roofline(n, 80, r64["kernel_ms"], "f64", SOD_TILE["fp64"]),
            "fallbacks": r64["fallbacks"]},
        "fp64_batched": None if not r64b else {
            "value": round(r64b["value"], 1), "ms_per_step": round(r64b["ms_per_step"], 4), "dtype": "f64",
            "kernel_ms": round(r64b["kernel_ms"], 4), "pass_ms": round(r64b["pass_ms"], 4),
            "tile": list(r64b["tile"]),
            "fallbacks": r64b["fallbacks"]},
        "fp32_session_loop": None if not r32s else {
            "value": round(r32s["value"], 1), "ms_per_step": round(r32s["ms_per_step"], 4),
            "kernel_ms": round(r32s["kernel_ms"], 4)},
        "fp32_kernel_ms": round(head["kernel_ms"], 4) if head else None,
        "fp32_pass_ms": round(head["pass_ms"], 4) if head and "pass_ms" in head else None,
        "fp32_over_fp64": round(head["value"] / r64["value"], 3) if head and r64 else None,
        "fp32_l1_rel": err["l1_rel"] if err else None,
        "fp32_error": err,
        "full_length": full,
        "batched": uneven,
        "served": served,
        "metric_face": metric,
        "adaptive": adaptive,
        "legacy": legacy,
        "checkpoint": checkpoint,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
