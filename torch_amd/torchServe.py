"""Serving loop: continuous batching of solve requests on one device.

The reference has no server; each ``Simulation`` runs to its end time on its
own (python/cfdSolver/torchSession.py:578-633). ``Server`` admits solve
requests as they arrive into one device ``Batch``, by HBM budget, keeps
every admitted state resident on the device, advances all of them a few
steps per round (one graph replay per step) and retires each request when
its run reaches the end time, so requests that arrive while others are on
the GPU overlap with them. Each request's result is the one the same
request gives alone (the batch keeps per-instance arithmetic).

A preparer thread builds each request's initial arrays when it arrives
(NumPy releases the interpreter lock), so host preparation overlaps the
device rounds; admission itself is the upload.
"""

import queue
import threading
import time

import numpy as np

from .torchBatch import Batch


class Request:
    """One solve: ``build()`` returns (primitive, options) for ``Simulation``-style
    arguments: lengths, t_end, and keyword options (reactive, kinetic_model,
    lambda_init, constants, boundaries, limiter, cfl, gamma)."""

    def __init__(self, ident, arrival, build, cells=None):
        self.ident = ident
        self.arrival = float(arrival)
        self.build = build
        self.cells = cells
        self.admitted = None
        self.completed = None
        self.steps = 0
        self.status = None
        self.result = None
        self.retries = 0
        self.skips = 0  # admission passes that found no budget for it

    @property
    def latency(self):
        return None if self.completed is None else self.completed - self.arrival


def _conservative(prim, opts):
    """Simulation's primitive-to-conservative conversion (torchSession.py:24-158)."""
    state = prim.copy()
    state[1:-1] *= prim[0]
    if opts.get("reactive"):
        lam = opts["lambda_init"]
        g = (4.5 - lam) / (3.5 - lam)
    else:
        g = opts.get("gamma", 1.4)
    state[-1] = prim[-1] / (g - 1) + 0.5 * prim[0] * np.sum(prim[1:-1] ** 2, axis=0)
    return state


class Server:
    """Continuous-batching solve server on one device.

    ``hbm_budget`` bounds the device bytes of admitted requests;
    ``round_steps`` is the number of steps between admission/retirement
    checks; ``keep_results`` downloads each finished state.

    ``policy`` orders admission among the requests waiting for budget:
    ``"fifo"`` admits in arrival order and stops at the first that does not
    fit; ``"backfill"`` admits, in arrival order, every waiting request that
    fits; ``"largest"`` does the same largest first (by cells). With
    backfilling, a request passed over in ``max_skips`` rounds reserves the
    budget: nothing behind it is admitted until it fits.
    """

    POLICIES = ("fifo", "backfill", "largest")

    def __init__(self, precision="fp32", device=0, hbm_budget=0, round_steps=4, max_steps=100000,
                 keep_results=False, policy="backfill", max_skips=16):
        if policy not in self.POLICIES:
            raise ValueError(f"policy must be one of {self.POLICIES}")
        self.policy = policy
        self.max_skips = int(max_skips)
        self.batch = Batch(precision=precision, device=device, hbm_budget=hbm_budget)
        self.round_steps = int(round_steps)
        self.max_steps = int(max_steps)
        self.keep_results = keep_results
        self.rounds = 0
        self.device_ms = 0.0
        self.timing = {}

    def payload(self, prim, opts):
        """Host-side admission data of a request: (config, state, lam, beta, t_end)."""
        from .torchSolver3D import config as _config

        dim = prim.ndim - 1
        counts = list(prim.shape[1:][::-1])
        lengths = list(opts.get("lengths", [1.0] * dim))
        bcs = list(opts.get("boundaries", ["Open"] * dim))
        reactive = bool(opts.get("reactive"))
        model = opts.get("kinetic_model", "arrhenius")
        cfg = _config(counts + [1] * (3 - dim), None, lengths + [1.0] * (3 - dim), dim=dim,
                      gamma=opts.get("gamma", 1.4), limiter=opts.get("limiter", "VanLeer"),
                      bcs=bcs + ["Open"] * (3 - dim), reactive=reactive, model=model,
                      gamma_mode=1 if reactive else 0, precision=self.batch.precision, cfl=opts.get("cfl", 0.4),
                      constants=opts.get("constants"))
        state = _conservative(prim, opts)
        lam = opts.get("lambda_init") if reactive else None
        beta = None
        if reactive and model == "znd":
            beta = opts.get("beta_init")
            beta = np.ones(prim.shape[1:]) if beta is None else beta
        return cfg, state, lam, beta, opts["t_end"]

    def _admit(self, req, payload):
        cfg, state, lam, beta, t_end = payload
        ident = self.batch._batch.add(cfg, state, 0.0, t_end, self.max_steps, lam=lam, beta=beta)
        if ident is not None:
            self.batch._meta[ident] = (cfg.dim + 2, 0, 0)
            req.cells = int(np.prod(state.shape[1:]))
        return ident

    def serve(self, requests, prebuild=False, clock=time.perf_counter):
        """Run the requests (arrival times in seconds from the start); returns them
        with admitted/completed times, steps and status filled in.

        With ``prebuild`` every request's initial arrays are built before the
        clock starts (they arrive with the request); otherwise the server
        builds them when the request arrives, in a preparer thread.
        """
        pending = sorted(requests, key=lambda r: r.arrival)
        built = {id(r): self.payload(*r.build()) for r in pending} if prebuild else {}
        # Host time of this call by phase (seconds): admissions, rounds right
        # after a membership change (they re-capture the step graph) and the
        # others, retirements (progress reads, downloads, removals).
        self.timing = dict(admit_s=0.0, changed_rounds=0, changed_s=0.0, steady_rounds=0, steady_s=0.0,
                           retire_s=0.0)
        ready = queue.Queue()
        start = clock()

        def prepare():
            for req in pending:
                wait = req.arrival - (clock() - start)
                if wait > 0:
                    time.sleep(wait)
                ready.put((req, built.pop(id(req)) if prebuild else self.payload(*req.build())))
            ready.put(None)

        worker = threading.Thread(target=prepare, daemon=True)
        worker.start()
        waiting, active = [], {}
        arrivals_done = False
        retired = False
        while not arrivals_done or waiting or active:
            # Collect prepared arrivals; wait for one only when idle.
            block = not active and not waiting and not arrivals_done
            while not arrivals_done:
                try:
                    item = ready.get(block=block)
                except queue.Empty:
                    break
                if item is None:
                    arrivals_done = True
                    break
                waiting.append(item)
                block = False
            # Admission while the budget allows, in the policy's order.
            if self.policy == "largest":
                waiting.sort(key=lambda item: (-(item[0].cells or 0), item[0].arrival))
            still, blocked = [], False
            t_admit = clock()
            changed = retired
            for req, payload in waiting:
                if blocked:
                    still.append((req, payload))
                    continue
                ident = self._admit(req, payload)
                if ident is None:
                    still.append((req, payload))
                    req.skips += 1
                    blocked = self.policy == "fifo" or req.skips > self.max_skips
                else:
                    req.admitted = clock() - start
                    active[ident] = req
                    changed = True
            self.timing["admit_s"] += clock() - t_admit
            waiting = still
            if waiting and not active:
                raise RuntimeError("a request does not fit the HBM budget on an empty device")
            if not active:
                continue
            t0 = clock()
            self.batch.run(self.round_steps)
            ran = clock() - t0
            self.device_ms += 1e3 * ran
            self.rounds += 1
            kind = "changed" if changed else "steady"
            self.timing[kind + "_rounds"] += 1
            self.timing[kind + "_s"] += ran
            now = clock() - start
            t_retire = clock()
            retired = False
            for ident in list(active):
                info = self.batch.info(ident)
                if info["active"]:
                    continue
                req = active.pop(ident)
                req.completed = now
                req.steps = info["steps"]
                req.status = info["status"]
                req.retries = info["retries"]
                if self.keep_results:
                    req.result = self.batch._batch.download(ident)
                self.batch.remove(ident)
                retired = True
            self.timing["retire_s"] += clock() - t_retire
        worker.join()
        return pending

    def close(self):
        self.batch.close()


def summary(requests):
    """p50/p99 time-to-solution and aggregate cell-updates/s of a served run."""
    done = [r for r in requests if r.completed is not None]
    lat = np.array([r.latency for r in done])
    first = min(r.arrival for r in requests)
    last = max(r.completed for r in done)
    work = sum(r.cells * r.steps for r in done)
    return dict(
        requests=len(requests), completed=len(done), failed=sum(1 for r in done if r.status),
        p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()),
        makespan_s=last - first, cell_updates_per_s=work / (last - first), cell_updates=work,
        steps=int(sum(r.steps for r in done)), retries=int(sum(r.retries for r in done)),
    )


def ignition_request(spec, arrival, limiter="VanLeer", cfl=0.4):
    """A Request for one served ignition solve (torch_amd.workloads.served_requests)."""
    from .workloads import ignition_problem

    def build():
        prim, lam = ignition_problem(spec)
        return prim, dict(reactive=True, kinetic_model=spec["model"], lambda_init=lam,
                          constants=spec["constants"], t_end=spec["t_run"], cfl=cfl, limiter=limiter)

    return Request(spec["index"], arrival, build, cells=spec["n"] ** 3)
