"""Batches of sessions advanced together on one device.

The reference steps every ``Simulation`` on its own
(python/cfdSolver/torchSession.py:374-522); a set of problems is a Python
loop over sessions. ``Batch`` keeps many sessions (inert of any dimension,
reactive in 3D) resident on the device and advances all of them with one
graph replay per step. Each instance keeps the single-session arithmetic
and bookkeeping (its own dt, time, step count, boundary and source
integrals, reactive dt halvings), so a session stepped in a batch ends where
it ends stepped alone.

Typical use with existing sessions::

    batch = Batch(precision="fp32")
    ids = [batch.add(sim) for sim in sessions]
    batch.run(100)
    for sim, i in zip(sessions, ids):
        batch.commit(i, sim)        # state, time, steps, integrals back
"""

import numpy as np

from . import _abi
from .torchExecution import DeviceExecution
from .torchSolver3D import config as _config


class Batch:
    """Sessions of any dimension and size stepped together (reactive: 3D).

    ``hbm_budget`` caps the device bytes the batch may hold (0: no cap);
    ``add`` returns None when an instance does not fit, which is the
    admission test of a serving loop.
    """

    def __init__(self, precision="fp32", device=0, hbm_budget=0):
        self.precision = str(precision).lower()
        self._execution = DeviceExecution(1, device=device, precision=self.precision)
        self._batch = _abi.DeviceBatch(self._execution.context, _abi.precision_code(self.precision), hbm_budget)
        self._meta = {}

    # -- admission -----------------------------------------------------------
    def add(self, sim, t_end=None, max_steps=None):
        """Admit a Simulation's committed state; returns the instance id (or None).

        The instance runs from ``sim.time`` to ``t_end`` (default: the
        session's) with at most ``max_steps`` further steps (default: what the
        session has left).
        """
        c = sim.config
        if sim.uses_metric_kernel:
            raise ValueError("batches hold Cartesian MUSCL-Hancock/HLLC sessions; metric/face sessions step alone")
        if c["reactive"] and c["dimension"] != 3:
            raise ValueError("batches hold inert sessions and reactive 3D sessions")
        cfg = sim._make_config()
        cfg.precision = _abi.precision_code(self.precision)
        t_end = c["t_end"] if t_end is None else float(t_end)
        max_steps = c["max_steps"] - sim.steps if max_steps is None else int(max_steps)
        ident = self._batch.add(cfg, sim.state, sim.time, t_end, max_steps, lam=sim.lam, beta=sim.beta)
        if ident is not None:
            # The device adds each step's terms to the session's running totals,
            # in the session's order.
            self._batch.seed(ident, sim.boundary_integral, sim.source_integral)
            self._meta[ident] = (c["dimension"] + 2, sim.steps, sim.fallbacks)
        return ident

    def add_problem(self, primitive, lengths, *, t_end, cfl=0.4, gamma=1.4, boundaries=None, limiter="VanLeer",
                    max_steps=100000, time=0.0, inlet_primitive=None, controller="Sound Speed Controlled",
                    variant="Variante 1 - Original", wavespeed="Pressure Based"):
        """Admit an inert problem from primitives without building a session.

        Takes the arguments of ``Simulation`` (torchSession.py:24-158) and
        converts the primitives to the conservative state the same way.
        """
        primitive = np.asarray(primitive, dtype=float)
        dim = primitive.ndim - 1
        if dim not in (1, 2, 3) or primitive.shape[0] != dim + 2:
            raise ValueError("Require primitives with shape (dimension+2, ... reversed spatial axes)")
        counts = list(primitive.shape[1:][::-1])
        lengths = list(np.atleast_1d(lengths))
        bcs = list(boundaries) if boundaries is not None else ["Open"] * dim
        cfg = _config(counts + [1] * (3 - dim), None, lengths + [1.0] * (3 - dim), dim=dim, gamma=gamma,
                      limiter=limiter, bcs=bcs + ["Open"] * (3 - dim), precision=self.precision, cfl=cfl,
                      inlet=inlet_primitive, controller=_abi.CONTROLLERS.index(controller),
                      variant=_abi.VARIANTS.index(variant), wavespeed=_abi.WAVESPEEDS.index(wavespeed))
        state = primitive.copy()
        state[1:-1] *= primitive[0]
        state[-1] = primitive[-1] / (gamma - 1) + 0.5 * primitive[0] * np.sum(primitive[1:-1] ** 2, axis=0)
        ident = self._batch.add(cfg, state, time, t_end, max_steps)
        if ident is not None:
            self._meta[ident] = (dim + 2, 0, 0)
        return ident

    def footprint(self, sim):
        cfg = sim._make_config()
        return _abi.DeviceBatch.footprint(cfg, _abi.precision_code(self.precision))

    def remove(self, ident):
        self._batch.remove(ident)
        self._meta.pop(ident, None)

    # -- stepping ------------------------------------------------------------
    def run(self, steps):
        """Advance every active instance by up to ``steps`` steps."""
        self._batch.run(steps)
        return self

    def tune(self, reps=3):
        """Pick the 3D tile shape and z chunk by timing the current instances."""
        return self._batch.tune(reps)

    def set_tiles(self, block_x, block_y, zchunk):
        self._batch.set_tiles(block_x, block_y, zchunk)

    def profile(self, enable=True):
        """Enable replay timing; returns (ms, steps) since the last call."""
        return self._batch.profile(enable)

    def time_pass(self, reps=5):
        """Average ms of one 3D second-order work-list launch (nothing advances)."""
        return self._batch.time_pass(reps)

    def time_step_kernel(self, steps=10):
        """Advance `steps` steps and return the mean ms of the batch's one 3D
        work-list launch within them (HIP events inside the step graph)."""
        return self._batch.time_step_kernel(steps)

    # -- results -------------------------------------------------------------
    def info(self, ident):
        """Progress of an instance. Its boundary and source integrals are the
        running totals: an instance admitted from a session starts from the
        session's totals, one admitted from primitives from zero."""
        p = self._batch.info(ident)
        nf = self._meta[ident][0]
        return dict(time=p.time, dt=p.dt, steps=p.steps, fallbacks=p.fallbacks, status=p.status,
                    active=bool(p.active), completed=bool(p.completed), cells=p.cells, retries=p.retries,
                    boundary_integral=np.array(p.outward[:nf]), source_integral=np.array(p.source[:nf]),
                    wave_max=np.array(p.wave_max[:]))

    def state(self, ident):
        """Committed conservative state of an instance."""
        return self._batch.download(ident)[0]

    def fractions(self, ident):
        """(lambda, beta) of a reactive instance (beta None unless ZND)."""
        _, lam, beta = self._batch.download(ident)
        return lam, beta

    def error(self, ident):
        """Raise the exception Simulation.step would have raised, if any."""
        status = self._batch.info(ident).status
        if status == _abi.NONPHYSICAL:
            raise RuntimeError("Hydrodynamic update produced a nonphysical state")
        if status == _abi.STALLED:
            raise RuntimeError("Integration stalled before requested final time")
        if status != _abi.OK:
            raise RuntimeError(f"batch instance stopped with status {status}")

    def commit(self, ident, sim):
        """Write an instance's progress back into its Simulation."""
        info = self.info(ident)
        _, start_steps, start_fallbacks = self._meta[ident]
        state, lam, beta = self._batch.download(ident)
        sim._upload(state, lam, beta)
        sim.time = info["time"]
        sim.steps = start_steps + info["steps"]
        sim.fallbacks = start_fallbacks + info["fallbacks"]
        sim.boundary_integral = info["boundary_integral"]
        sim.source_integral = info["source_integral"]
        return sim

    def close(self):
        if getattr(self, "_batch", None) is not None:
            self._batch.close()
            self._batch = None
        if getattr(self, "_execution", None) is not None:
            self._execution.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
