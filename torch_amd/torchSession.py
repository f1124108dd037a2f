"""Step-boundary simulation sessions on the GPU.

Mirrors the Cartesian path of python/cfdSolver/torchSession.py:19-694:
the ``Simulation`` constructor and its validation (:24-158, 190-271), the
``state`` layout (dim+2, ...reversed axes) float64 with ``lam``/``beta``,
``time``/``steps`` and the balance integrals (:285-331), ``step`` (:374-522),
``run`` with checkpoints (:578-633), ``save``/``load`` (:635-651, with the
schema-2 checkpoint layout of torch_amd.torchIO), ``preset`` (:653-694) and
the worker backend (:160-183). The state lives on the device between steps; ``step`` runs the
whole Strang-split step in libtorchhip (one host synchronisation per step,
two when reactive) and commits only on success, so a failed step leaves the
session unchanged.

Extensions beyond the reference's keywords: ``precision`` ("fp64" plain
method, "fp32" FP32 form), ``device``, ``checked`` (False admits CFL > 0.5
for workloads run outside the reference's validated range), ``constants``
(kinetics constants; the reference's 3D wrapper always uses its defaults)
and ``tile`` (3D kernel tile, normally chosen by the autotuner).
"""

import copy
import uuid

import numpy as np

from . import _abi
from . import torchGeometry as _geometry
from .presets import preset as _preset
from .torchExecution import DeviceExecution, validate_workers
from .torchSolver3D import LIMITERS, config as _config

try:
    import xxhash

    def _hasher():
        return xxhash.xxh3_128()
except ImportError:  # pragma: no cover - the image ships xxhash
    import hashlib

    def _hasher():
        return hashlib.blake2b(digest_size=16)


def _digest(arrays):
    """Content digest of the host copies (state, lam, beta; None skipped)."""
    h = _hasher()
    for a in arrays:
        if a is not None:
            h.update(memoryview(np.ascontiguousarray(a)).cast("B"))
        h.update(b"|")
    return h.digest()


_OPTIONS = {
    "controller": _abi.CONTROLLERS,
    "variant": _abi.VARIANTS,
    "wavespeed": _abi.WAVESPEEDS,
}


def _validate_grid(length, cells, cfl, end_time, checked):
    """torchSolver.py:1074-1088 (with checked=False the CFL cap is lifted)."""
    cfl_ok = 0 < cfl <= 0.5 if checked else cfl > 0
    if (not isinstance(cells, (int, np.integer)) or not np.isfinite([length, cfl, end_time]).all()
            or length <= 0 or cells < 2 or not cfl_ok or end_time < 0):
        raise ValueError("Require length > 0, cells >= 2, 0 < CFL <= 0.5 and final time >= 0")


def _physical(state):
    """torchSolver.py:1038-1046."""
    return bool(np.isfinite(state).all() and np.all(state[0] > 1e-12)
                and np.all(state[-1] - 0.5 * np.sum(state[1:-1] ** 2, axis=0) / state[0] > 1e-12))


class Simulation:
    """Cartesian simulation, resumable between steps, resident on an MI355X."""

    def __init__(self, primitive, lengths, *, t_end, cfl=0.4, gamma=1.4, boundaries=None, limiter="VanLeer",
                 reactive=False, kinetic_model="arrhenius", lambda_init=None, beta_init=None,
                 inlet_primitive=None, controller="Sound Speed Controlled", variant="Variante 1 - Original",
                 wavespeed="Pressure Based", max_steps=100000, geometry="cartesian", origins=None,
                 face_boundaries=None, workers=1, precision="fp64", device=0, checked=True, constants=None,
                 tile=None):
        primitive = np.array(primitive, dtype=float, copy=True)
        dim = primitive.ndim - 1
        if dim not in (1, 2, 3) or primitive.shape[0] != dim + 2:
            raise ValueError("Require primitives with shape (dimension+2, ... reversed spatial axes)")
        lengths = list(np.atleast_1d(lengths))
        counts = list(primitive.shape[1:][::-1])
        if len(lengths) != dim:
            raise ValueError("Lengths must be in x,y,z order, one per active axis")
        boundaries = list(boundaries) if boundaries is not None else ["Open"] * dim
        self.config = dict(
            dimension=dim, counts_xyz=counts, lengths_xyz=lengths, boundaries_xyz=boundaries, t_end=t_end,
            cfl=cfl, gamma=gamma, limiter=limiter, reactive=bool(reactive), kinetic_model=kinetic_model,
            inlet_primitive=list([1.0, 0.0, 100.0, 101325.0] if inlet_primitive is None else inlet_primitive),
            controller=controller, variant=variant, wavespeed=wavespeed, max_steps=max_steps,
            geometry=geometry, origins_xyz=list(np.atleast_1d(origins)) if origins is not None else [0.0] * dim,
        )
        if geometry != "cartesian" or face_boundaries is not None:
            self.config["face_boundaries"] = _geometry.session_faces(self.config, face_boundaries)
        self.precision = str(precision).lower()
        self.checked = bool(checked)
        self.constants = dict(constants or {})
        self.tile = tuple(tile) if tile is not None else (0, 0, 0)
        self._validate_config()
        if not np.isfinite(primitive).all() or np.any(primitive[[0, -1]] <= 0):
            raise ValueError("Initial density and pressure must be positive and primitives finite")
        lam = self._fraction(lambda_init, "lambda") if reactive else None
        beta = self._fraction(beta_init, "beta") if reactive and kinetic_model == "znd" else None
        if not reactive and lambda_init is not None or beta_init is not None and beta is None:
            raise ValueError("Composition requires reactive chemistry; beta requires ZND")
        g = (4.5 - lam) / (3.5 - lam) if reactive else gamma
        state = primitive.copy()
        state[1:-1] *= primitive[0]
        state[-1] = primitive[-1] / (g - 1) + 0.5 * primitive[0] * np.sum(primitive[1:-1] ** 2, axis=0)
        if not _physical(state):
            raise ValueError("Nonphysical initial conservative state")
        self.time, self.steps = 0.0, 0
        self.run_id = str(uuid.uuid4())
        self.fallbacks = 0
        self._host = (state, lam, beta)
        self.initial_integral = self._integral_of(state)
        self.boundary_integral = np.zeros(dim + 2)
        self.source_integral = np.zeros(dim + 2)
        self._execution = DeviceExecution(validate_workers(workers), device=device, precision=self.precision)
        self._device = None
        self._upload(state, lam, beta)

    # -- configuration -----------------------------------------------------
    def _validate_config(self):
        c = self.config
        for length, n in zip(c["lengths_xyz"], c["counts_xyz"]):
            _validate_grid(length, n, c["cfl"], c["t_end"], self.checked)
        _geometry.validate_geometry(c)
        if self.uses_metric_kernel:
            c["face_boundaries"] = _geometry.normalize_faces(c, c.get("face_boundaries"))
            if (c["controller"], c["variant"], c["wavespeed"]) != (_abi.CONTROLLERS[0], _abi.VARIANTS[0],
                                                                   _abi.WAVESPEEDS[0]):
                raise ValueError("Metric/face sessions use MUSCL-Rusanov/SSPRK2; legacy HLLC switches are unsupported")
        if not np.isfinite(c["gamma"]) or c["gamma"] <= 1:
            raise ValueError("gamma must be finite and greater than one")
        if c["limiter"] not in LIMITERS or c["kinetic_model"] not in ("arrhenius", "znd"):
            raise ValueError("Unknown limiter or chemistry model")
        if (not isinstance(c["max_steps"], (int, np.integer)) or isinstance(c["max_steps"], bool)
                or c["max_steps"] < 1):
            raise ValueError("Invalid maximum steps")
        for axis, boundary in enumerate(c["boundaries_xyz"]):
            if boundary not in ("Open", "Reflective", "Periodic") and not (
                    c["dimension"] == 2 and axis == 1 and boundary == "Inflow"):
                raise ValueError("Unsupported session boundary")
        for key, allowed in _OPTIONS.items():
            if c[key] not in allowed:
                raise ValueError(f"Unknown solver option: {c[key]}")
        if c["dimension"] == 2 and "Inflow" in c["boundaries_xyz"]:
            inlet = np.asarray(c["inlet_primitive"])
            if inlet.shape != (4,) or not np.isfinite(inlet).all() or inlet[0] <= 0 or inlet[3] <= 0 or inlet[2] < 0:
                raise ValueError("Invalid inlet primitives")
        _abi.precision_code(self.precision)

    def _fraction(self, values, name):
        shape = tuple(self.config["counts_xyz"][::-1])
        out = np.ones(shape) if values is None else np.array(values, dtype=float, copy=True)
        if out.shape != shape or not np.isfinite(out).all() or np.any((out < 0) | (out > 1)):
            raise ValueError(f"{name} must have the spatial shape and values in [0,1]")
        return out

    def _make_config(self):
        c = self.config
        dim = c["dimension"]
        return _config(
            c["counts_xyz"] + [1] * (3 - dim), None, c["lengths_xyz"] + [1.0] * (3 - dim), dim=dim,
            gamma=c["gamma"], limiter=c["limiter"], bcs=list(c["boundaries_xyz"]) + ["Open"] * (3 - dim),
            reactive=c["reactive"], model=c["kinetic_model"], gamma_mode=1 if c["reactive"] else 0,
            precision=self.precision, cfl=c["cfl"], inlet=c["inlet_primitive"], constants=self.constants,
            controller=_abi.CONTROLLERS.index(c["controller"]), variant=_abi.VARIANTS.index(c["variant"]),
            wavespeed=_abi.WAVESPEEDS.index(c["wavespeed"]), tile=self.tile)

    def _upload(self, state, lam, beta):
        if self._device is not None:
            self._device.close()
        self._device = _abi.DeviceState(self._execution.context, self._make_config())
        if self.uses_metric_kernel:
            c = self.config
            self._device.set_metric(c.get("geometry", "cartesian"), _geometry.factor_table(c), c["face_boundaries"])
        self._device.upload(state, lam, beta)
        self._host = (state, lam, beta)
        self._remember()

    def _fetch(self):
        if self._host is None:
            c = self.config
            znd = c["reactive"] and c["kinetic_model"] == "znd"
            self._host = self._device.download(c["reactive"], znd)
            self._remember()
        return self._host

    # The reference's state, lam and beta are plain arrays that callers edit
    # in place (testAdaptive.py:127-128, 288). Here they are host copies of
    # the device state, so every device operation first uploads copies that
    # changed since they were last uploaded or downloaded.
    def _remember(self):
        self._host_digest = None if self._host is None else _digest(self._host)

    def _push_edits(self):
        if self._host is not None and _digest(self._host) != getattr(self, "_host_digest", None):
            self._upload_edits(*self._host)
            self._remember()

    def _upload_edits(self, state, lam, beta):
        self._device.upload(state, lam, beta)

    # -- reference properties ------------------------------------------------
    @property
    def state(self):
        return self._fetch()[0]

    @state.setter
    def state(self, value):
        _, lam, beta = self._fetch()
        self._upload(np.array(value, dtype=float, copy=True), lam, beta)

    @property
    def lam(self):
        return self._fetch()[1]

    @property
    def beta(self):
        return self._fetch()[2]

    @property
    def workers(self):
        return self._execution.workers

    @property
    def execution(self):
        return self._execution

    def set_workers(self, workers):
        validate_workers(workers)
        state, lam, beta = self._fetch()
        old = self._execution
        self._device.close()
        self._device = None
        self._execution = DeviceExecution(workers, device=old.device, precision=old.precision)
        old.close()
        self._upload(state, lam, beta)
        return self

    def close(self):
        if self._device is not None:
            self._device.close()
            self._device = None
        if getattr(self, "_execution", None) is not None:
            self._execution.close()

    def to_adaptive(self, **options):
        """An adaptive (AMR) session continuing this one (torchSession.py:185-188)."""
        from .torchAdaptive import AdaptiveSimulation

        return AdaptiveSimulation.from_uniform(self, **options)

    @property
    def gamma(self):
        return (4.5 - self.lam) / (3.5 - self.lam) if self.config["reactive"] else self.config["gamma"]

    @property
    def spacing(self):
        return np.array(self.config["lengths_xyz"]) / self.config["counts_xyz"]

    def _integral_of(self, state):
        nf = self.config["dimension"] + 2
        if self.config.get("geometry", "cartesian") != "cartesian":
            return (state * self.metrics.volumes[None]).reshape(nf, -1).sum(axis=1)
        return state.reshape(nf, -1).sum(axis=1) * np.prod(self.spacing)

    @property
    def metrics(self):
        """Cell volumes and face areas (torchGeometry.Metrics, torchSession.py:297-301)."""
        if getattr(self, "_metrics", None) is None:
            self._metrics = _geometry.Metrics(self.config)
        return self._metrics

    @property
    def integral(self):
        return self._integral_of(self.state)

    @property
    def uses_metric_kernel(self):
        """torchSession.py:303-308."""
        return self.config.get("geometry", "cartesian") != "cartesian" or "face_boundaries" in self.config

    @property
    def primitive(self):
        s = self.state
        p = s.copy()
        p[1:-1] /= s[0]
        p[-1] = (self.gamma - 1) * (s[-1] - 0.5 * np.sum(s[1:-1] ** 2, axis=0) / s[0])
        return p

    @property
    def balance_residual(self):
        return self.integral - self.initial_integral + self.boundary_integral - self.source_integral

    @property
    def completed(self):
        return self.time == self.config["t_end"]

    def dt_limit(self):
        """Simulation._dt on the committed state (torchSession.py:333-360), or
        cfl_dt for metric sessions (torchFiniteVolume.py:66-98)."""
        self._push_edits()
        return self._device.metric_dt() if self.uses_metric_kernel else self._device.dt_limit()

    def device_totals(self):
        """Conserved totals reduced on the device (double)."""
        self._push_edits()
        return self._device.totals()

    # -- stepping ------------------------------------------------------------
    def step(self):
        """Advance one complete split step; failed steps leave session data intact."""
        if self.completed:
            return
        c = self.config
        if self.steps >= c["max_steps"]:
            raise RuntimeError("Maximum steps reached before requested final time")
        self._push_edits()
        if self.uses_metric_kernel:
            info = self._device.metric_step(self.time, c["t_end"])
        else:
            info = self._device.step(self.time, c["t_end"])
        nf = c["dimension"] + 2
        self._host = None
        self.boundary_integral += np.array(info.outward[:nf])
        self.source_integral += np.array(info.source[:nf])
        self.time = min(self.time + info.dt, c["t_end"])
        self.steps += 1
        self.fallbacks += int(info.fallback)
        self.last_dt = info.dt
        return info

    def run(self, *, step_limit=None, checkpoint=None, checkpoint_every=0, output_prefix=None, output_every=1,
            formats=("vtk", "xdmf")):
        """Run to the end time, or pause after ``step_limit`` more steps (torchSession.py:578-633).

        A checkpoint is written every ``checkpoint_every`` steps and at the
        end; checkpoints never shorten a step, so a restart takes the same
        step sequence. VTK/XDMF output is not part of the device backend.
        """
        for value in (step_limit, checkpoint_every, output_every):
            if value is not None and (not isinstance(value, int) or value < 0):
                raise ValueError("Step intervals must be nonnegative integers")
        if checkpoint_every and not checkpoint:
            raise ValueError("checkpoint_every requires a checkpoint destination")
        if output_prefix:
            raise NotImplementedError("VTK/XDMF output is not part of the device backend")
        from .torchIO import write_checkpoint

        count = 0
        while not self.completed and (step_limit is None or count < step_limit):
            self.step()
            count += 1
            if checkpoint and checkpoint_every and self.steps % checkpoint_every == 0:
                write_checkpoint(checkpoint, self)
        if checkpoint:
            write_checkpoint(checkpoint, self)
        return self

    def save(self, path):
        """Write a checkpoint of the committed state (torchSession.py:635-638)."""
        from .torchIO import write_checkpoint

        write_checkpoint(path, self)

    @classmethod
    def load(cls, path, *, t_end=None, precision=None, device=0, workers=1):
        """Resume a checkpoint, optionally with a later end time (torchSession.py:640-651)."""
        from .torchIO import read_checkpoint, record, load_content

        if t_end is None:
            return read_checkpoint(path, precision=precision, device=device, workers=workers)
        sim = record(load_content(path))
        if not np.isfinite(t_end) or t_end < sim.time:
            raise ValueError("New end time must be finite and not precede checkpoint")
        sim.config["t_end"] = t_end
        if precision is not None:
            sim.precision = str(precision).lower()
        sim._validate_config()
        return sim._attach(workers, device)

    def _attach(self, workers, device):
        """Give a host-side session shell (a validated checkpoint) its device state."""
        state, lam, beta = self._host
        self._execution = DeviceExecution(validate_workers(workers), device=device, precision=self.precision)
        self._device = None
        self._upload(state, lam, beta)
        return self

    @classmethod
    def preset(cls, name, counts, *, lengths=None, **options):
        counts = list(np.atleast_1d(counts))
        lengths = lengths if lengths is not None else [1.0] * len(counts)
        prim, defaults = _preset(name, counts, **options)
        return cls(prim, lengths, **defaults)

    def __deepcopy__(self, memo):
        state, lam, beta = (None if a is None else a.copy() for a in self._fetch())
        new = object.__new__(Simulation)
        for key, value in self.__dict__.items():
            if key not in ("_execution", "_device", "_host"):
                setattr(new, key, copy.deepcopy(value, memo))
        new._execution = copy.deepcopy(self._execution, memo)
        new._device = None
        new._upload(state, lam, beta)
        return new

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
