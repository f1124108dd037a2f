"""Step-boundary session driver (oracle restatement).

Restates python/cfdSolver/torchSession.py:19-576 (``Simulation.__init__``,
``_dt``, ``_reaction``, ``step`` and ``_metric_step``). Cartesian sessions:
Strang split R(dt/2) -> H(dt) -> R(dt/2) with up to 60 dt halvings while the
heated state violates the CFL bound, outward boundary integrals, source
integrals and a commit-at-end step. Sessions with a curvilinear geometry or
explicit face boundaries take the metric step instead (oracle.metric: SSPRK2
MUSCL-Rusanov, up to 20 dt halvings). ``checked=False`` skips the reference's 0 < CFL <= 0.5
validation (torchSolver.py:1074-1088) for workloads run outside it;
``constants`` passes kinetics constants the reference's 3D wrapper fixes.
"""

import numpy as np

from . import metric
from . import solver1d as one
from .execution import Execution
from . import solver2d as two
from . import solver3d as three


class Session:
    def __init__(self, primitive, lengths, *, t_end, cfl=0.4, gamma=1.4, boundaries=None,
                 limiter="VanLeer", reactive=False, kinetic_model="arrhenius",
                 lambda_init=None, beta_init=None, inlet_primitive=None,
                 controller="Sound Speed Controlled", variant="Variante 1 - Original",
                 wavespeed="Pressure Based", max_steps=100000, checked=True, constants=None,
                 geometry="cartesian", origins=None, face_boundaries=None, workers=None):
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
            dimension=dim, counts_xyz=counts, lengths_xyz=lengths, boundaries_xyz=boundaries,
            t_end=t_end, cfl=cfl, gamma=gamma, limiter=limiter, reactive=bool(reactive),
            kinetic_model=kinetic_model,
            inlet_primitive=list(two.INLET_DEFAULT if inlet_primitive is None else inlet_primitive),
            controller=controller, variant=variant, wavespeed=wavespeed, max_steps=max_steps,
            geometry=geometry, origins_xyz=list(np.atleast_1d(origins)) if origins is not None else [0.0] * dim,
        )
        self.faces = None
        if geometry != "cartesian" or face_boundaries is not None:
            metric.validate_geometry(self.config)
            self.faces = metric.default_faces(self.config, face_boundaries)
            self.config["face_boundaries"] = self.faces
        self.metrics = metric.Metrics(self.config) if self.faces is not None else None
        self.constants = dict(constants or {})
        for length, n in zip(lengths, counts):
            if checked:
                one_validate(length, n, cfl, t_end)
            elif not (np.isfinite([length, cfl, t_end]).all() and length > 0 and n >= 2 and cfl > 0):
                raise ValueError("Invalid grid")
        if not np.isfinite(primitive).all() or np.any(primitive[[0, -1]] <= 0):
            raise ValueError("Initial density and pressure must be positive and primitives finite")
        shape = tuple(counts[::-1])
        self.lam = self._fraction(lambda_init, shape) if reactive else None
        self.beta = self._fraction(beta_init, shape) if reactive and kinetic_model == "znd" else None
        g = self.gamma
        self.state = primitive.copy()
        self.state[1:-1] *= primitive[0]
        self.state[-1] = primitive[-1] / (g - 1) + 0.5 * primitive[0] * np.sum(primitive[1:-1] ** 2, axis=0)
        if not one.physical_state(self.state):
            raise ValueError("Nonphysical initial conservative state")
        self.time, self.steps = 0.0, 0
        self.initial_integral = self.integral.copy()
        self.boundary_integral = np.zeros(dim + 2)
        self.source_integral = np.zeros(dim + 2)
        self.fallbacks = 0
        # workers=None: the whole face grid as one batch; an int: the
        # reference's 8192-face batches through its worker pool
        # (torchSession.py:160-183, torchFlux.py:13, 66-76).
        self.execution = None if workers is None else Execution(workers)

    @staticmethod
    def _fraction(values, shape):
        out = np.ones(shape) if values is None else np.array(values, dtype=float, copy=True)
        if out.shape != shape or not np.isfinite(out).all() or np.any((out < 0) | (out > 1)):
            raise ValueError("fraction must have the spatial shape and values in [0,1]")
        return out

    @property
    def gamma(self):
        return one.gamma_calc(self.lam) if self.config["reactive"] else self.config["gamma"]

    @property
    def spacing(self):
        return np.array(self.config["lengths_xyz"]) / self.config["counts_xyz"]

    @property
    def integral(self):
        """torchSession.py:285-295 (volume-weighted on curvilinear grids)."""
        nf = self.config["dimension"] + 2
        if self.config["geometry"] != "cartesian":
            return (self.state * self.metrics.volumes[None]).reshape(nf, -1).sum(axis=1)
        return self.state.reshape(nf, -1).sum(axis=1) * np.prod(self.spacing)

    @property
    def balance_residual(self):
        return self.integral - self.initial_integral + self.boundary_integral - self.source_integral

    @property
    def completed(self):
        return self.time == self.config["t_end"]

    def dt_limit(self, state, gamma):
        """torchSession.py:333-360."""
        c = self.config
        h, bc = self.spacing, c["boundaries_xyz"]
        if c["dimension"] == 1:
            return one.cfl_dt(state, gamma, h[0], c["cfl"], bc[0], c["wavespeed"])
        if c["dimension"] == 2:
            return two.cfl_dt(state, gamma, h[0], h[1], c["cfl"], bc[0], bc[1], c["inlet_primitive"])
        return three.cfl_dt(state, gamma, h, c["cfl"])

    def reaction(self, state, lam, beta, dt):
        """torchSession.py:362-372."""
        dim, model = self.config["dimension"], self.config["kinetic_model"]
        if dim == 1:
            lam, state, beta = one.kinetics(lam, dt, state, model, beta, **self.constants)
            return state, lam, beta if model == "znd" else None
        if dim == 2:
            return two.reaction(state, lam, beta, dt, model, **self.constants)
        return three.reaction(state, lam, beta, dt, model, **self.constants)

    def hydro(self, old, gamma, dt):
        """Dimension-specific hydro step; returns (state, fluxes, fallback)."""
        c = self.config
        h, bc = self.spacing, c["boundaries_xyz"]
        if c["dimension"] == 1:
            state, flux, fb = one.hydro_step(old, gamma, h[0], dt, c["cfl"], c["limiter"], bc[0],
                                             c["controller"], c["variant"], c["wavespeed"])
            return state, [flux], fb
        if c["dimension"] == 2:
            state, fx, fy, fb = two.hydro(old, h[0], h[1], dt, gamma, c["limiter"], bc[0], bc[1],
                                          c["inlet_primitive"])
            return state, [fx, fy], fb
        batch = None if self.execution is None else 8192
        state, fluxes, fb = three.hydro(old, gamma, tuple(h), dt, c["limiter"], tuple(bc[::-1]),
                                        batch_size=batch, execution=self.execution)
        return state, fluxes, fb

    def advect(self, q, old, state, fluxes, dt):
        c = self.config
        h, bc = self.spacing, c["boundaries_xyz"]
        if c["dimension"] == 1:
            return one.advect(q, old, state, fluxes[0][0], dt / h[0], bc[0])
        if c["dimension"] == 2:
            return two.advect(q, old, state, fluxes[0][0], fluxes[1][0], dt / h[0], dt / h[1], bc[0], bc[1])
        return three.advect(q, old, state, fluxes, dt, tuple(h), tuple(bc[::-1]))

    def metric_step(self):
        """torchSession.py:523-576."""
        c = self.config
        if self.steps >= c["max_steps"]:
            raise RuntimeError("Maximum steps reached before requested final time")
        start = self.state
        dt = min(metric.cfl_dt(start, self.lam, c, self.metrics, self.faces), c["t_end"] - self.time)
        for _ in range(20):
            if not np.isfinite(dt) or dt <= 0 or self.time + dt == self.time:
                raise RuntimeError("Integration stalled before requested final time")
            old, lam, beta = start, self.lam, self.beta
            if c["reactive"]:
                old, lam, beta = self.reaction(start, self.lam, self.beta, dt / 2)
                allowed = metric.cfl_dt(old, lam, c, self.metrics, self.faces)
                if dt > allowed * (1 + 1e-12):
                    dt = min(dt / 2, allowed)
                    continue
            try:
                hydro, lam, beta, boundary, source, first = metric.advance(old, lam, beta, dt, c, self.metrics,
                                                                           self.faces)
            except RuntimeError:
                dt /= 2
                continue
            state = hydro
            if c["reactive"]:
                state, lam, beta = self.reaction(hydro, lam, beta, dt / 2)
                nf = c["dimension"] + 2
                source = source + ((old - start + state - hydro) * self.metrics.volumes[None]).reshape(nf, -1).sum(
                    axis=1)
            if not one.physical_state(state):
                raise RuntimeError("Reaction produced a nonphysical state")
            self.state, self.lam, self.beta = state, lam, beta
            self.boundary_integral += boundary
            self.source_integral += source
            self.time = min(self.time + dt, c["t_end"])
            self.steps += 1
            self.fallbacks += int(first)
            self.last_dt = dt
            return
        raise RuntimeError("Could not find a positive metric/face step")

    def step(self):
        """torchSession.py:374-522."""
        if self.completed:
            return
        if self.faces is not None:
            return self.metric_step()
        c = self.config
        dim = c["dimension"]
        h = self.spacing
        if self.steps >= c["max_steps"]:
            raise RuntimeError("Maximum steps reached before requested final time")
        dt = min(self.dt_limit(self.state, self.gamma), c["t_end"] - self.time)
        if not np.isfinite(dt) or dt <= 0 or self.time + dt == self.time:
            raise RuntimeError("Integration stalled before requested final time")
        start = self.state
        old, lam, beta = start, self.lam, self.beta
        if c["reactive"]:
            for _ in range(60):
                old, lam, beta = self.reaction(start, self.lam, self.beta, dt / 2)
                allowed = self.dt_limit(old, one.gamma_calc(lam))
                if dt <= allowed * (1 + 1e-12):
                    break
                dt = min(dt / 2, allowed)
            else:
                raise RuntimeError("Could not find stable reactive step")
        gamma = one.gamma_calc(lam) if c["reactive"] else c["gamma"]
        state, fluxes, fb = self.hydro(old, gamma, dt)
        if c["reactive"]:
            lam = self.advect(lam, old, state, fluxes, dt)
            if beta is not None:
                beta = self.advect(beta, old, state, fluxes, dt)
        outward = np.zeros(dim + 2)
        volume = np.prod(h)
        for direction, flux in enumerate(fluxes):
            axis = dim - direction
            faces = np.take(flux, -1, axis=axis) - np.take(flux, 0, axis=axis)
            outward += dt * volume / h[direction] * faces.reshape(dim + 2, -1).sum(axis=1)
        source = np.zeros(dim + 2)
        if c["reactive"]:
            hydro_state = state
            state, lam, beta = self.reaction(hydro_state, lam, beta, dt / 2)
            source = volume * (old - start + state - hydro_state).reshape(dim + 2, -1).sum(axis=1)
        if not one.physical_state(state):
            raise RuntimeError("Reaction produced a nonphysical state")
        self.state, self.lam, self.beta = state, lam, beta
        self.boundary_integral += outward
        self.source_integral += source
        self.time = min(self.time + dt, c["t_end"])
        self.steps += 1
        self.fallbacks += int(fb)
        self.last_dt = dt


def one_validate(length, cells, cfl, end_time):
    """torchSolver.py:1074-1088."""
    if (not isinstance(cells, (int, np.integer)) or not np.isfinite([length, cfl, end_time]).all()
            or length <= 0 or cells < 2 or not 0 < cfl <= 0.5 or end_time < 0):
        raise ValueError("Require length > 0, cells >= 2, 0 < CFL <= 0.5 and final time >= 0")
