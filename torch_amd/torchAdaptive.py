"""Adaptive (dyadic AMR) sessions on the GPU.

Mirrors python/cfdSolver/torchAdaptive.py: ``AdaptiveSimulation`` (:298-481,
created from a uniform session with ``Simulation.to_adaptive``,
torchSession.py:185-188) and ``remesh`` (:209-295). The per-step numerics —
face values with boundary planes, bounded least-squares reconstruction,
shared-face Rusanov fluxes, geometric sources, SSPRK2 with the first-order
retry, the CFL rate and the split chemistry — run on the device through
``torch_amr_*`` (torch_amd/csrc/fvmetric.hip, adapt.hip). The leaf tree and
remeshing are host data structures, as in the reference: a regrid reads the
committed state and its raw face values from the device, builds the next
leaf mesh (torchAdaptiveGeometry.LeafMesh) and installs it with the
restricted / prolonged state.
"""

import copy
import json

import numpy as np

from . import _abi
from . import torchGeometry as _geometry
from .torchAdaptiveGeometry import LeafMesh, children, settings, volume
from .torchExecution import DeviceExecution
from .torchSession import Simulation, _physical


def _conservative(p, gamma):
    """Conservative states of primitive face values (torchFiniteVolume.py:25-29)."""
    u = p.copy()
    u[1:-1] *= p[0]
    u[-1] = p[-1] / (gamma - 1) + 0.5 * p[0] * np.sum(p[1:-1] ** 2, axis=0)
    return u


def _indicator(mesh, left, right):
    """Largest normalised density/pressure jump over each leaf's faces."""
    a, b = left[[0, -1]], right[[0, -1]]
    jumps = np.max(abs(a - b) / np.maximum(abs(a) + abs(b), 1e-14), axis=0)
    out = np.zeros(len(mesh.keys))
    for idx in (mesh.left, mesh.right):
        inside = idx >= 0
        np.maximum.at(out, idx[inside], jumps[inside])
    return out


def _balanced(mesh, tag, max_level):
    """Extend refinement tags until face neighbours differ by at most one level."""
    i, j = mesh.left[mesh.interior], mesh.right[mesh.interior]
    for _ in range(max_level + 1):
        before = tag.copy()
        li, lj = mesh.levels[i] + tag[i], mesh.levels[j] + tag[j]
        tag[j[li > lj + 1]] = True
        tag[i[lj > li + 1]] = True
        if np.array_equal(before, tag):
            break
    return tag, i, j


def _coarsenable(mesh, tag, indicator, options, i, j):
    """Complete sibling groups below the coarsening threshold whose merge
    keeps the 2:1 balance with the (tagged) neighbours."""
    groups = {}
    for idx, key in enumerate(mesh.keys):
        if key[0] <= options["min_level"] or tag[idx] or indicator[idx] >= options["coarsen_threshold"]:
            continue
        parent = (int(key[0]) - 1, *(int(v) // 2 for v in key[1:]))
        groups.setdefault(parent, []).append(idx)
    accepted = {}
    for parent, ids in groups.items():
        if len(ids) != 2 ** mesh.dimension:
            continue
        member = np.zeros(len(tag), dtype=bool)
        member[ids] = True
        outside = np.concatenate((j[member[i] & ~member[j]], i[member[j] & ~member[i]]))
        if np.any(mesh.levels[outside] + tag[outside] > parent[0] + 1):
            continue
        accepted[parent] = ids
    return accepted


def remesh(mesh, state, fractions, config, faces, refine=None, coarsen=True):
    """One mesh transaction (torchAdaptive.py:209-295): jump-indicator (or
    given) refinement with 2:1 closure, sibling coarsening, the capacity
    limit, conservative restriction and limited linear prolongation.

    ``faces`` are the raw face values (left, right, gl, gr) of ``state`` on
    ``mesh``. Returns (mesh, state, fractions, budget_limited, changed)."""
    left, right, gl, gr = faces
    options = config["amr"]
    indicator = _indicator(mesh, left, right)
    tag = indicator > options["refine_threshold"] if refine is None else np.asarray(refine, dtype=bool)
    if tag.shape != indicator.shape:
        raise ValueError("Refinement mask must match leaves")
    tag = (tag | (mesh.levels < options["min_level"])) & (mesh.levels < options["max_level"])
    tag, i, j = _balanced(mesh, tag, options["max_level"])
    accepted = _coarsenable(mesh, tag, indicator, options, i, j) if coarsen else {}
    removed = {idx for ids in accepted.values() for idx in ids}
    grow = 2 ** mesh.dimension - 1
    budget_limited = len(tag) + int(tag.sum()) * grow - len(accepted) * grow > options["max_cells"]
    if budget_limited:
        tag[:] = False  # refinements wait; coarsening alone keeps the mesh valid
    if not np.any(tag) and not accepted:
        return mesh, state, fractions, budget_limited, False
    data = {}
    for parent, ids in accepted.items():
        w = mesh.volumes[ids]
        mean = np.sum(state[:, ids] * w, axis=1) / np.sum(w)
        data[parent] = (mean, [np.sum(state[0, ids] * q[ids] * w) / np.sum(state[0, ids] * w) for q in fractions])
    gradient = None
    if np.any(tag) and options["order"] == 2:
        gradient = mesh.gradients(state, _conservative(left, gl), _conservative(right, gr))
    lengths, counts = np.array(config["lengths_xyz"]), np.array(config["counts_xyz"])
    for idx, key_array in enumerate(mesh.keys):
        if idx in removed:
            continue
        key = tuple(int(v) for v in key_array)
        if not tag[idx]:
            data[key] = (state[:, idx], [q[idx] for q in fractions])
            continue
        group = children(key)
        low = mesh.origin + np.array([k[1:] for k in group]) * lengths / counts / 2 ** (key[0] + 1)
        high = low + lengths / counts / 2 ** (key[0] + 1)
        w = volume(config, low, high)
        values = np.repeat(state[:, idx, None], len(group), axis=1)
        if gradient is not None:
            change = (0.5 * (low + high) - mesh.centers[idx]) @ gradient[idx]
            change -= np.sum(change * w[:, None], axis=0) / np.sum(w)
            trial = values + change.T
            if _physical(trial):
                values = trial
        for k, value in zip(group, values.T):
            data[k] = (value, [q[idx] for q in fractions])
    candidate = LeafMesh(config, np.array(list(data), dtype=np.int64))
    out = np.stack([data[tuple(key)][0] for key in candidate.keys], axis=1)
    qs = [np.array([data[tuple(key)][1][a] for key in candidate.keys]) for a in range(len(fractions))]
    changed = not np.array_equal(mesh.keys, candidate.keys)
    return (candidate if changed else mesh), out, qs, budget_limited, changed


class AdaptiveSimulation(Simulation):
    """Leaf simulation over fixed root counts (torchAdaptive.py:298-481)."""

    @classmethod
    def from_uniform(cls, session, **options):
        if isinstance(session, cls):
            raise ValueError("Session is already adaptive")
        result = cls.__new__(cls)
        result.config = copy.deepcopy(session.config)
        result.config["amr"] = settings(**options)
        result.config["face_boundaries"] = _geometry.normalize_faces(result.config,
                                                                     result.config.get("face_boundaries"))
        result.precision, result.checked = session.precision, session.checked
        result.constants, result.tile = dict(session.constants), session.tile
        result._validate_config()
        result.mesh = LeafMesh(result.config)
        state, lam, beta = session._fetch()
        nf = result.config["dimension"] + 2
        state = state.reshape(nf, -1).copy()
        lam = None if lam is None else lam.ravel().copy()
        beta = None if beta is None else beta.ravel().copy()
        result.time, result.steps, result.run_id = session.time, session.steps, session.run_id
        result.fallbacks = 0
        for key in ("initial_integral", "boundary_integral", "source_integral"):
            setattr(result, key, getattr(session, key).copy())
        result.regrids, result.mesh_revision, result.budget_limited = 0, 0, False
        result._execution = DeviceExecution(session.workers, device=session.execution.device,
                                            precision=result.precision)
        result._device, result._leaves, result._leaves_key, result._mesh_on_device = None, None, None, None
        result._install(result.mesh, state, lam, beta)
        for _ in range(result.config["amr"]["max_level"]):
            result.regrid()
        return result

    # -- device plumbing -------------------------------------------------------
    def _solver_key(self):
        c = self.config
        return (c.get("geometry", "cartesian"), json.dumps(c["face_boundaries"], sort_keys=True, default=float),
                c["amr"]["order"], c["cfl"], c["gamma"], self.precision)

    def _install(self, mesh, state, lam, beta):
        key = self._solver_key()
        if self._leaves is not None and key != self._leaves_key:
            self._leaves.close()
            self._leaves = None
        if self._leaves is None:
            c = self.config
            self._leaves = _abi.DeviceLeaves(self._execution.context, self._make_config(),
                                             c.get("geometry", "cartesian"), c["face_boundaries"],
                                             c["amr"]["order"])
            self._leaves_key = key
        self._leaves.set_mesh(mesh)
        self._leaves.upload(state, lam, beta)
        self.mesh = self._mesh_on_device = mesh
        self._host = (state, lam, beta)

    def _sync(self):
        """Reinstall after the mesh or the solver options were replaced on
        the session (``s.mesh = LeafMesh(...)``, edited face boundaries)."""
        if self.mesh is not self._mesh_on_device or self._solver_key() != self._leaves_key:
            state, lam, beta = self._fetch()
            self._install(self.mesh, state, lam, beta)

    def _upload(self, state, lam, beta):
        self._install(self.mesh, state, lam, beta)

    def _fetch(self):
        if self._host is None:
            self._host = self._leaves.download()
        return self._host

    def _fraction(self, values, name):
        q = np.asarray(values, dtype=float)
        if q.shape != (len(self.mesh.keys),) or not np.isfinite(q).all() or np.any((q < 0) | (q > 1)):
            raise ValueError(f"Invalid adaptive {name}")
        return q.copy()

    def close(self):
        if getattr(self, "_leaves", None) is not None:
            self._leaves.close()
            self._leaves = None
        super().close()

    def set_workers(self, workers):
        state, lam, beta = self._fetch()
        old = self._execution
        if self._leaves is not None:
            self._leaves.close()
            self._leaves = None
        self._execution = DeviceExecution(workers, device=old.device, precision=old.precision)
        old.close()
        self._install(self.mesh, state, lam, beta)
        return self

    # -- reference properties ----------------------------------------------------
    @property
    def integral(self):
        return np.sum(self.state * self.mesh.volumes, axis=1)

    def dt_limit(self):
        """cfl of the committed leaves (torchAdaptive.py:149-166)."""
        self._sync()
        return self._leaves.dt()

    def device_totals(self):
        return self.integral

    # -- mesh transactions ---------------------------------------------------------
    def _fractions(self, lam, beta):
        return [q for q in (lam, beta) if q is not None]

    def regrid(self, refine=None, coarsen=True):
        """torchAdaptive.py:401-411."""
        self._sync()
        state, lam, beta = self._fetch()
        mesh, state, qs, budget, changed = remesh(self.mesh, state, self._fractions(lam, beta), self.config,
                                                  self._leaves.faces(), refine, coarsen)
        if changed:
            self._install(mesh, state, qs[0] if qs else None, qs[1] if len(qs) > 1 else None)
        self.regrids += 1
        self.mesh_revision += int(changed)
        self.budget_limited = budget

    def step(self):
        """AdaptiveSimulation.step (torchAdaptive.py:436-481); the step is
        committed only together with its regrid."""
        if self.completed:
            return
        c = self.config
        if self.steps >= c["max_steps"]:
            raise RuntimeError("Maximum steps reached")
        self._sync()
        do_regrid = (self.steps + 1) % c["amr"]["regrid_every"] == 0
        before = self._fetch() if do_regrid else None
        info = self._leaves.step(self.time, c["t_end"])
        self._host = None
        budget, changed = self.budget_limited, False
        if do_regrid:
            try:
                state, lam, beta = self._fetch()
                mesh, state, qs, budget, changed = remesh(self.mesh, state, self._fractions(lam, beta), c,
                                                          self._leaves.faces())
                if changed:
                    self._install(mesh, state, qs[0] if qs else None, qs[1] if len(qs) > 1 else None)
            except Exception:
                self._install(self.mesh, *before)
                raise
        nf = c["dimension"] + 2
        self.boundary_integral += np.array(info.outward[:nf])
        self.source_integral += np.array(info.source[:nf])
        self.time = min(self.time + info.dt, c["t_end"])
        self.steps += 1
        self.regrids += int(do_regrid)
        self.mesh_revision += int(changed)
        self.budget_limited = budget
        self.fallbacks += int(info.fallback)
        self.last_dt = info.dt
        return info

    def run(self, *, step_limit=None, checkpoint=None, checkpoint_every=0, output_prefix=None, output_every=1,
            formats=("vtk", "xdmf")):
        if checkpoint or output_prefix:
            raise NotImplementedError("Adaptive checkpoints and field output are not part of the device backend")
        return super().run(step_limit=step_limit)

    def save(self, path):
        raise NotImplementedError("Adaptive checkpoints are not part of the device backend")

    def __deepcopy__(self, memo):
        state, lam, beta = (None if a is None else a.copy() for a in self._fetch())
        new = object.__new__(type(self))
        for key, value in self.__dict__.items():
            if key not in ("_execution", "_device", "_leaves", "_host", "mesh", "_mesh_on_device", "_leaves_key"):
                setattr(new, key, copy.deepcopy(value, memo))
        new._execution = copy.deepcopy(self._execution, memo)
        new._device, new._leaves, new._leaves_key, new._mesh_on_device = None, None, None, None
        new._install(self.mesh, state, lam, beta)
        return new
