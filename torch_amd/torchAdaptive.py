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
from .torchAdaptiveGeometry import LeafMesh, canonical_order, settings, volume
from .torchExecution import DeviceExecution, validate_workers
from .torchSession import Simulation


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
    keeps the 2:1 balance with the (tagged) neighbours: (parents (G, d+1),
    member leaves (G, 2^d) in leaf order)."""
    d = mesh.dimension
    empty = (np.zeros((0, d + 1), dtype=np.int64), np.zeros((0, 2 ** d), dtype=np.int64))
    ids = np.flatnonzero((mesh.levels > options["min_level"]) & ~tag & (indicator < options["coarsen_threshold"]))
    if not len(ids):
        return empty
    parent = np.column_stack((mesh.levels[ids] - 1, mesh.keys[ids, 1:] // 2))
    codes = np.empty(len(ids), dtype=np.int64)
    for level in np.unique(parent[:, 0]):
        at = parent[:, 0] == level
        codes[at] = mesh._codes(level, parent[at, 1:])
    _, first, inverse, count = np.unique(codes, return_index=True, return_inverse=True, return_counts=True)
    inverse = inverse.ravel()
    complete = np.flatnonzero(count == 2 ** d)
    if not len(complete):
        return empty
    slot = np.full(len(count), -1)
    slot[complete] = np.arange(len(complete))
    member = ids[np.argsort(inverse, kind="stable")].reshape(-1)
    starts = np.concatenate(([0], np.cumsum(count)))
    groups = member[starts[complete][:, None] + np.arange(2 ** d)]
    # A group is rejected when a leaf across one of its faces, outside the
    # group, would end up more than one level finer than the merged parent.
    group = np.full(len(tag), -1)
    group[ids] = slot[inverse]
    limit = parent[first[complete], 0] + 1
    finer = mesh.levels + tag
    rejected = np.zeros(len(complete), dtype=bool)
    for a, b in ((i, j), (j, i)):
        crossing = (group[a] >= 0) & (group[a] != group[b])
        rejected[group[a][crossing][finer[b][crossing] > limit[group[a][crossing]]]] = True
    keep = ~rejected
    return parent[first[complete]][keep], groups[keep]


def remesh(mesh, state, fractions, config, faces, refine=None, coarsen=True):
    """One mesh transaction (torchAdaptive.py:209-295): jump-indicator (or
    given) refinement with 2:1 closure, sibling coarsening, the capacity
    limit, conservative restriction and limited linear prolongation.

    ``faces`` are the raw face values (left, right, gl, gr) of ``state`` on
    ``mesh``. Returns (mesh, state, fractions, budget_limited, changed)."""
    left, right, gl, gr = faces
    options = config["amr"]
    d = mesh.dimension
    nchild = 2 ** d
    indicator = _indicator(mesh, left, right)
    tag = indicator > options["refine_threshold"] if refine is None else np.asarray(refine, dtype=bool)
    if tag.shape != indicator.shape:
        raise ValueError("Refinement mask must match leaves")
    tag = (tag | (mesh.levels < options["min_level"])) & (mesh.levels < options["max_level"])
    tag, i, j = _balanced(mesh, tag, options["max_level"])
    if coarsen:
        parents, groups = _coarsenable(mesh, tag, indicator, options, i, j)
    else:
        parents, groups = np.zeros((0, d + 1), dtype=np.int64), np.zeros((0, nchild), dtype=np.int64)
    budget_limited = len(tag) + int(tag.sum()) * (nchild - 1) - len(parents) * (nchild - 1) > options["max_cells"]
    if budget_limited:
        tag[:] = False  # refinements wait; coarsening alone keeps the mesh valid
    if not np.any(tag) and not len(parents):
        return mesh, state, fractions, budget_limited, False
    removed = np.zeros(len(tag), dtype=bool)
    removed[groups.ravel()] = True
    keep = ~removed & ~tag
    keys, values, qs = [mesh.keys[keep]], [state[:, keep]], [[q[keep]] for q in fractions]
    if len(parents):
        # Conservative restriction: volume-weighted means, mass-weighted fractions.
        w = mesh.volumes[groups]
        keys.append(parents)
        values.append(np.sum(state[:, groups] * w, axis=2) / np.sum(w, axis=1))
        rho = state[0, groups]
        for out, q in zip(qs, fractions):
            out.append(np.sum(rho * q[groups] * w, axis=1) / np.sum(rho * w, axis=1))
    t = np.flatnonzero(tag)
    if len(t):
        # Prolongation: each tagged leaf's state on its children, plus the
        # limited conservative gradient at the child centres with the
        # volume-weighted mean removed, kept where every child stays physical.
        level = mesh.levels[t] + 1
        bits = np.array([[(c >> a) & 1 for a in range(d)] for c in range(nchild)])
        index = 2 * mesh.keys[t, 1:][:, None, :] + bits[None]
        lengths, counts = np.array(config["lengths_xyz"]), np.array(config["counts_xyz"])
        scale = (2.0 ** level)[:, None, None]
        low = mesh.origin + index * lengths / counts / scale
        high = low + lengths / counts / scale
        w = volume(config, low.reshape(-1, d), high.reshape(-1, d)).reshape(len(t), nchild)
        fine = np.repeat(state[:, t, None], nchild, axis=2)
        if options["order"] == 2:
            gradient = mesh.gradients(state, _conservative(left, gl), _conservative(right, gr))[t]
            change = np.matmul(0.5 * (low + high) - mesh.centers[t][:, None, :], gradient)
            change -= np.sum(change * w[:, :, None], axis=1, keepdims=True) / np.sum(w, axis=1)[:, None, None]
            trial = fine + change.transpose(2, 0, 1)
            internal = trial[-1] - 0.5 * np.sum(trial[1:-1] ** 2, axis=0) / trial[0]
            ok = (np.isfinite(trial).all(axis=(0, 2)) & np.all(trial[0] > 1e-12, axis=1)
                  & np.all(internal > 1e-12, axis=1))
            fine = np.where(ok[None, :, None], trial, fine)
        keys.append(np.column_stack((np.repeat(level, nchild), index.reshape(-1, d))))
        values.append(fine.reshape(state.shape[0], -1))
        for out, q in zip(qs, fractions):
            out.append(np.repeat(q[t], nchild))
    keys = np.concatenate(keys).astype(np.int64)
    order = canonical_order(keys)
    candidate = LeafMesh(config, keys[order])
    out = np.concatenate(values, axis=1)[:, order]
    qs = [np.concatenate(parts)[order] for parts in qs]
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
        self._remember()

    def _sync(self):
        """Reinstall after the mesh or the solver options were replaced on
        the session (``s.mesh = LeafMesh(...)``, edited face boundaries)."""
        if self.mesh is not self._mesh_on_device or self._solver_key() != self._leaves_key:
            state, lam, beta = self._fetch()
            self._install(self.mesh, state, lam, beta)
        else:
            self._push_edits()

    def _upload(self, state, lam, beta):
        self._install(self.mesh, state, lam, beta)

    def _fetch(self):
        if self._host is None:
            self._host = self._leaves.download()
            self._remember()
        return self._host

    def _upload_edits(self, state, lam, beta):
        self._leaves.upload(state, lam, beta)

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

    def _attach(self, workers, device):
        """Give a host-side shell (a validated schema-3 checkpoint) its device leaves."""
        state, lam, beta = self._host
        self._execution = DeviceExecution(validate_workers(workers), device=device, precision=self.precision)
        self._device, self._leaves, self._leaves_key, self._mesh_on_device = None, None, None, None
        self._install(self.mesh, state, lam, beta)
        return self

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
