"""Host side of the adaptive sessions: remeshing against the reference.

torch_amd.torchAdaptive.remesh (the mesh transaction AdaptiveSimulation runs
on the host between device steps) is driven here by the oracle's per-step
numerics (oracle/adaptive.py, itself pinned to the reference's rhs, cfl and
advance) and compared with the reference's multi-step adaptive runs in
tests/golden/adaptive.npz: the leaf keys after the initial refinement and
after every step (exact), the state, fractions and time (1e-12), the
boundary and source integrals. Also: refinement restriction conserves the
volume integral, and the capacity limit defers refinement.
"""

import json

import numpy as np
import pytest

from conftest import golden
from oracle import adaptive as oa
from oracle import solver1d as one


def run_config(d, tag):
    from torch_amd import torchAdaptiveGeometry as tag_geo
    from torch_amd import torchGeometry as geo

    spec = json.loads(str(d[tag + "_spec"]))
    p = d[tag + "_prim"]
    dim = p.ndim - 1
    opts = dict(spec["opts"])
    c = dict(dimension=dim, counts_xyz=list(p.shape[1:][::-1]), lengths_xyz=list(spec["lengths"]),
             boundaries_xyz=list(opts.get("boundaries", ["Open"] * dim)),
             origins_xyz=list(opts.get("origins", [0.0] * dim)), geometry=opts.get("geometry", "cartesian"),
             t_end=opts["t_end"], cfl=0.4, gamma=1.4, limiter="VanLeer", reactive=opts.get("reactive", False),
             kinetic_model=opts.get("kinetic_model", "arrhenius"), max_steps=100000)
    spec_faces = opts.get("face_boundaries")
    if c["geometry"] != "cartesian" or spec_faces is not None:
        c["face_boundaries"] = geo.session_faces(c, spec_faces)
    c["face_boundaries"] = geo.normalize_faces(c, c.get("face_boundaries"))
    c["amr"] = tag_geo.settings(**spec["amr"])
    return c, p


class HostLeaves:
    """AdaptiveSimulation's loop (torchAdaptive.py:298-481) with the oracle's
    numerics and the product's mesh transactions."""

    def __init__(self, c, p):
        from torch_amd.torchAdaptiveGeometry import LeafMesh

        self.c = c
        dim = c["dimension"]
        lam = np.ones(p[0].size) if c["reactive"] else None
        self.fractions = [lam] if lam is not None else []
        if c["reactive"] and c["kinetic_model"] == "znd":
            self.fractions.append(np.ones(p[0].size))
        g = one.gamma_calc(lam) if c["reactive"] else c["gamma"]
        prim = p.reshape(dim + 2, -1)
        state = prim.copy()
        state[1:-1] *= prim[0]
        state[-1] = prim[-1] / (g - 1) + 0.5 * prim[0] * np.sum(prim[1:-1] ** 2, axis=0)
        self.state, self.mesh = state, LeafMesh(c)
        self.time, self.steps, self.regrids = 0.0, 0, 0
        self.boundary, self.source = np.zeros(dim + 2), np.zeros(dim + 2)
        for _ in range(c["amr"]["max_level"]):
            self.regrid()

    def faces(self, state, fractions):
        left, right, gl, gr, _, _ = oa.face_values(self.mesh, state, fractions, self.c, False)
        return left, right, gl, gr

    def regrid(self, refine=None):
        from torch_amd.torchAdaptive import remesh

        self.mesh, self.state, self.fractions, _, _ = remesh(
            self.mesh, self.state, self.fractions, self.c, self.faces(self.state, self.fractions), refine)
        self.regrids += 1

    def react(self, state, fractions, dt):
        """torchAdaptive.py:418-434."""
        normal = np.stack((state[0], state[1], state[-1] - 0.5 * np.sum(state[2:-1] ** 2, axis=0) / state[0]))
        beta = fractions[1] if len(fractions) > 1 else None
        lam, reacted, beta = one.kinetics(fractions[0], dt, normal, self.c["kinetic_model"], beta)
        out = state.copy()
        out[-1] += reacted[-1] - normal[-1]
        return out, [lam] + ([beta] if len(fractions) > 1 else [])

    def step(self):
        from torch_amd.torchAdaptive import remesh

        c = self.c
        dt = min(oa.cfl(self.mesh, self.state, self.fractions, c), c["t_end"] - self.time)
        for _ in range(24):
            old, fractions = self.state, self.fractions
            if c["reactive"]:
                old, fractions = self.react(self.state, self.fractions, dt / 2)
                allowed = oa.cfl(self.mesh, old, fractions, c)
                if dt > allowed * (1 + 1e-12):
                    dt = min(dt / 2, allowed)
                    continue
            try:
                hydro, qs, boundary, src, _ = oa.advance(self.mesh, old, fractions, dt, c)
            except RuntimeError:
                dt /= 2
                continue
            state = hydro
            if c["reactive"]:
                state, qs = self.react(hydro, qs, dt / 2)
                src = src + np.sum((old - self.state + state - hydro) * self.mesh.volumes, axis=1)
            if (self.steps + 1) % c["amr"]["regrid_every"] == 0:
                self.mesh, state, qs, _, _ = remesh(self.mesh, state, qs, c,
                                                    self.faces_of(state, qs))
                self.regrids += 1
            self.state, self.fractions = state, qs
            self.boundary += boundary
            self.source += src
            self.time = min(self.time + dt, c["t_end"])
            self.steps += 1
            return
        raise RuntimeError("Could not find a positive adaptive step")

    def faces_of(self, state, fractions):
        return self.faces(state, fractions)


def rel(a, b):
    scale = np.max(np.abs(b), axis=tuple(range(1, np.ndim(b))), keepdims=True) if np.ndim(b) > 1 else np.max(
        np.abs(b))
    return float(np.max(np.abs(a - b) / np.maximum(scale, 1e-300)))


@pytest.fixture(scope="module")
def d():
    return golden("adaptive")


@pytest.mark.parametrize("tag", ["sod1d", "sod2d", "wave1d", "react2d", "znd2d", "order1", "periodic2d", "faces2d",
                                 "stag3d"])
def test_remesh_reproduces_the_reference_runs(d, tag):
    c, p = run_config(d, tag)
    run = HostLeaves(c, p)
    np.testing.assert_array_equal(run.mesh.keys, d[tag + "_keys0"])
    assert rel(run.state, d[tag + "_state0"]) <= 1e-13
    for k in range(1, int(d[tag + "_steps"]) + 1):
        run.step()
        np.testing.assert_array_equal(run.mesh.keys, d[f"{tag}_keys{k}"], err_msg=f"{tag} step {k}")
        assert rel(run.state, d[f"{tag}_state{k}"]) <= 1e-11, (tag, k)
        assert abs(run.time - float(d[f"{tag}_time{k}"])) <= 1e-12 * float(d[f"{tag}_time{k}"])
        if f"{tag}_lam{k}" in d:
            assert rel(run.fractions[0], d[f"{tag}_lam{k}"]) <= 1e-11
        if f"{tag}_beta{k}" in d:
            assert rel(run.fractions[1], d[f"{tag}_beta{k}"]) <= 1e-11
    assert run.regrids == int(d[tag + "_regrids"])
    for got, want in ((run.boundary, d[tag + "_boundary"]), (run.source, d[tag + "_source"])):
        assert np.max(np.abs(got - want)) <= 1e-11 * max(np.max(np.abs(want)), 1.0), tag


def test_restriction_and_prolongation_conserve_the_integral(d):
    from torch_amd.torchAdaptive import remesh

    c, p = run_config(d, "react2d")
    run = HostLeaves(c, p)
    before = np.sum(run.state * run.mesh.volumes, axis=1)
    lam_mass = np.sum(run.state[0] * run.fractions[0] * run.mesh.volumes)
    # refine everything that can be refined, then coarsen everything back
    mesh, state, qs, _, changed = remesh(run.mesh, run.state, run.fractions, c, run.faces(run.state, run.fractions),
                                         np.ones(len(run.mesh.keys), dtype=bool), False)
    assert changed
    np.testing.assert_allclose(np.sum(state * mesh.volumes, axis=1), before, rtol=1e-14, atol=1e-9)
    np.testing.assert_allclose(np.sum(state[0] * qs[0] * mesh.volumes), lam_mass, rtol=1e-13)


def test_capacity_limit_defers_refinement(d):
    from torch_amd.torchAdaptive import remesh
    from torch_amd.torchAdaptiveGeometry import settings

    c, p = run_config(d, "sod1d")
    c["amr"] = settings(max_level=1, regrid_every=2, max_cells=17)
    run = HostLeaves.__new__(HostLeaves)
    HostLeaves.__init__(run, c, p)
    keys = run.mesh.keys.copy()
    mesh, _, _, budget, changed = remesh(run.mesh, run.state, [], c, run.faces(run.state, []),
                                         np.ones(len(keys), dtype=bool))
    assert budget and not changed
    np.testing.assert_array_equal(mesh.keys, keys)
    with pytest.raises(ValueError):
        remesh(run.mesh, run.state, [], c, run.faces(run.state, []), np.ones(3, dtype=bool))
