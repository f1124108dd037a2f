"""Adaptive (AMR) leaves on the device against the reference's vectors and
the reference's own adaptive invariants.

Every case runs libtorchhip's adaptive kernels through the C ABI
(torch_amr_*). FP64 bars: 1e-12 relative per field for rhs, cfl and
advance on the reference's meshes; 1e-11 after multi-step runs with
regrids (leaf keys exact; Sod, entropy waves periodic in 1D and 2D,
reactive Arrhenius/ZND, first order, every face type, 3D stagnation inlet);
integrals to 1e-11 of their scale. The
invariant tests follow python/tests/testAdaptive.py:50-229 (rest states in
every geometry, poles and the radial origin, 2:1 closure across a periodic
seam, capacity and minimum level, second-order entropy advection).
"""

import json

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    assert a.shape == b.shape
    if b.ndim > 1:
        scale = np.max(np.abs(b), axis=tuple(range(1, b.ndim)), keepdims=True)
    else:
        scale = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b) / np.maximum(scale, 1e-300))) if b.size else 0.0


def sums_ok(got, want, bar):
    return np.max(np.abs(np.asarray(got) - want)) <= bar * max(np.max(np.abs(want)), 1.0)


@pytest.fixture(scope="module")
def d():
    return golden("adaptive")


@pytest.fixture(scope="module")
def ctx():
    from torch_amd import _abi

    c = _abi.Context(0)
    yield c
    c.close()


def amr_case(d, tag):
    from torch_amd import torchAdaptiveGeometry as tag_geo
    from torch_amd.torchAdaptiveGeometry import LeafMesh
    from torch_amd.torchGeometry import session_faces

    spec = json.loads(str(d[tag + "_spec"]))
    dim = len(spec["counts"])
    c = dict(dimension=dim, counts_xyz=spec["counts"], lengths_xyz=spec["lengths"], boundaries_xyz=["Open"] * dim,
             origins_xyz=spec["origins"], geometry=spec["geometry"], t_end=0.02, cfl=0.4, gamma=1.4,
             limiter="VanLeer", reactive=spec["reactive"], kinetic_model="arrhenius",
             inlet_primitive=[1.0, 0.0, 100.0, 101325.0], max_steps=100000)
    c["face_boundaries"] = session_faces(c, None)
    c["amr"] = tag_geo.settings(**spec["amr"])
    return LeafMesh(c, d[tag + "_keys"]), c


def leaves_for(ctx, c, precision="fp64"):
    from torch_amd import _abi
    from torch_amd.torchSolver3D import config

    dim = c["dimension"]
    cfg = config(c["counts_xyz"], None, c["lengths_xyz"], dim=dim, gamma=c["gamma"], reactive=c["reactive"],
                 model=c["kinetic_model"], gamma_mode=1 if c["reactive"] else 0, precision=precision, cfl=c["cfl"])
    return _abi.DeviceLeaves(ctx, cfg, c["geometry"], c["face_boundaries"], c["amr"]["order"])


def test_leaf_rhs_cfl_and_advance_match_the_reference(d, ctx):
    for tag in d["mesh_cases"]:
        print("case", tag, flush=True)
        mesh, c = amr_case(d, tag)
        state = d[tag + "_state"]
        lam = d[tag + "_lam"] if tag + "_lam" in d else None
        dev = leaves_for(ctx, c)
        try:
            dev.set_mesh(mesh)
            dev.upload(state, lam)
            for order, first in (("2", False), ("1", True)):
                change, species, boundary, source = dev.rhs(first)
                assert rel(change, d[f"{tag}_change{order}"]) <= 1e-12, (tag, order)
                assert sums_ok(boundary, d[f"{tag}_boundary{order}"], 1e-12), (tag, order)
                assert rel(source, d[f"{tag}_source{order}"]) <= 1e-12, (tag, order)
                if lam is not None:
                    assert rel(species[0], d[f"{tag}_species{order}"]) <= 1e-12, (tag, order)
            dt = dev.dt()
            assert abs(dt - float(d[tag + "_dt"])) <= 1e-12 * dt, tag
            info = dev.advance(0.5 * float(d[tag + "_dt"]))
            assert not info.fallback
            got, glam, _ = dev.download()
            assert rel(got, d[tag + "_adv_state"]) <= 1e-12, tag
            nf = c["dimension"] + 2
            assert sums_ok(info.outward[:nf], d[tag + "_adv_boundary"], 1e-12), tag
            assert sums_ok(info.source[:nf], d[tag + "_adv_source"], 1e-12), tag
            if lam is not None:
                assert rel(glam, d[tag + "_adv_lam"]) <= 1e-12, tag
        finally:
            dev.close()


def test_fp32_leaves_follow_the_reference(d, ctx):
    for tag in d["mesh_cases"][::3]:
        mesh, c = amr_case(d, tag)
        lam = d[tag + "_lam"] if tag + "_lam" in d else None
        dev = leaves_for(ctx, c, "fp32")
        try:
            dev.set_mesh(mesh)
            dev.upload(d[tag + "_state"], lam)
            change, _, _, _ = dev.rhs(False)
            assert rel(change, d[tag + "_change2"]) <= 2e-4, tag
            dt = dev.dt()
            assert abs(dt - float(d[tag + "_dt"])) <= 1e-5 * dt, tag
            dev.advance(0.5 * float(d[tag + "_dt"]))
            assert rel(dev.download()[0], d[tag + "_adv_state"]) <= 1e-5, tag
        finally:
            dev.close()


@pytest.mark.parametrize("tag", ["sod1d", "sod2d", "sod3d", "wave1d", "react2d", "znd2d", "order1", "periodic2d",
                                 "faces2d", "stag3d"])
def test_adaptive_runs_match_the_reference(d, tag):
    from torch_amd import Simulation

    print("run", tag, flush=True)
    spec = json.loads(str(d[tag + "_spec"]))
    opts = dict(spec["opts"])
    opts.pop("lengths", None)
    s = Simulation(d[tag + "_prim"], spec["lengths"], **opts).to_adaptive(**spec["amr"])
    try:
        np.testing.assert_array_equal(s.mesh.keys, d[tag + "_keys0"])
        assert rel(s.state, d[tag + "_state0"]) <= 1e-13
        for k in range(1, int(d[tag + "_steps"]) + 1):
            s.step()
            np.testing.assert_array_equal(s.mesh.keys, d[f"{tag}_keys{k}"], err_msg=f"{tag} step {k}")
            assert rel(s.state, d[f"{tag}_state{k}"]) <= 1e-11, (tag, k)
            assert abs(s.time - float(d[f"{tag}_time{k}"])) <= 1e-12 * float(d[f"{tag}_time{k}"])
            if f"{tag}_lam{k}" in d:
                assert rel(s.lam, d[f"{tag}_lam{k}"]) <= 1e-11, (tag, k)
            if f"{tag}_beta{k}" in d:
                assert rel(s.beta, d[f"{tag}_beta{k}"]) <= 1e-11, (tag, k)
        assert s.regrids == int(d[tag + "_regrids"])
        assert sums_ok(s.boundary_integral, d[tag + "_boundary"], 1e-11)
        assert sums_ok(s.source_integral, d[tag + "_source"], 1e-11)
    finally:
        s.close()


def uniform(dim=1, geometry="cartesian", reactive=False, **options):
    """The reference tests' base session (testAdaptive.py:31-46)."""
    from torch_amd import Simulation

    counts = [8, 4, 4][:dim]
    p = np.ones((dim + 2, *counts[::-1]))
    p[1:-1] = 0
    origins = [0.4, 0.3, 0.2][:dim] if geometry != "cartesian" else [0.0] * dim
    s = Simulation(p, [1.0] * dim, origins=origins, geometry=geometry, t_end=0.02, reactive=reactive)
    return s.to_adaptive(max_level=2, **options)


def test_rest_and_metric_partition_in_all_geometries():
    for dim in (1, 2, 3):
        for geometry in ("cartesian", "cylindrical", "annular", "spherical"):
            print("rest", dim, geometry, flush=True)
            s = uniform(dim, geometry)
            total, vol = s.integral.copy(), s.mesh.volumes.sum()
            s.regrid(refine=s.mesh.centers[:, 0] < 0.8, coarsen=False)
            assert s.mesh.levels.max() > 0
            assert abs(s.mesh.volumes.sum() - vol) <= 1e-12 * vol
            np.testing.assert_allclose(s.integral, total, atol=2e-14)
            s.run(step_limit=2)
            np.testing.assert_allclose(s.primitive[1:-1], 0, atol=1e-14)
            np.testing.assert_allclose(s.balance_residual, 0, atol=3e-14)
            s.close()


def test_adaptive_sod_along_every_axis_against_the_exact_solution():
    """testAdaptive.py:89-109 with the exact Riemann solution (oracle.exact)."""
    from oracle import exact
    from torch_amd import Simulation

    for dim in (1, 2, 3):
        for axis in range(dim):
            counts = [2] * dim
            counts[axis] = 32
            p = np.ones((dim + 2, *counts[::-1]))
            p[1:-1] = 0
            shape = [1] * dim
            shape[dim - 1 - axis] = 32
            right = np.broadcast_to((np.arange(32) >= 16).reshape(shape), counts[::-1])
            p[0, right] = 0.125
            p[-1, right] = 0.1
            s = Simulation(p, [1.0] * dim, t_end=0.08).to_adaptive(max_level=1, regrid_every=2)
            s.run()
            rho = exact.sample(s.mesh.centers[:, axis], s.time)[0]
            error = np.sum(abs(s.state[0] - rho) * s.mesh.volumes)
            assert error < 0.025, (dim, axis, error)
            assert s.mesh.levels.max() > 0
            np.testing.assert_allclose(s.balance_residual, 0, atol=1e-13)
            s.close()


def test_refined_rest_at_radial_origin_and_spherical_poles():
    from torch_amd import Simulation

    for dim in (1, 2):
        for geometry in ("cylindrical", "spherical"):
            p = np.ones((dim + 2, *([6] * dim)))
            p[1:-1] = 0
            lengths = [1.0] * dim
            if geometry == "spherical" and dim == 2:
                lengths[1] = np.pi
            s = Simulation(p, lengths, geometry=geometry, t_end=0.01).to_adaptive(max_level=1)
            s.regrid(refine=s.mesh.low[:, 0] == 0, coarsen=False)
            s.run()
            assert s.completed
            np.testing.assert_allclose(s.primitive[1:-1], 0, atol=2e-14)
            np.testing.assert_allclose(s.balance_residual, 0, atol=3e-14)
            s.close()


def test_two_to_one_closure_across_a_periodic_seam():
    from torch_amd.torchAdaptiveGeometry import LeafMesh

    for dim in (1, 2, 3):
        s = uniform(dim)
        s.config["face_boundaries"] = [{"type": "periodic"}] * (2 * dim)
        s.mesh = LeafMesh(s.config, s.mesh.keys)
        for _ in range(2):
            s.regrid(refine=s.mesh.low[:, 0] == 0, coarsen=False)
        m = s.mesh
        assert np.all(abs(m.levels[m.left[m.interior]] - m.levels[m.right[m.interior]]) < 2)
        assert m.levels[m.high[:, 0] == 1].max() > 0
        s.run()
        np.testing.assert_allclose(s.balance_residual, 0, atol=2e-14)
        s.close()


def test_capacity_and_minimum_level():
    s = uniform(max_cells=8)
    old = s.mesh.keys.copy()
    s.regrid(refine=np.ones(8, dtype=bool), coarsen=False)
    assert s.budget_limited
    np.testing.assert_array_equal(s.mesh.keys, old)
    s.close()
    s = uniform(min_level=1, max_cells=16)
    assert len(s.mesh.keys) == 16 and s.mesh.levels.min() == 1
    s.close()
    with pytest.raises(ValueError):
        uniform(min_level=2, max_cells=16)


def test_second_order_nonuniform_entropy_advection():
    from torch_amd import Simulation

    errors = []
    for n in (16, 32, 64):
        x = (np.arange(n) + 0.5) / n
        p = np.array([1 + 0.2 * np.sinc(1 / n) * np.sin(2 * np.pi * x), np.full(n, 0.5), np.ones(n)])
        s = Simulation(p, [1], boundaries=["Periodic"], t_end=0.1).to_adaptive(
            max_level=1, refine_threshold=1, coarsen_threshold=0, regrid_every=10000)
        s.regrid(refine=s.mesh.centers[:, 0] < 0.5, coarsen=False)
        s.run()
        x, width = s.mesh.centers[:, 0], s.mesh.widths[:, 0]
        exact = 1 + 0.2 * np.sinc(width) * np.sin(2 * np.pi * (x - 0.5 * s.time))
        errors.append(np.sum(abs(s.state[0] - exact) * width))
        np.testing.assert_allclose(s.balance_residual, 0, atol=4e-14)
        s.close()
    assert min(np.log2(np.array(errors[:-1]) / errors[1:])) > 1.8, errors


def test_mesh_validation_and_order_of_calls(d, ctx):
    mesh, c = amr_case(d, d["mesh_cases"][0])
    dev = leaves_for(ctx, c)
    try:
        with pytest.raises(ValueError, match="no mesh"):
            dev.dt()
        dev.set_mesh(mesh)
        with pytest.raises(ValueError, match="no state"):
            dev.rhs()
        bad = type("Bad", (), {})()
        for name in ("left", "right", "directions", "sides", "face_areas", "weights", "offset_left", "offset_right",
                     "displacement", "volumes", "divergence", "inverse"):
            setattr(bad, name, getattr(mesh, name).copy())
        bad.incidence = mesh.incidence
        bad.left[np.flatnonzero(bad.left >= 0)[0]] = len(mesh.volumes)
        with pytest.raises(ValueError):
            dev.set_mesh(bad)
        bad.left = mesh.left.copy()
        bad.volumes[0] = 0.0
        with pytest.raises(ValueError):
            dev.set_mesh(bad)
        bad.volumes = mesh.volumes.copy()
        start, faces, roles = mesh.incidence()
        cut = start[1] - 1  # the first leaf's last face side goes missing
        short = start.copy()
        short[1:] -= 1
        bad.incidence = lambda: (short, np.delete(faces, cut), np.delete(roles, cut))
        with pytest.raises(ValueError, match="misses"):
            dev.set_mesh(bad)
        bad.incidence = mesh.incidence
        dev.set_mesh(bad)  # the repaired copy installs
    finally:
        dev.close()


def test_exact_restart_with_a_changing_mesh(tmp_path):
    """testAdaptive.py:230-270: save after a step, resume, run both to the end."""
    from torch_amd import Simulation

    path = tmp_path / "restart.npz"
    for dim in (1, 2, 3):
        for model in ("inert", "arrhenius", "znd"):
            print("restart", dim, model, flush=True)
            p = np.ones((dim + 2, *([4] * dim)))
            p[1:-1] = 0.1
            p[0, ..., 2:] = 0.7
            reactive = model != "inert"
            s = Simulation(p, [1] * dim, t_end=0.08, reactive=reactive,
                           kinetic_model=model if reactive else "arrhenius").to_adaptive(max_level=1, regrid_every=1)
            s.step()
            s.save(path)
            loaded = Simulation.load(path)
            loaded.set_workers(2)
            s.run()
            loaded.run()
            for attr in ("state", "lam", "beta", "boundary_integral", "source_integral"):
                a, b = getattr(loaded, attr), getattr(s, attr)
                if b is None:
                    assert a is None
                else:
                    np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(loaded.mesh.keys, s.mesh.keys)
            assert (loaded.regrids, loaded.mesh_revision) == (s.regrids, s.mesh_revision)
            assert s.mesh_revision > 0
            loaded.close()
            s.close()


def test_a_failed_regrid_leaves_the_step_uncommitted(monkeypatch):
    """testAdaptive.py:272-291."""
    from torch_amd import torchAdaptive

    s = uniform()
    state, keys, time = s.state.copy(), s.mesh.keys.copy(), s.time

    def fail(*args, **kwargs):
        raise ValueError("failure")

    monkeypatch.setattr(torchAdaptive, "remesh", fail)
    s.config["amr"]["regrid_every"] = 1
    with pytest.raises(ValueError):
        s.step()
    np.testing.assert_array_equal(s.state, state)
    np.testing.assert_array_equal(s.mesh.keys, keys)
    assert s.time == time
    monkeypatch.undo()
    s.step()
    assert s.steps == 1
    s.close()


def test_steps_through_first_order_retries_and_halvings():
    """A rough 1D state at CFL 1.5 (unchecked): some steps retry at first
    order, some halve dt; the device run follows the oracle's step for step
    (the host-driven retry path behind the queued step)."""
    import test_adaptive_host as host
    from torch_amd import Simulation
    from torch_amd import torchAdaptiveGeometry as tag_geo
    from torch_amd import torchGeometry as geo

    rng = np.random.default_rng(0)
    n = 32
    prim = np.array([0.2 + 1.8 * rng.random(n), 2 * rng.standard_normal(n), 0.1 + 4.9 * rng.random(n)])
    c = dict(dimension=1, counts_xyz=[n], lengths_xyz=[1.0], boundaries_xyz=["Open"], origins_xyz=[0.0],
             geometry="cartesian", t_end=10.0, cfl=1.5, gamma=1.4, limiter="VanLeer", reactive=False,
             kinetic_model="arrhenius", max_steps=100000)
    c["face_boundaries"] = geo.normalize_faces(c, None)
    c["amr"] = tag_geo.settings(max_level=1, regrid_every=2)
    ref = host.HostLeaves(c, prim)
    s = Simulation(prim, [1.0], t_end=10.0, cfl=1.5, checked=False).to_adaptive(max_level=1, regrid_every=2)
    np.testing.assert_array_equal(s.mesh.keys, ref.mesh.keys)
    halvings = 0
    for k in range(8):
        ref.step()
        info = s.step()
        halvings += info.retries
        np.testing.assert_array_equal(s.mesh.keys, ref.mesh.keys, err_msg=f"step {k}")
        assert rel(s.state, ref.state) <= 1e-11, k
        assert s.time == pytest.approx(ref.time, rel=1e-12)
    assert s.fallbacks >= 1 and halvings >= 1
    assert sums_ok(s.boundary_integral, ref.boundary, 1e-11)
    s.close()
