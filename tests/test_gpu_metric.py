"""The metric/face scheme on the device against the reference's vectors,
the oracle and the reference's geometry invariants.

Every case runs libtorchhip's metric kernels through the C ABI (via the
torch_amd session mirror). FP64: 1e-12 relative per field for rhs, cfl_dt
and advance; 1e-11 after multi-step sessions (the same bars the oracle
meets against the reference); boundary/source integrals to 1e-12 of their
scale. FP32 reactive sessions on cylindrical and spherical grids take the
FP64 method's steps and dt halvings, within 2e-5 relative L1 per field.
"""

import copy
import json

import numpy as np
import pytest

from conftest import golden
from oracle import metric as om
from oracle.session import Session as OracleSession

pytestmark = pytest.mark.gpu


def close(actual, expected, rel=1e-12, what=""):
    actual, expected = np.asarray(actual, dtype=float), np.asarray(expected, dtype=float)
    assert actual.shape == expected.shape, what
    if expected.ndim > 1:
        scale = np.maximum(np.max(np.abs(expected), axis=tuple(range(1, expected.ndim)), keepdims=True), 1e-300)
    else:
        scale = max(np.max(np.abs(expected)) if expected.size else 0.0, 1e-300)
    err = np.max(np.abs(actual - expected) / scale) if expected.size else 0.0
    assert err <= rel, f"{what}: {err:.3e}"


def sums_close(actual, expected, rel=1e-12, what=""):
    scale = max(np.max(np.abs(expected)), 1.0)
    assert np.max(np.abs(np.asarray(actual) - expected)) <= rel * scale, what


@pytest.fixture(scope="module")
def d():
    return golden("metric")


def device_session(d, tag, precision="fp64"):
    from torch_amd import Simulation

    meta = json.loads(str(d[tag + "_meta"]))
    opts = dict(geometry=meta["geometry"], origins=meta["origin"], t_end=meta["t_end"], reactive=meta["reactive"],
                kinetic_model=meta["model"], limiter=meta["limiter"], face_boundaries=meta["faces"],
                precision=precision)
    if meta["reactive"]:
        opts["lambda_init"] = d[tag + "_lam0"]
    return Simulation(d[tag + "_prim"], meta["lengths"], **opts), meta


def test_rhs_cfl_and_advance_golden(d):
    names = list(d["rhs_cases"])
    for tag in names:
        sim, meta = device_session(d, tag)
        sim.state = d[tag + "_state"]
        nq = 1 if meta["reactive"] else 0
        for order, first in (("2", False), ("1", True)):
            change, species, outward, source = sim._device.metric_rhs(first, nq)
            close(change, d[tag + f"_change{order}"], what=f"{tag} change{order}")
            sums_close(outward, d[tag + f"_outward{order}"], what=f"{tag} outward{order}")
            close(source, d[tag + f"_source{order}"], what=f"{tag} source{order}")
            if nq:
                close(species[0], d[tag + f"_species{order}"], what=f"{tag} species{order}")
        dt = sim.dt_limit()
        assert abs(dt - float(d[tag + "_dt"])) <= 1e-12 * dt, tag
        info = sim._device.metric_advance(0.5 * float(d[tag + "_dt"]))
        sim._host = None
        close(sim.state, d[tag + "_adv_state"], what=f"{tag} advance")
        if nq:
            close(sim.lam, d[tag + "_adv_lam"], what=f"{tag} advance lam")
        sums_close(np.array(info.outward[: meta["dim"] + 2]), d[tag + "_adv_boundary"], 1e-11, f"{tag} boundary")
        sums_close(np.array(info.source[: meta["dim"] + 2]), d[tag + "_adv_source"], 1e-11, f"{tag} source")
        sim.close()


def test_first_order_retry_golden(d):
    from torch_amd import Simulation

    sim = Simulation(d["fallback_prim"], [1.0], t_end=1.0, face_boundaries={"x_min": "outflow", "x_max": "outflow"})
    info = sim._device.metric_advance(float(d["fallback_dt"]))
    sim._host = None
    assert info.fallback == 1
    close(sim.state, d["fallback_state"], what="fallback state")
    sums_close(np.array(info.outward[:3]), d["fallback_boundary"], 1e-11, "fallback boundary")


def test_sessions_golden(d):
    for tag in d["session_cases"]:
        sim, meta = device_session(d, tag)
        dts = []
        for _ in range(len(d[tag + "_dt"])):
            t0 = sim.time
            sim.step()
            dts.append(sim.time - t0)
        close(dts, d[tag + "_dt"], what=f"{tag} dt")
        close(sim.state, d[tag + "_state"], rel=1e-11, what=f"{tag} state")
        sums_close(sim.boundary_integral, d[tag + "_boundary"], 1e-10, f"{tag} boundary")
        sums_close(sim.source_integral, d[tag + "_source"], 1e-10, f"{tag} source")
        if sim.lam is not None:
            close(sim.lam, d[tag + "_lam"], rel=1e-11, what=f"{tag} lam")
        if sim.beta is not None:
            close(sim.beta, d[tag + "_beta"], rel=1e-11, what=f"{tag} beta")
        sim.close()


@pytest.mark.parametrize("geometry", om.GEOMETRIES)
def test_session_matches_oracle_every_limiter(geometry):
    """Random smooth states, each limiter, 4 steps against the oracle session."""
    from torch_amd import Simulation

    rng = np.random.default_rng(11)
    for dim in (1, 2, 3):
        counts = [10, 6, 5][:dim]
        origin = [0.4, 0.3, 0.2][:dim] if geometry != "cartesian" else [0.0] * dim
        p = np.ones((dim + 2, *counts[::-1]))
        p[0] = 1 + 0.3 * rng.random(p[0].shape)
        p[1:-1] = 0.2 * rng.standard_normal(p[1:-1].shape)
        p[-1] = 1 + 0.5 * rng.random(p[0].shape)
        faces = {"x_min": "wall", "x_max": {"type": "pressure_outlet", "pressure": 1.0,
                                            "backflow": [1.0, -0.1] + [0.0] * (dim - 1) + [1.0]}}
        for lim in ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos"):
            opts = dict(geometry=geometry, origins=origin, t_end=1.0, limiter=lim, face_boundaries=faces)
            ref = OracleSession(p, [0.8, 0.6, 0.4][:dim], **opts)
            sim = Simulation(p, [0.8, 0.6, 0.4][:dim], **opts)
            for _ in range(4):
                ref.step()
                sim.step()
            close(sim.state, ref.state, rel=1e-11, what=f"{dim}d {lim}")
            assert sim.time == pytest.approx(ref.time, rel=1e-12)
            sim.close()


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("geometry", om.GEOMETRIES)
def test_uniform_rest_preserved_at_origin_and_poles(dim, geometry):
    """testGeometry.py:112-138."""
    from torch_amd import Simulation

    origin = [0.3] * dim if dim == 3 or geometry == "annular" else [0.0] * dim
    lengths = [1.0, np.pi, 0.5][:dim] if geometry == "spherical" and dim < 3 else [0.5] * dim
    p = np.ones((dim + 2, *([6] * dim)))
    p[1:-1] = 0
    s = Simulation(p, lengths, geometry=geometry, origins=origin, t_end=0.02)
    old = s.state.copy()
    s.run()
    np.testing.assert_allclose(s.state, old, rtol=0, atol=3e-14)
    np.testing.assert_allclose(s.balance_residual, 0, atol=3e-14)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_closed_domains_conserve_mass_energy_species(dim):
    """testGeometry.py:296-339 on the spherical and annular grids."""
    from torch_amd import Simulation

    for geometry in ("spherical", "annular"):
        counts = [12, 6, 4][:dim]
        p = np.ones((dim + 2, *counts[::-1]))
        p[0] = 1 + 0.05 * np.linspace(0, 1, p[0].size).reshape(p.shape[1:])
        p[1:-1] = 0.05
        fractions = np.linspace(0.2, 0.8, p[0].size).reshape(p.shape[1:])
        walls = {a + side: "wall" for a in "xyz"[:dim] for side in ("_min", "_max")}
        s = Simulation(p, [0.8, 0.6, 0.4][:dim], geometry=geometry, origins=[0.4, 0.3, 0.2][:dim], t_end=0.01,
                       reactive=True, lambda_init=fractions, face_boundaries=walls)
        initial = s.integral.copy()
        species = np.sum(s.state[0] * s.lam * s.metrics.volumes)
        s.run()
        np.testing.assert_allclose(s.integral[[0, -1]], initial[[0, -1]], rtol=2e-13, atol=2e-13)
        assert abs(np.sum(s.state[0] * s.lam * s.metrics.volumes) - species) < 1e-12
        np.testing.assert_allclose(s.balance_residual, 0, atol=2e-13)


def test_uniform_inflow_and_pressure_outlet_on_every_face():
    """testGeometry.py:341-373."""
    from torch_amd import Simulation

    for dim in (1, 2, 3):
        for direction in range(dim):
            for side in (-1, 1):
                p = np.ones((dim + 2, *([5] * dim)))
                p[1:-1] = 0
                p[direction + 1] = -side * 0.2
                values = [1.0] + [0.0] * dim + [1.0]
                values[direction + 1] = -side * 0.2
                faces = {a + e: "periodic" for a in "xyz"[:dim] for e in ("_min", "_max")}
                axis = "xyz"[direction]
                faces[axis + ("_min" if side == -1 else "_max")] = dict(type="inflow", primitive=values,
                                                                      **{"lambda": 0.6})
                faces[axis + ("_max" if side == -1 else "_min")] = dict(type="pressure_outlet", pressure=1.0)
                for reactive in (False, True):
                    s = Simulation(p, [1.0] * dim, t_end=0.02, face_boundaries=faces, reactive=reactive,
                                   lambda_init=np.full(p.shape[1:], 0.6) if reactive else None)
                    initial = s.state.copy()
                    s.run()
                    np.testing.assert_allclose(s.state, initial, atol=1e-14, rtol=1e-14)
                    np.testing.assert_allclose(s.balance_residual, 0, atol=1e-13)


def test_boundary_errors_raise_value_error():
    from torch_amd import Simulation

    p = np.ones((3, 8))
    p[1] = 0.0
    # an outlet facing inflow with no backflow state
    p_in = p.copy()
    p_in[1] = -0.5
    s = Simulation(p_in, [1.0], t_end=0.1, face_boundaries={"x_min": "wall",
                                                            "x_max": {"type": "pressure_outlet", "pressure": 1.0}})
    with pytest.raises(ValueError, match="backflow"):
        s.step()
    assert s.steps == 0
    # a cold reservoir has no subsonic inlet state for this interior
    s = Simulation(p, [1.0], t_end=0.1, face_boundaries={
        "x_min": {"type": "stagnation_inlet", "total_pressure": 0.5, "total_temperature": 0.5, "gas_constant": 1.0},
        "x_max": "outflow"})
    with pytest.raises(ValueError, match="reservoir"):
        s.step()


def test_rusanov_flux_grid_matches_oracle():
    from torch_amd import rusanov_flux_grid

    rng = np.random.default_rng(5)
    for nv in (3, 4, 5):
        for d in range(nv - 2):
            left = np.abs(rng.standard_normal((nv, 7, 9))) + 0.5
            right = np.abs(rng.standard_normal((nv, 7, 9))) + 0.5
            left[1:-1] -= 0.7
            right[1:-1] -= 0.7
            gl, gr = 1.4, np.full((7, 9), 1.3)
            got = rusanov_flux_grid(left, right, gl, gr, d)
            ref = om.rusanov(left, right, np.broadcast_to(gl, (7, 9)), gr, d)
            close(got, ref, what=f"nv={nv} d={d}")


def test_metric_restart_is_exact(tmp_path, d):
    from torch_amd import read_checkpoint

    for tag in ("sess12", "sess13"):
        sim, _ = device_session(d, tag)
        continuous = copy.deepcopy(sim)
        for _ in range(4):
            continuous.step()
        sim.run(step_limit=2, checkpoint=tmp_path / f"{tag}.npz")
        resumed = read_checkpoint(tmp_path / f"{tag}.npz")
        assert resumed.uses_metric_kernel
        for _ in range(2):
            resumed.step()
        np.testing.assert_array_equal(resumed.state, continuous.state)
        np.testing.assert_array_equal(resumed.lam, continuous.lam)
        np.testing.assert_array_equal(resumed.source_integral, continuous.source_integral)
        assert resumed.time == continuous.time


def test_fp32_metric_session_tracks_fp64(d):
    for tag in ("sess4", "sess9"):
        a, _ = device_session(d, tag, "fp64")
        b, _ = device_session(d, tag, "fp32")
        for _ in range(5):
            a.step()
            b.step()
        err = np.abs(b.state - a.state).reshape(a.state.shape[0], -1).sum(axis=1) / np.abs(a.state).reshape(
            a.state.shape[0], -1).sum(axis=1)
        assert np.all(err < 1e-5), err


def test_session_step_through_the_first_order_retry():
    """A rough cylindrical state at CFL 0.9 (unchecked): some steps lose
    positivity at second order and retry at first order inside the step;
    every step equals the oracle's, fallbacks included."""
    from torch_amd import Simulation

    rng = np.random.default_rng(1)
    n = 64
    prim = np.array([0.2 + 1.8 * rng.random(n), 2 * rng.standard_normal(n), 0.1 + 4.9 * rng.random(n)])
    opts = dict(geometry="cylindrical", origins=[0.5], t_end=10.0, cfl=0.9, checked=False)
    dev = Simulation(prim, [1.0], **opts)
    ref = OracleSession(prim, [1.0], **opts)
    for k in range(10):
        ref.step()
        dev.step()
        assert dev.fallbacks == ref.fallbacks, k
        assert dev.time == pytest.approx(ref.time, rel=1e-12)
        close(dev.state, ref.state, rel=1e-11, what=f"step {k}")
    assert dev.fallbacks > 0
    sums_close(dev.boundary_integral, ref.boundary_integral, rel=1e-11)
    dev.close()


@pytest.mark.parametrize("geometry,dim,n", [("cylindrical", 2, 48), ("spherical", 3, 24)])
def test_fp32_reactive_metric_follows_the_fp64_steps(geometry, dim, n):
    """Request 13 of the served workload (Arrhenius constants, Gaussian hot
    spot) on a cylindrical (r, theta) grid from r = 0.5 and a spherical wedge
    r in [0.5, 1.5], theta in [0.5, 1.5], run to t_run in the FP32 form and
    the FP64 method. The FP32 attempts check the committed state's rates
    raised by the heating's exact change, and the advance factor
    exp(-rate dt) is evaluated to ~1e-13, so a far-field cell that reacts by
    1e-12 heats as it does in FP64: both take the same steps and dt halvings
    (without these the spherical FP32 run took 52 steps with 5 halvings
    against FP64's 98 and 97). Bar: 2e-5 relative L1 per field (measured
    2.3e-6 and 9.1e-6)."""
    from torch_amd import Simulation
    from torch_amd.workloads import served_requests

    spec = served_requests(seed=355, count=64)[13]
    c = (np.arange(n) + 0.5) / n
    grids = np.meshgrid(*([c] * dim), indexing="ij")[::-1]
    r2 = sum((g - spec["centre"][k]) ** 2 for k, g in enumerate(grids))
    prim = np.zeros((dim + 2,) + r2.shape)
    prim[0] = 1.0
    prim[-1] = 1.0 + (spec["t_hat"] - 1.0) * np.exp(-r2 / (2.0 * spec["sigma"] ** 2))
    out = {}
    for precision in ("fp64", "fp32"):
        sim = Simulation(prim, [1.0] * dim, geometry=geometry, origins=[0.5, 0.5, 0.0][:dim], t_end=spec["t_run"],
                         reactive=True, kinetic_model=spec["model"], lambda_init=np.ones(r2.shape),
                         constants=spec["constants"], cfl=0.4, precision=precision)
        try:
            retries = 0
            while not sim.completed:
                retries += sim.step().retries
            out[precision] = (sim.state.copy(), sim.steps, retries)
        finally:
            sim.close()
    (s32, n32, r32), (s64, n64, r64) = out["fp32"], out["fp64"]
    l1 = [float(np.abs(s32[k] - s64[k]).sum() / np.abs(s64[k]).sum()) for k in range(dim + 2)]
    print(f"{geometry} {dim}D: {n64} steps, {r64} halvings (FP32 {n32}, {r32}); L1 {['%.2e' % v for v in l1]}",
          flush=True)
    assert n32 == n64 and r32 == r64
    assert max(l1) <= 2e-5
