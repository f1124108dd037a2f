"""The metric/face oracle against the reference's vectors and invariants.

Golden vectors (tests/golden/metric.npz, written by make_metric_fixtures.py
from the reference itself) pin metrics, ghost states, rhs at both orders,
cfl_dt, SSPRK2 advance and multi-step sessions; the invariants restate the
reference's own geometry tests (python/tests/testGeometry.py:70-373): exact
volumes, uniform rest preserved at the radial origin and poles, closed
domains conserving mass, energy and species, uniform inflow/outflow
preserved on every face, and the characteristic boundary relations.
"""

import json

import numpy as np
import pytest

from conftest import golden
from oracle import metric
from oracle.session import Session

REL = 1e-12


def close(actual, expected, rel=REL, what=""):
    actual, expected = np.asarray(actual, dtype=float), np.asarray(expected, dtype=float)
    assert actual.shape == expected.shape, what
    scale = np.maximum(np.max(np.abs(expected), axis=tuple(range(1, expected.ndim)), keepdims=True)
                       if expected.ndim > 1 else np.max(np.abs(expected)), 1e-300)
    err = np.max(np.abs(actual - expected) / scale) if expected.size else 0.0
    assert err <= rel, f"{what}: {err:.3e}"


def session_of(d, tag, **extra):
    meta = json.loads(str(d[tag + "_meta"]))
    opts = dict(geometry=meta["geometry"], origins=meta["origin"], t_end=meta["t_end"], reactive=meta["reactive"],
                kinetic_model=meta["model"], limiter=meta["limiter"], face_boundaries=meta["faces"])
    if meta["reactive"]:
        opts["lambda_init"] = d[tag + "_lam0"]
    opts.update(extra)
    return Session(d[tag + "_prim"], meta["lengths"], **opts), meta


@pytest.fixture(scope="module")
def d():
    return golden("metric")


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("geometry", metric.GEOMETRIES)
def test_metrics_golden(d, dim, geometry):
    tag = f"metrics_{dim}d_{geometry}"
    s, _ = session_of(d, tag)
    np.testing.assert_array_equal(s.metrics.volumes, d[tag + "_volumes"])
    for k in range(dim):
        np.testing.assert_array_equal(s.metrics.areas[k], d[tag + f"_area{k}"])
        np.testing.assert_array_equal(s.metrics.divergence[k], d[tag + f"_div{k}"])


def test_ghost_golden(d):
    cases = json.loads(str(d["ghost_cases"]))
    assert len(cases) >= 12
    for case in cases:
        p, g, q = metric.ghost(np.array(case["inside"]), d["ghost_gamma"],
                               [np.array([0.5, 0.6, 0.7, 0.8])] if case["reactive"] else [],
                               case["face"], 0, case["side"], case["reactive"])
        close(p, case["p"], what=str(case["face"]))
        close(g, case["g"], what="gamma")
        for a, b in zip(q, case["q"]):
            close(a, b, what="q")


def test_rhs_cfl_and_advance_golden(d):
    names = list(d["rhs_cases"])
    assert len(names) >= 40
    for tag in names:
        s, meta = session_of(d, tag)
        state = d[tag + "_state"]
        fr = [d[tag + "_lam0"]] if meta["reactive"] else []
        for order, first in (("2", False), ("1", True)):
            change, species, outward, source = metric.rhs(state, fr, s.config, s.metrics, s.faces, first)
            close(change, d[tag + f"_change{order}"], what=f"{tag} change{order}")
            scale = np.max(np.abs(d[tag + f"_outward{order}"])) + 1e-300
            assert np.max(np.abs(outward - d[tag + f"_outward{order}"])) <= 1e-12 * max(scale, 1.0), tag
            close(source, d[tag + f"_source{order}"], what=f"{tag} source")
            if fr:
                close(species[0], d[tag + f"_species{order}"], what=f"{tag} species")
        lam = fr[0] if fr else None
        dt = metric.cfl_dt(state, lam, s.config, s.metrics, s.faces)
        assert abs(dt - float(d[tag + "_dt"])) <= 1e-12 * dt, tag
        res = metric.advance(state, lam, None, 0.5 * dt, s.config, s.metrics, s.faces)
        close(res[0], d[tag + "_adv_state"], what=f"{tag} advance")
        if lam is not None:
            close(res[1], d[tag + "_adv_lam"], what=f"{tag} advance lam")
        close(res[3], d[tag + "_adv_boundary"], rel=1e-11, what=f"{tag} boundary")


def test_first_order_retry_golden(d):
    """advance takes its first-order retry where the second-order stages lose positivity."""
    s = Session(d["fallback_prim"], [1.0], t_end=1.0, face_boundaries={"x_min": "outflow", "x_max": "outflow"})
    res = metric.advance(s.state, None, None, float(d["fallback_dt"]), s.config, s.metrics, s.faces)
    assert res[5], "expected the first-order retry"
    close(res[0], d["fallback_state"], what="fallback state")
    close(res[3], d["fallback_boundary"], rel=1e-11, what="fallback boundary")


def test_sessions_golden(d):
    for tag in d["session_cases"]:
        s, meta = session_of(d, tag)
        dts = []
        for _ in range(len(d[tag + "_dt"])):
            t0 = s.time
            s.step()
            dts.append(s.time - t0)
        close(dts, d[tag + "_dt"], what=f"{tag} dt")
        close(s.state, d[tag + "_state"], rel=1e-11, what=f"{tag} state")
        for name in ("boundary", "source"):
            ref = d[f"{tag}_{name}"]
            got = s.boundary_integral if name == "boundary" else s.source_integral
            assert np.max(np.abs(got - ref)) <= 1e-10 * max(np.max(np.abs(ref)), 1.0), (tag, name)
        if s.lam is not None:
            close(s.lam, d[tag + "_lam"], rel=1e-11, what=f"{tag} lam")


# -- invariants (testGeometry.py) -----------------------------------------------

def _example(dim, geometry, reactive=False, **extra):
    counts = [12, 6, 4][:dim]
    origin = [0.4, 0.3, 0.2][:dim] if geometry != "cartesian" else [0.1] * dim
    lengths = [0.8, 0.6, 0.4][:dim]
    p = np.ones((dim + 2, *counts[::-1]))
    p[1:-1] = 0.15
    faces = {"x_min": {"type": "inflow", "primitive": [1.0, 0.15] + [0.15] * (dim - 1) + [1.0]},
             "x_max": {"type": "pressure_outlet", "pressure": 1.0, "backflow": [1.0, -0.1] + [0.0] * (dim - 1) + [1.0]}}
    opts = dict(geometry=geometry, origins=origin, t_end=0.025, face_boundaries=faces, reactive=reactive)
    opts.update(extra)
    return Session(p, lengths, **opts), origin, lengths


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("geometry", metric.GEOMETRIES)
def test_exact_volumes(dim, geometry):
    """testGeometry.py:70-110: total volume in closed form."""
    s, origin, lengths = _example(dim, geometry)
    r0, r1 = origin[0], origin[0] + 0.8
    if geometry == "cartesian":
        expected = np.prod(lengths)
    elif geometry == "spherical":
        angular = 4 * np.pi if dim == 1 else (np.cos(0.3) - np.cos(0.9)) * (2 * np.pi if dim == 2 else 0.4)
        expected = (r1 ** 3 - r0 ** 3) / 3 * angular
    else:
        other = 2 * np.pi if dim == 1 else (2 * np.pi * 0.6 if geometry == "cylindrical" and dim == 2
                                             else 0.6 * (0.4 if dim == 3 else 1))
        expected = (r1 * r1 - r0 * r0) / 2 * other
    assert abs(s.metrics.volumes.sum() - expected) < 1e-13


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("geometry", metric.GEOMETRIES)
def test_uniform_rest_preserved_at_origin_and_poles(dim, geometry):
    """testGeometry.py:112-138: pressure forces balance the geometric source."""
    origin = [0.3] * dim if dim == 3 or geometry == "annular" else [0.0] * dim
    lengths = [1.0, np.pi, 0.5][:dim] if geometry == "spherical" and dim < 3 else [0.5] * dim
    p = np.ones((dim + 2, *([6] * dim)))
    p[1:-1] = 0
    s = Session(p, lengths, geometry=geometry, origins=origin, t_end=0.02)
    old = s.state.copy()
    while not s.completed:
        s.step()
    np.testing.assert_allclose(s.state, old, rtol=0, atol=3e-14)
    np.testing.assert_allclose(s.balance_residual, 0, atol=3e-14)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("geometry", metric.GEOMETRIES)
def test_closed_domains_conserve_mass_energy_species(dim, geometry):
    """testGeometry.py:296-339."""
    s, origin, lengths = _example(dim, geometry)
    p = np.ones_like(s.state)
    p[0] = 1 + 0.05 * np.linspace(0, 1, p[0].size).reshape(p.shape[1:])
    p[1:-1] = 0.05
    fractions = np.linspace(0.2, 0.8, p[0].size).reshape(p.shape[1:])
    walls = {a + side: "wall" for a in "xyz"[:dim] for side in ("_min", "_max")}
    s = Session(p, lengths, geometry=geometry, origins=origin, t_end=0.01, reactive=True, lambda_init=fractions,
                face_boundaries=walls)
    initial = s.integral.copy()
    species = np.sum(s.state[0] * s.lam * s.metrics.volumes)
    while not s.completed:
        s.step()
    np.testing.assert_allclose(s.integral[[0, -1]], initial[[0, -1]], rtol=2e-13, atol=2e-13)
    assert abs(np.sum(s.state[0] * s.lam * s.metrics.volumes) - species) < 1e-12
    np.testing.assert_allclose(s.balance_residual, 0, atol=2e-13)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_uniform_inflow_and_pressure_outlet_on_every_face(dim):
    """testGeometry.py:341-373."""
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
                s = Session(p, [1.0] * dim, t_end=0.02, face_boundaries=faces, reactive=reactive,
                            lambda_init=np.full(p.shape[1:], 0.6) if reactive else None)
                initial = s.state.copy()
                while not s.completed:
                    s.step()
                np.testing.assert_allclose(s.state, initial, atol=1e-14, rtol=1e-14)
                np.testing.assert_allclose(s.balance_residual, 0, atol=1e-13)


def test_characteristic_boundary_relations():
    """testGeometry.py:202-294: isentropic outlet, reservoir inlet, choking, backflow."""
    gamma = np.array([1.4])
    p = np.array([[1.2], [40.0], [101325.0]])
    a = np.sqrt(gamma * p[-1] / p[0])
    out, _, _ = metric.ghost(p, gamma, [], dict(type="pressure_outlet", pressure=90000.0), 0, 1, False)
    a_new = np.sqrt(gamma * out[-1] / out[0])
    np.testing.assert_allclose(out[-1], 90000.0)
    np.testing.assert_allclose(out[-1] / out[0] ** gamma, p[-1] / p[0] ** gamma)
    np.testing.assert_allclose(out[1] + 2 * a_new / (gamma - 1), p[1] + 2 * a / (gamma - 1))
    sup = p.copy()
    sup[1] = 600
    np.testing.assert_array_equal(metric.ghost(sup, gamma, [], dict(type="pressure_outlet", pressure=1.0),
                                               0, 1, False)[0], sup)
    face = dict(type="stagnation_inlet", total_pressure=110000.0, total_temperature=300.0, gas_constant=287.0)
    inside = np.array([[1.2], [20.0], [101325.0]])
    inlet, g, _ = metric.ghost(inside, gamma, [], face, 0, -1, False)
    speed, sound = inlet[1], np.sqrt(g * inlet[-1] / inlet[0])
    mach = speed / sound
    np.testing.assert_allclose(inlet[-1] * (1 + 0.5 * (g - 1) * mach ** 2) ** (g / (g - 1)), 110000.0, rtol=1e-13)
    np.testing.assert_allclose(inlet[-1] / (287 * inlet[0]) + speed ** 2 / (2 * 1.4 * 287 / 0.4), 300.0, rtol=1e-13)
    np.testing.assert_allclose(speed - 2 * sound / (g - 1),
                               inside[1] - 2 * np.sqrt(g * inside[-1] / inside[0]) / (g - 1))
    sonic, _, _ = metric.ghost(np.array([[1.0], [0.0], [1.0]]), gamma, [], face, 0, -1, False)
    assert abs(float(sonic[1, 0] / np.sqrt(1.4 * sonic[-1, 0] / sonic[0, 0])) - 1.0) < 1e-12
    with pytest.raises(ValueError):
        metric.ghost(np.array([[1.0], [-20.0], [101325.0]]), gamma, [],
                     dict(type="pressure_outlet", pressure=101325.0), 0, 1, False)
    back, _, q = metric.ghost(np.array([[1.0], [-20.0], [101325.0]]), gamma, [np.array([0.5])],
                              dict(type="pressure_outlet", pressure=101325.0, backflow=[2.0, -30.0, 101325.0],
                                   **{"lambda": 0.8}), 0, 1, True)
    np.testing.assert_array_equal(back[:, 0], [2.0, -30.0, 101325.0])
    assert q[0][0] == 0.8


def test_invalid_geometries_and_faces_rejected():
    """testGeometry.py:520-552."""
    p = np.ones((4, 4, 4))
    p[1:-1] = 0
    for opts in (dict(geometry="unknown"), dict(geometry="annular"),
                 dict(geometry="spherical", lengths=[1.0, 4.0]), dict(face_boundaries={"x_min": "periodic"}),
                 dict(face_boundaries={"bad": "wall"}),
                 dict(face_boundaries={"x_min": {"type": "inflow", "primitive": [1.0, -1.0, 0.0, 1.0]}}),
                 dict(geometry="annular", origins=[0.5, 0.0], lengths=[1.0, 2 * np.pi],
                      face_boundaries={"x_min": "periodic", "x_max": "periodic"})):
        kw = dict(opts)
        lengths = kw.pop("lengths", [1.0, 1.0])
        with pytest.raises(ValueError):
            Session(p, lengths, t_end=0.01, **kw)
    p3 = np.ones((5, 4, 4, 4))
    p3[1:-1] = 0
    for geo in ("spherical", "cylindrical"):
        with pytest.raises(ValueError):
            Session(p3, [1.0, 1.0, 1.0], geometry=geo, t_end=0.01)
