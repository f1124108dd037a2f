"""Write tests/golden/metric.npz: the reference's metric/face finite volumes.

Like make_fixtures.py, this executes the reference implementation (TORCH's
pure-NumPy ``cfdSolver``, read from /root/reference/python) in the
development container only, and stores small inputs and outputs:

- cell volumes, face areas and divergences (torchGeometry.Metrics) for every
  geometry in 1D/2D/3D;
- ghost states of the characteristic face types (torchBoundaries.ghost);
- rhs (change, species change, outward flux, integrated source) and cfl_dt
  (torchFiniteVolume) for every geometry, limiter and face type, with and
  without transported fractions, second and first order;
- SSPRK2 ``advance``, including a case (a rough 1D state at twice its CFL
  step) whose second-order stages lose positivity so the first-order retry
  is taken;
- multi-step metric sessions (Simulation.step -> _metric_step), inert and
  reactive (Arrhenius, ZND) with walls, inflow, pressure outlets with
  backflow, stagnation inlets and periodic faces.

Each case stores its inputs (primitives, lengths, origins, geometry, a JSON
face specification, options) so the tests can rebuild it without the
reference. Run: PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_metric_fixtures.py
"""

import json
import os
import sys

import numpy as np

sys.dont_write_bytecode = True
REFERENCE = "/root/reference/python"
OUT = os.path.dirname(os.path.abspath(__file__))
LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")
GEOMETRIES = ("cartesian", "cylindrical", "annular", "spherical")


def _load():
    sys.path.insert(0, REFERENCE)
    from cfdSolver.torchBoundaries import ghost  # noqa: E402
    from cfdSolver.torchFiniteVolume import advance, cfl_dt, rhs  # noqa: E402
    from cfdSolver.torchSession import Simulation  # noqa: E402

    return Simulation, rhs, cfl_dt, advance, ghost


def _grid(dim, geometry, counts=(12, 6, 4)):
    counts = list(counts[:dim])
    origin = [0.4, 0.3, 0.2][:dim] if geometry != "cartesian" else [0.1] * dim
    lengths = [0.8, 0.6, 0.4][:dim]
    centers = [o + (np.arange(n) + 0.5) * l / n for o, l, n in zip(origin, lengths, counts)]
    grids = np.meshgrid(*centers[::-1], indexing="ij")[::-1]
    return counts, origin, lengths, grids


def _primitive(dim, geometry, reactive=False, rng=None):
    """Smooth state with a pressure jump (reactive: 1e6 / 101325 Pa)."""
    counts, origin, lengths, grids = _grid(dim, geometry)
    p = np.ones((dim + 2, *counts[::-1]))
    p[0] = 1 + 0.05 * np.sin(3 * grids[0])
    if dim > 1:
        p[0] += 0.03 * np.cos(2 * grids[1])
    if dim == 3:
        p[0] += 0.01 * grids[2]
    p[1:-1] = 0.15
    if rng is not None:
        p[1:-1] += 0.1 * rng.standard_normal(p[1:-1].shape)
        p[0] *= 1 + 0.2 * rng.random(p[0].shape)
    p[-1] = np.where(grids[0] < 0.8, 1e6, 101325.0) if reactive else np.where(grids[0] < 0.8, 1.0, 0.4)
    return p, origin, lengths


def _faces(kind, dim, reactive):
    pr = 101325.0 if reactive else 1.0
    if kind == "inout":
        return {"x_min": {"type": "inflow", "primitive": [1.0, 0.15] + [0.15] * (dim - 1) + [pr], "lambda": 0.9},
                "x_max": {"type": "pressure_outlet", "pressure": pr, "backflow": [1.0, -0.1] + [0.0] * (dim - 1) + [pr]}}
    if kind == "walls":
        return {a + s: "wall" for a in "xyz"[:dim] for s in ("_min", "_max")}
    if kind == "stagnation":
        spec = {"x_min": {"type": "stagnation_inlet", "total_pressure": 1.3 * pr, "total_temperature": 300.0,
                          "gas_constant": pr / 300.0},
                "x_max": {"type": "pressure_outlet", "pressure": pr, "backflow": [1.0, -0.1] + [0.0] * (dim - 1) + [pr]}}
        if dim >= 2:
            spec["y_min"] = "wall"
            spec["y_max"] = "outflow"
        return spec
    if kind == "periodic":
        spec = {"x_min": "wall", "x_max": "outflow"}
        for a in "yz"[:dim - 1]:
            spec[a + "_min"] = spec[a + "_max"] = "periodic"
        return spec
    return None


def _session(Simulation, dim, geometry, faces, reactive=False, model="arrhenius", limiter="VanLeer", rng=None,
             lam=None):
    p, origin, lengths = _primitive(dim, geometry, reactive, rng)
    opts = dict(geometry=geometry, origins=origin, t_end=1e-3 if reactive else 0.2, reactive=reactive,
                kinetic_model=model, limiter=limiter, face_boundaries=faces)
    if reactive:
        opts["lambda_init"] = lam if lam is not None else np.full(p.shape[1:], 0.95)
    s = Simulation(p, lengths, **opts)
    meta = dict(dim=dim, geometry=geometry, origin=origin, lengths=lengths, faces=faces, reactive=reactive,
                model=model, limiter=limiter, t_end=opts["t_end"])
    return s, p, meta


def _store(data, tag, meta, p, lam=None):
    data[tag + "_meta"] = np.array(json.dumps(meta))
    data[tag + "_prim"] = p
    if lam is not None:
        data[tag + "_lam0"] = lam


def metric_cases(Simulation, data):
    for dim in (1, 2, 3):
        for geometry in GEOMETRIES:
            s, p, meta = _session(Simulation, dim, geometry, _faces("inout", dim, False))
            tag = f"metrics_{dim}d_{geometry}"
            _store(data, tag, meta, p)
            m = s.metrics
            data[tag + "_volumes"] = m.volumes
            for d, (a, v) in enumerate(zip(m.areas, m.divergence)):
                data[tag + f"_area{d}"] = a
                data[tag + f"_div{d}"] = v


def ghost_cases(ghost, data):
    gamma = np.array([1.4, 1.4, 1.4, 1.25])
    inside = np.array([[1.2, 1.0, 0.9, 1.1], [40.0, -20.0, 600.0, 5.0], [101325.0, 101325.0, 90000.0, 120000.0]])
    out = []
    for side in (-1, 1):
        for face in (dict(type="pressure_outlet", pressure=90000.0, backflow=[2.0, -30.0 * side, 101325.0],
                          **{"lambda": 0.8}),
                     dict(type="stagnation_inlet", total_pressure=130000.0, total_temperature=320.0,
                          gas_constant=287.0),
                     dict(type="wall"), dict(type="outflow"),
                     dict(type="inflow", primitive=[1.1, -side * 30.0, 100000.0], **{"lambda": 0.7})):
            for reactive in (False, True):
                src = inside.copy()
                if face["type"] == "stagnation_inlet":
                    src[1] = -side * np.array([5.0, 20.0, 1.0, 40.0])
                q = [np.array([0.5, 0.6, 0.7, 0.8])] if reactive else []
                try:
                    p, g, qs = ghost(src, gamma, q, face, 0, side, reactive)
                except ValueError:
                    continue
                out.append(dict(face=face, side=side, reactive=reactive, inside=src.tolist(),
                                p=p.tolist(), g=np.asarray(g).tolist(), q=[x.tolist() for x in qs]))
    data["ghost_cases"] = np.array(json.dumps(out))
    data["ghost_gamma"] = gamma


def rhs_cases(Simulation, rhs, cfl_dt, advance, data):
    rng = np.random.default_rng(2027)
    cases = []
    for dim in (1, 2, 3):
        for geometry in GEOMETRIES:
            for kind in ("inout", "walls", "stagnation", "periodic"):
                if kind == "periodic" and (dim == 1 or geometry != "cartesian"):
                    continue
                if kind == "stagnation" and geometry == "spherical" and dim >= 2:
                    continue
                cases.append((dim, geometry, kind, "VanLeer", False))
    for lim in LIMITERS:
        cases.append((3, "annular", "inout", lim, False))
        cases.append((2, "spherical", "walls", lim, False))
    for dim in (1, 2, 3):
        cases.append((dim, "cylindrical", "inout", "Superbee", True))
    cases.append((3, "cartesian", "periodic", "Unos", True))
    names = []
    for i, (dim, geometry, kind, lim, reactive) in enumerate(cases):
        lam = 0.5 + 0.4 * rng.random(tuple([12, 6, 4][:dim][::-1])) if reactive else None
        s, p, meta = _session(Simulation, dim, geometry, _faces(kind, dim, reactive), reactive=reactive,
                              limiter=lim, rng=rng, lam=lam)
        tag = f"rhs{i}"
        names.append(tag)
        _store(data, tag, meta, s.primitive, s.lam)
        data[tag + "_state"] = s.state
        fr = [s.lam] if reactive else []
        for order, first in (("2", False), ("1", True)):
            change, species, outward, source = rhs(s.state, fr, s.config, s.metrics, first_order=first)
            data[tag + f"_change{order}"] = change
            data[tag + f"_outward{order}"] = outward
            data[tag + f"_source{order}"] = source
            if species:
                data[tag + f"_species{order}"] = species[0]
        dt = cfl_dt(s.state, s.lam, s.config, s.metrics)
        data[tag + "_dt"] = np.array(dt)
        res = advance(s.state, s.lam, None, 0.5 * dt, s.config, s.metrics)
        data[tag + "_adv_state"] = res[0]
        if res[1] is not None:
            data[tag + "_adv_lam"] = res[1]
        data[tag + "_adv_boundary"] = res[3]
        data[tag + "_adv_source"] = res[4]
    data["rhs_cases"] = np.array(names)


def fallback_case(Simulation, rhs, cfl_dt, advance, data):
    """A rough 1D state advanced at twice its CFL step: the second-order
    SSPRK2 stages lose positivity, the first-order retry succeeds. The first
    seeded candidate (rng 3) for which that holds is kept."""
    from cfdSolver.torchSolver import physical_state

    rng = np.random.default_rng(3)
    n = 16
    for _ in range(400):
        p = np.ones((3, n))
        p[0] = np.exp(rng.uniform(-3, 1, n))
        p[1] = rng.uniform(-3, 3, n)
        p[2] = np.exp(rng.uniform(-6, 1, n))
        s = Simulation(p, [1.0], t_end=1.0, face_boundaries={"x_min": "outflow", "x_max": "outflow"})
        c, m = s.config, s.metrics
        for mult in (2.0, 2.5):
            dt = mult * cfl_dt(s.state, None, c, m)
            stage = s.state + dt * rhs(s.state, [], c, m)[0]
            second_ok = physical_state(stage) and physical_state(
                0.5 * (s.state + stage + dt * rhs(stage, [], c, m)[0]))
            if second_ok:
                continue
            try:
                res = advance(s.state, None, None, dt, c, m)
            except RuntimeError:
                continue
            data["fallback_prim"] = p
            data["fallback_dt"] = np.array(dt)
            data["fallback_state"] = res[0]
            data["fallback_boundary"] = res[3]
            return
    raise RuntimeError("no first-order retry candidate found")


def session_cases(Simulation, data):
    runs = []
    for dim in (1, 2, 3):
        for geometry in GEOMETRIES:
            runs.append((dim, geometry, "inout", False, "arrhenius", 5))
    runs += [(3, "cylindrical", "inout", True, "arrhenius", 4), (2, "spherical", "walls", True, "znd", 4),
             (1, "annular", "inout", True, "znd", 4), (3, "cartesian", "periodic", False, "arrhenius", 6),
             (2, "annular", "stagnation", False, "arrhenius", 5)]
    names = []
    for i, (dim, geometry, kind, reactive, model, steps) in enumerate(runs):
        s, p, meta = _session(Simulation, dim, geometry, _faces(kind, dim, reactive), reactive=reactive, model=model)
        tag = f"sess{i}"
        names.append(tag)
        _store(data, tag, meta, p, s.lam)
        dts = []
        for _ in range(steps):
            t0 = s.time
            s.step()
            dts.append(s.time - t0)
        data[tag + "_dt"] = np.array(dts)
        data[tag + "_state"] = s.state
        data[tag + "_boundary"] = s.boundary_integral
        data[tag + "_source"] = s.source_integral
        if s.lam is not None:
            data[tag + "_lam"] = s.lam
        if s.beta is not None:
            data[tag + "_beta"] = s.beta
    data["session_cases"] = np.array(names)


def main():
    Simulation, rhs, cfl_dt, advance, ghost = _load()
    data = {}
    metric_cases(Simulation, data)
    ghost_cases(ghost, data)
    rhs_cases(Simulation, rhs, cfl_dt, advance, data)
    fallback_case(Simulation, rhs, cfl_dt, advance, data)
    session_cases(Simulation, data)
    np.savez_compressed(os.path.join(OUT, "metric.npz"), **data)
    print("wrote", len(data), "arrays")


if __name__ == "__main__":
    main()
