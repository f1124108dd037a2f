"""Write tests/golden/adaptive.npz: the reference's adaptive (AMR) scheme.

Like make_fixtures.py, this executes the reference implementation (TORCH's
pure-NumPy ``cfdSolver``, read from /root/reference/python) in the
development container only, and stores small inputs and outputs:

- leaf meshes (torchAdaptiveGeometry.LeafMesh) after a refinement, in 1D/2D/3D
  and every geometry: keys, the face list, areas, offsets, volumes,
  divergences and gradient matrices;
- rhs (both orders), cfl and SSPRK2 advance (torchAdaptive.py:14-206) on
  those meshes with a nonuniform state, inert and with fractions;
- multi-step AdaptiveSimulation runs with regrids (Sod along x and y,
  an entropy wave on a periodic 1D mesh, a reactive spherical case, ZND
  with induction fractions, first-order reconstruction, a wave periodic in
  x and y, inflow / pressure-outlet / wall / outflow faces in 2D, a 3D
  stagnation inlet with periodic y and z): the leaf keys, state, fractions,
  time and balances after every step.

Each case stores what the tests need to rebuild it without the reference.
Run: PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_adaptive_fixtures.py
"""

import json
import os
import sys

import numpy as np

sys.dont_write_bytecode = True
REFERENCE = "/root/reference/python"
OUT = os.path.dirname(os.path.abspath(__file__))
GEOMETRIES = ("cartesian", "cylindrical", "annular", "spherical")
MESH_FIELDS = ("keys", "left", "right", "directions", "sides", "face_areas", "offset_left", "offset_right",
               "volumes", "divergence", "displacement", "weights", "inverse", "centers", "widths")


def base(Simulation, dim, geometry, reactive=False, **amr):
    counts = [8, 4, 4][:dim]
    p = np.ones((dim + 2, *counts[::-1]))
    p[1:-1] = 0
    origins = [0.4, 0.3, 0.2][:dim] if geometry != "cartesian" else [0.0] * dim
    spec = dict(counts=counts, origins=origins, lengths=[1.0] * dim, geometry=geometry, reactive=reactive, amr=amr)
    s = Simulation(p, [1.0] * dim, origins=origins, geometry=geometry, t_end=0.02, reactive=reactive)
    return s.to_adaptive(**amr), spec


def main():
    sys.path.insert(0, REFERENCE)
    from cfdSolver.torchAdaptive import advance, cfl, rhs  # noqa: E402
    from cfdSolver.torchSession import Simulation  # noqa: E402

    data = {}
    rng = np.random.default_rng(77)
    names = []
    for dim in (1, 2, 3):
        for geometry in GEOMETRIES:
            for reactive in ((False, True) if geometry == "spherical" else (False,)):
                s, spec = base(Simulation, dim, geometry, reactive, max_level=2)
                s.regrid(refine=s.mesh.centers[:, 0] < s.mesh.origin[0] + 0.3, coarsen=False)
                s.regrid(refine=s.mesh.centers[:, 0] < s.mesh.origin[0] + 0.15, coarsen=False)
                m = s.mesh
                n = len(m.keys)
                x = (m.centers[:, 0] - m.origin[0])
                prim = np.ones((dim + 2, n))
                prim[0] = 1 + 0.2 * np.sin(5 * x) + 0.05 * rng.random(n)
                prim[1:-1] = 0.1 * rng.standard_normal((dim, n))
                prim[-1] = 1 + 0.3 * np.cos(4 * x) + 0.05 * rng.random(n)
                lam = 0.3 + 0.6 * rng.random(n) if reactive else None
                g = (4.5 - lam) / (3.5 - lam) if reactive else 1.4
                state = prim.copy()
                state[1:-1] *= prim[0]
                state[-1] = prim[-1] / (g - 1) + 0.5 * prim[0] * np.sum(prim[1:-1] ** 2, axis=0)
                tag = f"mesh{len(names)}"
                names.append(tag)
                data[tag + "_spec"] = np.array(json.dumps(spec))
                for f in MESH_FIELDS:
                    data[f"{tag}_{f}"] = np.asarray(getattr(m, f))
                data[tag + "_state"] = state
                fr = []
                if reactive:
                    data[tag + "_lam"] = lam
                    fr = [lam]
                for order, first in (("2", False), ("1", True)):
                    change, species, boundary, source = rhs(m, state, fr, s.config, s.execution, first)
                    data[f"{tag}_change{order}"] = change
                    data[f"{tag}_boundary{order}"] = boundary
                    data[f"{tag}_source{order}"] = source
                    if fr:
                        data[f"{tag}_species{order}"] = species[0]
                dt = cfl(m, state, fr, s.config)
                data[tag + "_dt"] = np.array(dt)
                final, qs, b, src = advance(m, state, fr, 0.5 * dt, s.config, s.execution)
                data[tag + "_adv_state"] = final
                data[tag + "_adv_boundary"] = b
                data[tag + "_adv_source"] = src
                if qs:
                    data[tag + "_adv_lam"] = qs[0]
    data["mesh_cases"] = np.array(names)

    runs = []
    # Sod along x (2D) and y (3D) with regrids every 2 steps.
    for dim, axis in ((2, 0), (3, 1), (1, 0)):
        counts = [4] * dim
        counts[axis] = 16
        p = np.ones((dim + 2, *counts[::-1]))
        p[1:-1] = 0
        shape = [1] * dim
        shape[dim - 1 - axis] = 16
        right = np.broadcast_to((np.arange(16) >= 8).reshape(shape), counts[::-1])
        p[0, right] = 0.125
        p[-1, right] = 0.1
        runs.append((f"sod{dim}d", p, dict(t_end=0.08), dict(max_level=1, regrid_every=2), 8))
    x = (np.arange(16) + 0.5) / 16
    p = np.array([1 + 0.2 * np.sin(2 * np.pi * x), np.full(16, 0.5), np.ones(16)])
    runs.append(("wave1d", p, dict(boundaries=["Periodic"], t_end=0.1),
                 dict(max_level=1, refine_threshold=0.02, coarsen_threshold=0.005, regrid_every=3), 8))
    p = np.ones((4, 4, 8))
    p[1:-1] = 0.1
    p[-1] = np.where((np.arange(8) + 0.5) / 8 < 0.5, 1e6, 101325.0)[None, :]
    runs.append(("react2d", p, dict(geometry="spherical", origins=[0.4, 0.3], lengths=[0.8, 0.6], t_end=2e-5,
                                    reactive=True, kinetic_model="znd"), dict(max_level=1, regrid_every=2), 5))
    # ZND with induction fractions, cartesian.
    p = np.ones((4, 4, 8))
    p[1:-1] = 0.1
    p[-1] = np.where((np.arange(8) + 0.5) / 8 < 0.5, 1e6, 101325.0)[None, :]
    runs.append(("znd2d", p, dict(lengths=[0.8, 0.4], t_end=2e-5, reactive=True, kinetic_model="znd"),
                 dict(max_level=1, regrid_every=2), 5))
    # First-order reconstruction.
    runs.append(("order1", runs[0][1], dict(t_end=0.08), dict(max_level=1, regrid_every=2, order=1), 8))
    # An entropy wave periodic in x and y: refinement reaches both seams.
    x = (np.arange(16) + 0.5) / 16
    y = (np.arange(8) + 0.5) / 8
    p = np.ones((4, 8, 16))
    wrap_x = np.minimum(x, 1 - x)  # periodic distance from the x seam
    p[0] = 1 + 0.3 * np.exp(-60 * wrap_x[None, :] ** 2 - 20 * (y[:, None] - 0.5) ** 2)
    p[1], p[2] = 0.5, 0.25
    runs.append(("periodic2d", p, dict(boundaries=["Periodic", "Periodic"], t_end=0.1),
                 dict(max_level=2, refine_threshold=0.05, coarsen_threshold=0.01, regrid_every=3), 8))
    # Inflow, pressure outlet with backflow, wall and outflow faces.
    p = np.ones((4, 4, 16))
    p[0] = 1 + 0.3 * np.exp(-40 * (x[None, :] - 0.4) ** 2)
    p[1] = 0.2
    faces = {"x_min": {"type": "inflow", "primitive": [1.2, 0.3, 0.0, 1.2]},
             "x_max": {"type": "pressure_outlet", "pressure": 0.9, "backflow": [1.0, -0.1, 0.0, 0.9]},
             "y_min": "wall", "y_max": "outflow"}
    runs.append(("faces2d", p, dict(t_end=0.05, face_boundaries=faces),
                 dict(max_level=1, refine_threshold=0.02, coarsen_threshold=0.005, regrid_every=2), 8))
    # 3D stagnation inlet and pressure outlet, periodic in y and z.
    p = np.ones((5, 4, 4, 8))
    wrap_y = np.minimum((np.arange(4) + 0.5) / 4, 1 - (np.arange(4) + 0.5) / 4)
    p[0] = 1 + 0.3 * np.exp(-30 * wrap_y ** 2)[None, :, None] * np.exp(-30 * (x[::2] - 0.5) ** 2)[None, None, :]
    faces = {"x_min": {"type": "stagnation_inlet", "total_pressure": 1.3, "total_temperature": 300.0,
                       "gas_constant": 1.0 / 300.0},
             "x_max": {"type": "pressure_outlet", "pressure": 1.0, "backflow": [1.0, -0.1, 0.0, 0.0, 1.0]},
             "y_min": "periodic", "y_max": "periodic", "z_min": "periodic", "z_max": "periodic"}
    runs.append(("stag3d", p, dict(t_end=0.02, face_boundaries=faces),
                 dict(max_level=1, refine_threshold=0.05, coarsen_threshold=0.01, regrid_every=2), 5))
    run_names = []
    for tag, p, opts, amr, steps in runs:
        dim = p.ndim - 1
        o = dict(opts)
        lengths = o.pop("lengths", [1.0] * dim)
        s = Simulation(p, lengths, **o).to_adaptive(**amr)
        run_names.append(tag)
        data[tag + "_prim"] = p
        data[tag + "_spec"] = np.array(json.dumps(dict(opts=opts, amr=amr, lengths=lengths)))
        data[tag + "_keys0"] = s.mesh.keys
        data[tag + "_state0"] = s.state
        for k in range(steps):
            s.step()
            data[f"{tag}_keys{k + 1}"] = s.mesh.keys
            data[f"{tag}_state{k + 1}"] = s.state
            data[f"{tag}_time{k + 1}"] = np.array(s.time)
            if s.lam is not None:
                data[f"{tag}_lam{k + 1}"] = s.lam
            if s.beta is not None:
                data[f"{tag}_beta{k + 1}"] = s.beta
            if s.completed:
                break
        data[tag + "_steps"] = np.array(s.steps)
        data[tag + "_boundary"] = s.boundary_integral
        data[tag + "_source"] = s.source_integral
        data[tag + "_regrids"] = np.array(s.regrids)
    data["run_cases"] = np.array(run_names)
    np.savez_compressed(os.path.join(OUT, "adaptive.npz"), **data)
    print("wrote", len(data), "arrays")


if __name__ == "__main__":
    main()
