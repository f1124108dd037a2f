"""Write the golden input/output vectors under tests/golden/.

This script is the only place that executes the reference implementation
(TORCH's pure-NumPy ``cfdSolver`` package, read from /root/reference/python).
It runs in the development container, never on a GPU host, and writes small
``.npz`` files holding inputs and the reference's outputs. Tests compare the
oracle (and through it the HIP path) with these vectors.

Run:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_fixtures.py
"""

import os
import sys

import numpy as np

sys.dont_write_bytecode = True
REFERENCE = "/root/reference/python"
OUT = os.path.dirname(os.path.abspath(__file__))


def _load_reference():
    sys.path.insert(0, REFERENCE)
    import cfdSolver.torchFlux as flux  # noqa: E402
    import cfdSolver.torchSolver as one  # noqa: E402
    import cfdSolver.torchSolver2D as two  # noqa: E402
    import cfdSolver.torchSolver3D as three  # noqa: E402
    from cfdSolver.torchSession import Simulation  # noqa: E402

    return flux, one, two, three, Simulation


class _nan_hllc:
    """Replace the reference HLLC kernel with one returning NaN, as the
    reference's own fallback test does (python/tests/testSolver3D.py:301-320).
    """

    def __enter__(self):
        import cfdSolver.torchFlux as flux

        self._flux = flux
        self._saved = flux.hllc_flux
        flux.hllc_flux = lambda left, right, *a, **k: np.full_like(left, np.nan)

    def __exit__(self, *exc):
        self._flux.hllc_flux = self._saved


def _cons(prim, gamma):
    """Primitive (rho, velocities..., p) -> conservative, any dimension."""
    prim = np.asarray(prim, dtype=float)
    out = prim.copy()
    out[1:-1] *= prim[0]
    out[-1] = prim[-1] / (gamma - 1) + 0.5 * prim[0] * np.sum(
        prim[1:-1] ** 2, axis=0
    )
    return out


def hllc_cases(flux):
    rng = np.random.default_rng(20261203)
    data = {}
    for nv in (4, 5):
        m = 256
        prim_l = np.empty((nv, m))
        prim_r = np.empty((nv, m))
        for prim in (prim_l, prim_r):
            prim[0] = rng.uniform(0.1, 2.0, m)
            prim[1:-1] = rng.uniform(-2.0, 2.0, (nv - 2, m))
            prim[-1] = rng.uniform(0.05, 5.0, m)
        # Edge cases in the first columns: identical states, supersonic
        # both ways, a strong jump, zero velocity contact, floored states.
        prim_r[:, 0] = prim_l[:, 0]
        prim_l[1, 1], prim_r[1, 1] = 6.0, 6.0
        prim_l[1, 2], prim_r[1, 2] = -6.0, -6.0
        prim_l[:, 3] = [1.0] + [0.0] * (nv - 2) + [1000.0]
        prim_r[:, 3] = [1.0] + [0.0] * (nv - 2) + [0.01]
        prim_l[:, 4] = [1.0] + [0.0] * (nv - 2) + [1.0]
        prim_r[:, 4] = [0.125] + [0.0] * (nv - 2) + [1.0]
        prim_l[:, 5] = [1.0] + [-3.0] + [0.0] * (nv - 3) + [0.4]
        prim_r[:, 5] = [1.0] + [3.0] + [0.0] * (nv - 3) + [0.4]
        gl = rng.choice([1.4, 1.67, 1.3], m)
        gr = gl.copy()
        gr[m // 2:] = rng.uniform(1.2, 1.67, m - m // 2)
        ul = _cons(prim_l, gl)
        ur = _cons(prim_r, gr)
        # Floors: nonpositive density / internal energy in a few columns.
        ul[0, 6] = 1e-13
        ur[-1, 7] = 0.5 * np.sum(ur[1:-1, 7] ** 2) / ur[0, 7] - 1e-3
        out = flux.hllc_flux(ul, ur, gl, gamma_right=gr)
        data[f"nv{nv}_ul"], data[f"nv{nv}_ur"] = ul, ur
        data[f"nv{nv}_gl"], data[f"nv{nv}_gr"] = gl, gr
        data[f"nv{nv}_flux"] = out
        scalar = flux.hllc_flux(ul, ur, 1.4)
        data[f"nv{nv}_flux_g14"] = scalar
    # Grid driver: strided inputs, every direction, several batch sizes.
    for dim in (2, 3):
        shape = (dim + 2,) + tuple([5, 6, 7][:dim])
        p = np.empty(shape)
        p[0] = rng.uniform(0.2, 2, shape[1:])
        p[1:-1] = rng.uniform(-1, 1, (dim,) + shape[1:])
        p[-1] = rng.uniform(0.1, 3, shape[1:])
        g = rng.uniform(1.2, 1.6, shape[1:])
        left = _cons(p, g)
        right = _cons(p[:, ::-1] * 1.03, g[::-1])
        data[f"grid{dim}_left"], data[f"grid{dim}_right"] = left, right
        data[f"grid{dim}_gl"], data[f"grid{dim}_gr"] = g, g[::-1].copy()
        for direction in range(dim):
            data[f"grid{dim}_dir{direction}"] = flux.hllc_flux_grid(
                left, right, g, g[::-1], direction=direction, batch_size=17
            )
    np.savez_compressed(os.path.join(OUT, "hllc.npz"), **data)


def slope_cases(three):
    """Limited slopes of each limiter around the r branches."""
    rng = np.random.default_rng(7)
    padded = rng.normal(size=(5, 6, 7, 8))
    padded[:, 2, :, :] = padded[:, 3, :, :]  # plus == 0 rows
    padded[1, :, 3, :] = 1.0  # flat component
    padded[2, :, :, 4] = padded[2, :, :, 3] + 5e-15  # |plus| < 1e-14
    data = {"padded": padded}
    for limiter in three.LIMITERS:
        for axis in (1, 2, 3):
            data[f"{limiter}_axis{axis}"] = three._slope(padded, axis, limiter)
    np.savez_compressed(os.path.join(OUT, "slope.npz"), **data)


def base_state_3d(rng, nz, ny, nx, gamma=1.4):
    z, y, x = np.meshgrid(
        (np.arange(nz) + 0.5) / nz,
        (np.arange(ny) + 0.5) / ny,
        (np.arange(nx) + 0.5) / nx,
        indexing="ij",
    )
    p = np.empty((5, nz, ny, nx))
    inside = (x - 0.45) ** 2 + (y - 0.5) ** 2 + (z - 0.55) ** 2 <= 0.3**2
    p[0] = np.where(inside, 1.0, 0.3) + 0.1 * rng.random((nz, ny, nx))
    p[1] = 0.3 * np.sin(2 * np.pi * y) + 0.1 * rng.normal(size=x.shape)
    p[2] = -0.2 * np.cos(2 * np.pi * z) + 0.1 * rng.normal(size=x.shape)
    p[3] = 0.25 * np.sin(2 * np.pi * x) + 0.1 * rng.normal(size=x.shape)
    p[4] = np.where(inside, 2.0, 0.4) + 0.1 * rng.random((nz, ny, nx))
    return _cons(p, gamma)


def hydro3d_cases(three, one):
    rng = np.random.default_rng(355)
    nz, ny, nx = 5, 6, 10
    state = base_state_3d(rng, nz, ny, nx)
    spacing = (1.0 / nx, 1.1 / ny, 0.9 / nz)
    data = {"state": state, "spacing": np.array(spacing)}
    bcs = {
        "open": ("Open", "Open", "Open"),
        "periodic": ("Periodic", "Periodic", "Periodic"),
        "reflective": ("Reflective", "Reflective", "Reflective"),
        "mixed": ("Open", "Reflective", "Periodic"),  # x, y, z
    }
    dt = 0.004
    data["dt"] = np.array(dt)
    for limiter in three.LIMITERS:
        for key, (bx, by, bz) in bcs.items():
            res = three.muscl_hancock_step_3d(
                state, *spacing, dt, 1.4, limiter, bx, by, bz, True
            )
            tag = f"{limiter}_{key}"
            data[tag + "_U"] = res[0]
            if key in ("mixed", "reflective") or limiter == "VanLeer":
                data[tag + "_Fx"], data[tag + "_Fy"], data[tag + "_Fz"] = res[1:]
    # Variable gamma (reactive mixtures pass gamma(lambda) per cell).
    lam = rng.random((nz, ny, nx))
    gamma = one.gamma_calc(lam)
    vstate = base_state_3d(rng, nz, ny, nx, gamma)
    res, fluxes = three._hydro(
        vstate, gamma, spacing, dt, "Superbee", ("Periodic", "Reflective", "Open")
    )
    data["vargamma_state"], data["vargamma_lam"] = vstate, lam
    data["vargamma_U"] = res
    data["vargamma_Fx"], data["vargamma_Fy"], data["vargamma_Fz"] = fluxes
    # Species transport with the variable-gamma fluxes.
    q = rng.random((nz, ny, nx))
    data["advect_q"] = q
    data["advect_out"] = three._advect(
        q, vstate, res, fluxes, dt, spacing, ("Periodic", "Reflective", "Open")
    )
    # Whole-step first-order fallback on a rough random state (the
    # second-order attempt is unphysical, the first-order one is not).
    frng = np.random.default_rng(1)
    p = np.ones((5, 4, 4, 12))
    p[0] = frng.uniform(0.05, 1, (4, 4, 12))
    p[1:4] = frng.uniform(-3, 3, (3, 4, 4, 12))
    p[4] = frng.uniform(0.01, 1, (4, 4, 12))
    fstate = _cons(p, 1.4)
    fdt = 0.004
    calls = []
    original = three._hydro

    def tracked(*args, **kwargs):
        calls.append(kwargs.get("first_order", args[6] if len(args) > 6 else False))
        return original(*args, **kwargs)

    three._hydro = tracked
    try:
        fres, ffl = three._hydro(
            fstate, 1.4, (1 / 12, 0.25, 0.25), fdt, "Unos", ("Open",) * 3
        )
    finally:
        three._hydro = original
    data["fallback_state"], data["fallback_dt"] = fstate, np.array(fdt)
    data["fallback_U"] = fres
    data["fallback_is_first_order"] = np.array(True in calls)
    data["fallback_Fx"], data["fallback_Fy"], data["fallback_Fz"] = ffl
    # First-order path on the base state with every boundary kind, forced
    # the way the reference's own test forces it: NaN interface fluxes.
    with _nan_hllc():
        res = three.muscl_hancock_step_3d(
            state, *spacing, dt, 1.4, "VanLeer", "Reflective", "Periodic",
            "Open", True
        )
    data["forced_U"] = res[0]
    data["forced_Fx"], data["forced_Fy"], data["forced_Fz"] = res[1:]
    np.savez_compressed(os.path.join(OUT, "hydro3d.npz"), **data)


def reaction_cases(three, one):
    rng = np.random.default_rng(11)
    shape = (3, 4, 2)
    p = np.empty((5,) + shape)
    p[0] = rng.uniform(0.5, 2.0, shape)
    p[1:4] = rng.uniform(-50, 50, (3,) + shape)
    p[4] = rng.uniform(2e5, 4e6, shape)
    lam = rng.random(shape)
    beta = rng.random(shape)
    state = _cons(p, one.gamma_calc(lam))
    data = {"state": state, "lam": lam, "beta": beta}
    for model in ("arrhenius", "znd"):
        for dt in (1e-7, 1e-5):
            out, nl, nb = three.reaction_step_3d(
                state, lam, beta if model == "znd" else None, dt, model
            )
            tag = f"{model}_{dt:g}"
            data[tag + "_U"], data[tag + "_lam"] = out, nl
            if nb is not None:
                data[tag + "_beta"] = nb
    # Nondimensional constants passed straight to the kinetics.
    normal = np.stack((state[0].ravel(), state[1].ravel(),
                       (1 + rng.random(lam.size)) * 3.0))
    normal[2] = normal[2] + 0.5 * normal[1] ** 2 / normal[0]
    for model in ("arrhenius", "znd"):
        l2, s2, b2 = one.new_lambda_n_e_calc(
            lam.ravel(), 0.013, normal, model, beta.ravel(),
            q_heat=37.0, activation_energy=0.21, pre_exponential=55.0,
        )
        data[f"nondim_{model}_lam"], data[f"nondim_{model}_state"] = l2, s2
        data[f"nondim_{model}_beta"] = b2
    data["nondim_normal"] = normal
    np.savez_compressed(os.path.join(OUT, "reaction.npz"), **data)


def hydro2d_cases(two, one):
    rng = np.random.default_rng(99)
    ny, nx = 7, 10
    y, x = np.meshgrid((np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx,
                       indexing="ij")
    p = np.empty((4, ny, nx))
    inside = (x - 0.4) ** 2 + (y - 0.55) ** 2 <= 0.3**2
    p[0] = np.where(inside, 1.0, 0.3) + 0.1 * rng.random((ny, nx))
    p[1] = 0.3 * np.sin(2 * np.pi * y) + 0.1 * rng.normal(size=x.shape)
    p[2] = 0.2 * np.cos(2 * np.pi * x) + 0.1 * rng.normal(size=x.shape)
    p[3] = np.where(inside, 2.0, 0.4) + 0.1 * rng.random((ny, nx))
    state = _cons(p, 1.4)
    data = {"state": state}
    dx, dy, dt = 1.0 / nx, 1.2 / ny, 0.004
    data["spacing_dt"] = np.array([dx, dy, dt])
    inlet = [0.9, 0.05, 0.3, 0.8]
    data["inlet"] = np.array(inlet)
    bcs = {
        "open": ("Open", "Open"),
        "periodic": ("Periodic", "Periodic"),
        "reflective": ("Reflective", "Reflective"),
        "inflow": ("Periodic", "Inflow"),
    }
    for limiter in ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos"):
        for key, (bx, by) in bcs.items():
            u, fx, fy = two.muscl_hancock_step_2d(
                state, dx, dy, dt, 1.4, limiter, bx, by, True, inlet
            )
            tag = f"{limiter}_{key}"
            data[tag + "_U"], data[tag + "_Fx"], data[tag + "_Fy"] = u, fx, fy
    lam = rng.random((ny, nx))
    g = one.gamma_calc(lam)
    vstate = _cons(p, g)
    u, fx, fy = two.muscl_hancock_step_2d(
        vstate, dx, dy, dt, g, "Unos", "Reflective", "Inflow", True, inlet
    )
    data["vargamma_state"], data["vargamma_lam"] = vstate, lam
    data["vargamma_U"], data["vargamma_Fx"], data["vargamma_Fy"] = u, fx, fy
    q = rng.random((ny, nx))
    data["advect_q"] = q
    data["advect_out"] = two.advect_scalar_2d(
        q, vstate, u, fx[0], fy[0], dt / dx, dt / dy, "Reflective", "Inflow"
    )
    # CFL over the padded state, as the session computes it.
    for key, (bx, by) in bcs.items():
        padded = two.apply_boundary_conditions_2d(vstate, bx, by, inlet)
        gp = two.pad_scalar_2d(np.broadcast_to(g, (ny, nx)), bx, by,
                               inlet_value=1.4)
        data[f"cfl_{key}"] = np.array(two.compute_cfl_dt_2d(
            two.uconsv_to_uprim_2d(padded, gp), dx, dy, 0.4, gp))
    # First-order LLF path, forced with NaN interface fluxes.
    with _nan_hllc():
        fu, ffx, ffy = two.muscl_hancock_step_2d(
            vstate, dx, dy, dt, g, "VanLeer", "Reflective", "Inflow", True,
            inlet
        )
    data["forced_U"], data["forced_Fx"], data["forced_Fy"] = fu, ffx, ffy
    np.savez_compressed(os.path.join(OUT, "hydro2d.npz"), **data)


def session_cases(Simulation, one):
    """Multi-step session trajectories (dt, time, state, balances)."""
    data = {}

    def record(tag, sim, steps):
        dts = []
        for _ in range(steps):
            t0 = sim.time
            sim.step()
            dts.append(sim.time - t0)
        data[tag + "_dt"] = np.array(dts)
        data[tag + "_state"] = sim.state
        data[tag + "_boundary"] = sim.boundary_integral
        data[tag + "_source"] = sim.source_integral
        if sim.lam is not None:
            data[tag + "_lam"] = sim.lam
        if sim.beta is not None:
            data[tag + "_beta"] = sim.beta

    record("sod3d", Simulation.preset("3d_sod_x", [32, 4, 4], limiter="MINBEE",
                                      boundaries=["Open", "Reflective",
                                                  "Reflective"], t_end=0.2),
           20)
    record("explosion3d", Simulation.preset("3d_explosion", [8, 8, 8],
                                            limiter="Unos", t_end=0.2), 6)
    # Hot, fast-reacting mixture: the heated half-step raises the sound
    # speed enough to exercise the reactive dt-halving retry.
    z, y, x = np.meshgrid((np.arange(4) + 0.5) / 4, (np.arange(5) + 0.5) / 5,
                          (np.arange(6) + 0.5) / 6, indexing="ij")
    hot = np.zeros((5, 4, 5, 6))
    hot[0] = 1.0
    hot[4] = 1e8 * (1 + 0.5 * np.exp(-((x - 0.3) ** 2 + (y - 0.5) ** 2
                                       + (z - 0.5) ** 2) / 0.05))
    for model in ("arrhenius", "znd"):
        lam0 = 1.0 - 0.2 * x
        sim = Simulation(hot, [1.0, 1.0, 1.0], t_end=1e-5, reactive=True,
                         kinetic_model=model, lambda_init=lam0,
                         boundaries=["Open", "Periodic", "Reflective"])
        data[f"ignition3d_{model}_lam0"] = lam0
        record(f"ignition3d_{model}", sim, 4)
    data["ignition3d_prim"] = hot
    # 2D seeded quadrant problem.
    rng = np.random.default_rng(20261203)
    n = 24
    states = [
        (rng.uniform(0.1, 1.5), rng.uniform(-1, 1), rng.uniform(-1, 1),
         rng.uniform(0.03, 1.5))
        for _ in range(4)
    ]
    y, x = np.meshgrid((np.arange(n) + 0.5) / n, (np.arange(n) + 0.5) / n,
                       indexing="ij")
    quad = np.where(x >= 0.5, np.where(y >= 0.5, 0, 3), np.where(y >= 0.5, 1, 2))
    prim = np.empty((4, n, n))
    for k, s in enumerate(states):
        for c in range(4):
            prim[c][quad == k] = s[c]
    data["quadrant_states"] = np.array(states)
    record("quadrant2d", Simulation(prim, [1.0, 1.0], t_end=0.25,
                                    limiter="Superbee"), 12)
    rde = Simulation.preset("2d_rde_chamber", [12, 10])
    record("rde2d", rde, 3)
    # 1D legacy scheme, inert and reactive, several limiters and walls.
    for limiter in ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos"):
        record(f"sod1d_{limiter}", Simulation.preset(
            "sod_standard", [50], limiter=limiter), 8)
    record("sod1d_reflective", Simulation.preset(
        "sod_standard", [40], boundaries=["Reflective"]), 8)
    record("sod1d_periodic", Simulation.preset(
        "rarefaction_wave", [40], boundaries=["Periodic"]), 8)
    record("cj1d", Simulation.preset("detonation_cj", [40]), 5)
    record("cj1d_znd", Simulation.preset("detonation_cj", [40],
                                         kinetic_model="znd"), 5)
    # Legacy 1D Sod driver at CFL 0.5 to t=0.2 (final state only).
    for n in (50, 100):
        u, t = one.run_simulation(1.0, n, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1,
                                  c_cfl=0.5, t_corte=0.2)
        data[f"legacy_sod_{n}"] = u[:, :, -1]
        data[f"legacy_sod_{n}_steps"] = np.array(len(t) - 1)
    np.savez_compressed(os.path.join(OUT, "session.npz"), **data)


def main():
    flux, one, two, three, Simulation = _load_reference()
    hllc_cases(flux)
    slope_cases(three)
    hydro3d_cases(three, one)
    reaction_cases(three, one)
    hydro2d_cases(two, one)
    session_cases(Simulation, one)
    for name in sorted(os.listdir(OUT)):
        if name.endswith(".npz"):
            print(name, os.path.getsize(os.path.join(OUT, name)))


if __name__ == "__main__":
    main()
