"""The oracle against vectors written by the reference itself.

The fixtures under tests/golden/ were produced by running the reference
(tests/golden/make_fixtures.py). The oracle restates the same NumPy
arithmetic, so agreement is expected to the last bit or within a few ulps;
the tolerance here (rtol 1e-13 relative to each field's scale) is far below
the 1e-12 bar the HIP path is held to.
"""

import numpy as np
import pytest

from conftest import golden
from oracle import flux, presets, session, solver1d as one, solver2d as two, solver3d as three

LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")
TOL = 1e-13


def close(actual, expected, tol=TOL):
    actual, expected = np.asarray(actual), np.asarray(expected)
    assert actual.shape == expected.shape
    if expected.ndim == 0 or expected.shape[0] > 6:
        scale = np.max(np.abs(expected)) or 1.0
        np.testing.assert_allclose(actual, expected, rtol=0, atol=tol * scale)
        return
    for a, e in zip(actual, expected):  # per conserved field
        scale = np.max(np.abs(e)) or 1.0
        np.testing.assert_allclose(a, e, rtol=0, atol=tol * scale)


def cons(prim, gamma):
    out = prim.copy()
    out[1:-1] *= prim[0]
    out[-1] = prim[-1] / (gamma - 1) + 0.5 * prim[0] * np.sum(prim[1:-1] ** 2, axis=0)
    return out


@pytest.mark.parametrize("nv", [4, 5])
def test_hllc_random_and_edge_states(nv):
    d = golden("hllc")
    ul, ur, gl, gr = (d[f"nv{nv}_{k}"] for k in ("ul", "ur", "gl", "gr"))
    np.testing.assert_array_equal(flux.hllc(ul, ur, gl, gr), d[f"nv{nv}_flux"])
    np.testing.assert_array_equal(flux.hllc(ul, ur, 1.4), d[f"nv{nv}_flux_g14"])


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_grid_rotation(dim):
    d = golden("hllc")
    for direction in range(dim):
        out = flux.hllc_grid(d[f"grid{dim}_left"], d[f"grid{dim}_right"], d[f"grid{dim}_gl"],
                             d[f"grid{dim}_gr"], direction=direction)
        np.testing.assert_array_equal(out, d[f"grid{dim}_dir{direction}"])


def test_slopes_all_limiters():
    d = golden("slope")
    for limiter in LIMITERS:
        for axis in (1, 2, 3):
            np.testing.assert_array_equal(three.slope(d["padded"], axis, limiter),
                                          d[f"{limiter}_axis{axis}"])


BCS3 = {"open": ("Open",) * 3, "periodic": ("Periodic",) * 3,
        "reflective": ("Reflective",) * 3, "mixed": ("Open", "Reflective", "Periodic")}


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("key", list(BCS3))
def test_hydro3d_step(limiter, key):
    d = golden("hydro3d")
    bx, by, bz = BCS3[key]
    res = three.muscl_hancock_step_3d(d["state"], *d["spacing"], float(d["dt"]), 1.4, limiter,
                                      bx, by, bz, True)
    close(res[0], d[f"{limiter}_{key}_U"])
    if f"{limiter}_{key}_Fx" in d:
        for f, name in zip(res[1:], "xyz"):
            close(f, d[f"{limiter}_{key}_F{name}"])


def test_hydro3d_variable_gamma_and_species():
    d = golden("hydro3d")
    g = one.gamma_calc(d["vargamma_lam"])
    bcs = ("Periodic", "Reflective", "Open")
    res, fluxes, fb = three.hydro(d["vargamma_state"], g, tuple(d["spacing"]), float(d["dt"]),
                                  "Superbee", bcs)
    assert not fb
    close(res, d["vargamma_U"])
    for f, name in zip(fluxes, "xyz"):
        close(f, d[f"vargamma_F{name}"])
    q = three.advect(d["advect_q"], d["vargamma_state"], res, fluxes, float(d["dt"]),
                     tuple(d["spacing"]), bcs)
    close(q, d["advect_out"])


def test_hydro3d_fallbacks():
    d = golden("hydro3d")
    assert bool(d["fallback_is_first_order"])
    res, fluxes, fb = three.hydro(d["fallback_state"], 1.4, (1 / 12, 0.25, 0.25),
                                  float(d["fallback_dt"]), "Unos", ("Open",) * 3)
    assert fb
    close(res, d["fallback_U"])
    close(fluxes[0], d["fallback_Fx"])
    res, fluxes, fb = three.hydro(d["state"], 1.4, tuple(d["spacing"]), float(d["dt"]), "VanLeer",
                                  ("Open", "Periodic", "Reflective"), first_order=True)
    close(res, d["forced_U"])
    for f, name in zip(fluxes, "xyz"):
        close(f, d[f"forced_F{name}"])


def test_reaction_3d_and_nondimensional_kinetics():
    d = golden("reaction")
    for model in ("arrhenius", "znd"):
        for dt in (1e-7, 1e-5):
            tag = f"{model}_{dt:g}"
            out, lam, beta = three.reaction(d["state"], d["lam"],
                                            d["beta"] if model == "znd" else None, dt, model)
            close(out, d[tag + "_U"])
            close(lam, d[tag + "_lam"])
            if model == "znd":
                close(beta, d[tag + "_beta"])
        lam, st, beta = one.kinetics(d["lam"].ravel(), 0.013, d["nondim_normal"], model,
                                     d["beta"].ravel(), q_heat=37.0, activation_energy=0.21,
                                     pre_exponential=55.0)
        close(lam, d[f"nondim_{model}_lam"])
        close(st, d[f"nondim_{model}_state"])
        close(beta, d[f"nondim_{model}_beta"])


BCS2 = {"open": ("Open", "Open"), "periodic": ("Periodic", "Periodic"),
        "reflective": ("Reflective", "Reflective"), "inflow": ("Periodic", "Inflow")}


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("key", list(BCS2))
def test_hydro2d_step(limiter, key):
    d = golden("hydro2d")
    dx, dy, dt = d["spacing_dt"]
    u, fx, fy, fb = two.hydro(d["state"], dx, dy, dt, 1.4, limiter, *BCS2[key], d["inlet"])
    assert not fb
    close(u, d[f"{limiter}_{key}_U"])
    close(fx, d[f"{limiter}_{key}_Fx"])
    close(fy, d[f"{limiter}_{key}_Fy"])


def test_hydro2d_variable_gamma_species_cfl_and_fallback():
    d = golden("hydro2d")
    dx, dy, dt = d["spacing_dt"]
    g = one.gamma_calc(d["vargamma_lam"])
    u, fx, fy, fb = two.hydro(d["vargamma_state"], dx, dy, dt, g, "Unos", "Reflective", "Inflow",
                              d["inlet"])
    close(u, d["vargamma_U"])
    close(fx, d["vargamma_Fx"])
    close(fy, d["vargamma_Fy"])
    q = two.advect(d["advect_q"], d["vargamma_state"], u, fx[0], fy[0], dt / dx, dt / dy,
                   "Reflective", "Inflow")
    close(q, d["advect_out"])
    for key, (bx, by) in BCS2.items():
        assert two.cfl_dt(d["vargamma_state"], g, dx, dy, 0.4, bx, by, d["inlet"]) == pytest.approx(
            float(d[f"cfl_{key}"]), rel=1e-15)
    ny, nx = g.shape
    u_pad = two.pad(d["vargamma_state"], "Reflective", "Inflow", d["inlet"])
    gpad = two.pad_scalar(g, "Reflective", "Inflow", inlet_value=1.4)
    fx, fy = two.llf_fluxes(u_pad, gpad)
    out = d["vargamma_state"] - dt / dx * np.diff(fx, axis=2) - dt / dy * np.diff(fy, axis=1)
    close(out, d["forced_U"])
    close(fx, d["forced_Fx"])
    close(fy, d["forced_Fy"])


def run_session(prim, lengths, steps, **options):
    sim = session.Session(prim, lengths, **options)
    dts = []
    for _ in range(steps):
        t0 = sim.time
        sim.step()
        dts.append(sim.time - t0)
    return sim, np.array(dts)


def check_session(tag, sim, dts):
    d = golden("session")
    close(dts, d[tag + "_dt"])
    close(sim.state, d[tag + "_state"])
    close(sim.boundary_integral, d[tag + "_boundary"], 1e-12)
    close(sim.source_integral, d[tag + "_source"], 1e-12)
    if sim.lam is not None:
        close(sim.lam, d[tag + "_lam"])
    if sim.beta is not None:
        close(sim.beta, d[tag + "_beta"])


def test_session_3d_sod_and_explosion():
    prim, opt = presets.preset("3d_sod_x", [32, 4, 4], limiter="MINBEE",
                               boundaries=["Open", "Reflective", "Reflective"], t_end=0.2)
    opt.pop("reactive")
    check_session("sod3d", *run_session(prim, [1.0] * 3, 20, **opt))
    prim, opt = presets.preset("3d_explosion", [8, 8, 8], limiter="Unos", t_end=0.2)
    opt.pop("reactive")
    check_session("explosion3d", *run_session(prim, [1.0] * 3, 6, **opt))


@pytest.mark.parametrize("model", ["arrhenius", "znd"])
def test_session_3d_reactive_retry(model):
    d = golden("session")
    sim, dts = run_session(d["ignition3d_prim"], [1.0] * 3, 4, t_end=1e-5, reactive=True,
                           kinetic_model=model, lambda_init=d[f"ignition3d_{model}_lam0"],
                           boundaries=["Open", "Periodic", "Reflective"])
    check_session(f"ignition3d_{model}", sim, dts)


def test_session_2d_quadrant_and_rde():
    d = golden("session")
    n = 24
    y, x = np.meshgrid((np.arange(n) + 0.5) / n, (np.arange(n) + 0.5) / n, indexing="ij")
    quad = np.where(x >= 0.5, np.where(y >= 0.5, 0, 3), np.where(y >= 0.5, 1, 2))
    prim = np.empty((4, n, n))
    for k, s in enumerate(d["quadrant_states"]):
        for c in range(4):
            prim[c][quad == k] = s[c]
    check_session("quadrant2d", *run_session(prim, [1.0, 1.0], 12, t_end=0.25, limiter="Superbee"))
    prim, opt = presets.preset("2d_rde_chamber", [12, 10])
    check_session("rde2d", *run_session(prim, [1.0, 1.0], 3, **opt))


@pytest.mark.parametrize("limiter", LIMITERS)
def test_session_1d_limiters(limiter):
    prim, opt = presets.preset("sod_standard", [50], limiter=limiter)
    opt.pop("reactive")
    check_session(f"sod1d_{limiter}", *run_session(prim, [1.0], 8, **opt))


def test_session_1d_walls_and_reactive():
    prim, opt = presets.preset("sod_standard", [40], boundaries=["Reflective"])
    opt.pop("reactive")
    check_session("sod1d_reflective", *run_session(prim, [1.0], 8, **opt))
    prim, opt = presets.preset("rarefaction_wave", [40], boundaries=["Periodic"])
    opt.pop("reactive")
    check_session("sod1d_periodic", *run_session(prim, [1.0], 8, **opt))
    prim, opt = presets.preset("detonation_cj", [40])
    check_session("cj1d", *run_session(prim, [1.0], 5, **opt))
    prim, opt = presets.preset("detonation_cj", [40], kinetic_model="znd")
    check_session("cj1d_znd", *run_session(prim, [1.0], 5, **opt))


def test_legacy_1d_driver_state():
    d = golden("session")
    for n in (50, 100):
        prim, _ = presets.preset_1d("sod_standard", n)
        sim = session.Session(prim, [1.0], t_end=0.2, cfl=0.5)
        while not sim.completed:
            sim.step()
        # The legacy driver and the session share every kernel; the driver
        # clips dt the same way and stops at t=0.2.
        assert sim.steps == int(d[f"legacy_sod_{n}_steps"])
        close(sim.state, d[f"legacy_sod_{n}"])
