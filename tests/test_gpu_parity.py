"""HIP path against the reference's golden vectors and the oracle.

Every test calls libtorchhip through its C ABI (via the ctypes mirror in
torch_amd) on cuda device 0. The FP64 plain method is held to 1e-12
relative per conserved field (tests/parity.py); fallback decisions and step
counts must match exactly.
"""

import numpy as np
import pytest

from conftest import golden
from oracle import presets as opresets
from oracle import session as osession
from oracle import solver3d as o3
from parity import assert_fields

pytestmark = pytest.mark.gpu

LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")


@pytest.fixture(scope="module")
def tam():
    import torch_amd

    return torch_amd


@pytest.mark.parametrize("nv", [4, 5])
def test_hllc_faces_golden(tam, nv):
    d = golden("hllc")
    ul, ur, gl, gr = (d[f"nv{nv}_{k}"] for k in ("ul", "ur", "gl", "gr"))
    assert_fields(tam.hllc_flux(ul, ur, gl, gamma_right=gr), d[f"nv{nv}_flux"], what="hllc")
    assert_fields(tam.hllc_flux(ul, ur, 1.4), d[f"nv{nv}_flux_g14"], what="hllc g=1.4")


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_grid_golden(tam, dim):
    d = golden("hllc")
    for direction in range(dim):
        out = tam.hllc_flux_grid(d[f"grid{dim}_left"], d[f"grid{dim}_right"], d[f"grid{dim}_gl"],
                                 d[f"grid{dim}_gr"], direction=direction, batch_size=17)
        assert_fields(out, d[f"grid{dim}_dir{direction}"], what=f"grid dir {direction}")


BCS3 = {"open": ("Open",) * 3, "periodic": ("Periodic",) * 3,
        "reflective": ("Reflective",) * 3, "mixed": ("Open", "Reflective", "Periodic")}


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("key", list(BCS3))
def test_hydro3d_golden(tam, limiter, key):
    d = golden("hydro3d")
    bx, by, bz = BCS3[key]
    res = tam.muscl_hancock_step_3d(d["state"], *d["spacing"], float(d["dt"]), 1.4, limiter, bx, by, bz, True)
    assert_fields(res[0], d[f"{limiter}_{key}_U"], what="U'")
    if f"{limiter}_{key}_Fx" in d:
        for f, name in zip(res[1:], "xyz"):
            assert_fields(f, d[f"{limiter}_{key}_F{name}"], what="F" + name)


def test_hydro3d_variable_gamma_golden(tam):
    d = golden("hydro3d")
    from torch_amd.torchSolver3D import hydro

    g = (4.5 - d["vargamma_lam"]) / (3.5 - d["vargamma_lam"])
    res, fluxes = hydro(d["vargamma_state"], g, tuple(d["spacing"]), float(d["dt"]), "Superbee",
                        ("Periodic", "Reflective", "Open"))
    assert_fields(res, d["vargamma_U"], what="U'")
    for f, name in zip(fluxes, "xyz"):
        assert_fields(f, d[f"vargamma_F{name}"], what="F" + name)


def test_hydro3d_first_order_fallback_golden(tam):
    d = golden("hydro3d")
    res = tam.muscl_hancock_step_3d(d["fallback_state"], 1 / 12, 0.25, 0.25, float(d["fallback_dt"]), 1.4,
                                    "Unos", "Open", "Open", "Open", True)
    assert_fields(res[0], d["fallback_U"], what="fallback U'")
    for f, name in zip(res[1:], "xyz"):
        assert_fields(f, d[f"fallback_F{name}"], what="fallback F" + name)


def test_reaction3d_golden(tam):
    d = golden("reaction")
    for model in ("arrhenius", "znd"):
        for dt in (1e-7, 1e-5):
            tag = f"{model}_{dt:g}"
            out, lam, beta = tam.reaction_step_3d(d["state"], d["lam"], d["beta"] if model == "znd" else None,
                                                  dt, model)
            assert_fields(out, d[tag + "_U"], what=tag)
            assert_fields(lam.ravel(), d[tag + "_lam"].ravel(), what=tag + " lam")
            if model == "znd":
                assert_fields(beta.ravel(), d[tag + "_beta"].ravel(), what=tag + " beta")


BCS2 = {"open": ("Open", "Open"), "periodic": ("Periodic", "Periodic"),
        "reflective": ("Reflective", "Reflective"), "inflow": ("Periodic", "Inflow")}


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("key", list(BCS2))
def test_hydro2d_golden(tam, limiter, key):
    d = golden("hydro2d")
    dx, dy, dt = d["spacing_dt"]
    u, fx, fy = tam.muscl_hancock_step_2d(d["state"], dx, dy, dt, 1.4, limiter, *BCS2[key], True, d["inlet"])
    assert_fields(u, d[f"{limiter}_{key}_U"], what="U'")
    assert_fields(fx, d[f"{limiter}_{key}_Fx"], what="Fx")
    assert_fields(fy, d[f"{limiter}_{key}_Fy"], what="Fy")


def test_hydro2d_variable_gamma_inflow_golden(tam):
    d = golden("hydro2d")
    dx, dy, dt = d["spacing_dt"]
    g = (4.5 - d["vargamma_lam"]) / (3.5 - d["vargamma_lam"])
    u, fx, fy = tam.muscl_hancock_step_2d(d["vargamma_state"], dx, dy, dt, g, "Unos", "Reflective", "Inflow",
                                          True, d["inlet"])
    assert_fields(u, d["vargamma_U"], what="U'")
    assert_fields(fx, d["vargamma_Fx"], what="Fx")
    assert_fields(fy, d["vargamma_Fy"], what="Fy")


def run_device(prim, lengths, steps, **options):
    from torch_amd import Simulation

    sim = Simulation(prim, lengths, **options)
    dts = []
    for _ in range(steps):
        t0 = sim.time
        sim.step()
        dts.append(sim.time - t0)
    return sim, np.array(dts)


def check_session(tag, sim, dts):
    d = golden("session")
    assert_fields(dts, d[tag + "_dt"], what=tag + " dt")
    assert_fields(sim.state, d[tag + "_state"], what=tag + " state")
    scale = max(np.max(np.abs(sim.state)), 1.0) * np.prod(sim.spacing) * 1e3
    np.testing.assert_allclose(sim.boundary_integral, d[tag + "_boundary"], rtol=1e-10, atol=1e-12 * scale)
    np.testing.assert_allclose(sim.source_integral, d[tag + "_source"], rtol=1e-10, atol=1e-12 * scale)
    if sim.lam is not None:
        assert_fields(sim.lam.ravel(), d[tag + "_lam"].ravel(), what=tag + " lam")
    if sim.beta is not None:
        assert_fields(sim.beta.ravel(), d[tag + "_beta"].ravel(), what=tag + " beta")


def test_session3d_sod_and_explosion_golden(tam):
    prim, opt = opresets.preset("3d_sod_x", [32, 4, 4], limiter="MINBEE",
                                boundaries=["Open", "Reflective", "Reflective"], t_end=0.2)
    opt.pop("reactive")
    check_session("sod3d", *run_device(prim, [1.0] * 3, 20, **opt))
    prim, opt = opresets.preset("3d_explosion", [8, 8, 8], limiter="Unos", t_end=0.2)
    opt.pop("reactive")
    check_session("explosion3d", *run_device(prim, [1.0] * 3, 6, **opt))


@pytest.mark.parametrize("model", ["arrhenius", "znd"])
def test_session3d_reactive_retry_golden(tam, model):
    d = golden("session")
    sim, dts = run_device(d["ignition3d_prim"], [1.0] * 3, 4, t_end=1e-5, reactive=True, kinetic_model=model,
                          lambda_init=d[f"ignition3d_{model}_lam0"], boundaries=["Open", "Periodic", "Reflective"])
    check_session(f"ignition3d_{model}", sim, dts)


@pytest.mark.parametrize("model", ["arrhenius", "znd"])
def test_fp32_reactive_session_with_walls_tracks_fp64(tam, model):
    """The FP32 form's reactive 3D tiles (state, lambda and fractions staged
    in LDS a plane ahead, reflective sign flips applied on reading) against
    the FP64 method (register loads) on the ignition vectors' initial state,
    with reflective walls in x and z: the same steps and dt halvings, dt
    within 1e-6 relative, every field within 1e-5 relative L1, lambda within
    1e-5 absolute. A flip applied to the wrong cell or component breaks the
    wall-normal momentum at O(1)."""
    d = golden("session")
    opts = dict(t_end=1e-5, reactive=True, kinetic_model=model, lambda_init=d[f"ignition3d_{model}_lam0"],
                boundaries=["Reflective", "Periodic", "Reflective"])
    s64, dt64 = run_device(d["ignition3d_prim"], [1.0] * 3, 4, precision="fp64", **opts)
    s32, dt32 = run_device(d["ignition3d_prim"], [1.0] * 3, 4, precision="fp32", **opts)
    assert s32.steps == s64.steps and np.allclose(dt32, dt64, rtol=1e-6, atol=0)
    for c in range(5):
        ref = np.abs(s64.state[c]).sum()
        if ref > 0:
            assert np.abs(s32.state[c] - s64.state[c]).sum() / ref <= 1e-5, c
    assert np.max(np.abs(s32.lam - s64.lam)) <= 1e-5
    s32.close()
    s64.close()


def test_session2d_quadrant_and_rde_golden(tam):
    d = golden("session")
    n = 24
    y, x = np.meshgrid((np.arange(n) + 0.5) / n, (np.arange(n) + 0.5) / n, indexing="ij")
    quad = np.where(x >= 0.5, np.where(y >= 0.5, 0, 3), np.where(y >= 0.5, 1, 2))
    prim = np.empty((4, n, n))
    for k, s in enumerate(d["quadrant_states"]):
        for c in range(4):
            prim[c][quad == k] = s[c]
    check_session("quadrant2d", *run_device(prim, [1.0, 1.0], 12, t_end=0.25, limiter="Superbee"))
    prim, opt = opresets.preset("2d_rde_chamber", [12, 10])
    check_session("rde2d", *run_device(prim, [1.0, 1.0], 3, **opt))


@pytest.mark.parametrize("limiter", LIMITERS)
def test_session1d_limiters_golden(tam, limiter):
    prim, opt = opresets.preset("sod_standard", [50], limiter=limiter)
    opt.pop("reactive")
    check_session(f"sod1d_{limiter}", *run_device(prim, [1.0], 8, **opt))


def test_session1d_walls_and_reactive_golden(tam):
    for tag, name, extra in (("sod1d_reflective", "sod_standard", dict(boundaries=["Reflective"])),
                             ("sod1d_periodic", "rarefaction_wave", dict(boundaries=["Periodic"])),
                             ("cj1d", "detonation_cj", {}),
                             ("cj1d_znd", "detonation_cj", dict(kinetic_model="znd"))):
        n = 50 if tag.startswith("sod1d_l") else 40
        prim, opt = opresets.preset(name, [n], **extra)
        if not opt["reactive"]:
            opt.pop("reactive")
        check_session(tag, *run_device(prim, [1.0], 8 if tag.startswith("sod") else 5, **opt))


def test_session3d_sod_cfl08_against_oracle(tam):
    """3D Sod at the benchmark's CFL 0.8 / MINBEE / walls, 64x12x10, 25 steps."""
    prim, _ = opresets.preset("3d_sod_x", [64, 12, 10])
    opts = dict(t_end=0.2, cfl=0.8, limiter="MINBEE", boundaries=["Open", "Reflective", "Reflective"],
                checked=False)
    ref = osession.Session(prim, [1.0] * 3, **opts)
    dev, _ = run_device(prim, [1.0] * 3, 0, **opts)
    for _ in range(25):
        ref.step()
        dev.step()
        assert dev.fallbacks == ref.fallbacks
    assert dev.time == pytest.approx(ref.time, rel=1e-13)
    assert_fields(dev.state, ref.state, what="Sod CFL 0.8")
    np.testing.assert_allclose(dev.boundary_integral, ref.boundary_integral, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("limiter", LIMITERS)
def test_hydro3d_random_tiles_against_oracle(tam, limiter):
    """Uneven grid (tile edges inside the domain) for every tile shape."""
    rng = np.random.default_rng(7)
    nz, ny, nx = 37, 13, 70
    p = np.empty((5, nz, ny, nx))
    p[0] = rng.uniform(0.5, 1.5, (nz, ny, nx))
    p[1:4] = rng.uniform(-0.3, 0.3, (3, nz, ny, nx))
    p[4] = rng.uniform(0.5, 1.5, (nz, ny, nx))
    state = o3.to_cons(p, 1.4)
    bcs = ("Periodic", "Reflective", "Open")
    ref, rf, fb = o3.hydro(state, 1.4, (0.01, 0.02, 0.03), 1e-3, limiter, bcs)
    assert not fb
    from torch_amd.torchSolver3D import hydro
    from torch_amd import DeviceExecution

    for tile in ((64, 4), (64, 8), (32, 8), (32, 16)):
        ex = DeviceExecution()
        ex.tile = tile + (8,)
        out, fl = hydro(state, 1.4, (0.01, 0.02, 0.03), 1e-3, limiter, bcs, execution=ex)
        assert_fields(out, ref, what=f"tile {tile}")
        for a, b, name in zip(fl, rf, "xyz"):
            assert_fields(a, b, what=f"tile {tile} F{name}")
        ex.close()


@pytest.mark.parametrize("limiter", LIMITERS)
def test_hydro3d_fp32_tiles_agree_bitwise(tam, limiter):
    """The FP32 form's 768-thread tile (32x24) against its 256-thread tiles
    on the same uneven grid (tile edges inside the domain in x, y and z):
    the per-cell arithmetic does not depend on the tile, so states and face
    fluxes are bitwise equal."""
    rng = np.random.default_rng(11)
    nz, ny, nx = 37, 53, 70
    p = np.empty((5, nz, ny, nx))
    p[0] = rng.uniform(0.5, 1.5, (nz, ny, nx))
    p[1:4] = rng.uniform(-0.3, 0.3, (3, nz, ny, nx))
    p[4] = rng.uniform(0.5, 1.5, (nz, ny, nx))
    state = o3.to_cons(p, 1.4)
    bcs = ("Periodic", "Reflective", "Open")
    from torch_amd.torchSolver3D import hydro
    from torch_amd import DeviceExecution

    outs = []
    for tile in ((32, 24, 24), (32, 24, 8), (32, 8, 24), (64, 4, 16)):
        ex = DeviceExecution(precision="fp32")
        ex.tile = tile
        outs.append((tile,) + tuple(hydro(state, 1.4, (0.01, 0.02, 0.03), 1e-3, limiter, bcs, execution=ex)))
        ex.close()
    (t0, out0, fl0) = outs[0]
    for tile, out, fl in outs[1:]:
        assert np.array_equal(out, out0), (tile, t0)
        for a, b in zip(fl, fl0):
            assert np.array_equal(a, b), (tile, t0)
