"""The legacy-driver oracle against vectors the reference's own drivers wrote
(tests/golden/legacy.npz, made by make_legacy_fixtures.py): full 1D time
histories (inert and reactive) and the 2D/3D drivers' final states, step
counts, times and balances."""

import numpy as np
import pytest

from conftest import golden
from oracle import legacy

REL = 1e-12


def rel_err(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = np.max(np.abs(b), axis=tuple(range(1, b.ndim)), keepdims=True) if b.ndim > 1 else np.max(np.abs(b))
    return float(np.max(np.abs(a - b) / np.maximum(scale, 1e-300)))


INERT = {
    "sod60": dict(l=1.0, n=60, rho_l=1.0, u_l=0.0, p_l=1.0, rho_r=0.125, u_r=0.0, p_r=0.1, c_cfl=0.5, t_corte=0.2),
    "wall_superbee": dict(l=1.0, n=48, rho_l=1.0, u_l=0.5, p_l=1.0, rho_r=0.5, u_r=-0.25, p_r=0.3, c_cfl=0.4,
                          t_corte=0.1, limiter="Superbee", reflective_or_open="Reflective", divider=0.4),
    "periodic_enthalpy": dict(l=2.0, n=40, rho_l=1.0, u_l=-1.0, p_l=0.4, rho_r=1.0, u_r=1.0, p_r=0.4, c_cfl=0.45,
                              t_corte=0.1, limiter="Unos", reflective_or_open="Periodic",
                              wave_speed_calc_option="Enthalpy Based", w_controler="w=0",
                              hllc_variante="Variante_2"),
}
REACTIVE = {
    "cj_arrhenius": dict(l=1.0, n=40, rho_l=1.0, u_l=0.0, p_l=50 * 101325.0, rho_r=0.125, u_r=0.0,
                         p_r=0.1 * 101325.0, c_cfl=0.4, t_corte=2e-5),
    "cj_znd": dict(l=1.0, n=40, rho_l=1.0, u_l=0.0, p_l=50 * 101325.0, rho_r=0.125, u_r=0.0, p_r=0.1 * 101325.0,
                   c_cfl=0.4, t_corte=2e-5, kinetic_model="znd"),
}
DRIVER_2D = {"ex2d": dict(preset_name="2d_explosion", nx=12, ny=10, t_end=0.05, cfl=0.4),
             "quad2d": dict(preset_name="2d_riemann_quadrant", nx=14, ny=12, t_end=0.05, limiter="MINBEE"),
             "rde2d": dict(preset_name="2d_rde_chamber", nx=12, ny=10, t_end=2e-5)}
DRIVER_3D = {"ex3d": dict(preset_name="3d_explosion", nx=8, ny=6, nz=4, t_end=0.06),
             "sod3d": dict(preset_name="3d_sod_y", nx=4, ny=16, nz=4, t_end=0.05, limiter="Superbee",
                           bc_x="Reflective", bc_z="Periodic"),
             "ign3d": dict(preset_name="3d_ignition", nx=8, ny=4, nz=4, t_end=6e-5),
             "ign3d_znd": dict(preset_name="3d_ignition", nx=8, ny=4, nz=4, t_end=6e-5, kinetic_model="znd")}


@pytest.fixture(scope="module")
def d():
    return golden("legacy")


def check_history(d, tag, u, t, lam=None, diag=None):
    assert u.shape == d[tag + "_u"].shape, tag
    np.testing.assert_allclose(t, d[tag + "_t"], rtol=1e-12, atol=0)
    for k in range(u.shape[2]):
        assert rel_err(u[:, :, k], d[tag + "_u"][:, :, k]) <= 1e-11, (tag, k)
    if lam is not None:
        assert rel_err(lam, d[tag + "_lam"]) <= 1e-11, tag
    if diag is not None:
        for key in ("boundary", "source"):
            ref = d[f"{tag}_{key}"]
            got = diag[key + "_integral"]
            assert np.max(np.abs(got - ref)) <= 1e-11 * max(np.max(np.abs(ref)), 1.0), (tag, key)


@pytest.mark.parametrize("tag", list(INERT))
def test_run_simulation_history(d, tag):
    diag = {}
    u, t = legacy.run_simulation(**INERT[tag], diagnostics=diag)
    check_history(d, tag, u, t, diag=diag)


@pytest.mark.parametrize("tag", list(REACTIVE))
def test_run_simulation_wqr_history(d, tag):
    diag = {}
    u, t, lam = legacy.run_simulation_wqr(**REACTIVE[tag], lambda_init=d[tag + "_lam0"], diagnostics=diag)
    check_history(d, tag, u, t, lam, diag)


@pytest.mark.parametrize("tag", list(DRIVER_2D))
def test_run_torch_solver_2d(d, tag):
    r = legacy.run_torch_solver_2d(**DRIVER_2D[tag])
    assert r["steps"] == int(d[tag + "_steps"]) and r["time"] == pytest.approx(float(d[tag + "_time"]), rel=1e-14)
    assert rel_err(r["u_final"], d[tag + "_u_final"]) <= 1e-11
    assert r["net_boundary_mass"] == pytest.approx(float(d[tag + "_net_boundary_mass"]), rel=1e-10, abs=1e-13)
    assert r["exit_thrust_per_depth"] == pytest.approx(float(d[tag + "_exit_thrust_per_depth"]), rel=1e-12,
                                                       abs=1e-9)
    if tag + "_lambda_final" in d:
        assert rel_err(r["lambda_final"], d[tag + "_lambda_final"]) <= 1e-11


@pytest.mark.parametrize("tag", list(DRIVER_3D))
def test_run_torch_solver_3d(d, tag):
    r = legacy.run_torch_solver_3d(**DRIVER_3D[tag])
    assert r["steps"] == int(d[tag + "_steps"]) and r["time"] == pytest.approx(float(d[tag + "_time"]), rel=1e-14)
    assert rel_err(r["u_final"], d[tag + "_u_final"]) <= 1e-11
    for key in ("boundary_integral", "source_integral"):
        ref = d[f"{tag}_{key}"]
        assert np.max(np.abs(r[key] - ref)) <= 1e-10 * max(np.max(np.abs(ref)), 1.0), (tag, key)
    if tag + "_lambda_final" in d:
        assert rel_err(r["lambda_final"], d[tag + "_lambda_final"]) <= 1e-11
