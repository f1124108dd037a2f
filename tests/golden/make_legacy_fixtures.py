"""Write tests/golden/legacy.npz: the reference's legacy drivers.

Like make_fixtures.py, this executes the reference implementation (TORCH's
pure-NumPy ``cfdSolver``, read from /root/reference/python) in the
development container only, and stores small inputs and outputs:

- run_simulation (torchSolver.py:1135-1233): inert 1D with the full time
  history (every level of the state and the time array) and its balances;
- run_simulation_wqr (torchSolver.py:1236-1415): reactive 1D (Arrhenius and
  ZND) with the state and lambda histories;
- run_torch_solver_2d / run_torch_solver_3d (torchSolver2D.py:670-946,
  torchSolver3D.py:289-471): inert and reactive presets, final states,
  step counts, times and balances.

Run: PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_legacy_fixtures.py
"""

import os
import sys

import numpy as np

sys.dont_write_bytecode = True
REFERENCE = "/root/reference/python"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    sys.path.insert(0, REFERENCE)
    from cfdSolver.torchSolver import run_simulation, run_simulation_wqr  # noqa: E402
    from cfdSolver.torchSolver2D import run_torch_solver_2d  # noqa: E402
    from cfdSolver.torchSolver3D import run_torch_solver_3d  # noqa: E402

    data = {}
    inert = {
        "sod60": dict(l=1.0, n=60, rho_l=1.0, u_l=0.0, p_l=1.0, rho_r=0.125, u_r=0.0, p_r=0.1, c_cfl=0.5,
                      t_corte=0.2),
        "wall_superbee": dict(l=1.0, n=48, rho_l=1.0, u_l=0.5, p_l=1.0, rho_r=0.5, u_r=-0.25, p_r=0.3, c_cfl=0.4,
                              t_corte=0.1, limiter="Superbee", reflective_or_open="Reflective", divider=0.4),
        "periodic_enthalpy": dict(l=2.0, n=40, rho_l=1.0, u_l=-1.0, p_l=0.4, rho_r=1.0, u_r=1.0, p_r=0.4,
                                  c_cfl=0.45, t_corte=0.1, limiter="Unos", reflective_or_open="Periodic",
                                  wave_speed_calc_option="Enthalpy Based", w_controler="w=0",
                                  hllc_variante="Variante_2"),
    }
    for tag, kw in inert.items():
        diag = {}
        u, t = run_simulation(**kw, diagnostics=diag)
        data[f"{tag}_u"] = u
        data[f"{tag}_t"] = t
        data[f"{tag}_boundary"] = diag["boundary_integral"]
        data[f"{tag}_source"] = diag["source_integral"]
    reactive = {
        "cj_arrhenius": dict(l=1.0, n=40, rho_l=1.0, u_l=0.0, p_l=50 * 101325.0, rho_r=0.125, u_r=0.0,
                             p_r=0.1 * 101325.0, c_cfl=0.4, t_corte=2e-5),
        "cj_znd": dict(l=1.0, n=40, rho_l=1.0, u_l=0.0, p_l=50 * 101325.0, rho_r=0.125, u_r=0.0,
                       p_r=0.1 * 101325.0, c_cfl=0.4, t_corte=2e-5, kinetic_model="znd"),
    }
    for tag, kw in reactive.items():
        diag = {}
        lam0 = np.linspace(1.0, 0.6, kw["n"])
        u, t, lam = run_simulation_wqr(**kw, lambda_init=lam0, diagnostics=diag)
        data[f"{tag}_lam0"] = lam0
        data[f"{tag}_u"] = u
        data[f"{tag}_t"] = t
        data[f"{tag}_lam"] = lam
        data[f"{tag}_boundary"] = diag["boundary_integral"]
        data[f"{tag}_source"] = diag["source_integral"]
    for tag, kw in {"ex2d": dict(preset_name="2d_explosion", nx=12, ny=10, t_end=0.05, cfl=0.4),
                    "quad2d": dict(preset_name="2d_riemann_quadrant", nx=14, ny=12, t_end=0.05, limiter="MINBEE"),
                    "rde2d": dict(preset_name="2d_rde_chamber", nx=12, ny=10, t_end=2e-5)}.items():
        r = run_torch_solver_2d(**kw)
        for key in ("u_final", "time", "steps", "mass_conservation_error", "net_boundary_mass",
                    "exit_thrust_per_depth"):
            data[f"{tag}_{key}"] = np.asarray(r[key])
        if r["lambda_final"] is not None:
            data[f"{tag}_lambda_final"] = r["lambda_final"]
    for tag, kw in {"ex3d": dict(preset_name="3d_explosion", nx=8, ny=6, nz=4, t_end=0.06),
                    "sod3d": dict(preset_name="3d_sod_y", nx=4, ny=16, nz=4, t_end=0.05, limiter="Superbee",
                                  bc_x="Reflective", bc_z="Periodic"),
                    "ign3d": dict(preset_name="3d_ignition", nx=8, ny=4, nz=4, t_end=6e-5),
                    "ign3d_znd": dict(preset_name="3d_ignition", nx=8, ny=4, nz=4, t_end=6e-5,
                                      kinetic_model="znd")}.items():
        r = run_torch_solver_3d(**kw)
        for key in ("u_final", "time", "steps", "boundary_integral", "source_integral", "balance_residual"):
            data[f"{tag}_{key}"] = np.asarray(r[key])
        if r["lambda_final"] is not None:
            data[f"{tag}_lambda_final"] = r["lambda_final"]
        if r["beta_final"] is not None:
            data[f"{tag}_beta_final"] = r["beta_final"]
    np.savez_compressed(os.path.join(OUT, "legacy.npz"), **data)
    print("wrote", len(data), "arrays")


if __name__ == "__main__":
    main()
