"""Legacy drivers with time histories (oracle restatement).

TEST INFRASTRUCTURE ONLY (see the package docstring).

Restates the reference's stand-alone drivers on top of this package's
schemes: ``run_simulation`` / ``run_simulation_wqr`` (1D, every time level
kept; python/cfdSolver/torchSolver.py:918-969 initial states, 1074-1122
validation, 1135-1233 inert and 1236-1415 reactive loops) and
``run_torch_solver_2d`` / ``run_torch_solver_3d`` (torchSolver2D.py:670-946,
torchSolver3D.py:289-471). Their loops differ from Simulation.step in the
dt clamp (``if t + dt > t_end: dt = t_end - t``), the stall test
(``dt <= 1e-15``), the 1D reactive CFL test (hydro_dt_1d alone), the
running time (``t += dt``) and how the balances are summed.
"""

import numpy as np

from . import solver1d as one
from . import solver2d as two
from .presets import preset_2d, preset_3d
from .session import Session, one_validate


def validate_initial_1d(rho_l, u_l, p_l, rho_r, u_r, p_r, gamma, divider):
    """torchSolver.py:1091-1105."""
    values = [rho_l, u_l, p_l, rho_r, u_r, p_r, gamma, divider]
    if (not np.isfinite(values).all() or min(rho_l, rho_r, p_l, p_r) <= 0 or gamma <= 1
            or not 0 < divider < 100 or divider == 1):
        raise ValueError("Require finite primitives, positive density/pressure, gamma > 1 and divider in "
                         "(0,1) or (1,100) percent")


def validate_options_1d(limiter, controller, variant, waves, boundary):
    """torchSolver.py:1108-1118."""
    for value, allowed in ((limiter, ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")),
                           (controller, ("Sound Speed Controlled", "w=0")),
                           (variant, ("Variante 1 - Original", "Variante_2")),
                           (waves, ("Pressure Based", "Enthalpy Based")),
                           (boundary, ("Open", "Reflective", "Periodic"))):
        if value not in allowed:
            raise ValueError(f"Unknown solver option: {value}")


def two_state_primitive(n, rho_l, u_l, p_l, rho_r, u_r, p_r, divider):
    """torchSolver.py:918-945: left state up to round(ratio n), clamped to [1, n-1]."""
    ratio = divider / 100.0 if divider > 1.0 else divider
    cut = max(1, min(n - 1, int(np.round(ratio * n))))
    prim = np.zeros((3, n))
    prim[:, :cut] = np.array([rho_l, u_l, p_l])[:, None]
    prim[:, cut:] = np.array([rho_r, u_r, p_r])[:, None]
    return prim


def run_simulation(l, n, rho_l, u_l, p_l, rho_r, u_r, p_r, gamma=1.4, c_cfl=0.5, t_corte=0.2, limiter="VanLeer",
                   w_controler="Sound Speed Controlled", hllc_variante="Variante 1 - Original",
                   wave_speed_calc_option="Pressure Based", reflective_or_open="Open", divider=0.5,
                   diagnostics=None):
    """torchSolver.py:1135-1233."""
    one_validate(l, n, c_cfl, t_corte)
    validate_options_1d(limiter, w_controler, hllc_variante, wave_speed_calc_option, reflective_or_open)
    validate_initial_1d(rho_l, u_l, p_l, rho_r, u_r, p_r, gamma, divider)
    dx = l / float(n)
    u = one.to_cons(two_state_primitive(n, rho_l, u_l, p_l, rho_r, u_r, p_r, divider), gamma)
    if not one.physical_state(u):
        raise ValueError("Initial state must have positive density and internal energy")
    levels, times, t = [u], [0.0], 0.0
    boundary = np.zeros(3)
    while t < t_corte:
        dt = one.cfl_dt(u, gamma, dx, c_cfl, reflective_or_open, wave_speed_calc_option)
        if t + dt > t_corte:
            dt = t_corte - t
        if not np.isfinite(dt) or dt <= 1e-15:
            raise RuntimeError("Time integration stalled before the requested final time")
        u, flux, _ = one.hydro_step(u, gamma, dx, dt, c_cfl, limiter, reflective_or_open, w_controler,
                                    hllc_variante, wave_speed_calc_option)
        boundary += dt * (flux[:, -1] - flux[:, 0])
        t += dt
        times.append(t)
        levels.append(u)
    if diagnostics is not None:
        diagnostics.update(boundary_integral=boundary, source_integral=np.zeros(3))
    return np.stack(levels, axis=2), np.array(times)


def run_simulation_wqr(l, n, rho_l, u_l, p_l, rho_r, u_r, p_r, lambda_init=None, c_cfl=0.5, t_corte=0.2,
                       limiter="VanLeer", w_controler="Sound Speed Controlled",
                       hllc_variante="Variante 1 - Original", wave_speed_calc_option="Pressure Based",
                       reflective_or_open="Open", divider=0.5, kinetic_model="arrhenius", diagnostics=None):
    """torchSolver.py:1236-1415."""
    one_validate(l, n, c_cfl, t_corte)
    validate_options_1d(limiter, w_controler, hllc_variante, wave_speed_calc_option, reflective_or_open)
    validate_initial_1d(rho_l, u_l, p_l, rho_r, u_r, p_r, 1.4, divider)
    if kinetic_model not in ("arrhenius", "znd"):
        raise ValueError(f"Unknown kinetic model: {kinetic_model}")
    dx = l / float(n)
    lam = np.ones(n) if lambda_init is None else np.asarray(lambda_init, dtype=float)
    if lam.shape != (n,) or not np.isfinite(lam).all() or np.any((lam < 0) | (lam > 1)):
        raise ValueError("lambda_init must have one value in [0,1] per cell")
    u = one.to_cons(two_state_primitive(n, rho_l, u_l, p_l, rho_r, u_r, p_r, divider), one.gamma_calc(lam))
    if not one.physical_state(u):
        raise ValueError("Initial state must have positive density and internal energy")
    levels, lams, times, t = [u], [lam], [0.0], 0.0
    beta = np.ones(n)
    bc = reflective_or_open
    boundary, source = np.zeros(3), np.zeros(3)
    while t < t_corte:
        dt = one.cfl_dt(u, one.gamma_calc(lam), dx, c_cfl, bc, wave_speed_calc_option)
        if t + dt > t_corte:
            dt = t_corte - t
        if not np.isfinite(dt) or dt <= 1e-15:
            raise RuntimeError("Time integration stalled before the requested final time")
        for _ in range(60):
            lam_h, chem, beta_h = one.kinetics(lam, 0.5 * dt, u, kinetic_model, beta)
            allowed = one.dt_hydro(chem, one.gamma_calc(lam_h), dx, c_cfl)
            if dt <= allowed * (1.0 + 1e-12):
                break
            dt = min(0.5 * dt, allowed)
        else:
            raise RuntimeError("Could not find a CFL-stable reactive step")
        hydro, flux, _ = one.hydro_step(chem, one.gamma_calc(lam_h), dx, dt, c_cfl, limiter, bc, w_controler,
                                        hllc_variante, wave_speed_calc_option)
        lam_h = one.advect(lam_h, chem, hydro, flux[0], dt / dx, bc)
        beta_h = one.advect(beta_h, chem, hydro, flux[0], dt / dx, bc)
        lam_n, u_next, beta_n = one.kinetics(lam_h, 0.5 * dt, hydro, kinetic_model, beta_h)
        boundary += dt * (flux[:, -1] - flux[:, 0])
        source += dx * np.sum(chem - u + u_next - hydro, axis=1)
        t += dt
        times.append(t)
        levels.append(u_next)
        lams.append(lam_n)
        u, lam, beta = u_next, lam_n, beta_n
    if diagnostics is not None:
        diagnostics.update(boundary_integral=boundary, source_integral=source)
    return np.stack(levels, axis=2), np.array(times), np.stack(lams, axis=0)


def _legacy_loop(sim, t_end, max_steps):
    """Step a session with the legacy drivers' running time (t += dt, not
    min(t + dt, t_end)); returns (time, steps). The session's dt is
    min(dt_cfl, t_end - t), which is the 2D driver's clamp except where
    rounding separates the two tests."""
    time, steps = 0.0, 0
    while time < t_end and steps < max_steps:
        sim.time = time
        sim.step()
        time += sim.last_dt
        steps += 1
    return time, steps


def run_torch_solver_3d(preset_name="3d_explosion", *, primitive_init=None, nx=20, ny=20, nz=20, lx=1.0, ly=1.0,
                        lz=1.0, t_end=None, cfl=0.4, gamma=1.4, limiter="VanLeer", bc_x="Open", bc_y="Open",
                        bc_z="Open", is_reactive=False, kinetic_model="arrhenius", lambda_init=None, beta_init=None,
                        max_steps=100000):
    """torchSolver3D.py:289-471 (its loop is Simulation.step's with t += dt)."""
    if t_end is None:
        t_end = 1e-5 if preset_name == "3d_ignition" and primitive_init is None else 0.05
    if primitive_init is None:
        primitive_init, d = preset_3d(preset_name, nx, ny, nz)
        is_reactive = d["reactive"]
    sim = Session(primitive_init, [lx, ly, lz], t_end=t_end, cfl=cfl, gamma=gamma, limiter=limiter,
                  boundaries=[bc_x, bc_y, bc_z], reactive=is_reactive, kinetic_model=kinetic_model,
                  lambda_init=lambda_init, beta_init=beta_init, max_steps=max_steps)
    initial = sim.integral.copy()
    time, steps = _legacy_loop(sim, t_end, max_steps)
    if time < t_end:
        raise RuntimeError(f"Maximum steps reached at t={time}; requested {t_end}")
    return dict(u_final=sim.state, lambda_final=sim.lam, beta_final=sim.beta, time=time, steps=steps,
                initial_integral=initial, final_integral=sim.integral, boundary_integral=sim.boundary_integral,
                source_integral=sim.source_integral,
                balance_residual=sim.integral - initial + sim.boundary_integral - sim.source_integral)


def run_torch_solver_2d(preset_name="2d_explosion", nx=50, ny=50, lx=1.0, ly=1.0, t_end=0.2, cfl=0.45, gamma=1.4,
                        limiter="VanLeer", bc_x="Transmissive", bc_y="Transmissive", is_reactive=False,
                        kinetic_model="arrhenius", max_steps=100000, inlet_primitive=None,
                        ambient_pressure=101325.0):
    """torchSolver2D.py:670-946 for presets (its loop is Simulation.step's with t += dt)."""
    prim, d = preset_2d(preset_name, nx, ny)
    bcs = [{"Transmissive": "Open"}.get(b, b) for b in d["boundaries"]]
    sim = Session(prim, [lx, ly], t_end=t_end, cfl=cfl, gamma=gamma, limiter=limiter, boundaries=bcs,
                  reactive=d["reactive"], kinetic_model=kinetic_model, max_steps=max_steps,
                  inlet_primitive=inlet_primitive)
    dx, dy = lx / nx, ly / ny
    initial_mass = float(np.sum(sim.state[0]) * dx * dy)
    time, steps = _legacy_loop(sim, t_end, max_steps)
    if time < t_end:
        raise RuntimeError(f"Maximum steps reached at t={time}; requested {t_end}")
    final_mass = float(np.sum(sim.state[0]) * dx * dy)
    boundary_mass = sim.boundary_integral[0]
    prim = two.to_prim(sim.state, sim.gamma)
    return dict(u_final=sim.state, u_prim_final=prim, lambda_final=sim.lam, beta_final=sim.beta, time=time,
                steps=steps, net_boundary_mass=boundary_mass,
                mass_conservation_error=abs(final_mass - initial_mass + boundary_mass) / max(1e-12, initial_mass),
                exit_thrust_per_depth=float(dx * np.sum(prim[0, -1] * prim[2, -1] ** 2 + prim[3, -1]
                                                        - ambient_pressure)))

