"""Cartesian 2D step functions on the GPU.

Mirrors python/cfdSolver/torchSolver2D.py: ``muscl_hancock_step_2d``
(:274-466, including the Inflow y boundary, the predictor repair and the
first-order LLF fallback) and ``reaction_step_2d`` (:474-509). States are
(4, ny, nx) float64 host arrays in and out.
"""

import numpy as np

from . import _abi
from .torchExecution import device_of
from .torchSolver3D import config


def _gamma_array(gamma, shape):
    if np.isscalar(gamma) or np.ndim(gamma) == 0:
        return None
    return np.ascontiguousarray(np.broadcast_to(gamma, shape), dtype=float)


def muscl_hancock_step_2d(u_inner, dx, dy, dt, gamma=1.4, limiter="VanLeer", bc_x="Periodic", bc_y="Open",
                          return_fluxes=False, inlet_primitive=None, execution=None):
    u = np.ascontiguousarray(u_inner, dtype=float)
    if u.ndim != 3 or u.shape[0] != 4:
        raise ValueError("2D states have shape (4, ny, nx)")
    ny, nx = u.shape[1:]
    garr = _gamma_array(gamma, (ny, nx))
    ex = device_of(execution)
    cfg = config((nx, ny), (dx, dy), dim=2, gamma=gamma if garr is None else 1.4, limiter=limiter,
                 bcs=(bc_x, bc_y), gamma_mode=0 if garr is None else 2, precision=ex.precision,
                 inlet=inlet_primitive)
    dev = _abi.DeviceState(ex.context, cfg)
    try:
        dev.upload(u, garr)
        dev.hydro(dt, want_fluxes=return_fluxes)
        out, _, _ = dev.download()
        fluxes = dev.fluxes() if return_fluxes else []
    finally:
        dev.close()
    return (out, *fluxes) if return_fluxes else out


def reaction_step_2d(u_consv, lambda_val, beta_val, dt_chem, kinetic_model="arrhenius", q_heat=1.41584e8,
                     activation_energy=17.15 * 4184.0, pre_exponential=1.76e5, execution=None):
    u = np.ascontiguousarray(u_consv, dtype=float)
    lam = np.asarray(lambda_val, dtype=float)
    ny, nx = u.shape[1:]
    ex = device_of(execution)
    cfg = config((nx, ny), (1.0, 1.0), dim=2, reactive=True, model=kinetic_model, gamma_mode=1,
                 precision=ex.precision,
                 constants=dict(q_heat=q_heat, activation_energy=activation_energy,
                                pre_exponential=pre_exponential))
    dev = _abi.DeviceState(ex.context, cfg)
    try:
        dev.upload(u, lam, beta_val)
        dev.react(dt_chem)
        out, new_lam, new_beta = dev.download(True, True)
    finally:
        dev.close()
    return out, new_lam.reshape(lam.shape), None if beta_val is None else new_beta.reshape(lam.shape)


def run_torch_solver_2d(preset_name="2d_explosion", rho_init=None, u_init=None, v_init=None, p_init=None, nx=50,
                        ny=50, lx=1.0, ly=1.0, t_end=0.2, cfl=0.45, gamma=1.4, limiter="VanLeer",
                        bc_x="Transmissive", bc_y="Transmissive", is_reactive=False, kinetic_model="arrhenius",
                        max_steps=100000, inlet_primitive=None, ambient_pressure=101325.0, lambda_init=None,
                        beta_init=None, precision="fp64", device=0):
    """Legacy 2D driver (torchSolver2D.py:670-946) on a device session."""
    from .presets import preset_2d
    from .torchSession import Simulation
    from .torchSolver3D import _legacy_run

    if not np.isscalar(gamma) or not np.isfinite(gamma) or gamma <= 1:
        raise ValueError("gamma must be finite and greater than 1")
    if not np.isfinite(ambient_pressure) or ambient_pressure < 0:
        raise ValueError("Ambient pressure must be finite and nonnegative")
    fields = (rho_init, u_init, v_init, p_init)
    if any(f is not None for f in fields) and any(f is None for f in fields):
        raise ValueError("Custom initialization requires rho_init, u_init, v_init and p_init")
    if preset_name is None and rho_init is None:
        raise ValueError("Provide a preset or all four initial primitive fields")
    if preset_name is not None and rho_init is None:
        prim, d = preset_2d(preset_name, nx, ny)
        gamma = 1.4
        bc_x, bc_y = d["boundaries"]
        is_reactive = d["reactive"]
    else:
        prim = np.stack([np.broadcast_to(np.asarray(f, dtype=float), (ny, nx)) for f in fields])
    legacy_names = {"Transmissive": "Open"}
    sim = Simulation(prim, [lx, ly], t_end=t_end, cfl=cfl, gamma=gamma, limiter=limiter,
                     boundaries=[legacy_names.get(bc_x, bc_x), legacy_names.get(bc_y, bc_y)], reactive=is_reactive,
                     kinetic_model=kinetic_model, max_steps=max_steps, inlet_primitive=inlet_primitive,
                     lambda_init=lambda_init, beta_init=beta_init, precision=precision, device=device)
    dx, dy = lx / nx, ly / ny
    try:
        initial_mass = float(np.sum(sim.state[0]) * dx * dy)
        time, steps = _legacy_run(sim, t_end, max_steps, {
            "Integration stalled": "Time integration stalled before the requested final time",
            "Could not find stable reactive step": "Could not find a CFL-stable reactive step"})
        state, lam, beta = sim.state, sim.lam, sim.beta
        boundary_mass = float(sim.boundary_integral[0])
        g = sim.gamma
    finally:
        sim.close()
    final_mass = float(np.sum(state[0]) * dx * dy)
    p = state.copy()
    p[1:3] /= state[0]
    p[3] = (g - 1) * (state[3] - 0.5 * (state[1] ** 2 + state[2] ** 2) / state[0])
    x = np.linspace(0.5 * dx, lx - 0.5 * dx, nx)
    y = np.linspace(0.5 * dy, ly - 0.5 * dy, ny)
    X, Y = np.meshgrid(x, y)
    return {"x": x, "y": y, "X": X, "Y": Y, "u_final": state, "u_prim_final": p, "lambda_final": lam,
            "beta_final": beta, "time": time, "steps": steps,
            "mass_conservation_error": abs(final_mass - initial_mass + boundary_mass) / max(1e-12, initial_mass),
            "mass_change_rel": (final_mass - initial_mass) / initial_mass, "net_boundary_mass": boundary_mass,
            "exit_thrust_per_depth": float(dx * np.sum(p[0, -1] * p[2, -1] ** 2 + p[3, -1] - ambient_pressure))}
