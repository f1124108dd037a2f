"""Cartesian 3D step functions on the GPU.

Mirrors the public functions of python/cfdSolver/torchSolver3D.py: state
conversions (:27-44), ``muscl_hancock_step_3d`` (:206-222, the one-step
``_hydro`` of :122-203), ``reaction_step_3d`` (:225-243) and the presets
(:259-286). States are (5, nz, ny, nx) float64 host arrays in and out; the
update itself runs in libtorchhip (fused predictor/HLLC/update kernel plus
the device-side first-order fallback).
"""

import numpy as np

from . import _abi
from .presets import PRESETS_3D, get_3d_preset  # noqa: F401  (re-exported)
from .torchExecution import device_of

LIMITERS = _abi.LIMITERS


def uprim_to_uconsv_3d(primitive, gamma=1.4):
    primitive = np.asarray(primitive, dtype=float)
    state = primitive.copy()
    state[1:4] *= primitive[0]
    state[4] = primitive[4] / (gamma - 1) + 0.5 * primitive[0] * np.sum(primitive[1:4] ** 2, axis=0)
    return state


def uconsv_to_uprim_3d(state, gamma=1.4):
    state = np.asarray(state, dtype=float)
    primitive = state.copy()
    primitive[1:4] /= state[0]
    primitive[4] = (gamma - 1) * (state[4] - 0.5 * np.sum(state[1:4] ** 2, axis=0) / state[0])
    return primitive


def config(counts, spacing=None, lengths=None, *, dim, gamma=1.4, limiter="VanLeer", bcs=("Open",) * 3,
           reactive=False, model="arrhenius", gamma_mode=0, precision="fp64", cfl=0.4, inlet=None,
           constants=None, controller=0, variant=0, wavespeed=0, tile=(0, 0, 0)):
    """Fill a libtorchhip Config (counts, spacing and boundaries in x, y, z order)."""
    if limiter not in LIMITERS:
        raise ValueError(f"Unknown limiter: {limiter}")
    c = _abi.Config()
    c.dim = dim
    for d in range(3):
        c.n[d] = int(counts[d]) if d < dim else 1
        c.lengths[d] = float(lengths[d]) if lengths is not None and d < dim else 0.0
        c.spacing[d] = float(spacing[d]) if spacing is not None and d < dim else 0.0
        c.bc[d] = _abi.boundary_code(bcs[d], dim, d) if d < dim else 0
    c.gamma = float(gamma) if np.isscalar(gamma) else 1.4
    c.limiter = LIMITERS.index(limiter)
    inlet = [1.0, 0.0, 100.0, 101325.0] if inlet is None else list(inlet)
    for k in range(4):
        c.inlet_primitive[k] = float(inlet[k])
    c.reactive = int(bool(reactive))
    if model not in ("arrhenius", "znd"):
        raise ValueError("Unknown limiter or chemistry model")
    c.kinetic_model = 0 if model == "arrhenius" else 1
    k = dict(q_heat=141584000.0, activation_energy=17.15 * 4184.0, pre_exponential=1.76e5)
    k.update(constants or {})
    c.q_heat, c.activation_energy, c.pre_exponential = k["q_heat"], k["activation_energy"], k["pre_exponential"]
    c.controller, c.variant, c.wavespeed = controller, variant, wavespeed
    c.cfl = float(cfl)
    c.precision = _abi.precision_code(precision)
    c.block_x, c.block_y, c.zchunk = tile
    c.gamma_mode = gamma_mode
    return c


def _gamma_array(gamma, shape):
    if np.isscalar(gamma) or np.ndim(gamma) == 0:
        return None
    return np.ascontiguousarray(np.broadcast_to(gamma, shape), dtype=float)


def hydro(state, gamma, spacing, dt, limiter, boundaries, execution=None, return_fluxes=True):
    """One unsplit step; ``boundaries`` in (z, y, x) order like ``_hydro``.

    Returns (U', [Fx, Fy, Fz]) (fluxes only when requested).
    """
    state = np.ascontiguousarray(state, dtype=float)
    if state.ndim != 4 or state.shape[0] != 5:
        raise ValueError("3D states have shape (5, nz, ny, nx)")
    nz, ny, nx = state.shape[1:]
    garr = _gamma_array(gamma, state.shape[1:])
    ex = device_of(execution)
    bz, by, bx = boundaries
    tile = tuple(getattr(ex, "tile", None) or (0, 0, 0))
    cfg = config((nx, ny, nz), spacing, dim=3, gamma=gamma if garr is None else 1.4, limiter=limiter,
                 bcs=(bx, by, bz), gamma_mode=0 if garr is None else 2, precision=ex.precision, tile=tile)
    dev = _abi.DeviceState(ex.context, cfg)
    try:
        dev.upload(state, garr)
        dev.hydro(dt, want_fluxes=return_fluxes)
        result, _, _ = dev.download()
        fluxes = dev.fluxes() if return_fluxes else []
    finally:
        dev.close()
    return result, fluxes


def muscl_hancock_step_3d(state, dx, dy, dz, dt, gamma=1.4, limiter="VanLeer", bc_x="Open", bc_y="Open",
                          bc_z="Open", return_fluxes=False, execution=None):
    result, fluxes = hydro(state, gamma, (dx, dy, dz), dt, limiter, (bc_z, bc_y, bc_x), execution,
                           return_fluxes)
    return (result, *fluxes) if return_fluxes else result


def reaction_step_3d(state, lam, beta, dt, model="arrhenius", execution=None, constants=None):
    """Exponential-midpoint chemistry on the 1D-equivalent state; only E changes."""
    state = np.ascontiguousarray(state, dtype=float)
    lam = np.asarray(lam, dtype=float)
    nz, ny, nx = state.shape[1:]
    ex = device_of(execution)
    cfg = config((nx, ny, nz), (1.0, 1.0, 1.0), dim=3, reactive=True, model=model, gamma_mode=1,
                 precision=ex.precision, constants=constants)
    dev = _abi.DeviceState(ex.context, cfg)
    try:
        dev.upload(state, lam, beta)
        dev.react(dt)
        out, new_lam, new_beta = dev.download(True, True)
    finally:
        dev.close()
    return out, new_lam.reshape(lam.shape), None if beta is None else new_beta.reshape(lam.shape)


def _legacy_run(sim, t_end, max_steps, messages):
    """Step a device session with the legacy drivers' running time (t += dt);
    returns (time, steps). The session's dt is min(dt_cfl, t_end - t) and
    its stall test rejects dt <= 0 or t + dt == t: the 3D driver's rules, and
    the 2D driver's clamp except where rounding separates the two tests."""
    time, steps = 0.0, 0
    while time < t_end and steps < max_steps:
        sim.time = time
        try:
            sim.step()
        except RuntimeError as exc:
            for prefix, message in messages.items():
                if str(exc).startswith(prefix):
                    raise RuntimeError(message) from exc
            raise
        time += sim.last_dt
        steps += 1
    if time < t_end:
        raise RuntimeError(f"Maximum steps reached at t={time}; requested {t_end}")
    return time, steps


def run_torch_solver_3d(preset_name="3d_explosion", *, primitive_init=None, nx=20, ny=20, nz=20, lx=1.0, ly=1.0,
                        lz=1.0, t_end=None, cfl=0.4, gamma=1.4, limiter="VanLeer", bc_x="Open", bc_y="Open",
                        bc_z="Open", is_reactive=False, kinetic_model="arrhenius", lambda_init=None, beta_init=None,
                        max_steps=100000, precision="fp64", device=0):
    """Legacy 3D driver (torchSolver3D.py:289-471) on a device session:
    final state, coordinates and the five-equation balances."""
    from .torchSession import Simulation

    if t_end is None:
        t_end = 1e-5 if preset_name == "3d_ignition" and primitive_init is None else 0.05
    if not np.isscalar(gamma) or not np.isfinite(gamma) or gamma <= 1:
        raise ValueError("gamma must be finite and greater than one")
    if limiter not in LIMITERS or kinetic_model not in ("arrhenius", "znd"):
        raise ValueError("Invalid 3D limiter or kinetic model")
    if not isinstance(max_steps, (int, np.integer)) or max_steps < 1:
        raise ValueError("max_steps must be a positive integer")
    if primitive_init is None:
        cfg = get_3d_preset(preset_name, nx, ny, nz)
        primitive_init, is_reactive = cfg["primitive"], cfg["is_reactive"]
    primitive = np.array(primitive_init, dtype=float, copy=True)
    if primitive.shape != (5, nz, ny, nx) or not np.isfinite(primitive).all() or np.any(primitive[[0, 4]] <= 0):
        raise ValueError("Require finite (5,nz,ny,nx) primitives with positive density/pressure")
    if not is_reactive and (lambda_init is not None or beta_init is not None):
        raise ValueError("Composition requires is_reactive=True")
    if beta_init is not None and kinetic_model != "znd":
        raise ValueError("beta_init requires reactive ZND kinetics")
    sim = Simulation(primitive, [lx, ly, lz], t_end=t_end, cfl=cfl, gamma=gamma, limiter=limiter,
                     boundaries=[bc_x, bc_y, bc_z], reactive=is_reactive, kinetic_model=kinetic_model,
                     lambda_init=lambda_init, beta_init=beta_init, max_steps=max_steps, precision=precision,
                     device=device)
    try:
        initial = sim.integral.copy()
        time, steps = _legacy_run(sim, t_end, max_steps, {
            "Integration stalled": "3D integration stalled before final time",
            "Could not find stable reactive step": "Could not find a stable 3D reactive step"})
        state, lam, beta = sim.state, sim.lam, sim.beta
        final = sim.integral
        boundary, source = sim.boundary_integral.copy(), sim.source_integral.copy()
    finally:
        sim.close()
    g = (4.5 - lam) / (3.5 - lam) if is_reactive else gamma
    spacing = (lx / nx, ly / ny, lz / nz)
    x, y, z = [(np.arange(n) + 0.5) * h for n, h in zip((nx, ny, nz), spacing)]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    residual = final - initial + boundary - source
    return dict(u_final=state, u_prim_final=uconsv_to_uprim_3d(state, g), lambda_final=lam, beta_final=beta,
                time=time, steps=steps, x=x, y=y, z=z, X=X, Y=Y, Z=Z, initial_integral=initial,
                final_integral=final, boundary_integral=boundary, source_integral=source,
                balance_residual=residual, mass_conservation_error=abs(residual[0]) / initial[0],
                mass_change_rel=(final[0] - initial[0]) / initial[0])
