"""Legacy 1D drivers on the GPU, with every time level kept.

Mirrors python/cfdSolver/torchSolver.py: ``run_simulation`` (:1135-1233,
inert) and ``run_simulation_wqr`` / ``run_simulation_wQR`` (:1236-1418,
Arrhenius or two-stage ZND), with their validation (:1074-1122) and
two-state initial data (:918-969). The state stays on the device; each
step's level is downloaded as it is produced (the history the reference
returns: states (3, n, steps + 1), times, and for the reactive driver the
lambda levels (steps + 1, n)). The loop keeps the drivers' own rules, which
differ from Simulation.step: dt clamped with ``if t + dt > t_end``, a
stall below 1e-15, ``t += dt``, and in the reactive driver a CFL test of
the heated state against hydro_dt_1d alone (``torch_state_wave_maxima``,
which in the FP32 form measures the heating's exact change on the state
the half-step started from, so the test rejects what the FP64 one does).
"""

import numpy as np

from . import _abi
from .torchExecution import device_of
from .torchSolver3D import config

LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")


def validate_grid(length, cells, cfl, end_time):
    """torchSolver.py:1074-1088."""
    if (not isinstance(cells, (int, np.integer)) or not np.isfinite([length, cfl, end_time]).all()
            or length <= 0 or cells < 2 or not 0 < cfl <= 0.5 or end_time < 0):
        raise ValueError("Require length > 0, cells >= 2, 0 < CFL <= 0.5 and final time >= 0")


def validate_initial_1d(rho_l, u_l, p_l, rho_r, u_r, p_r, gamma, divider):
    """torchSolver.py:1091-1105."""
    values = [rho_l, u_l, p_l, rho_r, u_r, p_r, gamma, divider]
    if (not np.isfinite(values).all() or min(rho_l, rho_r, p_l, p_r) <= 0 or gamma <= 1
            or not 0 < divider < 100 or divider == 1):
        raise ValueError("Require finite primitives, positive density/pressure, gamma > 1 and divider in "
                         "(0,1) or (1,100) percent")


def validate_options_1d(limiter, controller, variant, waves, boundary):
    """torchSolver.py:1108-1118."""
    for value, allowed in ((limiter, LIMITERS), (controller, _abi.CONTROLLERS), (variant, _abi.VARIANTS),
                           (waves, _abi.WAVESPEEDS), (boundary, ("Open", "Reflective", "Periodic"))):
        if value not in allowed:
            raise ValueError(f"Unknown solver option: {value}")


def gamma_calc(lam):
    """torchSolver.py:71-78."""
    return (4.5 - lam) / (3.5 - lam)


def _initial(n, rho_l, u_l, p_l, rho_r, u_r, p_r, divider, gamma):
    """Conservative two-state data (torchSolver.py:918-969)."""
    ratio = divider / 100.0 if divider > 1.0 else divider
    cut = max(1, min(n - 1, int(np.round(ratio * n))))
    prim = np.zeros((3, n))
    prim[:, :cut] = np.array([rho_l, u_l, p_l])[:, None]
    prim[:, cut:] = np.array([rho_r, u_r, p_r])[:, None]
    u = prim.copy()
    u[1] = prim[1] * prim[0]
    u[2] = prim[2] / (gamma - 1) + 0.5 * prim[0] * prim[1] ** 2
    e = u[2] - 0.5 * u[1] ** 2 / u[0]
    if not (np.isfinite(u).all() and np.all(u[0] > 1e-12) and np.all(e > 1e-12)):
        raise ValueError("Initial state must have positive density and internal energy")
    return u


def _clamp(t, dt, t_end):
    if t + dt > t_end:
        dt = t_end - t
    if not np.isfinite(dt) or dt <= 1e-15:
        raise RuntimeError("Time integration stalled before the requested final time")
    return dt


def run_simulation(l, n, rho_l, u_l, p_l, rho_r, u_r, p_r, gamma=1.4, c_cfl=0.5, t_corte=0.2, limiter="VanLeer",
                   w_controler="Sound Speed Controlled", hllc_variante="Variante 1 - Original",
                   wave_speed_calc_option="Pressure Based", reflective_or_open="Open", divider=0.5,
                   diagnostics=None, execution=None):
    """Inert 1D run; returns (states (3, n, steps+1), times)."""
    validate_grid(l, n, c_cfl, t_corte)
    validate_options_1d(limiter, w_controler, hllc_variante, wave_speed_calc_option, reflective_or_open)
    validate_initial_1d(rho_l, u_l, p_l, rho_r, u_r, p_r, gamma, divider)
    dx = l / float(n)
    u = _initial(n, rho_l, u_l, p_l, rho_r, u_r, p_r, divider, gamma)
    ex = device_of(execution)
    cfg = config((n, 1, 1), None, (l, 1.0, 1.0), dim=1, gamma=gamma, limiter=limiter,
                 bcs=(reflective_or_open, "Open", "Open"), precision=ex.precision, cfl=c_cfl,
                 controller=_abi.CONTROLLERS.index(w_controler), variant=_abi.VARIANTS.index(hllc_variante),
                 wavespeed=_abi.WAVESPEEDS.index(wave_speed_calc_option))
    dev = _abi.DeviceState(ex.context, cfg)
    levels, times, t = [u], [0.0], 0.0
    boundary = np.zeros(3)
    try:
        dev.upload(u)
        while t < t_corte:
            dt = _clamp(t, dev.dt_limit(), t_corte)
            info = dev.hydro(dt)
            boundary += np.array(info.outward[:3])
            t += dt
            times.append(t)
            levels.append(dev.download()[0])
    finally:
        dev.close()
    if diagnostics is not None:
        diagnostics.update(boundary_integral=boundary, source_integral=np.zeros(3))
    return np.stack(levels, axis=2), np.array(times)


def run_simulation_wqr(l, n, rho_l, u_l, p_l, rho_r, u_r, p_r, lambda_init=None, c_cfl=0.5, t_corte=0.2,
                       limiter="VanLeer", w_controler="Sound Speed Controlled",
                       hllc_variante="Variante 1 - Original", wave_speed_calc_option="Pressure Based",
                       reflective_or_open="Open", divider=0.5, kinetic_model="arrhenius", diagnostics=None,
                       execution=None):
    """Reactive 1D run (Strang split R(dt/2) H(dt) R(dt/2)); returns
    (states (3, n, steps+1), times, lambda (steps+1, n))."""
    validate_grid(l, n, c_cfl, t_corte)
    validate_options_1d(limiter, w_controler, hllc_variante, wave_speed_calc_option, reflective_or_open)
    validate_initial_1d(rho_l, u_l, p_l, rho_r, u_r, p_r, 1.4, divider)
    if kinetic_model not in ("arrhenius", "znd"):
        raise ValueError(f"Unknown kinetic model: {kinetic_model}")
    dx = l / float(n)
    lam = np.ones(n) if lambda_init is None else np.asarray(lambda_init, dtype=float)
    if lam.shape != (n,) or not np.isfinite(lam).all() or np.any((lam < 0) | (lam > 1)):
        raise ValueError("lambda_init must have one value in [0,1] per cell")
    u = _initial(n, rho_l, u_l, p_l, rho_r, u_r, p_r, divider, gamma_calc(lam))
    beta = np.ones(n)
    ex = device_of(execution)
    cfg = config((n, 1, 1), None, (l, 1.0, 1.0), dim=1, limiter=limiter, bcs=(reflective_or_open, "Open", "Open"),
                 reactive=True, model=kinetic_model, gamma_mode=1, precision=ex.precision, cfl=c_cfl,
                 controller=_abi.CONTROLLERS.index(w_controler), variant=_abi.VARIANTS.index(hllc_variante),
                 wavespeed=_abi.WAVESPEEDS.index(wave_speed_calc_option))
    dev = _abi.DeviceState(ex.context, cfg)
    levels, lams, times, t = [u], [lam], [0.0], 0.0
    boundary, source = np.zeros(3), np.zeros(3)
    try:
        dev.upload(u, lam, beta)
        while t < t_corte:
            dt = _clamp(t, dev.dt_limit(), t_corte)
            start = dev.totals()
            for attempt in range(60):
                if attempt:
                    dev.upload(u, lam, beta)
                dev.react(0.5 * dt)
                allowed = c_cfl * dx / dev.wave_maxima()[1]
                if dt <= allowed * (1.0 + 1e-12):
                    break
                dt = min(0.5 * dt, allowed)
            else:
                raise RuntimeError("Could not find a CFL-stable reactive step")
            heated = dev.totals()
            info = dev.hydro(dt)
            hydro = dev.totals()
            dev.react(0.5 * dt)
            final = dev.totals()
            boundary += np.array(info.outward[:3])
            source += ((heated - start) + final) - hydro
            u, lam, beta = dev.download(True, True)
            t += dt
            times.append(t)
            levels.append(u)
            lams.append(lam)
    finally:
        dev.close()
    if diagnostics is not None:
        diagnostics.update(boundary_integral=boundary, source_integral=source)
    return np.stack(levels, axis=2), np.array(times), np.stack(lams, axis=0)


run_simulation_wQR = run_simulation_wqr
