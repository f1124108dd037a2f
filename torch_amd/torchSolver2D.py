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
