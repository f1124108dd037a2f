"""Directional HLLC and Rusanov interface fluxes on the GPU.

Mirrors python/cfdSolver/torchFlux.py:6-77 (``hllc_flux_grid``),
:80-122 (``rusanov_flux``, ``rusanov_flux_grid``) and :125-248
(``hllc_flux``): same signatures, validation and component order.
The faces are evaluated by libtorchhip's HLLC and Rusanov kernels;
``batch_size`` is validated for compatibility but the whole grid is one
launch (the reference's result does not depend on batching).
"""

import numpy as np

from . import _abi
from .torchExecution import device_of


def hllc_flux_grid(left, right, gamma, gamma_right=None, *, direction=0, batch_size=8192, execution=None):
    left = np.asarray(left, dtype=float)
    right = np.asarray(right, dtype=float)
    if left.shape != right.shape or left.ndim < 2 or left.shape[0] not in (4, 5):
        raise ValueError("HLLC requires matching four/five-component grids")
    if direction not in range(left.shape[0] - 2):
        raise ValueError("Invalid normal direction")
    if not isinstance(batch_size, int) or isinstance(batch_size, bool) or batch_size < 1:
        raise ValueError("HLLC batch size must be a positive integer")
    nv = left.shape[0]
    faces = left.shape[1:]
    gl = np.broadcast_to(gamma, faces).reshape(-1)
    gr = gl if gamma_right is None else np.broadcast_to(gamma_right, faces).reshape(-1)
    ex = device_of(execution)
    out = _abi.hllc_faces(ex.context, nv, direction, left.reshape(nv, -1), right.reshape(nv, -1), gl, gr,
                          _abi.precision_code(ex.precision))
    return out.reshape(left.shape)


def hllc_flux(u_l, u_r, gamma=1.4, wave_speed_mode="Pressure Based", gamma_right=None, execution=None):
    """Normal-basis HLLC for (4, m) or (5, m) states (rows [rho, m_n, m_t.., E])."""
    u_l = np.asarray(u_l, dtype=float)
    u_r = np.asarray(u_r, dtype=float)
    if u_l.ndim != 2 or u_l.shape != u_r.shape or u_l.shape[0] not in (4, 5):
        raise ValueError("HLLC requires matching (4,m) or (5,m) states")
    # In the normal basis the normal momentum is row 1, i.e. direction 0.
    return hllc_flux_grid(u_l, u_r, gamma, gamma if gamma_right is None else gamma_right,
                          direction=0, execution=execution)


def rusanov_flux_grid(left, right, gl, gr, direction, execution=None):
    """Rusanov flux of primitive face states (rho, velocities..., p) along
    ``direction`` (torchFlux.py:103-122); the reference's 8192-face worker
    batches are one launch here."""
    left = np.asarray(left, dtype=float)
    right = np.asarray(right, dtype=float)
    if left.shape != right.shape or left.ndim < 2 or left.shape[0] not in (3, 4, 5):
        raise ValueError("Rusanov requires matching (dim+2)-component face grids")
    if direction not in range(left.shape[0] - 2):
        raise ValueError("Invalid normal direction")
    nv = left.shape[0]
    faces = left.shape[1:]
    gl = np.broadcast_to(gl, faces).reshape(-1)
    gr = np.broadcast_to(gr, faces).reshape(-1)
    ex = device_of(execution)
    out = _abi.rusanov_faces(ex.context, nv, direction, left.reshape(nv, -1), right.reshape(nv, -1), gl, gr,
                             _abi.precision_code(ex.precision))
    return out.reshape(left.shape)


def rusanov_flux(left, right, gl, gr, direction, execution=None):
    """Independent normal Rusanov flux, including tangential momenta (torchFlux.py:80-100)."""
    return rusanov_flux_grid(left, right, gl, gr, direction, execution)
