"""Cartesian 2D MUSCL-Hancock + HLLC step (oracle restatement).

Restates python/cfdSolver/torchSolver2D.py:38-91, 124-530, 647-667. States
are (4, ny, nx) [rho, rho*u, rho*v, E]. The 2D scheme differs from 3D: the
predictor fluxes floor rho and p, the predicted state is repaired instead of
triggering a fallback, the CFL runs over the padded state, Unos lacks the
(1+2r)/3 bound, and the y axis also accepts an ``Inflow`` boundary.
"""

import numpy as np

from . import solver1d as one
from .flux import hllc_grid
from .solver3d import bc_kind, ghost_map

INLET_DEFAULT = [1.0, 0.0, 100.0, 101325.0]  # torchSolver2D.py:228-232


def to_prim(u, gamma=1.4):
    """Floored primitives (torchSolver2D.py:38-58)."""
    rho = np.maximum(u[0], 1e-12)
    vx, vy = u[1] / rho, u[2] / rho
    ke = 0.5 * rho * (vx**2 + vy**2)
    p = np.maximum((gamma - 1.0) * (u[3] - ke), 1e-12)
    return np.stack((rho, vx, vy, p))


def to_cons(prim, gamma=1.4):
    """Floored conversion (torchSolver2D.py:61-82)."""
    rho = np.maximum(prim[0], 1e-12)
    p = np.maximum(prim[3], 1e-12)
    ke = 0.5 * rho * (prim[1] ** 2 + prim[2] ** 2)
    return np.stack((rho, rho * prim[1], rho * prim[2], p / (gamma - 1.0) + ke))


def sound(prim, gamma=1.4):
    """torchSolver2D.py:85-91."""
    return np.sqrt(gamma * np.maximum(prim[3], 1e-12) / np.maximum(prim[0], 1e-12))


def inlet_state(inlet):
    inlet = np.asarray(INLET_DEFAULT if inlet is None else inlet, dtype=float)
    if (inlet.shape != (4,) or not np.isfinite(inlet).all() or inlet[0] <= 0
            or inlet[3] <= 0 or inlet[2] < 0):
        raise ValueError("Inlet requires [rho>0, u, v>=0, p>0] in SI")
    return to_cons(inlet[:, None, None], 1.4)[:, 0, 0]


def pad(u, bc_x, bc_y, inlet=None):
    """Two ghost layers (torchSolver2D.py:176-266), as x then y index maps.

    Reflective ghosts negate the normal momentum; the Inflow y boundary puts
    the fixed inlet state in both lower ghost rows and copies the last
    interior row (with its x ghosts) into both upper rows.
    """
    ny, nx = u.shape[1:]
    kx = bc_kind(bc_x)
    sx, gx = ghost_map(nx, kx)
    out = np.take(u, sx, axis=2)
    if kx == "reflective":
        out[1] *= np.where(gx, -1.0, 1.0)
    if bc_y.lower() == "inflow":
        state = inlet_state(inlet)
        sy = np.clip(np.arange(-2, ny + 2), 0, ny - 1)
        out = np.take(out, sy, axis=1)
        out[:, :2, :] = state[:, None, None]
        return out
    ky = bc_kind(bc_y) if bc_y.lower() != "inflow" else "open"
    sy, gy = ghost_map(ny, ky)
    out = np.take(out, sy, axis=1)
    if ky == "reflective":
        out[2] *= np.where(gy, -1.0, 1.0)[:, None]
    return out


def pad_scalar(q, bc_x, bc_y, inlet_value=1.0):
    """torchSolver2D.py:647-656."""
    inflow = bc_y.lower() == "inflow"
    ny, nx = q.shape
    sx, _ = ghost_map(nx, bc_kind(bc_x))
    sy, _ = ghost_map(ny, "open" if inflow else bc_kind(bc_y))
    out = np.asarray(q)[np.ix_(sy, sx)].astype(float)
    if inflow:
        out[:2, :] = inlet_value
    return out


def slopes(arr, axis, limiter):
    """torchSolver2D.py:124-168 (axis 1 = x, 0 = y of the spatial grid)."""
    lim = limiter.lower()
    if axis == 1:
        dm, dp = arr[:, :, 1:-1] - arr[:, :, :-2], arr[:, :, 2:] - arr[:, :, 1:-1]
    else:
        dm, dp = arr[:, 1:-1, :] - arr[:, :-2, :], arr[:, 2:, :] - arr[:, 1:-1, :]
    r = dm / np.where(np.abs(dp) < 1e-14, 1e-14, dp)
    if lim in ("minbee", "minmod"):
        phi = np.maximum(0.0, np.minimum(r, 1.0))
    elif lim == "superbee":
        phi = np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0)))
    elif lim == "unos":
        phi = np.maximum(0.0, np.minimum(np.minimum(2.0, 2.0 * r), (2.0 + r) / 3.0))
    elif lim in ("albada", "vanalbada", "van_albada"):
        rp = np.maximum(r, 0.0)
        phi = np.where(r > 0.0, (rp**2 + rp) / (rp**2 + 1.0), 0.0)
    else:  # vanleer and the reference's default branch
        rp = np.maximum(r, 0.0)
        phi = np.where(r > 0.0, 2.0 * rp / (1.0 + rp), 0.0)
    return phi * dp


def _floored_flux(u, gamma, normal):
    """Predictor fluxes with rho and p floors (torchSolver2D.py:326-352)."""
    rho = np.maximum(u[0], 1e-12)
    vx, vy = u[1] / rho, u[2] / rho
    p = np.maximum((gamma - 1.0) * (u[3] - 0.5 * rho * (vx**2 + vy**2)), 1e-12)
    f = np.empty_like(u)
    if normal == 0:
        f[0], f[1], f[2], f[3] = rho * vx, rho * vx**2 + p, rho * vx * vy, vx * (u[3] + p)
    else:
        f[0], f[1], f[2], f[3] = rho * vy, rho * vx * vy, rho * vy**2 + p, vy * (u[3] + p)
    return f


def llf_fluxes(u_pad, gpad):
    """First-order LLF fluxes on floored primitives (torchSolver2D.py:420-457)."""
    prim = to_prim(u_pad, gpad)
    _, ux, vy, p = prim
    a = sound(prim, gpad)
    fx = np.stack((u_pad[1], u_pad[1] * ux + p, u_pad[2] * ux, ux * (u_pad[3] + p)))
    fy = np.stack((u_pad[2], u_pad[1] * vy, u_pad[2] * vy + p, vy * (u_pad[3] + p)))
    sx, sy = np.abs(ux) + a, np.abs(vy) + a
    ax = np.maximum(sx[2:-2, 1:-2], sx[2:-2, 2:-1])
    ay = np.maximum(sy[1:-2, 2:-2], sy[2:-1, 2:-2])
    flux_x = 0.5 * (fx[:, 2:-2, 1:-2] + fx[:, 2:-2, 2:-1] - ax * (u_pad[:, 2:-2, 2:-1] - u_pad[:, 2:-2, 1:-2]))
    flux_y = 0.5 * (fy[:, 1:-2, 2:-2] + fy[:, 2:-1, 2:-2] - ay * (u_pad[:, 2:-1, 2:-2] - u_pad[:, 1:-2, 2:-2]))
    return flux_x, flux_y


def hydro(u, dx, dy, dt, gamma=1.4, limiter="VanLeer", bc_x="Periodic", bc_y="Open", inlet=None):
    """One unsplit 2D step (torchSolver2D.py:274-466).

    Returns (U', Fx, Fy, used_fallback).
    """
    ny, nx = u.shape[1:]
    u_pad = pad(u, bc_x, bc_y, inlet)
    gpad = pad_scalar(np.broadcast_to(gamma, (ny, nx)), bc_x, bc_y, inlet_value=1.4)
    g = gpad[1:ny + 3, 1:nx + 3]
    active = u_pad[:, 1:ny + 3, 1:nx + 3]
    dxs = slopes(u_pad[:, 1:ny + 3, :], 1, limiter)
    dys = slopes(u_pad[:, :, 1:nx + 3], 0, limiter)
    ubar = (active - (0.5 * dt / dx) * (_floored_flux(active + 0.5 * dxs, g, 0) - _floored_flux(active - 0.5 * dxs, g, 0))
            - (0.5 * dt / dy) * (_floored_flux(active + 0.5 * dys, g, 1) - _floored_flux(active - 0.5 * dys, g, 1)))
    ubar[0] = np.maximum(ubar[0], 1e-12)
    kin = 0.5 * (ubar[1] ** 2 + ubar[2] ** 2) / ubar[0]
    pbar = (g - 1.0) * (ubar[3] - kin)
    ubar[3] = np.where(pbar < 1e-12, 1e-12 / (g - 1.0) + kin, ubar[3])
    fx = hllc_grid(ubar[:, 1:ny + 1, :-1] + 0.5 * dxs[:, 1:ny + 1, :-1],
                   ubar[:, 1:ny + 1, 1:] - 0.5 * dxs[:, 1:ny + 1, 1:],
                   g[1:ny + 1, :-1], g[1:ny + 1, 1:], direction=0)
    fy = hllc_grid(ubar[:, :-1, 1:nx + 1] + 0.5 * dys[:, :-1, 1:nx + 1],
                   ubar[:, 1:, 1:nx + 1] - 0.5 * dys[:, 1:, 1:nx + 1],
                   g[:-1, 1:nx + 1], g[1:, 1:nx + 1], direction=1)
    out = u - (dt / dx) * (fx[:, :, 1:] - fx[:, :, :-1]) - (dt / dy) * (fy[:, 1:, :] - fy[:, :-1, :])
    fell_back = False
    if not one.physical_state(out):
        fell_back = True
        fx, fy = llf_fluxes(u_pad, gpad)
        out = u - dt / dx * np.diff(fx, axis=2) - dt / dy * np.diff(fy, axis=1)
    if not one.physical_state(out):
        raise RuntimeError("Hydrodynamic update produced a nonphysical state")
    return out, fx, fy, fell_back


def muscl_hancock_step_2d(u, dx, dy, dt, gamma=1.4, limiter="VanLeer", bc_x="Periodic", bc_y="Open",
                          return_fluxes=False, inlet_primitive=None):
    """torchSolver2D.py:274-466 signature."""
    out, fx, fy, _ = hydro(u, dx, dy, dt, gamma, limiter, bc_x, bc_y, inlet_primitive)
    return (out, fx, fy) if return_fluxes else out


def cfl_dt(state, gamma, dx, dy, cfl, bc_x, bc_y, inlet=None):
    """Session 2D time step over the padded state (torchSession.py:346-355,
    torchSolver2D.py:512-530)."""
    ny, nx = state.shape[1:]
    padded = pad(state, bc_x, bc_y, inlet)
    gp = pad_scalar(np.broadcast_to(gamma, (ny, nx)), bc_x, bc_y, inlet_value=1.4)
    prim = to_prim(padded, gp)
    a = sound(prim, gp)
    total = np.max((np.abs(prim[1]) + a) / dx) + np.max((np.abs(prim[2]) + a) / dy)
    if total <= 1e-12:
        return 1e-3
    return float(cfl / total)


def advect(q, old, new, mass_x, mass_y, dt_dx, dt_dy, bc_x, bc_y):
    """torchSolver2D.py:659-667."""
    p = pad_scalar(q, bc_x, bc_y)
    fx = mass_x * np.where(mass_x >= 0, p[2:-2, 1:-2], p[2:-2, 2:-1])
    fy = mass_y * np.where(mass_y >= 0, p[1:-2, 2:-2], p[2:-1, 2:-2])
    rho_q = old[0] * q - dt_dx * np.diff(fx, axis=1) - dt_dy * np.diff(fy, axis=0)
    return np.clip(rho_q / np.maximum(new[0], 1e-12), 0.0, 1.0)


def reaction(u, lam, beta, dt, model="arrhenius", **constants):
    """torchSolver2D.py:474-509."""
    shape = lam.shape
    flat = u.reshape(4, -1)
    normal = np.stack((flat[0], flat[1], flat[3] - 0.5 * flat[2] ** 2 / flat[0]))
    nl, reacted, nb = one.kinetics(lam.ravel(), dt, normal, model,
                                   None if beta is None else beta.ravel(), **constants)
    result = np.array(u, dtype=float, copy=True)
    result[3] += (reacted[2] - normal[2]).reshape(shape)
    return result, nl.reshape(shape), nb.reshape(shape) if beta is not None else None
