"""Cartesian 3D MUSCL-Hancock + HLLC step (oracle restatement).

Restates python/cfdSolver/torchSolver3D.py:27-256. States are
(5, nz, ny, nx) float64 [rho, rho*u, rho*v, rho*w, E], x fastest.
Boundaries are given in (z, y, x) order as in the reference's ``_hydro``.
Ghost layers are built from integer index maps (the same maps the HIP
kernels apply at load time) instead of materialised ``np.pad`` copies.
"""

import numpy as np

from . import solver1d as one
from .flux import hllc_grid

LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")
PERIODIC = ("periodic", "wrap")
OPEN = ("open", "transmissive", "outflow")
REFLECTIVE = ("reflective", "wall", "slip")


def bc_kind(name):
    mode = name.lower()
    if mode in PERIODIC:
        return "periodic"
    if mode in OPEN:
        return "open"
    if mode in REFLECTIVE:
        return "reflective"
    raise ValueError(f"Unknown 3D boundary: {name}")


def ghost_map(n, kind, width=2):
    """Source index and 'is ghost' flag for padded positions -width..n+width-1.

    Equivalent to np.pad modes wrap / edge / symmetric used at
    torchSolver3D.py:47-70 (MATLAB statement of the same maps:
    matlab/solvers/torchSolver3D.m:124-126).
    """
    i = np.arange(-width, n + width)
    if kind == "periodic":
        src = np.mod(i, n)
    elif kind == "open":
        src = np.clip(i, 0, n - 1)
    else:
        src = np.where(i < 0, -i - 1, np.where(i >= n, 2 * n - 1 - i, i))
    return src, (i < 0) | (i >= n)


def pad(values, boundaries, vector=False, width=2):
    """Ghost layers on the spatial axes (torchSolver3D.py:47-70).

    Reflective ghosts negate the momentum component normal to that wall
    (component 3 - spatial_axis) when ``vector``; scalars keep their sign.
    """
    out = values
    lead = 1 if vector else 0
    for axis, name in enumerate(boundaries):  # z, y, x
        kind = bc_kind(name)
        n = values.shape[axis + lead]
        src, is_ghost = ghost_map(n, kind, width)
        out = np.take(out, src, axis=axis + lead)
        if vector and kind == "reflective":
            shape = [1] * out.ndim
            shape[axis + lead] = src.size
            sign = np.where(is_ghost, -1.0, 1.0).reshape(shape[1:])
            out = out.copy()
            out[3 - axis] = out[3 - axis] * sign
    return out


def to_prim(state, gamma):
    """torchSolver3D.py:37-44 (no floors)."""
    prim = state.copy()
    prim[1:4] = state[1:4] / state[0]
    prim[4] = (gamma - 1) * (state[4] - 0.5 * np.sum(state[1:4] ** 2, axis=0) / state[0])
    return prim


def to_cons(prim, gamma):
    """torchSolver3D.py:27-34."""
    prim = np.asarray(prim, dtype=float)
    state = prim.copy()
    state[1:4] *= prim[0]
    state[4] = prim[4] / (gamma - 1) + 0.5 * prim[0] * np.sum(prim[1:4] ** 2, axis=0)
    return state


def directional_flux(state, gamma, m):
    """Physical flux along momentum component m (torchSolver3D.py:77-83)."""
    prim = to_prim(state, gamma)
    vel, p = prim[m], prim[4]
    f = state * vel
    f[m] += p
    f[4] = vel * (state[4] + p)
    return f


def limiter_phi(r, limiter):
    """torchSolver3D.py:91-110."""
    pos = np.maximum(r, 0)
    if limiter == "VanLeer":
        return 2 * pos / (1 + pos)
    if limiter == "Albada":
        return (pos**2 + pos) / (pos**2 + 1)
    if limiter == "MINBEE":
        return np.maximum(0, np.minimum(r, 1))
    if limiter == "Superbee":
        return np.maximum(0, np.maximum(np.minimum(2 * r, 1), np.minimum(r, 2)))
    if limiter == "Unos":
        return np.maximum(0, np.minimum(np.minimum(np.minimum(2, 2 * r), (2 + r) / 3), (1 + 2 * r) / 3))
    raise ValueError(f"Unknown limiter: {limiter}")


def slope(padded, axis, limiter):
    """Limited conservative slope on the active region (torchSolver3D.py:86-111).

    ``axis`` is the array axis of ``padded`` (3 = x, 2 = y, 1 = z). The
    active region drops one ghost layer on every spatial axis.
    """
    def along(lo, hi):
        index = [slice(None), slice(1, -1), slice(1, -1), slice(1, -1)]
        index[axis] = slice(lo, hi)
        return padded[tuple(index)]

    centre = along(1, -1)
    minus = centre - along(0, -2)
    plus = along(2, None) - centre
    r = minus / np.where(abs(plus) < 1e-14, 1e-14, plus)
    return limiter_phi(r, limiter) * plus


def faces(array, axis, active):
    """Left/right face operands (torchSolver3D.py:114-119)."""
    left = [slice(None)] + [slice(1, -1) if active else slice(2, -2)] * 3
    right = list(left)
    left[axis] = slice(None, -1) if active else slice(1, -2)
    right[axis] = slice(1, None) if active else slice(2, -1)
    return array[tuple(left)], array[tuple(right)]


def first_order_fluxes(padded, gp):
    """Rusanov/LLF fluxes on cell averages (torchSolver3D.py:161-176)."""
    fluxes = []
    for m, axis in enumerate((3, 2, 1), 1):
        left, right = faces(padded, axis, False)
        gl, gr = faces(gp[None], axis, False)
        pl, pr = to_prim(left, gl[0]), to_prim(right, gr[0])
        with np.errstate(invalid="ignore"):
            speed = np.maximum(
                abs(pl[m]) + np.sqrt(gl[0] * pl[4] / pl[0]),
                abs(pr[m]) + np.sqrt(gr[0] * pr[4] / pr[0]),
            )
        fluxes.append(0.5 * (directional_flux(left, gl[0], m) + directional_flux(right, gr[0], m)
                             - speed * (right - left)))
    return fluxes


def predictor(padded, gp, spacing, dt, limiter):
    """Slopes and half-step predicted active states (torchSolver3D.py:134-151).

    Returns (slopes, predicted, physical) where ``physical`` is False when
    any predicted face state P +- s/2 fails ``physical_state``.
    """
    active = padded[:, 1:-1, 1:-1, 1:-1]
    ga = gp[1:-1, 1:-1, 1:-1]
    slopes = [slope(padded, a, limiter) for a in (3, 2, 1)]
    predicted = active.copy()
    for m, (s, h) in enumerate(zip(slopes, spacing), 1):
        predicted -= 0.5 * dt / h * (
            directional_flux(active + 0.5 * s, ga, m) - directional_flux(active - 0.5 * s, ga, m)
        )
    ok = all(one.physical_state(predicted + sign * 0.5 * s) for s in slopes for sign in (-1, 1))
    return slopes, predicted, ok


def hydro(state, gamma, spacing, dt, limiter, boundaries, first_order=False, batch_size=None, execution=None):
    """One unsplit step with the global first-order fallback (torchSolver3D.py:122-203).

    Returns (U', [Fx, Fy, Fz], used_first_order). Raises RuntimeError when
    the first-order result is also unphysical.
    """
    padded = pad(state, boundaries, True)
    gp = pad(np.broadcast_to(gamma, state.shape[1:]), boundaries)
    if not first_order:
        slopes, predicted, ok = predictor(padded, gp, spacing, dt, limiter)
        if not ok:
            return hydro(state, gamma, spacing, dt, limiter, boundaries, True, batch_size, execution)
        ga = gp[1:-1, 1:-1, 1:-1]
        fluxes = []
        for m, axis in enumerate((3, 2, 1), 1):
            pl, pr = faces(predicted, axis, True)
            sl, sr = faces(slopes[m - 1], axis, True)
            gl, gr = faces(ga[None], axis, True)
            fluxes.append(hllc_grid(pl + 0.5 * sl, pr - 0.5 * sr, gl[0], gr[0], direction=m - 1,
                                    batch_size=batch_size, execution=execution))
    else:
        fluxes = first_order_fluxes(padded, gp)
    result = state.copy()
    for flux, h, axis in zip(fluxes, spacing, (3, 2, 1)):
        result -= dt / h * np.diff(flux, axis=axis)
    if not one.physical_state(result):
        if not first_order:
            return hydro(state, gamma, spacing, dt, limiter, boundaries, True, batch_size, execution)
        raise RuntimeError("3D hydrodynamic update produced a nonphysical state")
    return result, fluxes, first_order


def muscl_hancock_step_3d(state, dx, dy, dz, dt, gamma=1.4, limiter="VanLeer",
                          bc_x="Open", bc_y="Open", bc_z="Open", return_fluxes=False):
    """torchSolver3D.py:206-222."""
    result, fluxes, _ = hydro(state, gamma, (dx, dy, dz), dt, limiter, (bc_z, bc_y, bc_x))
    return (result, *fluxes) if return_fluxes else result


def advect(q, old, new, fluxes, dt, spacing, boundaries):
    """Upwind transport of rho*q with the hydro mass fluxes (torchSolver3D.py:246-256)."""
    qp = pad(q, boundaries)
    rhoq = old[0] * q
    for flux, h, axis in zip(fluxes, spacing, (3, 2, 1)):
        left, right = faces(qp[None], axis, False)
        scalar = flux[0] * np.where(flux[0] >= 0, left[0], right[0])
        rhoq -= dt / h * np.diff(scalar, axis=axis - 1)
    fraction = rhoq / new[0]
    if np.any(fraction < -1e-10) or np.any(fraction > 1 + 1e-10):
        raise RuntimeError("3D scalar transport left the physical interval")
    return np.clip(fraction, 0, 1)


def reaction(state, lam, beta, dt, model="arrhenius", **constants):
    """Chemistry on the 1D-equivalent state (torchSolver3D.py:225-243).

    ``constants`` (q_heat, activation_energy, pre_exponential) default to the
    reference's; the reference's 3D wrapper never overrides them.
    """
    flat = state.reshape(5, -1)
    normal = np.stack((flat[0], flat[1], flat[4] - 0.5 * np.sum(flat[2:4] ** 2, axis=0) / flat[0]))
    new_lam, reacted, new_beta = one.kinetics(
        lam.ravel(), dt, normal, model, None if beta is None else beta.ravel(), **constants)
    result = state.copy()
    result[4] += (reacted[2] - normal[2]).reshape(lam.shape)
    return (result, new_lam.reshape(lam.shape),
            None if beta is None else new_beta.reshape(lam.shape))


def cfl_dt(state, gamma, spacing, cfl):
    """cfl / sum_m max(|u_m| + a)/h_m over the interior (torchSession.py:356-360)."""
    p = to_prim(state, gamma)
    a = np.sqrt(gamma * p[-1] / p[0])
    return cfl / sum(np.max(abs(p[m]) + a) / h for m, h in enumerate(spacing, 1))


def wave_maxima(state, gamma):
    """The three maxima max(|u_m| + a) that ``cfl_dt`` combines."""
    p = to_prim(state, gamma)
    a = np.sqrt(gamma * p[-1] / p[0])
    return np.array([np.max(abs(p[m]) + a) for m in (1, 2, 3)])
