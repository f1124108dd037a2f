"""Kinetics, state checks and the legacy 1D scheme (oracle restatement).

Restates python/cfdSolver/torchSolver.py. Conservative 1D states are
(3, n) arrays [rho, rho*u, E].
"""

import numpy as np

R_UNIVERSAL = 8314.46  # torchSolver.py:99 (J/(kmol K))
R_ARRHENIUS = 8.314  # torchSolver.py:769
Q_HEAT = 141584000.0  # torchSolver.py:715
E_ACT = 17.15 * 4184.0  # torchSolver.py:716
A_PRE = 1.76e5  # torchSolver.py:717
LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")


def gamma_calc(lam):
    """torchSolver.py:71-78."""
    return (4.5 - lam) / (3.5 - lam)


def physical_state(state):
    """torchSolver.py:1038-1046: finite, rho > 1e-12, internal energy > 1e-12."""
    state = np.asarray(state)
    return bool(
        np.isfinite(state).all()
        and np.all(state[0] > 1e-12)
        and np.all(
            state[-1] - 0.5 * np.sum(state[1:-1] ** 2, axis=0) / state[0]
            > 1e-12
        )
    )


def to_prim(u, gamma):
    """Floored [rho, u, p] (torchSolver.py:117-127)."""
    rho = np.maximum(u[0], 1e-12)
    vel = u[1] / rho
    p = np.maximum((u[2] - 0.5 * rho * (vel**2)) * (gamma - 1.0), 1e-12)
    return np.vstack((rho, vel, p))


def to_cons(prim, gamma):
    """torchSolver.py:130-142."""
    rho, vel, p = prim
    safe = np.where(np.abs(rho) < 1e-14, 1e-14, rho)
    return np.vstack((rho, rho * vel, rho * (0.5 * (vel**2) + p / ((gamma - 1.0) * safe))))


def sound(prim, gamma):
    """torchSolver.py:81-87."""
    return np.sqrt(gamma * np.maximum(prim[2], 0.0) / np.maximum(prim[0], 1e-12))


def euler_flux(u, gamma):
    """torchSolver.py:145-156 (floored primitives, raw energy)."""
    rho, vel, p = to_prim(u, gamma)
    return np.vstack((rho * vel, rho * (vel**2) + p, vel * (u[2] + p)))


def temperature(gamma, lam, state):
    """torchSolver.py:90-114 (mixture molar mass from lambda)."""
    prim = to_prim(state, gamma)
    rho = np.maximum(prim[0], 1e-12)
    molar = lam * (2.0 / 3.0) * 2.016 + lam * (1.0 / 3.0) * 32.0 + (1.0 - lam) * 18.016
    return (prim[2] * molar) / (rho * R_UNIVERSAL)


def kinetics(lam0, dt, state0, model="arrhenius", beta0=None, *,
             q_heat=Q_HEAT, activation_energy=E_ACT, pre_exponential=A_PRE):
    """Exponential-midpoint Arrhenius / two-stage ZND (torchSolver.py:708-789).

    Returns (lambda, state, beta) like ``new_lambda_n_e_calc``. Both
    sub-steps restart from (lambda0, beta0, state0); only the rate is taken
    at the half-step.
    """
    if model not in ("arrhenius", "znd"):
        raise ValueError(f"Unknown kinetic model: {model}")
    consts = [dt, q_heat, activation_energy, pre_exponential]
    if not np.isfinite(consts).all() or min(consts) < 0:
        raise ValueError("Reaction time and constants must be finite and nonnegative")
    lam0 = np.asarray(lam0, dtype=float)
    if np.any(~np.isfinite(lam0)) or np.any((lam0 < 0) | (lam0 > 1)):
        raise ValueError("Reactant fraction must lie in [0,1]")
    state0 = np.asarray(state0, dtype=float)
    if state0.shape != (3, lam0.size) or lam0.ndim != 1 or not physical_state(state0):
        raise ValueError("Reaction requires a physical (3,n) state and an (n,) reactant fraction")
    beta0 = np.ones_like(lam0) if beta0 is None else np.asarray(beta0, dtype=float)
    if beta0.shape != lam0.shape or not np.isfinite(beta0).all() or np.any((beta0 < 0) | (beta0 > 1)):
        raise ValueError("Induction fraction must have one value in [0,1] per cell")

    def rate(lam, state):
        t = np.maximum(temperature(gamma_calc(lam), lam, state), 1e-12)
        return pre_exponential * np.exp(-activation_energy / (R_ARRHENIUS * t))

    def advance(k, duration):
        if model == "znd":
            active = np.maximum(0.0, duration - beta0 / np.maximum(k, 1e-300))
            lam = lam0 * np.exp(-5.0 * k * active)
            beta = np.clip(beta0 - k * duration, 0.0, 1.0)
        else:
            lam = lam0 * np.exp(-k * duration)
            beta = np.zeros_like(beta0)
        state = np.copy(state0)
        state[2] += q_heat * state[0] * (lam0 - lam)
        return lam, state, beta

    lam_mid, state_mid, _ = advance(rate(lam0, state0), 0.5 * dt)
    return advance(rate(lam_mid, state_mid), dt)


# ---------------------------------------------------------------------------
# Legacy 1D MUSCL-Hancock (torchSolver.py:164-700, 1049-1132)
# ---------------------------------------------------------------------------

def ghost(u, bc):
    """One ghost cell per side (torchSolver.py:164-186)."""
    if bc == "Reflective":
        left = np.array([[u[0, 0]], [-u[1, 0]], [u[2, 0]]])
        right = np.array([[u[0, -1]], [-u[1, -1]], [u[2, -1]]])
    elif bc == "Periodic":
        left, right = u[:, -1:], u[:, :1]
    elif bc == "Open":
        left, right = u[:, :1], u[:, -1:]
    else:
        raise ValueError(f"Unknown boundary condition option: {bc}")
    return np.hstack((left, u, right))


def ghost_scalar(q, bc):
    """torchSolver.py:1121-1122 (no sign change; reflective walls use edge)."""
    return np.pad(q, 1, mode="wrap" if bc == "Periodic" else "edge")


def _zero_bad(x):
    return np.where(np.isfinite(x), x, 0.0)


def _split_gamma(gamma):
    if np.isscalar(gamma):
        return gamma, gamma
    return gamma[:-1], gamma[1:]


def wave_speeds(padded, gamma, mode="Pressure Based"):
    """Cell-average wave speeds S_L, S*, S_R per face (torchSolver.py:283-404)."""
    gl, gr = _split_gamma(gamma)
    with np.errstate(divide="ignore", invalid="ignore"):
        pl_prim = to_prim(padded[:, :-1], gl)
        pr_prim = to_prim(padded[:, 1:], gr)
        rl, ul, pl = np.maximum(pl_prim[0], 1e-12), pl_prim[1], pl_prim[2]
        rr, ur, pr = np.maximum(pr_prim[0], 1e-12), pr_prim[1], pr_prim[2]
        al, ar = sound(pl_prim, gl), sound(pr_prim, gr)
        el_tot = rl * (0.5 * (ul**2) + pl / ((gl - 1.0) * rl))
        er_tot = rr * (0.5 * (ur**2) + pr / ((gr - 1.0) * rr))
    if mode == "Pressure Based":
        # PVRS estimate of p* (torchSolver.py:283-310) on unsplit primitives.
        prim = to_prim(padded, gamma)
        a_l = sound(prim[:, :-1], gl)
        a_r = sound(prim[:, 1:], gr)
        p_s = np.maximum(
            0.0,
            0.5 * (prim[2, :-1] + prim[2, 1:])
            - 0.5 * (prim[1, 1:] - prim[1, :-1]) * (0.5 * (prim[0, 1:] + prim[0, :-1]))
            * (0.5 * (a_l + a_r)),
        )
        ql = np.ones_like(p_s)
        qr = np.ones_like(p_s)
        ml, mr = p_s > pl, p_s > pr
        g_l = gl[ml] if isinstance(gl, np.ndarray) else gl
        g_r = gr[mr] if isinstance(gr, np.ndarray) else gr
        with np.errstate(divide="ignore", invalid="ignore"):
            ql[ml] = np.sqrt(1.0 + (g_l + 1.0) / (2.0 * g_l) * (p_s[ml] / pl[ml] - 1.0))
            qr[mr] = np.sqrt(1.0 + (g_r + 1.0) / (2.0 * g_r) * (p_s[mr] / pr[mr] - 1.0))
        s_l = ul - al * ql
        s_r = ur + ar * qr
    elif mode == "Enthalpy Based":
        with np.errstate(divide="ignore", invalid="ignore"):
            sl, sr = np.sqrt(rl), np.sqrt(rr)
            den = np.maximum(sl + sr, 1e-12)
            u_avg = (sl * ul + sr * ur) / den
            h_avg = (sl * ((el_tot + pl) / rl) + sr * ((er_tot + pr) / rr)) / den
            g_avg = 0.5 * (gl + gr) if isinstance(gl, np.ndarray) else gl
            a_avg = np.sqrt(np.maximum((g_avg - 1.0) * (h_avg - 0.5 * (u_avg**2)), 1e-10))
        s_l = np.minimum(ul - al, u_avg - a_avg)
        s_r = np.maximum(ur + ar, u_avg + a_avg)
    else:
        raise ValueError(f"Unknown wave speed calculation option: {mode}")
    den = rl * (s_l - ul) - rr * (s_r - ur)
    den = np.where(np.abs(den) < 1e-14, 1e-14, den)
    s_s = (pr - pl + rl * ul * (s_l - ul) - rr * ur * (s_r - ur)) / den
    s_s = np.clip(s_s, s_l, s_r)
    return np.vstack((s_l, s_s, s_r))


def limited_slope(delta, dup, ddown, w, limiter):
    """Sweby-bounded limited slope (torchSolver.py:201-275)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        r = _zero_bad(dup / ddown)
        den = 1.0 - w + (1.0 + w) * r
        eps_r = _zero_bad(2.0 / np.where(np.abs(den) < 1e-12, 1e-12, den))
        eps = np.zeros_like(delta)
        pos = r > 0
        if limiter == "VanLeer":
            v = _zero_bad((2.0 * r) / (1.0 + r))
            eps = np.where(pos, np.minimum(v, eps_r), eps)
        elif limiter == "Albada":
            v = _zero_bad((r * (1.0 + r)) / (1.0 + r**2))
            eps = np.where(pos, np.minimum(v, eps_r), eps)
        elif limiter == "MINBEE":
            eps = np.where(pos & (r < 1.0), r, eps)
            eps = np.where(r >= 1.0, np.minimum(1.0, eps_r), eps)
        elif limiter == "Superbee":
            v = _zero_bad(np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0))))
            eps = np.where(pos, np.minimum(v, eps_r), eps)
        elif limiter == "Unos":
            eps = np.minimum(np.maximum(0.0, np.minimum(np.minimum(2.0, 2.0 * r), (2.0 + r) / 3.0)), eps_r)
        else:
            raise ValueError(f"Unknown slope limiter option: {limiter}")
    eps = np.maximum(0.0, _zero_bad(eps))
    slope = eps * delta
    bound = 2.0 * np.minimum(np.abs(dup), np.abs(ddown))
    return np.where(dup * ddown > 0, np.sign(ddown) * np.minimum(np.abs(slope), bound), 0.0)


def reconstruct(padded, dt, dx, gamma, limiter, controller, bc):
    """Half-step evolved sub-cell states with ghosts (torchSolver.py:412-508).

    Returns (3, 2n+2): [ghost, l0, r0, l1, r1, ..., ghost].
    """
    inner, lo, hi = padded[:, 1:-1], padded[:, :-2], padded[:, 2:]
    delta = 0.5 * (hi - lo)
    dup, ddown = inner - lo, hi - inner
    g_in = gamma if np.isscalar(gamma) else gamma[1:-1]
    if controller == "Sound Speed Controlled":
        prim = to_prim(inner, g_in)
        with np.errstate(divide="ignore", invalid="ignore"):
            mach = np.abs(prim[1]) / np.maximum(sound(prim, g_in), 1e-12)
        mach = _zero_bad(mach)
        w = np.clip((1.0 - mach) / (1.0 + mach), -1.0, 1.0)
    elif controller == "w=0":
        w = np.zeros(inner.shape[1])
    else:
        raise ValueError(f"Unknown w_controler option: {controller}")
    d = limited_slope(delta, dup, ddown, w, limiter)
    ul, ur = inner - 0.5 * d, inner + 0.5 * d
    change = -(dt / (2.0 * dx)) * (euler_flux(ur, g_in) - euler_flux(ul, g_in))
    out = []
    for s in (ul + change, ur + change):
        p = (g_in - 1.0) * (s[2] - 0.5 * s[1] ** 2 / np.maximum(s[0], 1e-12))
        bad = (s[0] <= 1e-12) | (p <= 1e-12) | ~np.isfinite(s).all(axis=0)
        out.append(np.where(bad, inner, s))
    sub = np.empty((3, 2 * inner.shape[1]))
    sub[:, 0::2], sub[:, 1::2] = out
    return ghost(sub, bc)


def hllc_1d(sub, waves, gamma, variant="Variante 1 - Original"):
    """Legacy HLLC on predicted interface states (torchSolver.py:534-676)."""
    ul, ur = sub[:, 0::2], sub[:, 1::2]
    gl, gr = _split_gamma(gamma)
    pl_prim, pr_prim = to_prim(ul, gl), to_prim(ur, gr)
    rl, vl, pl = np.maximum(pl_prim[0], 1e-12), pl_prim[1], pl_prim[2]
    rr, vr, pr = np.maximum(pr_prim[0], 1e-12), pr_prim[1], pr_prim[2]
    el = rl * (0.5 * (vl**2) + pl / ((gl - 1.0) * rl))
    er = rr * (0.5 * (vr**2) + pr / ((gr - 1.0) * rr))
    fl, fr = euler_flux(ul, gl), euler_flux(ur, gr)
    s_l, s_r = waves[0], waves[2]
    den = rl * (s_l - vl) - rr * (s_r - vr)
    den = np.where(abs(den) < 1e-14, 1e-14, den)
    s_s = np.clip((pr - pl + rl * vl * (s_l - vl) - rr * vr * (s_r - vr)) / den, s_l, s_r)
    eps = 1e-12
    n = s_l.shape[0]
    if variant == "Variante 1 - Original":
        ok_l, ok_r = np.abs(s_l - s_s) > eps, np.abs(s_r - s_s) > eps
        dl = np.where(ok_l, s_l - s_s, -eps)
        dr = np.where(ok_r, s_r - s_s, eps)
        dlu = np.where(np.abs(s_l - vl) < eps, -eps, s_l - vl)
        dru = np.where(np.abs(s_r - vr) < eps, eps, s_r - vr)
        tl = el / rl + (s_s - vl) * (s_s + (pl / (rl * dlu)))
        tr = er / rr + (s_s - vr) * (s_s + (pr / (rr * dru)))
        usl = rl * ((s_l - vl) / dl) * np.vstack((np.ones(n), s_s, tl))
        usr = rr * ((s_r - vr) / dr) * np.vstack((np.ones(n), s_s, tr))
        usl = np.where(ok_l, usl, ul)
        usr = np.where(ok_r, usr, ur)
        fsl, fsr = fl + s_l * (usl - ul), fr + s_r * (usr - ur)
    elif variant == "Variante_2":
        ds = np.vstack((np.zeros(n), np.ones(n), s_s))
        dl = np.where(np.abs(s_l - s_s) < eps, np.sign(s_l - s_s + 1e-30) * eps, s_l - s_s)
        dr = np.where(np.abs(s_r - s_s) < eps, np.sign(s_r - s_s + 1e-30) * eps, s_r - s_s)
        fsl = (s_s * (s_l * ul - fl) + s_l * (pl + rl * (s_l - vl) * (s_s - vl)) * ds) / dl
        fsr = (s_s * (s_r * ur - fr) + s_r * (pr + rr * (s_r - vr) * (s_s - vr)) * ds) / dr
    else:
        raise ValueError(f"Unknown HLLC variant: {variant}")
    out = np.zeros((3, n))
    out = np.where(s_l > 0, fl, out)
    out = np.where((s_l <= 0) & (s_s > 0), fsl, out)
    out = np.where((s_s <= 0) & (s_r >= 0), fsr, out)
    out = np.where(s_r < 0, fr, out)
    return out


def dt_waves(waves, dx, cfl):
    """torchSolver.py:684-689."""
    top = np.max(np.abs(waves))
    return (cfl * dx) / top if top > 0 else 1e-6


def dt_hydro(padded, gamma, dx, cfl):
    """torchSolver.py:1049-1051."""
    prim = to_prim(padded, gamma)
    return cfl * dx / np.max(np.abs(prim[1]) + sound(prim, gamma))


def update(padded, flux, dt, dx, gamma):
    """Conservative update with global LLF fallback (torchSolver.py:1054-1071).

    Returns (state, flux, used_fallback).
    """
    out = padded[:, 1:-1] + (-(dt / dx) * (flux[:, 1:] - flux[:, :-1]))
    fell_back = False
    if not physical_state(out):
        fell_back = True
        prim = to_prim(padded, gamma)
        speed = np.abs(prim[1]) + sound(prim, gamma)
        alpha = np.maximum(speed[:-1], speed[1:])
        pf = euler_flux(padded, gamma)
        flux = 0.5 * (pf[:, :-1] + pf[:, 1:] - alpha * np.diff(padded))
        out = padded[:, 1:-1] + (-(dt / dx) * (flux[:, 1:] - flux[:, :-1]))
    if not physical_state(out):
        raise RuntimeError("Hydrodynamic update produced a nonphysical state")
    return out, flux, fell_back


def advect(q, old, new, mass_flux, dt_dx, bc):
    """Upwind species transport (torchSolver.py:1125-1132)."""
    padded = ghost_scalar(q, bc)
    flux = mass_flux * np.where(mass_flux >= 0, padded[:-1], padded[1:])
    rho_q = old[0] * q - dt_dx * np.diff(flux)
    return np.clip(rho_q / np.maximum(new[0], 1e-12), 0.0, 1.0)


def padded_gamma(gamma, bc):
    return gamma if np.isscalar(gamma) else ghost_scalar(gamma, bc)


def cfl_dt(state, gamma, dx, cfl, bc, wavespeed="Pressure Based"):
    """Session 1D time step (torchSession.py:338-345)."""
    padded = ghost(state, bc)
    gp = padded_gamma(gamma, bc)
    waves = wave_speeds(padded, gp, wavespeed)
    return min(dt_waves(waves, dx, cfl), dt_hydro(padded, gp, dx, cfl))


def hydro_step(state, gamma, dx, dt, cfl, limiter, bc,
               controller="Sound Speed Controlled",
               variant="Variante 1 - Original", wavespeed="Pressure Based"):
    """One 1D hydro step as the session runs it (torchSession.py:413-437).

    Returns (state, flux, used_fallback). ``cfl`` is passed through to the
    reference's reconstruction, which does not use it.
    """
    padded = ghost(state, bc)
    gp = padded_gamma(gamma, bc)
    waves = wave_speeds(padded, gp, wavespeed)
    sub = reconstruct(padded, dt, dx, gp, limiter, controller, bc)
    flux = hllc_1d(sub, waves, gp, variant)
    return update(padded, flux, dt, dx, gp)
