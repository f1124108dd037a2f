"""Exact solution of the 1D Euler Riemann problem (ideal gas).

Test oracle only. Implements the textbook exact solver (E. F. Toro,
"Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4):
Newton iteration on the pressure function f(p) = f_L(p) + f_R(p) + du for
the star pressure, then sampling of the self-similar solution at x/t. It is
independent of the reference code and is used to check the oracle against
the reference's published Sod L1 errors (README.md:770-777) and the 3D Sod
profiles (python/tests/testSolver3D.py:104-120).
"""

import numpy as np


def _f(p, rho, pk, ak, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(a / (p + b))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + b))
    ratio = p / pk  # rarefaction
    f = 2.0 * ak / (g - 1.0) * (ratio ** ((g - 1.0) / (2.0 * g)) - 1.0)
    return f, 1.0 / (rho * ak) * ratio ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g=1.4, tol=1e-15, iters=200):
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    if ur - ul >= 2.0 * (al + ar) / (g - 1.0):
        raise ValueError("vacuum is generated")
    p = max(tol, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar))
    for _ in range(iters):
        fl, dl = _f(p, rl, pl, al, g)
        fr, dr = _f(p, rr, pr, ar, g)
        step = (fl + fr + ur - ul) / (dl + dr)
        new = max(tol, p - step)
        if abs(new - p) <= tol * 0.5 * (new + p):
            p = new
            break
        p = new
    fl, _ = _f(p, rl, pl, al, g)
    fr, _ = _f(p, rr, pr, ar, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(x, t, x0=0.5, g=1.4, rl=1.0, ul=0.0, pl=1.0, rr=0.125, ur=0.0, pr=0.1):
    """(rho, u, p) at positions x and time t > 0."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    out = np.empty((3, len(x)))
    for i, xi in enumerate(np.asarray(x, dtype=float)):
        s = (xi - x0) / t
        if s <= us:  # left of the contact
            if ps > pl:
                sh = ul - al * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sh:
                    out[:, i] = (rl, ul, pl)
                else:
                    r = rl * (ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)
                    out[:, i] = (r, us, ps)
            else:
                head = ul - al
                tail = us - al * (ps / pl) ** ((g - 1) / (2 * g))
                if s <= head:
                    out[:, i] = (rl, ul, pl)
                elif s >= tail:
                    out[:, i] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)
                    out[:, i] = (rl * c ** (2 / (g - 1)), 2 / (g + 1) * (al + (g - 1) / 2 * ul + s),
                                 pl * c ** (2 * g / (g - 1)))
        else:  # right of the contact
            if ps > pr:
                sh = ur + ar * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sh:
                    out[:, i] = (rr, ur, pr)
                else:
                    r = rr * (ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)
                    out[:, i] = (r, us, ps)
            else:
                head = ur + ar
                tail = us + ar * (ps / pr) ** ((g - 1) / (2 * g))
                if s >= head:
                    out[:, i] = (rr, ur, pr)
                elif s <= tail:
                    out[:, i] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) - (g - 1) / ((g + 1) * ar) * (ur - s)
                    out[:, i] = (rr * c ** (2 / (g - 1)), 2 / (g + 1) * (-ar + (g - 1) / 2 * ur + s),
                                 pr * c ** (2 * g / (g - 1)))
    return out
