"""HLLC interface flux (oracle restatement).

Restates python/cfdSolver/torchFlux.py:125-248 (``hllc_flux``) and the
component rotation of ``hllc_flux_grid`` (torchFlux.py:36-77). Rows of the
normal-basis arrays are [rho, m_normal, m_tangential..., E].
"""

import numpy as np

FLOOR = 1e-12
GUARD = 1e-14


def _guard(x):
    # torchFlux.py:190,196-197,205-206: |x| < 1e-14 is replaced by +1e-14.
    return np.where(np.abs(x) < GUARD, GUARD, x)


def hllc(ul, ur, gamma_l, gamma_r=None):
    """Normal HLLC flux for (nv, m) states, nv in {4, 5}.

    torchFlux.py:125-248. Region masks are applied in the order L, *L, *R,
    R with the last assignment winning; a face that matches no mask (NaN
    wave speeds) is NaN here (the reference leaves np.empty contents).
    """
    gl = gamma_l
    gr = gamma_l if gamma_r is None else gamma_r
    gbar = 0.5 * (gl + gr)
    rl = np.maximum(ul[0], FLOOR)
    rr = np.maximum(ur[0], FLOOR)
    unl, unr = ul[1] / rl, ur[1] / rr
    utl, utr = ul[2:-1] / rl, ur[2:-1] / rr
    kel = 0.5 * rl * (unl**2 + np.sum(utl**2, axis=0))
    ker = 0.5 * rr * (unr**2 + np.sum(utr**2, axis=0))
    pl = np.maximum((gl - 1.0) * (ul[-1] - kel), FLOOR)
    pr = np.maximum((gr - 1.0) * (ur[-1] - ker), FLOOR)
    al = np.sqrt(gl * pl / rl)
    ar = np.sqrt(gr * pr / rr)

    def physical(u, rho, un, ut, p):
        f = np.empty_like(u)
        f[0] = rho * un
        f[1] = rho * un**2 + p
        f[2:-1] = rho * un * ut
        f[-1] = un * (u[-1] + p)
        return f

    fl = physical(ul, rl, unl, utl, pl)
    fr = physical(ur, rr, unr, utr, pr)

    sql, sqr = np.sqrt(rl), np.sqrt(rr)
    inv = 1.0 / (sql + sqr)
    u_roe = (sql * unl + sqr * unr) * inv
    v_roe = (sql * utl + sqr * utr) * inv
    hl = (ul[-1] + pl) / rl
    hr = (ur[-1] + pr) / rr
    h_roe = (sql * hl + sqr * hr) * inv
    ke_roe = 0.5 * (u_roe**2 + np.sum(v_roe**2, axis=0))
    a_roe = np.sqrt(np.maximum((gbar - 1.0) * (h_roe - ke_roe), FLOOR))

    s_l = np.minimum(unl - al, u_roe - a_roe)
    s_r = np.maximum(unr + ar, u_roe + a_roe)
    s_star = (
        pr - pl + rl * unl * (s_l - unl) - rr * unr * (s_r - unr)
    ) / _guard(rl * (s_l - unl) - rr * (s_r - unr))

    def star(u, rho, un, ut, p, s):
        fac = (s - un) / _guard(s - s_star)
        rs = np.maximum(rho * fac, FLOOR)
        out = np.empty_like(u)
        out[0] = rs
        out[1] = rs * s_star
        out[2:-1] = rs * ut
        out[-1] = rs * (
            u[-1] / rho + (s_star - un) * (s_star + p / (rho * _guard(s - un)))
        )
        return out

    fsl = fl + s_l * (star(ul, rl, unl, utl, pl, s_l) - ul)
    fsr = fr + s_r * (star(ur, rr, unr, utr, pr, s_r) - ur)
    close_l = np.abs(s_l - s_star) < FLOOR
    close_r = np.abs(s_r - s_star) < FLOOR
    fsl = np.where(close_l, fl, fsl)
    fsr = np.where(close_r, fr, fsr)

    out = np.full_like(ul, np.nan)
    for mask, value in (
        (0.0 <= s_l, fl),
        ((s_l <= 0.0) & (0.0 <= s_star), fsl),
        ((s_star <= 0.0) & (0.0 <= s_r), fsr),
        (0.0 >= s_r, fr),
    ):
        out = np.where(mask, value, out)
    return out


def normal_order(nv, direction):
    """Row order [rho, normal, tangentials (ascending), E] (torchFlux.py:37-39)."""
    normal = direction + 1
    return [0, normal] + [m for m in range(1, nv - 1) if m != normal] + [nv - 1]


def hllc_grid(left, right, gamma_l, gamma_r=None, direction=0, batch_size=None, execution=None):
    """HLLC over Cartesian-ordered face grids (torchFlux.py:6-77).

    Every face is independent, so the result does not depend on batching
    (the reference's own test, testPerformance.py:23-52). By default the
    whole grid is one batch. With ``batch_size`` the faces go through the
    reference's worker path (torchFlux.py:44-76): flattened in C order, cut
    into batches of that many faces, each gathered, rotated to the normal
    basis and solved by one task of ``execution.map`` (oracle.execution),
    and scattered back in order.
    """
    nv = left.shape[0]
    if left.shape != right.shape or nv not in (4, 5):
        raise ValueError("HLLC requires matching four/five-component grids")
    if direction not in range(nv - 2):
        raise ValueError("Invalid normal direction")
    order = normal_order(nv, direction)
    gl = np.broadcast_to(gamma_l, left.shape[1:]).reshape(-1)
    gr = gl if gamma_r is None else np.broadcast_to(
        gamma_r, left.shape[1:]).reshape(-1)
    if batch_size is None:
        flux = hllc(left[order].reshape(nv, -1), right[order].reshape(nv, -1), gl,
                    gr)
        out = np.empty_like(flux)
        out[order] = flux
        return out.reshape(left.shape)
    rows_l = [left[m].reshape(-1) for m in order]
    rows_r = [right[m].reshape(-1) for m in order]
    out = np.empty((nv, gl.size))

    def solve(bounds):
        a, b = bounds
        ul = np.array([row[a:b] for row in rows_l])
        ur = np.array([row[a:b] for row in rows_r])
        return a, b, hllc(ul, ur, gl[a:b], gr[a:b])

    spans = ((a, min(a + batch_size, gl.size)) for a in range(0, gl.size, batch_size))
    for a, b, value in (map(solve, spans) if execution is None else execution.map(solve, spans)):
        out[order, a:b] = value
    return out.reshape(left.shape)
