"""Adaptive (dyadic AMR) finite volumes, per-step numerics (oracle restatement).

TEST INFRASTRUCTURE ONLY (see the package docstring).

Restates python/cfdSolver/torchAdaptive.py:14-206 on a leaf mesh given as
arrays (any object with the attributes of the reference's LeafMesh: left,
right, directions, sides, face_areas, offset_left, offset_right,
displacement, weights, inverse, volumes, divergence): boundary-plane face
values (:14-77), bounded least-squares gradients and second-order
reconstruction with the wall mirror after it (torchAdaptiveGeometry.py:
334-372, torchAdaptive.py:59-77), geometric sources (:80-103), the
shared-face Rusanov rhs with its scatter in face order (:106-146), cfl
(:149-166) and SSPRK2 with the first-order retry (:169-206). The tree itself
and remeshing are host data structures, pinned separately.
"""

import numpy as np

from . import metric
from .solver1d import gamma_calc, physical_state


def bounded_gradients(mesh, values, left, right):
    d = mesh.displacement.shape[1]
    diff = right - left
    vector = np.zeros((values.shape[1], d, values.shape[0]))
    part = mesh.weights[:, None, None] * mesh.displacement[:, :, None] * diff.T[:, None, :]
    low, high = values.T.copy(), values.T.copy()
    for idx, other in ((mesh.left, right), (mesh.right, left)):
        mask = idx >= 0
        np.add.at(vector, idx[mask], part[mask])
        np.minimum.at(low, idx[mask], other.T[mask])
        np.maximum.at(high, idx[mask], other.T[mask])
    grad = np.einsum("nij,njk->nik", mesh.inverse, vector)
    theta = np.ones_like(values.T)
    for idx, offset in ((mesh.left, mesh.offset_left), (mesh.right, mesh.offset_right)):
        mask = idx >= 0
        ids = idx[mask]
        change = np.einsum("nd,ndv->nv", offset[mask], grad[ids])
        allowed = np.where(change >= 0, high[ids] - values.T[ids], low[ids] - values.T[ids])
        ratio = np.ones_like(change)
        np.divide(allowed, change, out=ratio, where=abs(change) > 1e-14)
        np.minimum.at(theta, ids, np.clip(ratio, 0, 1))
    return grad * theta[:, None, :]


def face_values(mesh, state, fractions, c, reconstruct=True):
    g = np.broadcast_to(gamma_calc(fractions[0]) if c["reactive"] else c["gamma"], state.shape[1:])
    p = metric.to_prim(state, g)
    li, ri = np.maximum(mesh.left, 0), np.maximum(mesh.right, 0)
    left, right = p[:, li], p[:, ri]
    gl, gr = g[li].copy(), g[ri].copy()
    ql = [q[li].copy() for q in fractions]
    qr = [q[ri].copy() for q in fractions]
    d = c["dimension"]
    for k in range(d):
        for side in (-1, 1):
            mask = (mesh.directions == k) & (mesh.sides == side)
            if not mask.any():
                continue
            face = c["face_boundaries"][2 * k + (side == 1)]
            if side == -1:
                pp, gg, qs = metric.ghost(right[:, mask], gr[mask], [q[mask] for q in qr], face, k, side,
                                          c["reactive"])
                left[:, mask], gl[mask] = pp, gg
                for a, b in zip(ql, qs):
                    a[mask] = b
            else:
                pp, gg, qs = metric.ghost(left[:, mask], gl[mask], [q[mask] for q in ql], face, k, side,
                                          c["reactive"])
                right[:, mask], gr[mask] = pp, gg
                for a, b in zip(qr, qs):
                    a[mask] = b
    if reconstruct and c["amr"]["order"] == 2:
        grad = bounded_gradients(mesh, p, left, right)
        for values, idx, offset in ((left, mesh.left, mesh.offset_left), (right, mesh.right, mesh.offset_right)):
            mask = idx >= 0
            values[:, mask] += np.einsum("nd,ndv->vn", offset[mask], grad[idx[mask]])
        for k in range(d):
            for side in (-1, 1):
                if c["face_boundaries"][2 * k + (side == 1)]["type"] != "wall":
                    continue
                mask = (mesh.directions == k) & (mesh.sides == side)
                if side == -1:
                    left[:, mask] = right[:, mask]
                    left[k + 1, mask] *= -1
                    gl[mask] = gr[mask]
                else:
                    right[:, mask] = left[:, mask]
                    right[k + 1, mask] *= -1
                    gr[mask] = gl[mask]
    return left, right, gl, gr, ql, qr


def source(mesh, c, p):
    d, geometry = c["dimension"], c.get("geometry", "cartesian")
    out = np.zeros_like(p)
    if geometry == "cartesian":
        return out
    rho, pr = p[0], p[-1]
    radial = mesh.divergence[0]
    out[1] = pr * radial
    if geometry == "spherical":
        inv_r = radial / 2
        out[1] += rho * np.sum(p[2:-1] ** 2, axis=0) * inv_r
        if d >= 2:
            cot_r = mesh.divergence[1]
            out[2] = pr * cot_r - rho * p[1] * p[2] * inv_r
            if d == 3:
                out[2] += rho * p[3] ** 2 * cot_r
                out[3] = -rho * p[3] * (p[1] * inv_r + p[2] * cot_r)
    elif d >= 2 and not (geometry == "cylindrical" and d == 2):
        out[1] += rho * p[2] ** 2 * radial
        out[2] = -rho * p[1] * p[2] * radial
    return out


def rhs(mesh, state, fractions, c, first_order=False):
    left, right, gl, gr, ql, qr = face_values(mesh, state, fractions, c, not first_order)
    flux = np.empty_like(left)
    for k in range(c["dimension"]):
        ids = np.flatnonzero(mesh.directions == k)
        flux[:, ids] = metric.rusanov(left[:, ids], right[:, ids], gl[ids], gr[ids], k)
    g = gamma_calc(fractions[0]) if c["reactive"] else c["gamma"]
    src = source(mesh, c, metric.to_prim(state, g))
    change = src * mesh.volumes
    species = [np.zeros_like(q) for q in fractions]
    weighted = flux * mesh.face_areas
    boundary = np.zeros(state.shape[0])
    for idx, sign, side in ((mesh.left, -1, 1), (mesh.right, 1, -1)):
        mask = idx >= 0
        np.add.at(change.T, idx[mask], sign * weighted.T[mask])
        boundary += side * np.sum(weighted[:, mesh.sides == side], axis=1)
        for q, a, b in zip(species, ql, qr):
            moved = weighted[0] * np.where(flux[0] >= 0, a, b)
            np.add.at(q, idx[mask], sign * moved[mask])
    return change / mesh.volumes, [q / mesh.volumes for q in species], boundary, np.sum(src * mesh.volumes, axis=1)


def cfl(mesh, state, fractions, c):
    left, right, gl, gr, _, _ = face_values(mesh, state, fractions, c)
    rate = np.zeros(state.shape[1])
    for k in range(c["dimension"]):
        mask = mesh.directions == k
        speed = np.maximum(abs(left[k + 1, mask]) + np.sqrt(gl[mask] * left[-1, mask] / left[0, mask]),
                           abs(right[k + 1, mask]) + np.sqrt(gr[mask] * right[-1, mask] / right[0, mask]))
        weighted = speed * mesh.face_areas[mask]
        for idx in (mesh.left[mask], mesh.right[mask]):
            active = idx >= 0
            np.add.at(rate, idx[active], weighted[active])
    return c["cfl"] / np.max(0.5 * rate / mesh.volumes)


def advance(mesh, state, fractions, dt, c):
    """Returns (state, fractions, boundary, source, first_order_used)."""
    held = [state[0] * q for q in fractions]
    for first in (False, True):
        k1, q1, b1, s1 = rhs(mesh, state, fractions, c, first)
        stage = state + dt * k1
        if not physical_state(stage):
            continue
        sq = [(h + dt * k) / stage[0] for h, k in zip(held, q1)]
        if any(not np.isfinite(q).all() or np.any((q < -1e-12) | (q > 1 + 1e-12)) for q in sq):
            continue
        sq = [np.clip(q, 0, 1) for q in sq]
        k2, q2, b2, s2 = rhs(mesh, stage, sq, c, first)
        final = 0.5 * (state + stage + dt * k2)
        if not physical_state(final):
            continue
        fq = [0.5 * (h + stage[0] * f + dt * k) / final[0] for h, f, k in zip(held, sq, q2)]
        if any(not np.isfinite(q).all() or np.any((q < -1e-12) | (q > 1 + 1e-12)) for q in fq):
            continue
        return final, [np.clip(q, 0, 1) for q in fq], 0.5 * dt * (b1 + b2), 0.5 * dt * (s1 + s2), first
    raise RuntimeError("Adaptive update lost positivity")
