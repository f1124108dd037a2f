"""Metric/face finite volumes (oracle restatement).

TEST INFRASTRUCTURE ONLY (see the package docstring): the checker for the
HIP metric kernels, never shipped or measured.

Restates the reference's second finite-volume family, used by sessions with
a non-Cartesian geometry or explicit face boundaries:

- geometry and metrics: python/cfdSolver/torchGeometry.py:12-31 (component
  names), :34-42 (cell edges), :45-90 (validation), :93-213 (cell volumes,
  face areas as outer products of per-axis factors, geometric sources);
- face boundaries: torchBoundaries.py:17-174 (normalisation/validation of
  the six face types), :177-246 (ghost states: wall, inflow, stagnation
  inlet with the Riemann-invariant reservoir solve and choking, pressure
  outlet with backflow);
- the scheme: torchFiniteVolume.py:16-122 (primitive reconstruction with a
  boundary plane per side, limited slopes with the positivity mask and
  periodic slope wrap, wall face mirroring, the Rusanov flux, area-weighted
  divergence, species upwinding, boundary and source integrals), :125-229
  (rhs), :232-276 (SSPRK2 with the first-order retry) and :66-98 (cfl_dt);
- the Rusanov flux of torchFlux.py:80-100.
"""

import numpy as np

from .solver1d import gamma_calc, physical_state

GEOMETRIES = ("cartesian", "cylindrical", "annular", "spherical")
FACE_TYPES = ("outflow", "wall", "periodic", "inflow", "pressure_outlet", "stagnation_inlet")


def component_names(c):
    """torchGeometry.py:12-31."""
    g, dim = c.get("geometry", "cartesian"), c["dimension"]
    if g == "cartesian":
        axes = list("xyz"[:dim])
    elif g == "spherical":
        axes = ["r", "theta", "phi"][:dim]
    elif g == "cylindrical" and dim == 2:
        axes = ["r", "z"]
    else:
        axes = ["r", "theta", "z"][:dim]
    return ["rho"] + ["momentum_" + a for a in axes] + ["energy"]


def edges(c):
    """Cell edges per logical axis (torchGeometry.py:34-42)."""
    dim = c["dimension"]
    origin = c.get("origins_xyz", [0.0] * dim)
    return [np.linspace(o, o + l, n + 1) for o, l, n in zip(origin, c["lengths_xyz"], c["counts_xyz"])]


def _azimuth_axis(g, dim):
    if g == "spherical" and dim == 3:
        return 2
    if (g == "annular" and dim >= 2) or (g == "cylindrical" and dim == 3):
        return 1
    return None


def validate_geometry(c):
    """torchGeometry.py:45-90."""
    dim, g = c["dimension"], c.get("geometry", "cartesian")
    if g not in GEOMETRIES:
        raise ValueError("Unknown geometry")
    origin = np.asarray(c.get("origins_xyz", [0.0] * dim), dtype=float)
    if origin.shape != (dim,) or not np.isfinite(origin).all():
        raise ValueError("One finite origin per logical coordinate is required")
    if g == "cartesian":
        return
    if origin[0] < 0 or (g == "annular" and origin[0] <= 0):
        raise ValueError("Radius must be nonnegative; an annulus needs a positive inner radius")
    if g == "cylindrical" and dim == 3 and origin[0] == 0:
        raise ValueError("3D cylindrical grids require a positive inner radius")
    lengths = c["lengths_xyz"]
    if g == "spherical" and dim >= 2:
        top = origin[1] + lengths[1]
        if origin[1] < 0 or top > np.pi + 1e-14:
            raise ValueError("Spherical theta must lie in [0, pi]")
        if dim == 3 and (origin[0] == 0 or origin[1] <= 0 or top >= np.pi):
            raise ValueError("3D spherical grids require r>0 and 0<theta<pi (shell/wedge)")
    a = _azimuth_axis(g, dim)
    if a is not None:
        if lengths[a] > 2 * np.pi + 1e-14:
            raise ValueError("Azimuthal extent cannot exceed 2*pi")
        if lengths[a] / c["counts_xyz"][a] >= np.pi:
            raise ValueError("Azimuthal cells must span less than pi radians")


def metric_factors(c):
    """Per-axis factors whose outer products are the cell volumes and the
    face areas of each direction (torchGeometry.py:113-182).

    Returns (volume_factors, area_factors): lists indexed by logical axis;
    area_factors[d][a] has n_a + 1 entries when a == d.
    """
    dim, g = c["dimension"], c.get("geometry", "cartesian")
    ax = edges(c)
    n = c["counts_xyz"]
    widths = [np.diff(e) for e in ax]
    if g == "cartesian":
        vol = list(widths)
        area = [[np.ones(n[a] + 1) if a == d else widths[a] for a in range(dim)] for d in range(dim)]
        return vol, area
    r = ax[0]
    dr = np.diff(r)
    half_sq = np.diff(r ** 2) / 2
    if g == "spherical":
        third_cube = np.diff(r ** 3) / 3
        if dim == 1:
            return [4 * np.pi * third_cube], [[4 * np.pi * r ** 2]]
        theta = ax[1]
        cap = -np.diff(np.cos(theta))
        scale = 1 if dim == 3 else 2 * np.pi
        tail = [widths[2]] if dim == 3 else []
        vol = [scale * third_cube, cap] + tail
        area = [[scale * r ** 2, cap] + tail, [scale * half_sq, np.sin(theta)] + tail]
        if dim == 3:
            area.append([half_sq, widths[1], np.ones(n[2] + 1)])
        return vol, area
    if dim == 1:
        return [2 * np.pi * half_sq], [[2 * np.pi * r]]
    if g == "cylindrical" and dim == 2:
        return ([2 * np.pi * half_sq, widths[1]],
                [[2 * np.pi * r, widths[1]], [2 * np.pi * half_sq, np.ones(n[1] + 1)]])
    rest = widths[2:]
    vol = [half_sq, widths[1]] + rest
    area = [[r, widths[1]] + rest, [dr, np.ones(n[1] + 1)] + rest]
    if dim == 3:
        area.append([half_sq, widths[1], np.ones(n[2] + 1)])
    return vol, area


def outer(factors):
    """Outer product over logical axes in (reversed) array order: starts from
    ones and multiplies by axis 0's factor, then axis 1's, ..."""
    dim = len(factors)
    out = np.ones(tuple(len(f) for f in factors)[::-1])
    for a, f in enumerate(factors):
        shape = [1] * dim
        shape[dim - 1 - a] = len(f)
        out *= np.asarray(f, dtype=float).reshape(shape)
    return out


class Metrics:
    """Cell volumes, face areas and area divergences (torchGeometry.py:93-213)."""

    def __init__(self, c):
        validate_geometry(c)
        self.config = c
        dim = c["dimension"]
        vol, area = metric_factors(c)
        self.volumes = outer(vol)
        self.areas = [outer(f) for f in area]
        self.divergence = [np.diff(a, axis=dim - 1 - d) / self.volumes for d, a in enumerate(self.areas)]
        if not np.isfinite(self.volumes).all() or np.any(self.volumes <= 0):
            raise ValueError("Invalid cell volumes")

    def source(self, prim):
        """Geometric momentum sources of curvilinear coordinates (:184-213)."""
        c = self.config
        dim, g = c["dimension"], c.get("geometry", "cartesian")
        out = np.zeros_like(prim)
        if g == "cartesian":
            return out
        rho, p = prim[0], prim[-1]
        radial = self.divergence[0]
        out[1] = p * radial
        if g == "spherical":
            inv_r = radial / 2
            out[1] += rho * np.sum(prim[2:-1] ** 2, axis=0) * inv_r
            if dim >= 2:
                cot_r = self.divergence[1]
                out[2] = p * cot_r - rho * prim[1] * prim[2] * inv_r
                if dim == 3:
                    out[2] += rho * prim[3] ** 2 * cot_r
                    out[3] = -rho * prim[3] * (prim[1] * inv_r + prim[2] * cot_r)
        elif dim >= 2 and not (g == "cylindrical" and dim == 2):
            out[1] += rho * prim[2] ** 2 * radial
            out[2] = -rho * prim[1] * prim[2] * radial
        return out


# -- face boundaries --------------------------------------------------------
_LEGACY = {"Open": "outflow", "Reflective": "wall", "Periodic": "periodic", "Inflow": "inflow"}


def _base_faces(c):
    faces = []
    for name in c["boundaries_xyz"]:
        kind = _LEGACY[name]
        low = dict(type=kind)
        if kind == "inflow":
            low["primitive"] = list(c["inlet_primitive"])
        faces += [low, dict(type="outflow") if kind == "inflow" else dict(low)]
    return faces


def normalize_faces(c, spec=None):
    """Six-type face list x_min, x_max, y_min, ... (torchBoundaries.py:17-174)."""
    dim = c["dimension"]
    if spec is None:
        faces = _base_faces(c)
    elif isinstance(spec, dict):
        names = [a + s for a in "xyz"[:dim] for s in ("_min", "_max")]
        if set(spec) - set(names):
            raise ValueError("Face keys must use logical x_min,x_max,y_min,...")
        faces = _base_faces(c)
        for i, name in enumerate(names):
            if name in spec:
                v = spec[name]
                faces[i] = dict(type=v) if isinstance(v, str) else dict(v)
    else:
        faces = [dict(v) for v in spec]
    if len(faces) != 2 * dim:
        raise ValueError("Require one boundary per face")
    for i, face in enumerate(faces):
        kind = face.get("type")
        if kind not in FACE_TYPES:
            raise ValueError("Unknown face boundary type")
        allowed = {"type", "lambda", "beta"}
        if kind == "inflow":
            allowed |= {"primitive"}
        if kind == "pressure_outlet":
            allowed |= {"pressure", "backflow"}
        if kind == "stagnation_inlet":
            allowed |= {"total_pressure", "total_temperature", "gas_constant"}
        if set(face) - allowed:
            raise ValueError("Unknown boundary parameters")
        for key in ("lambda", "beta"):
            v = face.get(key, 1.0)
            if not np.isscalar(v) or not np.isfinite(v) or not 0 <= v <= 1:
                raise ValueError("Boundary fractions must lie in [0,1]")
        for key in ("primitive", "backflow"):
            if key in face:
                p = np.asarray(face[key], dtype=float)
                if p.shape != (dim + 2,) or not np.isfinite(p).all() or np.any(p[[0, -1]] <= 0):
                    raise ValueError("Invalid boundary primitive state")
                inward = 1 if i % 2 == 0 else -1
                if p[i // 2 + 1] * inward <= 0:
                    raise ValueError("Inlet/backflow velocity must point into the domain")
                face[key] = p.tolist()
        if kind == "inflow" and "primitive" not in face:
            raise ValueError("Inflow requires primitive state")
        needed = {"pressure_outlet": ("pressure",),
                  "stagnation_inlet": ("total_pressure", "total_temperature", "gas_constant")}.get(kind, ())
        for key in needed:
            if key == "gas_constant":
                face.setdefault(key, 287.0)
            v = face.get(key)
            if v is None or not np.isscalar(v) or not np.isfinite(v) or v <= 0:
                raise ValueError("Boundary pressures, temperatures and gas constant must be positive")
    for d in range(dim):
        if (faces[2 * d]["type"] == "periodic") != (faces[2 * d + 1]["type"] == "periodic"):
            raise ValueError("Periodic boundaries must be paired on an axis")
    g = c.get("geometry", "cartesian")
    origin = c.get("origins_xyz", [0.0] * dim)
    if g != "cartesian":
        if any(f["type"] == "periodic" for f in faces[:2]):
            raise ValueError("Radius cannot be periodic")
        if origin[0] == 0 and faces[0]["type"] != "wall":
            raise ValueError("The radial origin requires a reflective wall")
        if g == "spherical" and dim >= 2:
            poles = ((2, origin[1] == 0),
                     (3, np.isclose(origin[1] + c["lengths_xyz"][1], np.pi, atol=1e-14, rtol=0)))
            for i, at_pole in poles:
                if at_pole and faces[i]["type"] != "wall":
                    raise ValueError("A spherical pole requires a reflective wall")
            if any(f["type"] == "periodic" for f in faces[2:4]):
                raise ValueError("Polar theta cannot be periodic")
        a = _azimuth_axis(g, dim)
        if (a is not None and faces[2 * a]["type"] == "periodic"
                and not np.isclose(c["lengths_xyz"][a], 2 * np.pi, atol=1e-14, rtol=0)):
            raise ValueError("Azimuthal periodicity requires a complete 2*pi domain")
    return faces


def default_faces(c, spec):
    """Session defaults for curvilinear grids: walls at the radial origin and
    at spherical poles unless given (torchSession.py:93-121)."""
    dim, g = c["dimension"], c.get("geometry", "cartesian")
    if (c["boundaries_xyz"] == ["Open"] * dim and g != "cartesian" and (spec is None or isinstance(spec, dict))):
        given = dict(spec or {})
        spec = {}
        origin = c["origins_xyz"]
        if origin[0] == 0:
            spec["x_min"] = "wall"
        if g == "spherical" and dim >= 2:
            if origin[1] == 0:
                spec["y_min"] = "wall"
            if np.isclose(origin[1] + c["lengths_xyz"][1], np.pi, rtol=0, atol=1e-14):
                spec["y_max"] = "wall"
        spec.update(given)
    return normalize_faces(c, spec)


def ghost(prim, gamma, fractions, face, direction, side, reactive):
    """Boundary-plane state of one face; side -1 low, +1 high (torchBoundaries.py:177-246)."""
    p = prim.copy()
    g = np.array(gamma, dtype=float, copy=True)
    qs = [np.array(q, dtype=float, copy=True) for q in fractions]
    kind = face["type"]
    nrm = direction + 1
    if kind == "wall":
        p[nrm] *= -1
    elif kind in ("inflow", "stagnation_inlet"):
        g[...] = gamma_calc(face.get("lambda", 1.0)) if reactive else gamma
        for q, key in zip(qs, ("lambda", "beta")):
            q[...] = face.get(key, 1.0)
        if kind == "inflow":
            p[...] = np.asarray(face["primitive"], dtype=float).reshape((-1,) + (1,) * (p.ndim - 1))
        else:
            a_in = np.sqrt(gamma * prim[-1] / prim[0])
            riemann = -side * prim[nrm] - 2 * a_in / (gamma - 1)
            a0sq = g * face["gas_constant"] * face["total_temperature"]
            k = (g + 1) / (g - 1)
            disc = riemann ** 2 - k * (0.5 * (g - 1) * riemann ** 2 - a0sq)
            if np.any(disc < 0):
                raise ValueError("No subsonic reservoir-inlet solution")
            a = (-riemann + np.sqrt(disc)) / k
            un = riemann + 2 * a / (g - 1)
            choked = un >= a
            a = np.where(choked, np.sqrt(2 * a0sq / (g + 1)), a)
            un = np.where(choked, a, un)
            if np.any(un < 0) or np.any(a <= 0):
                raise ValueError("Reservoir cannot supply inflow against the interior pressure")
            p[1:-1] = 0
            p[nrm] = -side * un
            p[-1] = face["total_pressure"] * (a * a / a0sq) ** (g / (g - 1))
            p[0] = g * p[-1] / (a * a)
    elif kind == "pressure_outlet":
        a = np.sqrt(g * p[-1] / p[0])
        un = side * p[nrm]
        sub = un < a
        rho = p[0] * (face["pressure"] / p[-1]) ** (1 / g)
        a_new = np.sqrt(g * face["pressure"] / rho)
        vel = un + 2 * (a - a_new) / (g - 1)
        back = sub & ((vel < 0) | (un < 0))
        if np.any(back) and "backflow" not in face:
            raise ValueError("Pressure outlet backflow requires an explicit backflow primitive state")
        p[0] = np.where(sub, rho, p[0])
        p[-1] = np.where(sub, face["pressure"], p[-1])
        p[nrm] = np.where(sub, side * vel, p[nrm])
        if np.any(back):
            fixed = np.asarray(face["backflow"], dtype=float).reshape((-1,) + (1,) * (p.ndim - 1))
            p = np.where(back[None], fixed, p)
            if reactive:
                g = np.where(back, gamma_calc(face.get("lambda", 1.0)), g)
            for q, key in zip(qs, ("lambda", "beta")):
                q[...] = np.where(back, face.get(key, 1.0), q)
    return p, g, qs


def boundary_plane(prim, gamma, fractions, c, faces, d, side):
    """The plane beyond one end of axis d (torchFiniteVolume.py:49-63)."""
    axis = c["dimension"] - 1 - d
    face = faces[2 * d + (side == 1)]
    if face["type"] == "periodic":
        idx = -1 if side == -1 else 0
    else:
        idx = 0 if side == -1 else -1
    p = np.take(prim, idx, axis=axis + 1)
    g = np.take(gamma, idx, axis=axis)
    qs = [np.take(q, idx, axis=axis) for q in fractions]
    if face["type"] != "periodic":
        p, g, qs = ghost(p, g, qs, face, d, side, c["reactive"])
    return p, g, qs


def padded(prim, gamma, fractions, c, faces, d):
    """Primitives, gamma and fractions with one boundary plane each side along d."""
    axis = c["dimension"] - 1 - d
    lo = boundary_plane(prim, gamma, fractions, c, faces, d, -1)
    hi = boundary_plane(prim, gamma, fractions, c, faces, d, 1)
    pp = np.concatenate((np.expand_dims(lo[0], axis + 1), prim, np.expand_dims(hi[0], axis + 1)), axis=axis + 1)
    gg = np.concatenate((np.expand_dims(lo[1], axis), gamma, np.expand_dims(hi[1], axis)), axis=axis)
    qq = [np.concatenate((np.expand_dims(a, axis), q, np.expand_dims(b, axis)), axis=axis)
          for q, a, b in zip(fractions, lo[2], hi[2])]
    return pp, gg, qq


def to_prim(state, gamma):
    p = state.copy()
    p[1:-1] /= state[0]
    p[-1] = (gamma - 1) * (state[-1] - 0.5 * np.sum(state[1:-1] ** 2, axis=0) / state[0])
    return p


def to_cons(p, gamma):
    u = p.copy()
    u[1:-1] *= p[0]
    u[-1] = p[-1] / (gamma - 1) + 0.5 * p[0] * np.sum(p[1:-1] ** 2, axis=0)
    return u


def _pair(a, axis):
    lo = [slice(None)] * a.ndim
    hi = list(lo)
    lo[axis], hi[axis] = slice(None, -1), slice(1, None)
    return a[tuple(lo)], a[tuple(hi)]


def cell_gamma(state, lam, c):
    g = gamma_calc(lam) if c["reactive"] else c["gamma"]
    return np.broadcast_to(g, state.shape[1:])


def cfl_dt(state, lam, c, metrics, faces):
    """torchFiniteVolume.py:66-98: cfl / max over cells of the summed
    area-weighted face speeds over the volume."""
    gamma = cell_gamma(state, lam, c)
    p = to_prim(state, gamma)
    a = np.sqrt(gamma * p[-1] / p[0])
    rate = np.zeros(state.shape[1:])
    for d, area in enumerate(metrics.areas):
        axis = c["dimension"] - 1 - d
        speed = abs(p[d + 1]) + a
        fs = np.empty_like(area)
        inner = [slice(None)] * c["dimension"]
        inner[axis] = slice(1, -1)
        fs[tuple(inner)] = np.maximum(*_pair(speed, axis))
        for idx, side in ((0, -1), (-1, 1)):
            bp, bg, _ = boundary_plane(p, gamma, [], c, faces, d, side)
            end = [slice(None)] * c["dimension"]
            end[axis] = idx
            fs[tuple(end)] = np.maximum(np.take(speed, idx, axis=axis), abs(bp[d + 1]) + np.sqrt(bg * bp[-1] / bp[0]))
        rate += np.maximum(*_pair(area * fs, axis)) / metrics.volumes
    return c["cfl"] / np.max(rate)


def limited(minus, plus, name):
    """phi(r) * plus with r = minus / (|plus| < 1e-14 ? 1e-14 : plus) (torchFiniteVolume.py:101-122)."""
    r = minus / np.where(abs(plus) < 1e-14, 1e-14, plus)
    pos = np.maximum(r, 0.0)
    if name == "VanLeer":
        phi = 2 * pos / (1 + pos)
    elif name == "Albada":
        phi = (pos * pos + pos) / (pos * pos + 1)
    elif name == "MINBEE":
        phi = np.maximum(0, np.minimum(r, 1))
    elif name == "Superbee":
        phi = np.maximum(0, np.maximum(np.minimum(2 * r, 1), np.minimum(r, 2)))
    else:
        phi = np.maximum(0, np.minimum(np.minimum(2, 2 * r), np.minimum((2 + r) / 3, (1 + 2 * r) / 3)))
    return phi * plus


def rusanov(left, right, gl, gr, d):
    """Rusanov flux of primitive face states along d (torchFlux.py:80-100)."""
    ul, ur = to_cons(left, gl), to_cons(right, gr)
    fl, fr = ul * left[d + 1], ur * right[d + 1]
    fl[d + 1] += left[-1]
    fr[d + 1] += right[-1]
    fl[-1] += left[d + 1] * left[-1]
    fr[-1] += right[d + 1] * right[-1]
    s = np.maximum(abs(left[d + 1]) + np.sqrt(gl * left[-1] / left[0]),
                   abs(right[d + 1]) + np.sqrt(gr * right[-1] / right[0]))
    return 0.5 * (fl + fr - s * (ur - ul))


def rhs(state, fractions, c, metrics, faces, first_order=False):
    """Semi-discrete change, species change, outward boundary flux and
    integrated geometric source (torchFiniteVolume.py:125-229)."""
    dim = c["dimension"]
    gamma = cell_gamma(state, fractions[0] if fractions else None, c)
    p = to_prim(state, gamma)
    source = metrics.source(p)
    change = source.copy()
    species = [np.zeros_like(q) for q in fractions]
    outward = np.zeros(dim + 2)
    for d, area in enumerate(metrics.areas):
        axis = dim - 1 - d
        pp, gg, qq = padded(p, gamma, fractions, c, faces, d)
        slope = np.zeros_like(pp)
        if not first_order:
            mid, before, after = ([slice(None)] * (dim + 1) for _ in range(3))
            mid[axis + 1], before[axis + 1], after[axis + 1] = slice(1, -1), slice(None, -2), slice(2, None)
            slope[tuple(mid)] = limited(pp[tuple(mid)] - pp[tuple(before)], pp[tuple(after)] - pp[tuple(mid)],
                                        c["limiter"])
            bad = np.any((pp[[0, -1]] - 0.5 * abs(slope[[0, -1]])) <= 1e-12, axis=0)
            slope *= (~bad)[None]
            if faces[2 * d]["type"] == "periodic":
                for target, src in ((0, -2), (-1, 1)):
                    sl = [slice(None)] * (dim + 1)
                    sl[axis + 1] = target
                    slope[tuple(sl)] = np.take(slope, src, axis=axis + 1)
        pl, pr = _pair(pp, axis + 1)
        sl, sr = _pair(slope, axis + 1)
        left, right = pl + 0.5 * sl, pr - 0.5 * sr
        gl, gr = (np.array(x) for x in _pair(gg, axis))
        for end, hi in ((0, 0), (-1, 1)):
            if faces[2 * d + hi]["type"] != "wall":
                continue
            idx = [slice(None)] * (dim + 1)
            idx[axis + 1] = end
            sidx = tuple(idx[1:])
            if hi == 0:
                left[tuple(idx)] = right[tuple(idx)]
                gl[sidx] = gr[sidx]
                idx[0] = d + 1
                left[tuple(idx)] *= -1
            else:
                right[tuple(idx)] = left[tuple(idx)]
                gr[sidx] = gl[sidx]
                idx[0] = d + 1
                right[tuple(idx)] *= -1
        flux = rusanov(left, right, gl, gr, d)
        weighted = flux * area[None]
        change -= np.diff(weighted, axis=axis + 1) / metrics.volumes[None]
        outward += (np.take(weighted, -1, axis=axis + 1) - np.take(weighted, 0, axis=axis + 1)).reshape(
            dim + 2, -1).sum(axis=1)
        for j, q in enumerate(qq):
            ql, qr = _pair(q, axis)
            moved = flux[0] * np.where(flux[0] >= 0, ql, qr) * area
            species[j] -= np.diff(moved, axis=axis) / metrics.volumes
    integrated = (source * metrics.volumes[None]).reshape(dim + 2, -1).sum(axis=1)
    return change, species, outward, integrated


def advance(state, lam, beta, dt, c, metrics, faces):
    """SSPRK2 with a first-order retry (torchFiniteVolume.py:232-276).

    Returns (state, lam, beta, boundary, source, first_order_used)."""
    fractions = [q for q in (lam, beta) if q is not None]
    for first_order in (False, True):
        held = [state[0] * q for q in fractions]
        k1, q1, b1, s1 = rhs(state, fractions, c, metrics, faces, first_order)
        stage = state + dt * k1
        if not physical_state(stage):
            continue
        stage_q = [(h + dt * k) / stage[0] for h, k in zip(held, q1)]
        if any(np.any((q < -1e-12) | (q > 1 + 1e-12)) for q in stage_q):
            continue
        stage_q = [np.clip(q, 0, 1) for q in stage_q]
        k2, q2, b2, s2 = rhs(stage, stage_q, c, metrics, faces, first_order)
        result = 0.5 * (state + stage + dt * k2)
        if not physical_state(result):
            continue
        final = [(0.5 * (h + stage[0] * f + dt * k)) / result[0] for h, f, k in zip(held, stage_q, q2)]
        if any(np.any((q < -1e-12) | (q > 1 + 1e-12)) for q in final):
            continue
        final = [np.clip(q, 0, 1) for q in final]
        return (result, final[0] if final else None, final[1] if len(final) > 1 else None,
                0.5 * dt * (b1 + b2), 0.5 * dt * (s1 + s2), first_order)
    raise RuntimeError("Metric/face update lost positivity; reduce CFL")
