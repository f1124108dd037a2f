"""Grid metrics and face boundaries of the metric/face scheme (host side).

Mirrors python/cfdSolver/torchGeometry.py (component names :12-31, cell
edges :34-42, validation :45-90, ``Metrics`` :93-213) and the
normalisation of face boundaries in torchBoundaries.py:17-174 plus the
session defaults of torchSession.py:93-121. Volumes and face areas are
outer products of per-axis factors; the device evaluates those products
itself from the factor table built here (``factor_table``), in the same
axis order, so host and device metrics agree bit for bit.
"""

import numpy as np

GEOMETRIES = ("cartesian", "cylindrical", "annular", "spherical")
FACE_TYPES = ("outflow", "wall", "periodic", "inflow", "pressure_outlet", "stagnation_inlet")
_FROM_LEGACY = {"Open": "outflow", "Reflective": "wall", "Periodic": "periodic", "Inflow": "inflow"}


def component_names(c):
    g, dim = c.get("geometry", "cartesian"), c["dimension"]
    axes = {"cartesian": list("xyz"), "spherical": ["r", "theta", "phi"]}.get(
        g, ["r", "z"] if g == "cylindrical" and dim == 2 else ["r", "theta", "z"])
    return ["rho"] + ["momentum_" + a for a in axes[:dim]] + ["energy"]


def edges(c):
    dim = c["dimension"]
    origins = c.get("origins_xyz", [0.0] * dim)
    return [np.linspace(o, o + length, n + 1) for o, length, n in zip(origins, c["lengths_xyz"], c["counts_xyz"])]


def azimuth_axis(geometry, dim):
    if geometry == "spherical" and dim == 3:
        return 2
    return 1 if (geometry == "annular" and dim >= 2) or (geometry == "cylindrical" and dim == 3) else None


def validate_geometry(c):
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
        if origin[1] < 0 or origin[1] + lengths[1] > np.pi + 1e-14:
            raise ValueError("Spherical theta must lie in [0, pi]")
        if dim == 3 and (origin[0] == 0 or origin[1] <= 0 or origin[1] + lengths[1] >= np.pi):
            raise ValueError("3D spherical grids require r>0 and 0<theta<pi (shell/wedge)")
    a = azimuth_axis(g, dim)
    if a is not None and lengths[a] > 2 * np.pi + 1e-14:
        raise ValueError("Azimuthal extent cannot exceed 2*pi")
    if a is not None and lengths[a] / c["counts_xyz"][a] >= np.pi:
        raise ValueError("Azimuthal cells must span less than pi radians")


def factors(c):
    """(volume factor per axis, area factors per direction and axis)."""
    dim, g = c["dimension"], c.get("geometry", "cartesian")
    e = edges(c)
    n = c["counts_xyz"]
    width = [np.diff(x) for x in e]
    faces = [np.ones(k + 1) for k in n]
    if g == "cartesian":
        return width, [[faces[a] if a == d else width[a] for a in range(dim)] for d in range(dim)]
    r = e[0]
    sq = np.diff(r ** 2) / 2
    if g == "spherical":
        cube = np.diff(r ** 3) / 3
        if dim == 1:
            return [4 * np.pi * cube], [[4 * np.pi * r ** 2]]
        k = 1 if dim == 3 else 2 * np.pi
        phi = [width[2]] if dim == 3 else []
        solid = -np.diff(np.cos(e[1]))
        area = [[k * r ** 2, solid] + phi, [k * sq, np.sin(e[1])] + phi]
        if dim == 3:
            area.append([sq, width[1], faces[2]])
        return [k * cube, solid] + phi, area
    if dim == 1:
        return [2 * np.pi * sq], [[2 * np.pi * r]]
    if g == "cylindrical" and dim == 2:
        return [2 * np.pi * sq, width[1]], [[2 * np.pi * r, width[1]], [2 * np.pi * sq, faces[1]]]
    rest = width[2:]
    area = [[r, width[1]] + rest, [np.diff(r), faces[1]] + rest]
    if dim == 3:
        area.append([sq, width[1], faces[2]])
    return [sq, width[1]] + rest, area


def factor_table(c):
    """Flat factor table of torch_state_set_metric: volume factors of axes
    0..dim-1, then area factors of each direction over each axis."""
    vol, area = factors(c)
    parts = list(vol) + [f for per_dir in area for f in per_dir]
    return np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=float) for p in parts]))


def _outer(parts):
    dim = len(parts)
    out = np.ones(tuple(len(p) for p in parts)[::-1])
    for a, p in enumerate(parts):
        shape = [1] * dim
        shape[dim - 1 - a] = len(p)
        out *= np.asarray(p, dtype=float).reshape(shape)
    return out


class Metrics:
    """Host copy of the cell volumes and face areas (torchGeometry.py:93-182)."""

    def __init__(self, c):
        validate_geometry(c)
        vol, area = factors(c)
        self.volumes = _outer(vol)
        self.areas = [_outer(a) for a in area]
        dim = c["dimension"]
        self.divergence = [np.diff(a, axis=dim - 1 - d) / self.volumes for d, a in enumerate(self.areas)]
        if not np.isfinite(self.volumes).all() or np.any(self.volumes <= 0):
            raise ValueError("Invalid cell volumes")


def _legacy_faces(c):
    out = []
    for name in c["boundaries_xyz"]:
        kind = _FROM_LEGACY[name]
        low = dict(type=kind)
        if kind == "inflow":
            low["primitive"] = list(c["inlet_primitive"])
        out += [low, dict(type="outflow") if kind == "inflow" else dict(low)]
    return out


def normalize_faces(c, spec=None):
    """torchBoundaries.py:17-174."""
    dim = c["dimension"]
    if spec is None:
        faces = _legacy_faces(c)
    elif isinstance(spec, dict):
        keys = [a + side for a in "xyz"[:dim] for side in ("_min", "_max")]
        if set(spec) - set(keys):
            raise ValueError("Face keys must use logical x_min,x_max,y_min,...")
        faces = _legacy_faces(c)
        for i, key in enumerate(keys):
            if key in spec:
                faces[i] = dict(type=spec[key]) if isinstance(spec[key], str) else dict(spec[key])
    else:
        faces = [dict(v) for v in spec]
    if len(faces) != 2 * dim:
        raise ValueError("Require one boundary per face")
    extra = {"inflow": {"primitive"}, "pressure_outlet": {"pressure", "backflow"},
             "stagnation_inlet": {"total_pressure", "total_temperature", "gas_constant"}}
    required = {"pressure_outlet": ("pressure",),
                "stagnation_inlet": ("total_pressure", "total_temperature", "gas_constant")}
    for i, face in enumerate(faces):
        kind = face.get("type")
        if kind not in FACE_TYPES:
            raise ValueError("Unknown face boundary type")
        if set(face) - ({"type", "lambda", "beta"} | extra.get(kind, set())):
            raise ValueError("Unknown boundary parameters")
        for key in ("lambda", "beta"):
            v = face.get(key, 1.0)
            if not np.isscalar(v) or not np.isfinite(v) or not 0 <= v <= 1:
                raise ValueError("Boundary fractions must lie in [0,1]")
        for key in ("primitive", "backflow"):
            if key not in face:
                continue
            p = np.asarray(face[key], dtype=float)
            if p.shape != (dim + 2,) or not np.isfinite(p).all() or np.any(p[[0, -1]] <= 0):
                raise ValueError("Invalid boundary primitive state")
            if p[i // 2 + 1] * (-1 if i % 2 == 0 else 1) >= 0:
                raise ValueError("Inlet/backflow velocity must point into the domain")
            face[key] = p.tolist()
        if kind == "inflow" and "primitive" not in face:
            raise ValueError("Inflow requires primitive state")
        for key in required.get(kind, ()):
            if key == "gas_constant":
                face.setdefault(key, 287.0)
            v = face.get(key)
            if v is None or not np.isscalar(v) or not np.isfinite(v) or v <= 0:
                raise ValueError("Boundary pressures, temperatures and gas constant must be positive")
    for d in range(dim):
        if (faces[2 * d]["type"] == "periodic") != (faces[2 * d + 1]["type"] == "periodic"):
            raise ValueError("Periodic boundaries must be paired on an axis")
    g = c.get("geometry", "cartesian")
    if g != "cartesian":
        origin = c.get("origins_xyz", [0.0] * dim)
        if faces[0]["type"] == "periodic" or faces[1]["type"] == "periodic":
            raise ValueError("Radius cannot be periodic")
        if origin[0] == 0 and faces[0]["type"] != "wall":
            raise ValueError("The radial origin requires a reflective wall")
        if g == "spherical" and dim >= 2:
            top = np.isclose(origin[1] + c["lengths_xyz"][1], np.pi, atol=1e-14, rtol=0)
            if (origin[1] == 0 and faces[2]["type"] != "wall") or (top and faces[3]["type"] != "wall"):
                raise ValueError("A spherical pole requires a reflective wall")
            if faces[2]["type"] == "periodic" or faces[3]["type"] == "periodic":
                raise ValueError("Polar theta cannot be periodic")
        a = azimuth_axis(g, dim)
        if (a is not None and faces[2 * a]["type"] == "periodic"
                and not np.isclose(c["lengths_xyz"][a], 2 * np.pi, atol=1e-14, rtol=0)):
            raise ValueError("Azimuthal periodicity requires a complete 2*pi domain")
    return faces


def session_faces(c, spec):
    """Faces of a new session: walls at a radial origin and at spherical poles
    unless given, when the legacy boundaries are all Open (torchSession.py:93-121)."""
    dim, g = c["dimension"], c.get("geometry", "cartesian")
    if c["boundaries_xyz"] == ["Open"] * dim and g != "cartesian" and (spec is None or isinstance(spec, dict)):
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
