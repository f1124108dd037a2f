"""Checkpoints of device sessions in the reference's schema-2 layout, and of
adaptive sessions in its schema-3 leaf layout.

Follows python/cfdSolver/torchIO.py: the metadata record and its checks
(:66-119), the datasets written per session (:192-214: ``metadata`` as a
UTF-8 JSON byte vector, ``state/<component>`` as flat little-endian float64
with x varying fastest, ``state/lambda`` and ``state/beta`` when present),
the atomic replace (:47-63) and the validating reader (:233-391), including
the accepted schema/solver families and the upgrade to solver 1.4 on the
next save.

The reference frames these datasets in HDF5 through h5py, which this image
does not have. The container here is a NumPy ``.npz`` archive holding the
same named datasets; it executes nothing when read (``allow_pickle=False``).
A path ending in ``.h5``/``.hdf5`` is written and read through h5py, with the
reference's dataset options, when h5py is importable, and is refused with a
ValueError when it is not.

The session state is downloaded at a step boundary, so a checkpoint holds
the committed state. The device-side options that change the arithmetic
(precision, kinetic constants, the CFL validation switch) ride in an extra
``device`` record that the reference's reader ignores; a restored session
therefore takes the same steps as the uninterrupted one, bit for bit.
"""

import json
import os
import tempfile
from pathlib import Path

import numpy as np

SCHEMA_VERSION = 2
SOLVER_VERSION = "1.4"
ADAPTIVE_SCHEMA = 3
ADAPTIVE_SOLVER = "2.0"
# (schema_version, solver_version) pairs the reader accepts (torchIO.py:272-273).
FAMILIES = ((1, "1.1"), (2, "1.2"), (2, "1.4"))
REQUIRED = {"config", "time", "steps", "run_id", "initial_integral", "boundary_integral", "source_integral",
            "current_integral"}


def component_names(config):
    """Component dataset names (torchGeometry.py:12-31)."""
    from .torchGeometry import component_names as names

    return names(config)


def _plain(value):
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, np.generic):
        return value.item()
    raise TypeError(f"Cannot serialize {type(value).__name__}")


def _json(value):
    return json.dumps(value, default=_plain, allow_nan=False, separators=(",", ":"))


def _atomic(path, write):
    """Write to a temporary file beside ``path``, fsync it, then rename over ``path``."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, temp = tempfile.mkstemp(prefix="." + path.name + ".", suffix=".tmp", dir=path.parent)
    os.close(fd)
    try:
        write(Path(temp))
        with open(temp, "rb") as handle:
            os.fsync(handle.fileno())
        os.replace(temp, path)
    finally:
        if os.path.exists(temp):
            os.unlink(temp)


def _container(path):
    suffix = Path(path).suffix.lower()
    if suffix in (".h5", ".hdf5"):
        try:
            import h5py  # noqa: F401
        except ImportError:
            raise ValueError("HDF5 checkpoints need h5py, which is not installed; use a .npz path") from None
        return "hdf5"
    return "npz"


def metadata(session):
    """The checkpoint record of a session, after the reference's checks (torchIO.py:66-119)."""
    from .torchSession import _physical

    session._validate_config()
    c = session.config
    dim = c["dimension"]
    state = session.state
    if state.shape != (dim + 2, *c["counts_xyz"][::-1]) or not _physical(state):
        raise ValueError("Cannot save a nonphysical or mismatched state")
    if (not np.isfinite(session.time) or not 0 <= session.time <= c["t_end"]
            or not isinstance(session.steps, int) or not 0 <= session.steps <= c["max_steps"]):
        raise ValueError("Cannot save invalid integration counters")
    if c["reactive"]:
        if session.lam is None or c["kinetic_model"] == "znd" and session.beta is None:
            raise ValueError("Missing reactive composition")
        session._fraction(session.lam, "lambda")
        if c["kinetic_model"] == "znd":
            session._fraction(session.beta, "beta")
    for value in (session.initial_integral, session.boundary_integral, session.source_integral):
        if np.shape(value) != (dim + 2,) or not np.isfinite(value).all():
            raise ValueError("Cannot save invalid conservation balances")
    return dict(
        format="TORCH checkpoint", schema_version=SCHEMA_VERSION, solver_version=SOLVER_VERSION,
        geometry=c.get("geometry", "cartesian"), units="SI", layout="x-fastest", config=c,
        run_id=session.run_id, time=session.time, steps=session.steps,
        initial_integral=session.initial_integral, boundary_integral=session.boundary_integral,
        source_integral=session.source_integral, current_integral=session._integral_of(state),
        device=dict(precision=session.precision, constants=session.constants, checked=session.checked),
    )


def _adaptive(session):
    from .torchAdaptive import AdaptiveSimulation

    return isinstance(session, AdaptiveSimulation)


def adaptive_metadata(session):
    """The schema-3 record of an adaptive session, after the reference's
    checks (torchAdaptiveIO.py:86-153)."""
    from .torchSession import _physical

    session._validate_config()
    c = session.config
    state = session.state
    if state.shape != (c["dimension"] + 2, len(session.mesh.keys)) or not _physical(state):
        raise ValueError("Invalid adaptive state")
    if (not np.isfinite(session.time) or not 0 <= session.time <= c["t_end"]
            or not isinstance(session.steps, int) or not 0 <= session.steps <= c["max_steps"]):
        raise ValueError("Invalid adaptive counters")
    want_lam = bool(c["reactive"])
    want_beta = want_lam and c["kinetic_model"] == "znd"
    if (session.lam is not None) != want_lam or (session.beta is not None) != want_beta:
        raise ValueError("Adaptive composition/model mismatch")
    for key in ("steps", "regrids", "mesh_revision"):
        value = getattr(session, key)
        if not isinstance(value, int) or isinstance(value, bool) or value < 0:
            raise ValueError("Invalid adaptive counters")
    if session.mesh_revision > session.regrids or not isinstance(session.budget_limited, bool):
        raise ValueError("Invalid adaptive regrid metadata")
    if not isinstance(session.run_id, str) or not session.run_id or session.initial_integral[0] <= 0:
        raise ValueError("Invalid adaptive run identity")
    if session.lam is not None:
        session._fraction(session.lam, "lambda")
    if session.beta is not None:
        session._fraction(session.beta, "beta")
    for value in (session.initial_integral, session.boundary_integral, session.source_integral):
        if np.shape(value) != (state.shape[0],) or not np.isfinite(value).all():
            raise ValueError("Invalid adaptive conservation balances")
    return dict(
        format="TORCH checkpoint", schema_version=ADAPTIVE_SCHEMA, solver_version=ADAPTIVE_SOLVER,
        geometry=c["geometry"], units="SI", layout="leaf-level-z-y-x", config=c, run_id=session.run_id,
        time=session.time, steps=session.steps, initial_integral=session.initial_integral,
        boundary_integral=session.boundary_integral, source_integral=session.source_integral,
        current_integral=session.integral, cell_count=len(session.mesh.keys), regrids=session.regrids,
        mesh_revision=session.mesh_revision, budget_limited=bool(session.budget_limited),
        device=dict(precision=session.precision, constants=session.constants, checked=session.checked),
    )


def content(session):
    """Dataset name -> array of a session's checkpoint (torchIO.py:192-214;
    adaptive sessions: torchAdaptiveIO.py:156-185, leaf levels and indices
    as little-endian int64)."""
    if _adaptive(session):
        data = {"metadata": np.frombuffer(_json(adaptive_metadata(session)).encode("utf-8"), dtype=np.uint8)}
        for name, row in zip(component_names(session.config), session.state):
            data["state/" + name] = row.astype("<f8")
        for name, value in (("lambda", session.lam), ("beta", session.beta)):
            if value is not None:
                data["state/" + name] = value.astype("<f8")
        data["mesh/levels"] = session.mesh.levels.astype("<i8")
        data["mesh/indices"] = session.mesh.keys[:, 1:].ravel().astype("<i8")
        return data
    data = {"metadata": np.frombuffer(_json(metadata(session)).encode("utf-8"), dtype=np.uint8)}
    for name, row in zip(component_names(session.config), session.state):
        data["state/" + name] = row.ravel().astype("<f8")
    for name, value in (("lambda", session.lam), ("beta", session.beta)):
        if value is not None:
            data["state/" + name] = value.ravel().astype("<f8")
    return data


def write_checkpoint(path, session):
    """Replace a checkpoint atomically on the same filesystem (torchIO.py:217-227)."""
    data = content(session)
    kind = _container(path)

    def write(temp):
        if kind == "hdf5":
            import h5py

            with h5py.File(temp, "w") as handle:
                for name, value in data.items():
                    handle.create_dataset(name, data=value, fletcher32=name != "metadata")
                handle.flush()
        else:
            with open(temp, "wb") as handle:
                np.savez(handle, **data)

    _atomic(path, write)
    return str(path)


def load_content(path):
    """Dataset name -> array of a checkpoint file, without interpreting it."""
    try:
        if _container(path) == "hdf5":
            import h5py

            with h5py.File(path, "r") as handle:
                out = {"metadata": np.asarray(handle["metadata"])}
                for group in ("state", "mesh"):
                    for name in handle.get(group, ()):
                        out[group + "/" + name] = np.asarray(handle[group + "/" + name])
                return out
        with np.load(path, allow_pickle=False) as archive:
            return {name: archive[name] for name in archive.files}
    except (OSError, KeyError, ValueError) as exc:
        raise ValueError(f"Invalid or incomplete TORCH checkpoint: {exc}") from exc


def record(data):
    """Validate checkpoint datasets like read_checkpoint (torchIO.py:233-391).

    Returns a host-side session shell (a ``Simulation`` with its config,
    counters, integrals and host arrays set, and no device state), so every
    check runs before the device is touched.
    """
    from .torchSession import Simulation, _physical

    try:
        raw = np.asarray(data["metadata"])
        if raw.ndim != 1 or raw.dtype != np.dtype("uint8") or raw.size > 1_000_000:
            raise ValueError("Invalid checkpoint metadata encoding")
        meta = json.loads(raw.tobytes().decode("utf-8"))
        if not isinstance(meta, dict) or not REQUIRED.issubset(meta):
            raise ValueError("Incomplete checkpoint metadata")
        if meta.get("schema_version") == ADAPTIVE_SCHEMA:
            return adaptive_record(data, meta)
        expected = dict(format="TORCH checkpoint", units="SI", layout="x-fastest")
        if any(meta.get(k) != v for k, v in expected.items()):
            raise ValueError("Unsupported checkpoint schema, solver version, geometry or units")
        version = (meta.get("schema_version"), meta.get("solver_version"))
        if version not in FAMILIES:
            raise ValueError("Unsupported checkpoint schema/solver family")
        config = meta["config"]
        if not isinstance(config, dict):
            raise ValueError("Incomplete checkpoint metadata")
        if version[0] == 1 and (meta.get("geometry") != "cartesian"
                                or any(k in config for k in ("geometry", "origins_xyz", "face_boundaries"))):
            raise ValueError("Schema 1 only supports legacy Cartesian configurations")
        if version[0] == 2 and meta.get("geometry") != config.get("geometry", "cartesian"):
            raise ValueError("Checkpoint geometry/configuration mismatch")
        device = meta.get("device") or {}
        sim = Simulation.__new__(Simulation)
        sim._device = None
        sim._execution = None
        sim.config = config
        for key in ("counts_xyz", "lengths_xyz", "boundaries_xyz", "origins_xyz"):
            if key in config:
                config[key] = list(np.atleast_1d(config[key]))
        sim.precision = str(device.get("precision", "fp64")).lower()
        sim.checked = bool(device.get("checked", True))
        sim.constants = dict(device.get("constants") or {})
        sim.tile = (0, 0, 0)
        if not isinstance(config.get("dimension"), int) or config["dimension"] not in (1, 2, 3):
            raise ValueError("Invalid dimension")
        for key in ("counts_xyz", "lengths_xyz", "boundaries_xyz"):
            if len(config[key]) != config["dimension"]:
                raise ValueError("Incomplete mesh/boundary metadata")
        sim._validate_config()
        dim = config["dimension"]
        shape = tuple(config["counts_xyz"][::-1])
        size = int(np.prod(shape))

        def array(name):
            value = np.asarray(data["state/" + name])
            if value.shape != (size,) or value.dtype.kind != "f" or value.dtype.itemsize != 8:
                raise ValueError("Checkpoint dataset has incompatible shape or precision")
            return np.asarray(value, dtype=float).reshape(shape)

        state = np.stack([array(name) for name in component_names(config)])
        lam = array("lambda") if config["reactive"] else None
        beta = array("beta") if config["reactive"] and config["kinetic_model"] == "znd" else None
        names = set(component_names(config)) | ({"lambda"} if lam is not None else set()) | (
            {"beta"} if beta is not None else set())
        if {k[len("state/"):] for k in data if k.startswith("state/")} != names:
            raise ValueError("Unexpected or missing composition datasets")
    except (KeyError, TypeError, UnicodeError, json.JSONDecodeError) as exc:
        raise ValueError(f"Invalid or incomplete TORCH checkpoint: {exc}") from exc
    sim.time, sim.steps, sim.run_id = meta["time"], meta["steps"], meta["run_id"]
    if (not isinstance(sim.run_id, str) or not sim.run_id or not isinstance(sim.time, (int, float))
            or isinstance(sim.time, bool) or not np.isfinite(sim.time) or not 0 <= sim.time <= config["t_end"]):
        raise ValueError("Invalid checkpoint run identity or time")
    if (not isinstance(sim.steps, int) or isinstance(sim.steps, bool) or sim.steps < 0
            or sim.steps > config["max_steps"]):
        raise ValueError("Invalid checkpoint step counter")
    for name in ("initial_integral", "boundary_integral", "source_integral", "current_integral"):
        value = np.asarray(meta[name], dtype=float)
        if value.shape != (dim + 2,) or not np.isfinite(value).all():
            raise ValueError("Invalid integrated diagnostics")
        if name != "current_integral":
            setattr(sim, name, value)
        elif not np.allclose(value, sim._integral_of(state), rtol=1e-12, atol=1e-12):
            raise ValueError("Checkpoint state/integral mismatch")
    if not _physical(state) or sim.initial_integral[0] <= 0:
        raise ValueError("Nonphysical checkpoint state")
    if lam is not None:
        sim._fraction(lam, "lambda")
    if beta is not None:
        sim._fraction(beta, "beta")
    sim.fallbacks = 0
    sim._host = (state, lam, beta)
    return sim


def adaptive_record(data, meta):
    """Validate an adaptive (schema 3) checkpoint like the reference's reader
    (torchAdaptiveIO.py:188-300); returns a host-side AdaptiveSimulation
    shell (mesh, state, counters and balances; no device state)."""
    from .torchAdaptive import AdaptiveSimulation
    from .torchAdaptiveGeometry import LeafMesh, settings

    required = {"cell_count", "regrids", "mesh_revision", "budget_limited", "config", "time", "steps", "run_id",
                "initial_integral", "boundary_integral", "source_integral", "current_integral"}
    if not required <= set(meta) or any(meta.get(k) != v for k, v in dict(
            format="TORCH checkpoint", solver_version=ADAPTIVE_SOLVER, units="SI",
            layout="leaf-level-z-y-x").items()):
        raise ValueError("Unsupported adaptive checkpoint metadata")
    sim = AdaptiveSimulation.__new__(AdaptiveSimulation)
    sim._device = sim._leaves = sim._leaves_key = sim._mesh_on_device = sim._execution = None
    sim.config = c = meta["config"]
    if not isinstance(c, dict) or meta.get("geometry") != c.get("geometry") or "amr" not in c \
            or "face_boundaries" not in c:
        raise ValueError("Adaptive geometry/config mismatch")
    device = meta.get("device") or {}
    sim.precision = str(device.get("precision", "fp64")).lower()
    sim.checked = bool(device.get("checked", True))
    sim.constants = dict(device.get("constants") or {})
    sim.tile = (0, 0, 0)
    c["amr"] = settings(**c["amr"])
    for key in ("counts_xyz", "lengths_xyz", "origins_xyz", "boundaries_xyz"):
        c[key] = list(np.atleast_1d(c[key]))
    sim._validate_config()
    n, d = meta["cell_count"], c["dimension"]
    if not isinstance(n, int) or isinstance(n, bool) or not 0 < n <= c["amr"]["max_cells"]:
        raise ValueError("Invalid adaptive cell count")

    def array(path, length, kind):
        value = np.asarray(data[path])
        if value.shape != (length,) or value.dtype.kind != kind or value.dtype.itemsize != 8:
            raise ValueError("Invalid adaptive dataset layout or precision")
        if not np.isfinite(value).all():
            raise ValueError("Nonfinite adaptive dataset")
        return value

    levels = array("mesh/levels", n, "i")
    indices = array("mesh/indices", n * d, "i").reshape(n, d)
    keys = np.column_stack((levels, indices)).astype(np.int64)
    sim.mesh = LeafMesh(c, keys)
    if not np.array_equal(sim.mesh.keys, keys) or np.any(levels < c["amr"]["min_level"]):
        raise ValueError("Noncanonical or undersized adaptive leaves")
    names = component_names(c)
    expected = set(names) | ({"lambda"} if c["reactive"] else set()) | (
        {"beta"} if c["reactive"] and c["kinetic_model"] == "znd" else set())
    if {k[len("state/"):] for k in data if k.startswith("state/")} != expected:
        raise ValueError("Unexpected adaptive state fields")
    state = np.stack([array("state/" + name, n, "f").astype(float) for name in names])
    lam = array("state/lambda", n, "f").astype(float) if c["reactive"] else None
    beta = array("state/beta", n, "f").astype(float) if "beta" in expected else None
    for key in ("steps", "regrids", "mesh_revision"):
        value = meta[key]
        if not isinstance(value, int) or isinstance(value, bool) or value < 0:
            raise ValueError("Invalid adaptive step/regrid counters")
        setattr(sim, key, value)
    sim.time, sim.run_id, sim.budget_limited = meta["time"], meta["run_id"], meta["budget_limited"]
    if not isinstance(sim.budget_limited, bool) or sim.mesh_revision > sim.regrids:
        raise ValueError("Invalid adaptive regrid metadata")
    for key in ("initial_integral", "boundary_integral", "source_integral"):
        value = np.asarray(meta[key], dtype=float)
        if value.shape != (d + 2,) or not np.isfinite(value).all():
            raise ValueError("Invalid adaptive balances")
        setattr(sim, key, value)
    if not isinstance(sim.run_id, str) or not sim.run_id or sim.initial_integral[0] <= 0:
        raise ValueError("Invalid adaptive run identity")
    sim.fallbacks = 0
    sim._host = (state, lam, beta)
    adaptive_metadata(sim)
    current = np.asarray(meta["current_integral"], dtype=float)
    if current.shape != (d + 2,) or not np.isfinite(current).all() or not np.allclose(
            current, sim.integral, rtol=1e-12, atol=1e-12):
        raise ValueError("Adaptive state/integral mismatch")
    return sim


def read_checkpoint(path, *, precision=None, device=0, workers=1):
    """Validate a checkpoint and return a session resident on ``device``.

    ``precision`` overrides the saved one ("fp64" for checkpoints without a
    device record, such as the reference's own).
    """
    sim = record(load_content(path))
    if precision is not None:
        sim.precision = str(precision).lower()
        sim._validate_config()
    return sim._attach(workers, device)
