"""ctypes binding of libtorchhip.so (declared in include/torch_hip.h).

The library is loaded from this package directory. When it is missing or
cannot be loaded, every entry point raises: there is no CPU fallback.
"""

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libtorchhip.so")

OK, NONPHYSICAL, INVALID, HIP_ERROR, STALLED, NO_STABLE_STEP, SPECIES_RANGE, REACTION_STATE, BUDGET = (
    0, 2, 3, 4, 5, 6, 7, 8, 9)

LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")
BOUNDARY_CODES = {"open": 0, "transmissive": 0, "outflow": 0, "periodic": 1, "wrap": 1,
                  "reflective": 2, "wall": 2, "slip": 2, "inflow": 3}
CONTROLLERS = ("Sound Speed Controlled", "w=0")
VARIANTS = ("Variante 1 - Original", "Variante_2")
WAVESPEEDS = ("Pressure Based", "Enthalpy Based")
PRECISIONS = {"fp64": 0, "float64": 0, "fp32": 1, "float32": 1}

EXPORTS = (
    "torch_ctx_create", "torch_ctx_destroy", "torch_last_error", "torch_state_create",
    "torch_state_destroy", "torch_state_upload", "torch_state_download", "torch_state_dt",
    "torch_state_step", "torch_state_hydro", "torch_state_fluxes", "torch_state_react",
    "torch_state_totals", "torch_state_profile", "torch_hllc_faces", "torch_hllc_faces_static",
    "torch_batch_footprint", "torch_batch_create", "torch_batch_destroy", "torch_batch_add",
    "torch_batch_remove", "torch_batch_seed", "torch_batch_run", "torch_batch_info", "torch_batch_download",
    "torch_batch_tune", "torch_batch_set_tiles", "torch_batch_profile", "torch_batch_graph_counts",
    "torch_batch_time_pass", "torch_batch_time_step_kernel",
    "torch_state_set_metric", "torch_state_metric_dt", "torch_state_metric_rhs", "torch_state_metric_advance",
    "torch_state_metric_step", "torch_rusanov_faces", "torch_state_wave_maxima",
    "torch_amr_create", "torch_amr_destroy", "torch_amr_set_mesh", "torch_amr_upload", "torch_amr_download",
    "torch_amr_faces", "torch_amr_dt", "torch_amr_rhs", "torch_amr_advance", "torch_amr_step",
)

GEOMETRY_CODES = {"cartesian": 0, "cylindrical": 1, "annular": 2, "spherical": 3}
FACE_CODES = {"outflow": 0, "wall": 1, "periodic": 2, "inflow": 3, "pressure_outlet": 4, "stagnation_inlet": 5}


class Config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("n", ctypes.c_int * 3), ("lengths", ctypes.c_double * 3),
        ("gamma", ctypes.c_double), ("limiter", ctypes.c_int), ("bc", ctypes.c_int * 3),
        ("inlet_primitive", ctypes.c_double * 4), ("reactive", ctypes.c_int),
        ("kinetic_model", ctypes.c_int), ("q_heat", ctypes.c_double),
        ("activation_energy", ctypes.c_double), ("pre_exponential", ctypes.c_double),
        ("controller", ctypes.c_int), ("variant", ctypes.c_int), ("wavespeed", ctypes.c_int),
        ("cfl", ctypes.c_double), ("precision", ctypes.c_int), ("block_x", ctypes.c_int),
        ("block_y", ctypes.c_int), ("zchunk", ctypes.c_int), ("spacing", ctypes.c_double * 3),
        ("gamma_mode", ctypes.c_int),
    ]


class StepInfo(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("fallback", ctypes.c_int), ("retries", ctypes.c_int),
        ("outward", ctypes.c_double * 5), ("source", ctypes.c_double * 5),
        ("wave_max", ctypes.c_double * 3),
    ]


class Face(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int), ("has_backflow", ctypes.c_int), ("primitive", ctypes.c_double * 5),
        ("pressure", ctypes.c_double), ("total_pressure", ctypes.c_double),
        ("total_temperature", ctypes.c_double), ("gas_constant", ctypes.c_double),
        ("lam", ctypes.c_double), ("beta", ctypes.c_double),
    ]


def faces_of(faces):
    """Normalized face dicts (torchBoundaries.normalize_faces) -> Face array."""
    out = (Face * 6)()
    for i, f in enumerate(faces):
        o = out[i]
        o.type = FACE_CODES[f["type"]]
        state = f.get("primitive", f.get("backflow"))
        o.has_backflow = int("backflow" in f)
        if state is not None:
            for k, v in enumerate(state):
                o.primitive[k] = float(v)
        o.pressure = float(f.get("pressure", 0.0))
        o.total_pressure = float(f.get("total_pressure", 0.0))
        o.total_temperature = float(f.get("total_temperature", 0.0))
        o.gas_constant = float(f.get("gas_constant", 0.0))
        o.lam = float(f.get("lambda", 1.0))
        o.beta = float(f.get("beta", 1.0))
    return out


class BatchProgress(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("dt", ctypes.c_double), ("steps", ctypes.c_int),
        ("fallbacks", ctypes.c_int), ("status", ctypes.c_int), ("active", ctypes.c_int),
        ("completed", ctypes.c_int), ("cells", ctypes.c_long), ("retries", ctypes.c_int),
        ("outward", ctypes.c_double * 5), ("source", ctypes.c_double * 5), ("wave_max", ctypes.c_double * 3),
    ]


_lib = None
_lock = threading.Lock()
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def library_path():
    return LIBRARY


def load():
    """Load libtorchhip.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIBRARY):
                raise RuntimeError(
                    f"HIP library {LIBRARY} is missing; build it with `python -m torch_amd.build`")
            lib = ctypes.CDLL(LIBRARY)
            vp = ctypes.c_void_p
            sig = {
                "torch_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
                "torch_ctx_destroy": (ctypes.c_int, [vp]),
                "torch_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
                "torch_state_create": (ctypes.c_int, [vp, ctypes.POINTER(Config), ctypes.POINTER(vp)]),
                "torch_state_destroy": (ctypes.c_int, [vp]),
                "torch_state_upload": (ctypes.c_int, [vp, _dp, _dp, _dp]),
                "torch_state_download": (ctypes.c_int, [vp, _dp, _dp, _dp]),
                "torch_state_dt": (ctypes.c_int, [vp, _dp]),
                "torch_state_wave_maxima": (ctypes.c_int, [vp, _dp]),
                "torch_state_step": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double,
                                                    ctypes.POINTER(StepInfo)]),
                "torch_state_hydro": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int,
                                                     ctypes.POINTER(StepInfo)]),
                "torch_state_fluxes": (ctypes.c_int, [vp, _dp, _dp, _dp]),
                "torch_state_react": (ctypes.c_int, [vp, ctypes.c_double]),
                "torch_state_totals": (ctypes.c_int, [vp, _dp]),
                "torch_state_profile": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_long)]),
                "torch_hllc_faces": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_long, _dp, _dp, _dp, _dp, _dp]),
                "torch_hllc_faces_static": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                           ctypes.c_long, _dp, _dp, _dp, _dp, _dp]),
                "torch_batch_footprint": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int,
                                                         ctypes.POINTER(ctypes.c_size_t)]),
                "torch_batch_create": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(vp)]),
                "torch_batch_destroy": (ctypes.c_int, [vp]),
                "torch_batch_add": (ctypes.c_int, [vp, ctypes.POINTER(Config), _dp, _dp, _dp, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
                "torch_batch_remove": (ctypes.c_int, [vp, ctypes.c_int]),
                "torch_batch_seed": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp]),
                "torch_batch_run": (ctypes.c_int, [vp, ctypes.c_int]),
                "torch_batch_info": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(BatchProgress)]),
                "torch_batch_download": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp]),
                "torch_batch_tune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                    _dp]),
                "torch_batch_set_tiles": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
                "torch_batch_profile": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_long)]),
                "torch_batch_graph_counts": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_long),
                                                            ctypes.POINTER(ctypes.c_long)]),
                "torch_batch_time_pass": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
                "torch_batch_time_step_kernel": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
                "torch_state_set_metric": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.c_long,
                                                          ctypes.POINTER(Face)]),
                "torch_state_metric_dt": (ctypes.c_int, [vp, _dp]),
                "torch_state_metric_rhs": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
                "torch_state_metric_advance": (ctypes.c_int, [vp, ctypes.c_double, ctypes.POINTER(StepInfo)]),
                "torch_state_metric_step": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double,
                                                           ctypes.POINTER(StepInfo)]),
                "torch_rusanov_faces": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_long, _dp, _dp, _dp, _dp, _dp]),
                "torch_amr_create": (ctypes.c_int, [vp, ctypes.POINTER(Config), ctypes.c_int, ctypes.POINTER(Face),
                                                    ctypes.c_int, ctypes.POINTER(vp)]),
                "torch_amr_destroy": (ctypes.c_int, [vp]),
                "torch_amr_set_mesh": (ctypes.c_int, [vp, ctypes.c_long, ctypes.c_long, _ip, _ip, _ip, _ip,
                                                      _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_long,
                                                      _lp, _ip, _ip]),
                "torch_amr_upload": (ctypes.c_int, [vp, _dp, _dp, _dp]),
                "torch_amr_download": (ctypes.c_int, [vp, _dp, _dp, _dp]),
                "torch_amr_faces": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp]),
                "torch_amr_dt": (ctypes.c_int, [vp, _dp]),
                "torch_amr_rhs": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
                "torch_amr_advance": (ctypes.c_int, [vp, ctypes.c_double, ctypes.POINTER(StepInfo)]),
                "torch_amr_step": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double,
                                                  ctypes.POINTER(StepInfo)]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(lib, name)
                fn.restype, fn.argtypes = res, args
            _lib = lib
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().torch_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code):
    """Map a status code to the reference's exception types."""
    if code == OK:
        return
    msg = last_error()
    if code == INVALID:
        raise ValueError(msg)
    raise RuntimeError(msg)


def ptr(array):
    """Pointer to a C-contiguous float64 array (None passes NULL)."""
    if array is None:
        return None
    if array.dtype != np.float64 or not array.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return array.ctypes.data_as(_dp)


def boundary_code(name, dim, axis):
    code = BOUNDARY_CODES.get(str(name).lower())
    if code is None or (code == 3 and not (dim == 2 and axis == 1)):
        raise ValueError("Unsupported session boundary")
    return code


def precision_code(precision):
    code = PRECISIONS.get(str(precision).lower())
    if code is None:
        raise ValueError(f"Unknown precision: {precision}")
    return code


class Context:
    """One device context (HIP stream + lock) of libtorchhip."""

    def __init__(self, device=0):
        lib = load()
        handle = ctypes.c_void_p()
        check(lib.torch_ctx_create(int(device), ctypes.byref(handle)))
        self.handle = handle
        self.device = int(device)

    def close(self):
        if getattr(self, "handle", None):
            load().torch_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceState:
    """A device-resident state created from a Config."""

    def __init__(self, context, config):
        self.context = context
        self.config = config
        handle = ctypes.c_void_p()
        check(load().torch_state_create(context.handle, ctypes.byref(config), ctypes.byref(handle)))
        self.handle = handle
        self.dim = config.dim
        counts = [config.n[d] for d in range(config.dim)]
        self.shape = tuple(counts[::-1])
        self.nf = config.dim + 2

    def upload(self, state, lam=None, beta=None):
        state = np.ascontiguousarray(state, dtype=np.float64)
        if state.shape != (self.nf,) + self.shape:
            raise ValueError(f"state must have shape {(self.nf,) + self.shape}")
        lam = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64)
        beta = None if beta is None else np.ascontiguousarray(beta, dtype=np.float64)
        check(load().torch_state_upload(self.handle, ptr(state), ptr(lam), ptr(beta)))

    def download(self, want_lam=False, want_beta=False):
        state = np.empty((self.nf,) + self.shape)
        lam = np.empty(self.shape) if want_lam else None
        beta = np.empty(self.shape) if want_beta else None
        check(load().torch_state_download(self.handle, ptr(state), ptr(lam), ptr(beta)))
        return state, lam, beta

    def dt_limit(self):
        out = ctypes.c_double()
        check(load().torch_state_dt(self.handle, ctypes.byref(out)))
        return out.value

    def wave_maxima(self):
        out = np.zeros(3)
        check(load().torch_state_wave_maxima(self.handle, ptr(out)))
        return out

    def step(self, time, t_end):
        info = StepInfo()
        check(load().torch_state_step(self.handle, float(time), float(t_end), ctypes.byref(info)))
        return info

    def hydro(self, dt, want_fluxes=False):
        info = StepInfo()
        check(load().torch_state_hydro(self.handle, float(dt), int(bool(want_fluxes)), ctypes.byref(info)))
        return info

    def fluxes(self):
        nz, ny, nx = (1,) * (3 - self.dim) + self.shape
        shapes = []
        for d in range(self.dim):
            full = [nz + (d == 2), ny + (d == 1), nx + (d == 0)]
            shapes.append((self.nf,) + tuple(full[3 - self.dim:]))
        arrays = [np.empty(s) for s in shapes] + [None] * (3 - self.dim)
        check(load().torch_state_fluxes(self.handle, *(ptr(a) for a in arrays)))
        return arrays[: self.dim]

    def react(self, dt):
        check(load().torch_state_react(self.handle, float(dt)))

    def totals(self):
        out = np.zeros(5)
        check(load().torch_state_totals(self.handle, ptr(out)))
        return out[: self.nf]

    def profile(self, enable=True):
        """Enable kernel timing; returns (kernel_ms, launches) since the last call."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        check(load().torch_state_profile(self.handle, int(enable), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- metric/face scheme ------------------------------------------------
    def set_metric(self, geometry, factors, faces):
        factors = np.ascontiguousarray(factors, dtype=np.float64)
        check(load().torch_state_set_metric(self.handle, GEOMETRY_CODES[geometry], ptr(factors), factors.size,
                                            faces_of(faces)))

    def metric_dt(self):
        out = ctypes.c_double()
        check(load().torch_state_metric_dt(self.handle, ctypes.byref(out)))
        return out.value

    def metric_rhs(self, first_order=False, nq=0):
        change = np.empty((self.nf,) + self.shape)
        species = np.empty((max(nq, 1),) + self.shape)
        outward, source = np.empty(self.nf), np.empty(self.nf)
        check(load().torch_state_metric_rhs(self.handle, int(bool(first_order)), ptr(change),
                                            ptr(species) if nq else None, ptr(outward), ptr(source)))
        return change, [species[j] for j in range(nq)], outward, source

    def metric_advance(self, dt):
        info = StepInfo()
        check(load().torch_state_metric_advance(self.handle, float(dt), ctypes.byref(info)))
        return info

    def metric_step(self, time, t_end):
        info = StepInfo()
        check(load().torch_state_metric_step(self.handle, float(time), float(t_end), ctypes.byref(info)))
        return info

    def close(self):
        if getattr(self, "handle", None):
            load().torch_state_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _typed(array, dtype, ctype):
    array = np.ascontiguousarray(array, dtype=dtype)
    return array, array.ctypes.data_as(ctypes.POINTER(ctype))


class DeviceLeaves:
    """Device-resident adaptive leaves (torch_amr_*): one leaf mesh at a time
    with its state, lambda and beta."""

    def __init__(self, context, config, geometry, faces, order):
        self.context = context
        self.config = config
        self.dim = config.dim
        self.nf = config.dim + 2
        self.nq = (2 if config.kinetic_model == 1 else 1) if config.reactive else 0
        handle = ctypes.c_void_p()
        check(load().torch_amr_create(context.handle, ctypes.byref(config), GEOMETRY_CODES[geometry],
                                      faces_of(faces), int(order), ctypes.byref(handle)))
        self.handle = handle
        self.nleaf = 0
        self.nface = 0

    def set_mesh(self, mesh):
        """Install a torchAdaptiveGeometry.LeafMesh (its face list, metrics,
        gradient matrices and incidence lists)."""
        start, faces, roles = mesh.incidence()
        keep = []
        ints = [_typed(a, np.int32, ctypes.c_int) for a in (mesh.left, mesh.right, mesh.directions, mesh.sides)]
        reals = [_typed(a, np.float64, ctypes.c_double) for a in (
            mesh.face_areas, mesh.weights, mesh.offset_left, mesh.offset_right, mesh.displacement,
            mesh.volumes, mesh.divergence, mesh.inverse)]
        starts = _typed(start, np.int64, ctypes.c_long)
        entries = [_typed(a, np.int32, ctypes.c_int) for a in (faces, roles)]
        keep.extend(ints + reals + [starts] + entries)
        nleaf, nface = len(mesh.volumes), len(mesh.left)
        check(load().torch_amr_set_mesh(self.handle, nleaf, nface, *(p for _, p in ints), *(p for _, p in reals),
                                        len(faces), starts[1], *(p for _, p in entries)))
        self.nleaf, self.nface = nleaf, nface

    def upload(self, state, lam=None, beta=None):
        state = np.ascontiguousarray(state, dtype=np.float64)
        if state.shape != (self.nf, self.nleaf):
            raise ValueError(f"leaf state must have shape {(self.nf, self.nleaf)}")
        lam = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64)
        beta = None if beta is None else np.ascontiguousarray(beta, dtype=np.float64)
        for q in (lam, beta):
            if q is not None and q.shape != (self.nleaf,):
                raise ValueError("leaf fractions must have one value per leaf")
        check(load().torch_amr_upload(self.handle, ptr(state), ptr(lam), ptr(beta)))

    def download(self):
        state = np.empty((self.nf, self.nleaf))
        lam = np.empty(self.nleaf) if self.nq > 0 else None
        beta = np.empty(self.nleaf) if self.nq > 1 else None
        check(load().torch_amr_download(self.handle, ptr(state), ptr(lam), ptr(beta)))
        return state, lam, beta

    def faces(self):
        """Raw face values (left, right, gl, gr) of the committed state."""
        left, right = np.empty((self.nf, self.nface)), np.empty((self.nf, self.nface))
        gl, gr = np.empty(self.nface), np.empty(self.nface)
        check(load().torch_amr_faces(self.handle, ptr(left), ptr(right), ptr(gl), ptr(gr)))
        return left, right, gl, gr

    def dt(self):
        out = ctypes.c_double()
        check(load().torch_amr_dt(self.handle, ctypes.byref(out)))
        return out.value

    def rhs(self, first_order=False):
        change = np.empty((self.nf, self.nleaf))
        species = np.empty((max(self.nq, 1), self.nleaf))
        outward, source = np.empty(self.nf), np.empty(self.nf)
        check(load().torch_amr_rhs(self.handle, int(bool(first_order)), ptr(change),
                                   ptr(species) if self.nq else None, ptr(outward), ptr(source)))
        return change, [species[j] for j in range(self.nq)], outward, source

    def advance(self, dt):
        info = StepInfo()
        check(load().torch_amr_advance(self.handle, float(dt), ctypes.byref(info)))
        return info

    def step(self, time, t_end):
        info = StepInfo()
        check(load().torch_amr_step(self.handle, float(time), float(t_end), ctypes.byref(info)))
        return info

    def close(self):
        if getattr(self, "handle", None):
            load().torch_amr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hllc_faces(context, nv, direction, left, right, gl, gr, precision=0):
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    m = left.shape[1]
    gl = np.ascontiguousarray(np.broadcast_to(gl, (m,)), dtype=np.float64)
    gr = np.ascontiguousarray(np.broadcast_to(gr, (m,)), dtype=np.float64)
    out = np.empty_like(left)
    check(load().torch_hllc_faces(context.handle, int(precision), int(nv), int(direction), m, ptr(left),
                                  ptr(right), ptr(gl), ptr(gr), ptr(out)))
    return out


def hllc_faces_static(context, nv, direction, left, right, gl, gr, precision=0):
    """hllc_faces through the allocation-free entry point (at most 65536 faces)."""
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    m = left.shape[1]
    gl = np.ascontiguousarray(np.broadcast_to(gl, (m,)), dtype=np.float64)
    gr = np.ascontiguousarray(np.broadcast_to(gr, (m,)), dtype=np.float64)
    out = np.empty_like(left)
    check(load().torch_hllc_faces_static(context.handle, int(precision), int(nv), int(direction), m, ptr(left),
                                         ptr(right), ptr(gl), ptr(gr), ptr(out)))
    return out


def rusanov_faces(context, nv, direction, left, right, gl, gr, precision=0):
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    m = left.shape[1]
    gl = np.ascontiguousarray(np.broadcast_to(gl, (m,)), dtype=np.float64)
    gr = np.ascontiguousarray(np.broadcast_to(gr, (m,)), dtype=np.float64)
    out = np.empty_like(left)
    check(load().torch_rusanov_faces(context.handle, int(precision), int(nv), int(direction), m, ptr(left),
                                     ptr(right), ptr(gl), ptr(gr), ptr(out)))
    return out


class DeviceBatch:
    """A device batch of inert instances (torch_batch_* in include/torch_hip.h)."""

    def __init__(self, context, precision=1, hbm_budget=0):
        self.context = context
        self.precision = int(precision)
        handle = ctypes.c_void_p()
        check(load().torch_batch_create(context.handle, self.precision, int(hbm_budget), ctypes.byref(handle)))
        self.handle = handle
        self.shapes = {}

    @staticmethod
    def footprint(config, precision):
        out = ctypes.c_size_t()
        check(load().torch_batch_footprint(ctypes.byref(config), int(precision), ctypes.byref(out)))
        return out.value

    def add(self, config, state, time, t_end, max_steps, lam=None, beta=None):
        """Admit an instance; returns its id, or None when the HBM budget is full."""
        state = np.ascontiguousarray(state, dtype=np.float64)
        counts = [config.n[d] for d in range(config.dim)]
        shape = (config.dim + 2,) + tuple(counts[::-1])
        if state.shape != shape:
            raise ValueError(f"state must have shape {shape}")
        lam = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64)
        beta = None if beta is None else np.ascontiguousarray(beta, dtype=np.float64)
        for a in (lam, beta):
            if a is not None and a.shape != shape[1:]:
                raise ValueError(f"lambda/beta must have shape {shape[1:]}")
        out = ctypes.c_int()
        code = load().torch_batch_add(self.handle, ctypes.byref(config), ptr(state), ptr(lam), ptr(beta),
                                      float(time), float(t_end), int(max_steps), ctypes.byref(out))
        if code == BUDGET:
            return None
        check(code)
        self.shapes[out.value] = (shape, bool(config.reactive), config.kinetic_model == 1)
        return out.value

    def remove(self, ident):
        check(load().torch_batch_remove(self.handle, int(ident)))
        self.shapes.pop(ident, None)

    def seed(self, ident, outward, source):
        """Start an instance's boundary/source integrals at a session's totals."""
        nf = self.shapes[ident][0][0]
        out = np.ascontiguousarray(np.asarray(outward, dtype=np.float64).reshape(nf))
        src = np.ascontiguousarray(np.asarray(source, dtype=np.float64).reshape(nf))
        check(load().torch_batch_seed(self.handle, int(ident), ptr(out), ptr(src)))

    def run(self, steps):
        check(load().torch_batch_run(self.handle, int(steps)))

    def info(self, ident):
        out = BatchProgress()
        check(load().torch_batch_info(self.handle, int(ident), ctypes.byref(out)))
        return out

    def download(self, ident):
        """(state, lam, beta) of an instance; lam/beta None unless reactive (ZND)."""
        shape, reactive, znd = self.shapes[ident]
        state = np.empty(shape)
        lam = np.empty(shape[1:]) if reactive else None
        beta = np.empty(shape[1:]) if znd else None
        check(load().torch_batch_download(self.handle, int(ident), ptr(state), ptr(lam), ptr(beta)))
        return state, lam, beta

    def tune(self, reps=3):
        bx, by, zc, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(load().torch_batch_tune(self.handle, int(reps), ctypes.byref(bx), ctypes.byref(by),
                                      ctypes.byref(zc), ctypes.byref(ms)))
        return (bx.value, by.value, zc.value), ms.value

    def set_tiles(self, block_x, block_y, zchunk):
        check(load().torch_batch_set_tiles(self.handle, int(block_x), int(block_y), int(zchunk)))

    def time_pass(self, reps=5):
        """Average ms of the 3D second-order work list (nothing advances)."""
        ms = ctypes.c_double()
        check(load().torch_batch_time_pass(self.handle, int(reps), ctypes.byref(ms)))
        return ms.value

    def time_step_kernel(self, steps=10):
        """Advance `steps` ordinary steps, timing the one 3D work list of each
        with HIP events inside the step graph; its mean ms per launch."""
        ms = ctypes.c_double()
        check(load().torch_batch_time_step_kernel(self.handle, int(steps), ctypes.byref(ms)))
        return ms.value

    def profile(self, enable=True):
        """Enable replay timing; returns (ms, steps) since the last call."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        check(load().torch_batch_profile(self.handle, int(enable), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def graph_counts(self):
        """(instantiations, in-place updates) of the step graph so far."""
        built, updated = ctypes.c_long(), ctypes.c_long()
        check(load().torch_batch_graph_counts(self.handle, ctypes.byref(built), ctypes.byref(updated)))
        return built.value, updated.value

    def close(self):
        if getattr(self, "handle", None):
            load().torch_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
