"""poremaps_amd -- MI355X-native POREMAPS Stokes permeability solver.

Python view of the C ABI in ``include/poremaps_amd.h`` (ctypes over the
in-tree ``libporemaps_amd.so``).  The solver itself is C++/HIP; this module
only moves host buffers across the boundary, for tests, benchmarks and
scripting.  There is no CPU fallback: every call goes to the GPU library, and
importing the binding fails loudly when the library has not been built.

The call sequence mirrors the reference's ``POREMAPS input.inp`` run
(src/main_stokes_solver.cc:195-411): parameters, geometry, ``calc_flow``,
field output.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libporemaps_amd.so")
CLI_PATH = os.path.join(_HERE, "bin", "POREMAPS")

PM_OK = 0
MODES = {"reference": 0, "jacobi": 1, "sor3": 2}
PRECISIONS = {"f64": 0, "f32dc": 1}
FIELDS = {"velx": 0, "vely": 1, "velz": 2, "press": 3, "nhood": 4}
# Relaxation (velocities, pressure) of the fast modes: the symplectic sweep
# is stable while omega * omega_p stays below about 19 (wave CFL of the
# pressure-velocity coupling); omega speeds up the viscous relaxation of the
# velocities, omega_p the slow Darcy diffusion of the pressure.  OMEGA,
# OMEGA_P: the pair for open media; AUTO (omega = 0, the batch and serving
# default): the library picks the pair per geometry from its porosity,
# (0.6, 30) below 0.20 and (2.5, 7) otherwise.
OMEGA, OMEGA_P = 2.5, 7.0
AUTO = 0.0
# serving: requests of at most this many voxels (128^3) run in their own batch
SMALL_CELLS = 128 ** 3


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("boundary_method", C.c_int), ("voxel_size", C.c_double),
        ("max_iter", C.c_int), ("it_eval", C.c_int), ("it_write", C.c_int),
        ("solving_algorithm", C.c_int), ("eps", C.c_double), ("porosity", C.c_double),
        ("dom_interest", C.c_int * 6),
        ("mode", C.c_int), ("precision", C.c_int), ("omega", C.c_double), ("device", C.c_int),
        ("omega_p", C.c_double), ("z_slabs", C.c_int), ("dom_decomposition", C.c_int * 3),
        ("walls", C.c_int),
    ]


class LogRow(C.Structure):
    _fields_ = [("iteration", C.c_int)] + [
        (n, C.c_double) for n in ("conv", "tps", "wmax_velz", "wmean_velz", "k13", "k23", "k33",
                                  "wk13", "wk23", "wk33")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Result(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("conv", C.c_double), ("porosity", C.c_double),
                ("seconds", C.c_double), ("glups", C.c_double), ("n_log_rows", C.c_int),
                ("last", LogRow), ("restarts", C.c_int)]


class InpExtra(C.Structure):
    _fields_ = [("dom_decomposition", C.c_int * 3), ("geometry_file", C.c_char * 400),
                ("log_file", C.c_char * 400), ("write_output", C.c_int * 4)]


LOG_FN = C.CFUNCTYPE(C.c_int, C.POINTER(LogRow), C.c_void_p)
PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p)

_lib = None


class PoremapsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m poremaps_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.pm_last_error.restype = C.c_char_p
        L.pm_io_last_error.restype = C.c_char_p
        L.pm_params_default.argtypes = [P(Params)]
        L.pm_ctx_create.argtypes = [P(Params), P(C.c_void_p)]
        L.pm_load_geometry.argtypes = [C.c_void_p, C.c_void_p]
        L.pm_solve.argtypes = [C.c_void_p, LOG_FN, PROGRESS_FN, C.c_void_p, P(Result)]
        L.pm_get_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.pm_get_field_halo.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.pm_porosity.argtypes = [C.c_void_p, P(C.c_double)]
        L.pm_permeability.argtypes = [C.c_void_p, C.c_void_p]
        L.pm_step.argtypes = [C.c_void_p, C.c_int]
        L.pm_device_bytes.argtypes = [P(Params)]
        L.pm_device_bytes.restype = C.c_int64
        L.pm_destroy.argtypes = [C.c_void_p]
        L.pm_read_inp.argtypes = [C.c_char_p, P(Params), P(InpExtra)]
        L.pm_write_raw_f64.argtypes = [C.c_char_p, C.c_void_p, C.c_int64, C.c_double, C.c_int]
        L.pm_write_raw_i32.argtypes = [C.c_char_p, C.c_void_p, C.c_int64]
        L.pm_write_log_row.argtypes = [C.c_char_p, P(LogRow), C.c_int]
        L.pm_rank_map.argtypes = [C.c_int, C.c_int, C.c_int, P(C.c_int), C.c_void_p]
        _lib = L
    return _lib


def _check(rc, io=False):
    if rc != PM_OK:
        L = lib()
        msg = (L.pm_io_last_error() if io else L.pm_last_error()) or b""
        raise PoremapsError(rc, msg.decode(errors="replace"))


def make_params(nx, ny, nz, *, boundary_method=0, voxel_size=1.0, max_iter=100_000_000,
                it_eval=100, it_write=100, solving_algorithm=2, eps=1e-6, porosity=-1.0,
                dom_interest=(0, 0, 0, 0, 0, 0), mode="reference", precision="f64", omega=1.0,
                device=0, omega_p=0.0, z_slabs=0, dom_decomposition=(0, 0, 0), walls=0) -> Params:
    p = Params()
    lib().pm_params_default(C.byref(p))
    p.nx, p.ny, p.nz = nx, ny, nz
    p.boundary_method = boundary_method
    p.voxel_size = voxel_size
    p.max_iter, p.it_eval, p.it_write = max_iter, it_eval, it_write
    p.solving_algorithm = solving_algorithm
    p.eps, p.porosity = eps, porosity
    p.dom_interest = (C.c_int * 6)(*dom_interest)
    p.mode = MODES[mode] if isinstance(mode, str) else mode
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
    p.omega = omega
    p.omega_p = omega_p
    p.z_slabs = z_slabs
    p.dom_decomposition = (C.c_int * 3)(*dom_decomposition)
    p.walls = walls
    p.device = device
    return p


def rank_map(shape_zyx, dims) -> np.ndarray:
    """The reference's domain_decomp_ contents: the rank of every voxel under
    dom_decomposition dims (x, y, z), shape (nz, ny, nx) int32."""
    nz, ny, nx = shape_zyx
    out = np.empty((nz, ny, nx), np.int32)
    _check(lib().pm_rank_map(nx, ny, nz, (C.c_int * 3)(*dims), out.ctypes.data), io=True)
    return out


def device_bytes(params: Params) -> int:
    return int(lib().pm_device_bytes(C.byref(params)))


class Solver:
    """One geometry on one GPU (a pm_ctx)."""

    def __init__(self, raw: np.ndarray, **kw):
        raw = np.ascontiguousarray(raw, dtype=np.uint8)
        nz, ny, nx = raw.shape
        self.shape = (nz, ny, nx)
        self.params = make_params(nx, ny, nz, **kw)
        h = C.c_void_p()
        _check(lib().pm_ctx_create(C.byref(self.params), C.byref(h)))
        self._h = h
        _check(lib().pm_load_geometry(self._h, raw.ctypes.data))

    def close(self):
        if getattr(self, "_h", None):
            lib().pm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve(self, log=None, progress=None):
        """Iterate to eps; returns (Result, list of log-row dicts)."""
        rows = []

        def _log(r, _u):
            rows.append(r.contents.as_dict())
            return 0 if log is None else int(bool(log(rows[-1])))

        def _prog(it, conv, secs, _u):
            return 0 if progress is None else int(bool(progress(it, conv, secs)))

        lf, pf = LOG_FN(_log), PROGRESS_FN(_prog)
        res = Result()
        _check(lib().pm_solve(self._h, lf, pf if progress else PROGRESS_FN(0), None, C.byref(res)))
        return res, rows

    def step(self, n: int):
        _check(lib().pm_step(self._h, int(n)))

    def field(self, name: str) -> np.ndarray:
        nz, ny, nx = self.shape
        dt = np.int32 if name == "nhood" else np.float64
        out = np.empty((nz, ny, nx), dt)
        _check(lib().pm_get_field(self._h, FIELDS[name], out.ctypes.data))
        return out

    def field_halo(self, name: str) -> np.ndarray:
        nz, ny, nx = self.shape
        out = np.empty((nz + 4, ny + 4, nx + 4), np.float64)
        _check(lib().pm_get_field_halo(self._h, FIELDS[name], out.ctypes.data))
        return out

    def state_halo(self):
        """(p, u, v, w) with halos, the reference's array layout."""
        return tuple(self.field_halo(n) for n in ("press", "velx", "vely", "velz"))

    @property
    def porosity(self) -> float:
        v = C.c_double()
        _check(lib().pm_porosity(self._h, C.byref(v)))
        return v.value

    def relaxation(self):
        """(omega, omega_p) in use: the pair chosen from the porosity when the
        solver was made with omega=AUTO, after any restart."""
        L = lib()
        L.pm_relaxation.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        w, wp = C.c_double(), C.c_double()
        _check(L.pm_relaxation(self._h, C.byref(w), C.byref(wp)))
        return w.value, wp.value

    def kernel_choice(self):
        """(sweep kernel form in use, {form: ms per sweep} measured by the
        autotuner, empty until it has finished).  The dict also holds the
        FP32+DC Jacobi layout timings ("layout-compact", "layout-dense", ms
        per sweep) and, for the packed 3-colour SOR, "packed-far": the number
        of entries whose links fit no packed record."""
        L = lib()
        L.pm_kernel_choice.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
        name, tried = C.create_string_buffer(64), C.create_string_buffer(1024)
        _check(L.pm_kernel_choice(self._h, name, 64, tried, 1024))
        ms = {}
        for item in filter(None, tried.value.decode().split(";")):
            k, v = item.split("=")
            ms[k] = float(v)
        return name.value.decode(), ms

    def permeability(self) -> dict:
        out = (C.c_double * 8)()
        _check(lib().pm_permeability(self._h, out))
        keys = ("k13", "k23", "k33", "wk13", "wk23", "wk33", "wmax_velz", "wmean_velz")
        return dict(zip(keys, list(out)))


class BatchParams(C.Structure):
    _fields_ = [("precision", C.c_int), ("omega", C.c_double), ("it_eval", C.c_int), ("it_write", C.c_int),
                ("max_iter", C.c_int), ("eps", C.c_double), ("device", C.c_int), ("use_graph", C.c_int),
                ("budget_bytes", C.c_int64), ("omega_p", C.c_double)]


SAMPLE_STATES = {0: "running", 1: "converged", 2: "max_iter", 3: "diverged", 4: "rejected"}
SAMPLE_REJECTED = 4


def _batch_lib():
    L = lib()
    if not getattr(L, "_batch_declared", False):
        P = C.POINTER
        L.pm_batch_params_default.argtypes = [P(BatchParams)]
        L.pm_batch_create.argtypes = [P(BatchParams), P(C.c_void_p)]
        L.pm_batch_sample_bytes.argtypes = [P(BatchParams), C.c_int, C.c_int, C.c_int]
        L.pm_batch_sample_bytes.restype = C.c_int64
        L.pm_batch_add.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.c_double, P(C.c_int)]
        L.pm_batch_step.argtypes = [C.c_void_p, P(C.c_int)]
        L.pm_batch_launch.argtypes = [C.c_void_p]
        L.pm_batch_ready.argtypes = [C.c_void_p]
        L.pm_batch_finish.argtypes = [C.c_void_p, P(C.c_int)]
        L.pm_batch_status.argtypes = [C.c_void_p, C.c_int, P(C.c_int), P(Result)]
        L.pm_batch_get_field.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.pm_batch_remove.argtypes = [C.c_void_p, C.c_int]
        L.pm_batch_memory.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64)]
        L.pm_batch_destroy.argtypes = [C.c_void_p]
        L.pm_batch_set_eps.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L._batch_declared = True
    return L


class Batch:
    """Many geometries on one GPU (include/poremaps_amd_batch.h)."""

    def __init__(self, *, precision="f32dc", omega=AUTO, it_eval=100, it_write=100, max_iter=100_000_000,
                 eps=1e-6, device=0, use_graph=True, budget_bytes=0, omega_p=0.0):
        L = _batch_lib()
        p = _batch_params(precision, omega, it_eval, it_write, max_iter, eps, device, use_graph, budget_bytes,
                          omega_p)
        self.params = p
        h = C.c_void_p()
        _check(L.pm_batch_create(C.byref(p), C.byref(h)))
        self._h = h
        self.shapes = {}

    def sample_bytes(self, shape) -> int:
        nz, ny, nx = shape
        return int(_batch_lib().pm_batch_sample_bytes(C.byref(self.params), nx, ny, nz))

    def add(self, raw: np.ndarray, *, boundary_method=0, voxel_size=1.0, porosity=-1.0) -> int:
        raw = np.ascontiguousarray(raw, dtype=np.uint8)
        nz, ny, nx = raw.shape
        slot = C.c_int()
        _check(_batch_lib().pm_batch_add(self._h, raw.ctypes.data, nx, ny, nz, boundary_method, voxel_size,
                                         porosity, C.byref(slot)))
        self.shapes[slot.value] = raw.shape
        return slot.value

    def set_eps(self, slot: int, eps: float):
        _check(_batch_lib().pm_batch_set_eps(self._h, slot, eps))

    def set_dom_interest(self, slot: int, box):
        """x0 x1 y0 y1 z0 z1 (interior indices, inclusive) of the sample's
        sub-domain permeability (k13, k23, k33)."""
        L = _batch_lib()
        L.pm_batch_set_dom_interest.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        _check(L.pm_batch_set_dom_interest(self._h, slot, (C.c_int * 6)(*box)))

    def step(self) -> int:
        n = C.c_int()
        _check(_batch_lib().pm_batch_step(self._h, C.byref(n)))
        return n.value

    def launch(self):
        """Enqueue one interval on the batch's stream and return (finish()
        completes it; ready() tells whether the GPU is done with it)."""
        _check(_batch_lib().pm_batch_launch(self._h))

    def ready(self) -> bool:
        return bool(_batch_lib().pm_batch_ready(self._h))

    def finish(self) -> int:
        n = C.c_int()
        _check(_batch_lib().pm_batch_finish(self._h, C.byref(n)))
        return n.value

    def status(self, slot: int):
        st, res = C.c_int(), Result()
        _check(_batch_lib().pm_batch_status(self._h, slot, C.byref(st), C.byref(res)))
        return SAMPLE_STATES[st.value], res

    def field(self, slot: int, name: str) -> np.ndarray:
        dt = np.int32 if name == "nhood" else np.float64
        out = np.empty(self.shapes[slot], dt)
        _check(_batch_lib().pm_batch_get_field(self._h, slot, FIELDS[name], out.ctypes.data))
        return out

    def remove(self, slot: int):
        _check(_batch_lib().pm_batch_remove(self._h, slot))
        self.shapes.pop(slot, None)

    def memory(self):
        u, b = C.c_int64(), C.c_int64()
        _check(_batch_lib().pm_batch_memory(self._h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def close(self):
        if getattr(self, "_h", None):
            _batch_lib().pm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Request(C.Structure):
    _fields_ = [("raw", C.c_void_p), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("boundary_method", C.c_int), ("voxel_size", C.c_double), ("porosity", C.c_double),
                ("eps", C.c_double), ("arrival_s", C.c_double), ("dom_interest", C.c_int * 6)]


class Response(C.Structure):
    _fields_ = [("k", C.c_double * 9), ("arrival_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("sweeps", C.c_int * 3), ("state", C.c_int * 3),
                ("k_dom", C.c_double * 9)]

    @property
    def tensor(self) -> np.ndarray:
        return np.array(self.k[:], dtype=np.float64).reshape(3, 3)

    @property
    def tensor_dom(self) -> np.ndarray:
        """The dom_interest sub-domain tensor (zeros without a box)."""
        return np.array(self.k_dom[:], dtype=np.float64).reshape(3, 3)

    @property
    def latency_s(self) -> float:
        return self.done_s - self.arrival_s


SERVE_PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p)


class ServeParams(C.Structure):
    _fields_ = [("batch", BatchParams), ("realtime", C.c_int), ("progress", SERVE_PROGRESS_FN),
                ("user", C.c_void_p), ("max_solves", C.c_int), ("small_cells", C.c_int64)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("completed", C.c_int), ("latency_p50_s", C.c_double),
                ("latency_p99_s", C.c_double), ("latency_mean_s", C.c_double), ("latency_max_s", C.c_double),
                ("solves_per_hour", C.c_double), ("max_running", C.c_int), ("peak_bytes", C.c_int64),
                ("intervals", C.c_int), ("admit_s", C.c_double), ("step_s", C.c_double), ("rejected", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _serve_lib():
    L = _batch_lib()
    if not getattr(L, "_serve_declared", False):
        P = C.POINTER
        L.pm_tensor.argtypes = [P(BatchParams), C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                C.c_double, P(C.c_double), P(Result)]
        L.pm_serve.argtypes = [P(ServeParams), P(Request), C.c_int, P(Response), P(ServeStats)]
        L._serve_declared = True
    return L


def _batch_params(precision="f32dc", omega=AUTO, it_eval=100, it_write=100, max_iter=100_000_000, eps=1e-6,
                  device=0, use_graph=True, budget_bytes=0, omega_p=0.0) -> BatchParams:
    p = BatchParams()
    _batch_lib().pm_batch_params_default(C.byref(p))
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
    p.omega, p.it_eval, p.it_write, p.max_iter = omega, it_eval, it_write, max_iter
    p.eps, p.device, p.use_graph, p.budget_bytes = eps, device, int(use_graph), budget_bytes
    p.omega_p = omega_p
    return p


def tensor(raw: np.ndarray, *, boundary_method=0, voxel_size=1.0, porosity=-1.0, **batch_kw):
    """Full 3x3 permeability tensor k[i, j] (velocity axis i, driving axis j;
    0 = x, 1 = y, 2 = z) from three rotated solves that share the GPU
    (include/poremaps_amd_serve.h).  Returns (k, [Result] * 3) with the
    results in the order flow along z, x, y."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    nz, ny, nx = raw.shape
    p = _batch_params(**batch_kw)
    k = (C.c_double * 9)()
    res = (Result * 3)()
    _check(_serve_lib().pm_tensor(C.byref(p), raw.ctypes.data, nx, ny, nz, boundary_method, voxel_size, porosity,
                                  k, res))
    return np.array(k[:]).reshape(3, 3), list(res)


def serve(requests, *, realtime=True, progress=None, max_solves=0, small_cells=SMALL_CELLS, **batch_kw):
    """Serve permeability-tensor requests: `requests` is a sequence of
    (raw, boundary_method, voxel_size, eps, arrival_s[, dom_interest]) (eps
    <= 0: the batch's eps; dom_interest: x0 x1 y0 y1 z0 z1 of the physical
    axes, for Response.tensor_dom).  ``progress(completed, running, t_s)`` is called as requests
    complete and about once a second; a true return stops serving.
    Requests of at most ``small_cells`` voxels run in a second batch that
    shares the GPU time with the first (0: one batch).  Returns (responses,
    stats) with one Response per request, in request order."""
    raws = [np.ascontiguousarray(r[0], dtype=np.uint8) for r in requests]
    n = len(raws)
    req = (Request * max(n, 1))()
    for i, (raw, r) in enumerate(zip(raws, requests)):
        _, bc, dx, eps, t = r[:5]
        box = tuple(r[5]) if len(r) > 5 else (0,) * 6
        nz, ny, nx = raw.shape
        req[i] = Request(raw.ctypes.data, nx, ny, nz, bc, dx, -1.0, eps, t, (C.c_int * 6)(*box))
    resp = (Response * max(n, 1))()
    st = ServeStats()
    cb = SERVE_PROGRESS_FN(lambda d, r, t, _u: int(bool(progress(d, r, t)))) if progress else SERVE_PROGRESS_FN(0)
    sp = ServeParams(_batch_params(**batch_kw), int(realtime), cb, None, int(max_solves), int(small_cells))
    _check(_serve_lib().pm_serve(C.byref(sp), req, n, resp, C.byref(st)))
    return list(resp)[:n], st


def read_inp(path: str):
    """Parse an ordered .inp (src/read_problem_params.cc:75-90)."""
    p, x = Params(), InpExtra()
    _check(lib().pm_read_inp(path.encode(), C.byref(p), C.byref(x)), io=True)
    return p, x
