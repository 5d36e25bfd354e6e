"""CPU oracle for the POREMAPS Stokes sweep -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP implementation in ``poremaps_amd``.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product path never does.

The numerical work happens in ``libpm_oracle.so`` (plain C, built by
``oracle/Makefile`` with ``-ffp-contract=off``), a clean-room restatement of
the reference's ``src/solver.cc``, ``src/geometry.cc``, ``src/initialize.cc``,
``src/parallelization.cc`` (one-rank periodic semantics),
``src/boundary_conditions.cc``, ``src/evaluation.cc`` and ``src/calc_flow.cc``
(file:line citations sit beside each C function), plus the build-side fast
modes (Jacobi, 3-colour SOR, defect correction) whose semantics the GPU fast
kernels must reproduce bit for bit.

Arrays use the reference's halo-2 layout: shape (nz+4, ny+4, nx+4), x fastest.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpm_oracle.so")
_lib = None

RE = 0.01
C2 = 1.5e6
DT = 1.0e-4


def build() -> str:
    """Compile libpm_oracle.so with the committed Makefile; returns its path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_I = C.c_int
_D = C.c_double
_F = C.c_float
_VP = C.c_void_p


class OrcParams(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int), ("it_eval", C.c_int), ("it_write", C.c_int),
        ("solver", C.c_int), ("eps", C.c_double), ("porosity", C.c_double),
        ("dom", C.c_int * 6), ("dx", C.c_double),
    ]


def _declare(L):
    L.orc_geom_halo.argtypes = [_u8p, _I, _I, _I, _I, _u8p]
    L.orc_geom_halo_walls.argtypes = [_u8p, _I, _I, _I, _I, _I, _u8p]
    L.orc_eval_geometry.argtypes = [_u8p, _I, _I, _I, _i32p]
    L.orc_porosity.argtypes = [_u8p, _I, _I, _I]
    L.orc_porosity.restype = _D
    L.orc_init_fields.argtypes = [_u8p, _I, _I, _I, _I, _f64p, _f64p, _f64p, _f64p]
    L.orc_exchange_f64.argtypes = [_f64p, _I, _I, _I, _I]
    L.orc_pins.argtypes = [_u8p, _I, _I, _I, _f64p]
    L.orc_sweep_ref.argtypes = [_I, _u8p, _i32p, _I, _I, _I, _f64p, _f64p, _f64p, _f64p]
    L.orc_sweeps_ref_slabs.argtypes = [_I, _u8p, _i32p, _I, _I, _I, _I, _f64p, _f64p, _f64p, _f64p, _I, _I]
    L.orc_sweeps_ref_slabs.restype = _I
    L.orc_sweeps_ref_blocks.argtypes = [_I, _u8p, _i32p, _I, _I, _I, _I, _f64p, _f64p, _f64p, _f64p, _VP, _I,
                                        C.POINTER(C.c_int)]
    L.orc_sweeps_ref_blocks.restype = _I
    L.orc_convergence.argtypes = [_u8p, _f64p, _I, _I, _I, C.POINTER(C.c_double)]
    L.orc_convergence.restype = _D
    L.orc_permeability.argtypes = [_u8p, _f64p, _f64p, _f64p, _f64p, _I, _I, _I,
                                   C.POINTER(C.c_int), _D, _D, C.POINTER(C.c_double)]
    L.orc_calc_flow.argtypes = [_u8p, _i32p, _I, _I, _I, _I, C.POINTER(OrcParams),
                                _f64p, _f64p, _f64p, _f64p, _VP, _I, C.POINTER(C.c_int)]
    L.orc_calc_flow.restype = _I
    L.orc_pack_code.argtypes = [_i32p, _I, _I, _I, _u8p]
    for t, p in (("f64", _f64p), ("f32", _f32p)):
        s = _D if t == "f64" else _F
        getattr(L, f"orc_sweep_jacobi_{t}").argtypes = (
            [_u8p, _I, _I, _I, _I, _D, _D] + [p] * 10 + [s] + [_VP] * 4)
        getattr(L, f"orc_sweep_colour_{t}").argtypes = (
            [_u8p, _I, _I, _I, _D, _D] + [p] * 5 + [s] + [_VP] * 4)
        getattr(L, f"orc_ghost_{t}").argtypes = [_u8p, _I, _I, _I, _I] + [p] * 4 + [_VP] + [s] * 3
    L.orc_outlet_plane.argtypes = [_u8p, _I, _I, _I, _I] + [_f64p] * 4 + [_D, _f64p]
    L.orc_residual.argtypes = [_u8p, _I, _I, _I] + [_f64p] * 5 + [_f32p] * 4
    L.orc_fold.argtypes = [_u8p, _I, _I, _I, _I] + [_f64p] * 4 + [_f32p] * 4
    L.orc_solve_fast.argtypes = [_I, _I, _u8p, _u8p, _I, _I, _I, _I, _D, _D, _D, _I, _I] + [_f64p] * 4 + [
        C.POINTER(C.c_double)]
    L.orc_solve_fast.restype = _I
    L.orc_set_slip.argtypes = [_I, _I]
    L.orc_d2_ref.argtypes = [_I] + [_D] * 5
    L.orc_d2_ref.restype = _D
    L.orc_d2_f64.argtypes = [_I] + [_D] * 5
    L.orc_d2_f64.restype = _D
    L.orc_d2_f32.argtypes = [_I] + [_F] * 5
    L.orc_d2_f32.restype = _F


def _dims(a):
    nz, ny, nx = (s - 4 for s in a.shape)
    return nx, ny, nz


# ----------------------------------------------------------------------------
# reference-faithful pieces
# ----------------------------------------------------------------------------

def geom_halo(raw: np.ndarray, bc: int, walls: int = 0) -> np.ndarray:
    """raw (nz, ny, nx) uint8 0/1 -> haloed geometry (nz+4, ny+4, nx+4);
    walls = 1: no-slip walls on the non-periodic x / y faces (solid halos)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    nz, ny, nx = raw.shape
    G = np.empty((nz + 4, ny + 4, nx + 4), np.uint8)
    lib().orc_geom_halo_walls(raw, nx, ny, nz, bc, walls, G)
    return G


def eval_geometry(G: np.ndarray) -> np.ndarray:
    code = np.empty(G.shape, np.int32)
    lib().orc_eval_geometry(np.ascontiguousarray(G), *_dims(G), code)
    return code


def porosity(G: np.ndarray) -> float:
    return lib().orc_porosity(np.ascontiguousarray(G), *_dims(G))


def init_fields(G: np.ndarray, bc: int):
    p, u, v, w = (np.empty(G.shape, np.float64) for _ in range(4))
    lib().orc_init_fields(G, *_dims(G), bc, p, u, v, w)
    return p, u, v, w


def exchange(f: np.ndarray, bc: int) -> None:
    lib().orc_exchange_f64(f, *_dims(f), bc)


def pins(G: np.ndarray, p: np.ndarray) -> None:
    lib().orc_pins(G, *_dims(G), p)


def sweep_ref(solver: int, G, code, p, u, v, w) -> None:
    lib().orc_sweep_ref(solver, G, code, *_dims(G), p, u, v, w)


def sweeps_ref_slabs(solver: int, G, code, p, u, v, w, bc: int, nsweeps: int, threads: int) -> None:
    """nsweeps reference sweeps, exchanges and pins with the interior cut into
    `threads` z-slabs (one thread each), as mpirun -np P with
    dom_decomposition 1 1 P runs them; threads = 1 is the serial sweep."""
    if lib().orc_sweeps_ref_slabs(solver, G, code, *_dims(G), bc, p, u, v, w, nsweeps, threads) != 0:
        raise MemoryError("orc_sweeps_ref_slabs: allocation failed")


def sweeps_ref_blocks(solver: int, G, code, p, u, v, w, bc: int, nsweeps: int, dims, outlet=None) -> None:
    """nsweeps reference sweeps, exchanges and pins as `mpirun -np P*Q*R`
    with dom_decomposition P Q R runs them (dims = (P, Q, R), x, y, z).
    outlet: the outlet plane before its last pins ((ny+4, nx+4) float64,
    updated in place), carried between calls; None for one call from a state
    that has not been swept since its pins."""
    d = (C.c_int * 3)(*dims)
    ptr = None
    if outlet is not None:
        assert outlet.dtype == np.float64 and outlet.flags.c_contiguous and outlet.shape == G.shape[1:]
        ptr = outlet.ctypes.data
    if lib().orc_sweeps_ref_blocks(solver, G, code, *_dims(G), bc, p, u, v, w, ptr, nsweeps, d) != 0:
        raise MemoryError("orc_sweeps_ref_blocks: allocation failed")


def convergence(G, w, kprev: float):
    k = C.c_double(kprev)
    conv = lib().orc_convergence(G, w, *_dims(G), C.byref(k))
    return conv, k.value


def permeability(G, u, v, w, p, dom, dx: float, phi: float):
    """Returns (k13, k23, k33, wk13, wk23, wk33, wmax_velz, wmean_velz)."""
    d = (C.c_int * 6)(*dom)
    out = (C.c_double * 8)()
    lib().orc_permeability(G, u, v, w, p, *_dims(G), d, dx, phi, out)
    return tuple(out)


def calc_flow(G, code, bc: int, state, *, solver=2, eps=1e-6, max_iter=10**8,
              it_eval=100, it_write=100, porosity=-1.0, dom=(0,) * 6, dx=1.0,
              max_rows=100000):
    """Reference solver loop on ``state`` (p, u, v, w, modified in place).
    Returns (sweeps, rows[n, 11])."""
    p, u, v, w = state
    prm = OrcParams(max_iter, it_eval, it_write, solver, eps, porosity,
                    (C.c_int * 6)(*dom), dx)
    rows = np.zeros((max_rows, 11), np.float64)
    n = C.c_int(0)
    t = lib().orc_calc_flow(G, code, *_dims(G), bc, C.byref(prm), p, u, v, w,
                            rows.ctypes.data, max_rows, C.byref(n))
    return t, rows[: min(n.value, max_rows)].copy()


# ----------------------------------------------------------------------------
# build-side fast modes
# ----------------------------------------------------------------------------

def d2(case, ll, l, c, r, rr, kind="ref"):
    """get_2nd_derivation (src/solver.cc:57-91); kind ref | f64 | f32."""
    return getattr(lib(), f"orc_d2_{kind}")(case, ll, l, c, r, rr)


def pack_code(code: np.ndarray) -> np.ndarray:
    pc = np.empty(code.shape, np.uint8)
    lib().orc_pack_code(np.ascontiguousarray(code, dtype=np.int32), *_dims(code), pc)
    return pc


def _rptrs(R):
    if R is None:
        return [None] * 4
    assert all(r.dtype == np.float32 and r.flags.c_contiguous for r in R)
    return [r.ctypes.data for r in R]


def _tag(a):
    return "f64" if a.dtype == np.float64 else "f32"


def sweep_jacobi(pc, bc, omega, src, dst, q0, q1, pin_out, R=None, omega_p=None) -> None:
    """Out-of-place symplectic Jacobi sweep from src=(p,u,v,w) into dst;
    omega relaxes the velocities, omega_p (default omega) the pressure."""
    fn = getattr(lib(), f"orc_sweep_jacobi_{_tag(src[0])}")
    fn(pc, *_dims(pc), bc, omega, omega if omega_p is None else omega_p, *src, *dst, q0, q1, pin_out,
       *_rptrs(R))


def sweep_colour(pc, omega, state, q, pin_out, R=None, omega_p=None) -> None:
    fn = getattr(lib(), f"orc_sweep_colour_{_tag(state[0])}")
    fn(pc, *_dims(pc), omega, omega if omega_p is None else omega_p, *state, q, pin_out, *_rptrs(R))


def ghost(G, bc, state, q, pin_in, pin_hi, pin_out) -> None:
    fn = getattr(lib(), f"orc_ghost_{_tag(state[0])}")
    fn(G, *_dims(G), bc, *state, None if q is None else q.ctypes.data, pin_in, pin_hi, pin_out)


def solution_pins(G):
    nz = G.shape[0] - 4
    return float(nz + 2), 2.0, 3.0


def outlet_plane_init(p):
    """q of a freshly initialised solution state: p on the plane z = nz+1."""
    return np.ascontiguousarray(p[-3]).copy()


def outlet_plane(pc, bc, X, pin_out=3.0):
    q = np.zeros(pc.shape[1:], np.float64)
    lib().orc_outlet_plane(pc, *_dims(pc), bc, *X, pin_out, q)
    return q


def residual(pc, X, q):
    R = tuple(np.zeros(pc.shape, np.float32) for _ in range(4))
    lib().orc_residual(pc, *_dims(pc), *X, q, *R)
    return R


@contextlib.contextmanager
def slip(bc, on=True):
    """The build's slip walls (walls = 2) for the fast-mode sweeps and ghost
    fills run inside: the non-periodic x / y faces of boundary method bc."""
    lib().orc_set_slip(1 if on else 0, bc)
    try:
        yield
    finally:
        lib().orc_set_slip(0, bc)


def solve_fast(raw, bc, mode="jacobi", precision="f32dc", omega=2.5, omega_p=None, eps=1e-6, it_eval=100,
               max_iter=10 ** 7, walls=0):
    """A whole fast-mode solve of the geometry raw (nz, ny, nx) to the stop
    rule: returns (sweeps, kprev, (p, u, v, w) halo-2 FP64 state, G).
    walls: 0 the reference's open faces, 1 no-slip, 2 slip walls."""
    G = geom_halo(raw, bc, walls)
    pc = pack_code(eval_geometry(G))
    st = [a.copy() for a in init_fields(G, bc)]
    pins(G, st[0])
    k = C.c_double(0.0)
    with slip(bc, walls == 2):
        t = lib().orc_solve_fast(1 if mode == "jacobi" else 2, 1 if precision == "f32dc" else 0, G, pc, *_dims(G),
                                 bc, omega, omega if omega_p is None else omega_p, eps, it_eval, max_iter, *st,
                                 C.byref(k))
    if t < 0:
        raise MemoryError("orc_solve_fast: allocation failed")
    return t, k.value, st, G


def fold(G, bc, X, d) -> None:
    lib().orc_fold(G, *_dims(G), bc, *X, *d)
