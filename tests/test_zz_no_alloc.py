"""pm_permeability_noalloc against pm_permeability.

The entry point that allocates nothing reduces fields the caller keeps in
device memory, with its partial sums in a fixed device array of the library.
It runs the same reduction kernels over the same rows in the same order as
pm_permeability, so its eight values must equal pm_permeability's bit for
bit.  Two calls in a row, the second on a smaller domain (the array keeps the
first call's partials; the second must not see them), with and without a
dom_interest box, with the porosity given and computed.
"""
import ctypes as C

import numpy as np
import pytest

import poremaps_amd as pm
from conftest import random_geometry

pytestmark = pytest.mark.gpu


def _noalloc(dev, shape, dom, dx, phi):
    L = pm.lib()
    L.pm_permeability_noalloc.argtypes = [C.c_void_p] * 5 + [C.c_int] * 3 + [C.POINTER(C.c_int), C.c_double,
                                                                           C.c_double, C.c_void_p, C.c_void_p]
    nz, ny, nx = shape
    out = (C.c_double * 8)()
    rc = L.pm_permeability_noalloc(*dev, nx, ny, nz, (C.c_int * 6)(*dom), dx, phi, out,
                                   None)
    assert rc == 0, L.pm_last_error().decode()
    return np.array(out[:])


class DeviceArrays:
    """Device copies of host arrays, made with the HIP runtime the library
    itself links (the test is the caller that owns device memory)."""

    def __init__(self, arrays):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []
        for a in arrays:
            a = np.ascontiguousarray(a)
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
            self.ptrs.append(p)
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _case(shape, bc, seed, dom, phi, sweeps):
    raw = random_geometry(shape, 0.4, seed)
    print("[case] noalloc", shape, bc, dom, phi, flush=True)
    with pm.Solver(raw, boundary_method=bc, voxel_size=2e-5, mode="jacobi", precision="f64", omega=2.0,
                   dom_interest=dom, porosity=phi) as s:
        s.step(sweeps)
        want = np.array(list(s.permeability().values()), dtype=np.float64)
        fields = [s.field(n) for n in ("press", "velx", "vely", "velz")]
    dev = DeviceArrays([raw] + fields)
    try:
        got = _noalloc(dev.ptrs, shape, dom, 2e-5, phi)
    finally:
        dev.close()
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    assert np.isfinite(got).all() and got[5] != 0.0


@pytest.mark.parametrize("dom,phi", [((0, 0, 0, 0, 0, 0), -1.0), ((2, 20, 1, 15, 3, 25), -1.0),
                                     ((1, 5, 2, 6, 1, 7), 0.5)])
def test_noalloc_equals_pm_permeability_twice(dom, phi):
    # first a larger domain, then a smaller one on the same device array
    _case((30, 18, 26), 0, 5, dom, phi, 40)
    small = tuple(min(d, n - 1) for d, n in zip(dom, (9, 9, 7, 7, 11, 11))) if any(dom) else dom
    _case((11, 7, 9), 2, 6, small, phi, 25)


def test_noalloc_rejects_bad_sizes():
    L = pm.lib()
    L.pm_permeability_noalloc.argtypes = [C.c_void_p] * 5 + [C.c_int] * 3 + [C.POINTER(C.c_int), C.c_double,
                                                                           C.c_double, C.c_void_p, C.c_void_p]
    out = (C.c_double * 8)()
    one = C.c_void_p(1)
    assert L.pm_permeability_noalloc(one, one, one, one, one, 2048, 2048, 1024, None, 1.0, -1.0, out, None) != 0
    assert b"exceed" in L.pm_last_error()
    assert L.pm_permeability_noalloc(one, one, one, one, one, 0, 4, 4, None, 1.0, -1.0, out, None) != 0
