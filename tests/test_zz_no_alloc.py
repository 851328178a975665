"""The allocation-free compute entry point (pis_compute_potential_noalloc) against the
allocating one (pis_ctx_create + pis_set_state + pis_compute_potential), through the C ABI.

Both run the same kernels on the same input, so forces and PE must be the same bits, in
FP64 and in mixed precision; checked on two calls in a row, the second on a smaller system
(the scratch array keeps the first call's contents).  Device buffers for the caller's
input and output come from hipMalloc through ctypes.
"""
import ctypes as C

import numpy as np
import pytest

from pis_amd import _capi, workloads as W
from pis_amd.engine import _DescHolder

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceArray:
    """A device copy of a numpy array (hipMalloc), read back with get()."""

    def __init__(self, hip, a):
        self.hip = hip
        self.a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, _H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.a.nbytes, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _lib():
    return _capi.lib()


def allocating(spec, precision):
    lib = _lib()
    h = _DescHolder(spec, precision)
    ctx = C.c_void_p()
    _capi.check(lib.pis_ctx_create(C.byref(h.desc), C.byref(ctx)))
    try:
        pos = np.ascontiguousarray(spec.positions, dtype=np.float64)
        _capi.check(lib.pis_set_state(ctx, pos.ctypes.data_as(C.POINTER(C.c_double)), None, None))
        f = np.zeros((spec.n, 3))
        pe = C.c_double(0.0)
        _capi.check(lib.pis_compute_potential(ctx, f.ctypes.data_as(C.POINTER(C.c_double)), C.byref(pe)))
    finally:
        lib.pis_ctx_destroy(ctx)
    return f, pe.value


def noalloc(spec, precision, types=None):
    """(status, forces, pe) of the allocation-free call on device copies of the input."""
    lib = _lib()
    hip = _hip()
    h = _DescHolder(spec, precision)
    h.desc.type_ids = None  # the types come from device memory
    d_pos = DeviceArray(hip, np.asarray(spec.positions, dtype=np.float64))
    d_types = DeviceArray(hip, np.asarray(spec.type_ids if types is None else types, dtype=np.uint32))
    d_f = DeviceArray(hip, np.zeros((spec.n, 3)))
    d_pe = DeviceArray(hip, np.zeros(1))
    try:
        st = lib.pis_compute_potential_noalloc(C.byref(h.desc), d_pos.p, d_types.p, d_f.p, d_pe.p, None)
        return st, d_f.get(), float(d_pe.get()[0])
    finally:
        for d in (d_pos, d_types, d_f, d_pe):
            d.free()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_equals_the_allocating_entry_point_twice(precision):
    systems = [
        W.lj_liquid(12, 0.8442, 0.72, seed=21),  # 6,912 atoms, tile-relative staging
        W.ka_mixture(2000, 1.2, 1.0, seed=22),   # smaller, two types
    ]
    for spec in systems:
        print(f"noalloc {precision}: N={spec.n}", flush=True)
        f_ref, pe_ref = allocating(spec, precision)
        st, f, pe = noalloc(spec, precision)
        assert st == 0, _capi.lib().pis_last_error()
        assert np.array_equal(f, f_ref)
        assert pe == pe_ref
        assert np.all(np.isfinite(f)) and np.abs(f).max() > 0.0


def test_noalloc_rejects_what_does_not_fit_and_bad_types():
    big = W.lj_liquid(40, 0.8442, 0.72, seed=23)  # 256,000 atoms: beyond the 64 MiB scratch
    st, _, _ = noalloc(big, "mixed")
    assert st == 10  # PIS_ERR_CAPACITY, before any device work
    msg = _capi.lib().pis_last_error()
    print(msg.decode(), flush=True)
    assert b"scratch" in msg
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=24)
    types = np.asarray(spec.type_ids, dtype=np.uint32).copy()
    types[7] = 3  # one species only
    st, _, _ = noalloc(spec, "fp64", types)
    assert st == 4  # PIS_ERR_INVALID_ATOM_TYPE
    # and the next call on the same scratch is unaffected
    f_ref, pe_ref = allocating(spec, "fp64")
    st, f, pe = noalloc(spec, "fp64")
    assert st == 0 and np.array_equal(f, f_ref) and pe == pe_ref
