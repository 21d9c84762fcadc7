"""The allocation-free upward continuation (tcsem_upcont_plane_noalloc) against the allocating
entry point (tcsem_upcont_plane), through the C-ABI on caller device buffers.

Two calls in a row share the library's scratch array: the first on a larger air plane, the second
on a smaller one, so the second call runs over the first one's leftovers.  Both must equal the
allocating entry point bit for bit (same kernels, same inputs)."""
import ctypes as C

import numpy as np
import pytest

import tcsem_fdtd_amd as T
from tests import decks

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """A hipMalloc'd buffer of the test (the entry point under test allocates nothing)."""
    _hip = None

    def __init__(self, nbytes):
        if DeviceBuffer._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            DeviceBuffer._hip = h
        self.p = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self._hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def get(self, shape, dtype):
        a = np.empty(shape, dtype=dtype)
        assert a.nbytes == self.nbytes
        assert self._hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        if self.p:
            self._hip.hipFree(self.p)
            self.p = C.c_void_p()


def _context(setup_kw, precision):
    st = decks.small_setup(T, **setup_kw)
    gi = T.grid_info(st)
    Z = T.grid_arrays(st)["Z"][:, None, None] * np.ones((1, gi.M, gi.L))
    return T.Context(st, decks.layered_cond((gi.N, gi.M, gi.L), Z, seed=2), precision)


@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_noalloc_upcont_equals_the_allocating_entry_point(precision):
    dt = np.float64 if precision == T.FP64 else np.float32
    rng = np.random.default_rng(31)
    # the larger plane first, then the smaller one over its leftovers in the scratch array
    for kw in ({"nx": 80, "ny": 64, "nz": 8}, {}):
        with _context(kw, precision) as c:
            L, M = c.L, c.M
            bz0 = rng.standard_normal((M, L))
            bx_ref, by_ref = c.upcont_plane(bz0)
            bufs = [DeviceBuffer(M * L * dt().itemsize), DeviceBuffer(M * (L + 1) * dt().itemsize),
                    DeviceBuffer((M + 1) * L * dt().itemsize)]
            try:
                bufs[0].put(bz0.astype(dt))
                bufs[1].put(np.full((M, L + 1), np.nan, dt))
                bufs[2].put(np.full((M + 1, L), np.nan, dt))
                rc = T.lib().tcsem_upcont_plane_noalloc(c.h, bufs[0].p, bufs[1].p, bufs[2].p)
                assert rc == 0, T.lib().tcsem_last_error().decode()
                bx = bufs[1].get((M, L + 1), dt).astype(np.float64)
                by = bufs[2].get((M + 1, L), dt).astype(np.float64)
            finally:
                for b in bufs:
                    b.free()
            assert np.array_equal(bx, bx_ref), (kw, np.abs(bx - bx_ref).max())
            assert np.array_equal(by, by_ref), (kw, np.abs(by - by_ref).max())


def test_noalloc_upcont_rejects_planes_beyond_its_limit(monkeypatch):
    """Planes that take the line chain (FFT lengths above 2048, forced here) have no
    allocation-free path: a clear size error, and the buffers are not touched."""
    monkeypatch.setenv("TCSEM_UPCONT_CHAIN", "lines")
    with _context({}, T.FP64) as c:
        bufs = [DeviceBuffer(8 * c.M * c.L), DeviceBuffer(8 * c.M * (c.L + 1)), DeviceBuffer(8 * (c.M + 1) * c.L)]
        try:
            bufs[1].put(np.full((c.M, c.L + 1), 7.0))
            rc = T.lib().tcsem_upcont_plane_noalloc(c.h, bufs[0].p, bufs[1].p, bufs[2].p)
            assert rc == -6 and b"allocation-free" in T.lib().tcsem_last_error()
            assert np.all(bufs[1].get((c.M, c.L + 1), np.float64) == 7.0)
        finally:
            for b in bufs:
                b.free()
