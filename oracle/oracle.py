"""ctypes loader for the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: the checker for the HIP product.  Only tests/, the smoke check in
__graft_entry__.py and the cpu_baseline leg of bench.py import this module; the product
package never does.  Build with ``make -C oracle``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

EX, EY, EZ, BX, BY, BZ, BXAIR, BYAIR = range(8)


class Setup(C.Structure):
    """FDTD_setup.in (ModelingSetup.cpp:51-59); mirrors tcsem_setup in include/tcsem.h."""
    _fields_ = [("sfa1", C.c_double), ("sfa2", C.c_double), ("n_src", C.c_int),
                ("sigma_t", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
                ("z0", C.c_double), ("dxmin", C.c_double), ("dymin", C.c_double),
                ("dzmin", C.c_double), ("arg1", C.c_double), ("arg2", C.c_double),
                ("ALx2", C.c_int), ("ALx1", C.c_int), ("Nx", C.c_int), ("ARx1", C.c_int),
                ("ARx2", C.c_int), ("Ay2", C.c_int), ("Ay1", C.c_int), ("Ny", C.c_int),
                ("Az2", C.c_int), ("Az1", C.c_int), ("Nz", C.c_int)]


class RxTime(C.Structure):
    _fields_ = [("T0", C.c_double), ("Tn", C.c_double), ("Nt", C.c_int), ("is_log", C.c_int),
                ("is_comp", C.c_int * 6)]


class Tx(C.Structure):
    _fields_ = [("id", C.c_int), ("ort", C.c_int), ("bipolen", C.c_double), ("x", C.c_double),
                ("y", C.c_double), ("z", C.c_double), ("azi_deg", C.c_double),
                ("dip_deg", C.c_double), ("bipolen1", C.c_double), ("x1", C.c_double),
                ("y1", C.c_double), ("z1", C.c_double), ("azi1_deg", C.c_double),
                ("dip1_deg", C.c_double), ("nr", C.c_int), ("rx_ids", C.POINTER(C.c_int))]


class TxPos(C.Structure):
    """orc_txpos: located transmitter cells (LoadRxTxPos.cpp:143-428)."""
    _fields_ = ([("ort", C.c_int)] + [(n, C.c_double) for n in
                ("BipoLen", "x", "y", "z", "azi", "dip")] +
                [(n, C.c_int) for n in ("ixa", "iya", "iza", "ixb", "iyb", "izb", "ictx",
                                        "icty", "ictz")] +
                [(n, C.c_double) for n in ("BipoLen1", "x1", "y1", "z1", "azi1", "dip1")] +
                [(n, C.c_int) for n in ("ixa1", "iya1", "iza1", "ixb1", "iyb1", "izb1",
                                        "ictx1", "icty1", "ictz1")])


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def lib():
    """Load (building on first use if needed) liboracle.so."""
    global _lib
    if _lib is not None:
        return _lib
    src = os.path.join(_HERE, "tcsem_oracle.c")
    if not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    L = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "orc_create": (P, [C.POINTER(Setup), dp, dp, dp, dp]),
        "orc_destroy": (None, [P]),
        "orc_info": (None, [P, ip]),
        "orc_grid": (None, [P, dp, dp, dp, dp, dp, dp, dp, dp]),
        "orc_maps": (None, [P, ip, ip, ip, ip, ip, ip]),
        "orc_get_field": (C.c_int, [P, C.c_int, dp]),
        "orc_set_field": (C.c_int, [P, C.c_int, dp]),
        "orc_zero_fields": (None, [P]),
        "orc_waveform": (C.c_int, [C.c_int, C.c_double, dp, dp, dp]),
        "orc_search_insert": (C.c_int, [dp, C.c_int, C.c_double]),
        "orc_times": (None, [C.POINTER(RxTime), dp]),
        "orc_spline_ir2r": (None, [P, dp, dp]),
        "orc_fft2": (None, [C.c_int, C.c_int, dp, dp, C.c_int]),
        "orc_upcont_filter": (None, [P, dp, dp, dp]),
        "orc_bilinear": (None, [P, dp, dp, dp, dp]),
        "orc_upcont_plane": (None, [P, dp, dp, dp]),
        "orc_upcont": (None, [P]),
        "orc_spline_ir2r_cpu": (None, [P, dp, dp]),
        "orc_upcont_filter_cpu": (None, [P, dp, dp, dp]),
        "orc_bilinear_cpu": (None, [P, dp, dp, dp, dp]),
        "orc_upcont_plane_cpu": (None, [P, dp, dp, dp]),
        "orc_set_cpu_twins": (None, [P, C.c_int]),
        "orc_update_e": (None, [P, C.c_double, C.c_double]),
        "orc_update_b": (None, [P, C.c_double]),
        "orc_update_bz": (None, [P]),
        "orc_locate_tx": (C.c_int, [P, C.POINTER(Tx), C.POINTER(TxPos)]),
        "orc_add_source": (None, [P, C.POINTER(TxPos), C.c_double]),
        "orc_step": (None, [P, C.c_double, C.c_double, C.c_double, C.POINTER(TxPos), C.c_double]),
        "orc_spline4": (C.c_double, [dp, dp, C.c_double]),
        "orc_rx_stencil": (None, [P, C.c_double, C.c_double, C.c_double, C.c_int, ip, dp]),
        "orc_rx_values": (None, [P, C.c_int, dp, C.c_int, ip, dp]),
        "orc_run": (C.c_int, [P, C.POINTER(RxTime), C.c_int, dp, C.POINTER(Tx), C.c_int, dp]),
        "orc_schedule": (C.c_int, [C.POINTER(Setup), C.c_double, C.POINTER(RxTime), C.c_int,
                                   dp, dp, dp, dp]),
        "orc_mindx": (C.c_double, [P]),
        "orc_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def make_setup(**kw):
    s = Setup()
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def make_rxtime(T0, Tn, Nt, is_log, comps):
    r = RxTime()
    r.T0, r.Tn, r.Nt, r.is_log = T0, Tn, Nt, is_log
    for i in range(6):
        r.is_comp[i] = int(comps[i])
    return r


def make_tx(ort, bipolen, x, y, z, rx_ids, azi=0.0, dip=0.0, bipolen1=0.0, x1=0.0, y1=0.0,
            z1=0.0, azi1=0.0, dip1=0.0, txid=1):
    t = Tx()
    t.id, t.ort, t.bipolen, t.x, t.y, t.z = txid, ort, bipolen, x, y, z
    t.azi_deg, t.dip_deg = azi, dip
    t.bipolen1, t.x1, t.y1, t.z1, t.azi1_deg, t.dip1_deg = bipolen1, x1, y1, z1, azi1, dip1
    ids = np.ascontiguousarray(rx_ids, dtype=np.int32)
    t.nr = len(ids)
    t.rx_ids = _ip(ids)
    t._keep = ids
    return t


class Model:
    """One FP64 oracle model (fields, grid, up-continuation work buffers)."""

    def __init__(self, setup, cond=None, dx=None, dy=None, dz=None):
        self.L_ = lib()
        self.setup = setup
        c = None if cond is None else np.ascontiguousarray(cond, dtype=np.float64)
        self._keep = [c]
        args = []
        for a in (c, dx, dy, dz):
            if a is None:
                args.append(None)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                self._keep.append(a)
                args.append(_dp(a))
        self.h = self.L_.orc_create(C.byref(setup), *args)
        if not self.h:
            raise ValueError("oracle rejected the setup")
        info = (C.c_int * 8)()
        self.L_.orc_info(self.h, info)
        (self.L, self.M, self.N, self.AnLx, self.AnRx, self.Any, self.L1, self.M1) = list(info)

    def __del__(self):
        if getattr(self, "h", None):
            self.L_.orc_destroy(self.h)
            self.h = None

    def shape(self, comp):
        L, M, N = self.L, self.M, self.N
        return {EX: (N + 1, M + 1, L), EY: (N + 1, M, L + 1), EZ: (N, M + 1, L + 1),
                BX: (N, M, L + 1), BY: (N, M + 1, L), BZ: (N + 1, M, L),
                BXAIR: (M, L + 1), BYAIR: (M + 1, L)}[comp]

    def grid(self):
        L, M, N, L1, M1 = self.L, self.M, self.N, self.L1, self.M1
        out = [np.zeros(n) for n in (L, M, N, L, M, N, L1, M1)]
        self.L_.orc_grid(self.h, *[_dp(a) for a in out])
        return dict(zip(("dx", "dy", "dz", "X", "Y", "Z", "Xn", "Yn"), out))

    def maps(self):
        L, M, L1, M1 = self.L, self.M, self.L1, self.M1
        out = [np.zeros(n, dtype=np.int32) for n in (L1, M1, L, M, L + 1, M + 1)]
        self.L_.orc_maps(self.h, *[_ip(a) for a in out])
        return dict(zip(("xN2o", "yN2o", "xo2N", "yo2N", "xBx2N", "yBy2N"), out))

    def get(self, comp):
        a = np.zeros(self.shape(comp))
        self.L_.orc_get_field(self.h, comp, _dp(a))
        return a

    def set(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(comp))
        self.L_.orc_set_field(self.h, comp, _dp(a))

    def zero(self):
        self.L_.orc_zero_fields(self.h)

    def spline_ir2r(self, bz0):
        bz0 = np.ascontiguousarray(bz0, dtype=np.float64)
        out = np.zeros((self.M1, self.L1))
        self.L_.orc_spline_ir2r(self.h, _dp(bz0), _dp(out))
        return out

    def upcont_filter(self, bzair):
        bzair = np.ascontiguousarray(bzair, dtype=np.float64)
        bx, by = np.zeros((self.M1, self.L1)), np.zeros((self.M1, self.L1))
        self.L_.orc_upcont_filter(self.h, _dp(bzair), _dp(bx), _dp(by))
        return bx, by

    def bilinear(self, bxr, byr):
        bxr = np.ascontiguousarray(bxr, dtype=np.float64)
        byr = np.ascontiguousarray(byr, dtype=np.float64)
        bxa, bya = np.zeros((self.M, self.L + 1)), np.zeros((self.M + 1, self.L))
        self.L_.orc_bilinear(self.h, _dp(bxr), _dp(byr), _dp(bxa), _dp(bya))
        return bxa, bya

    def upcont_plane(self, bz0):
        bz0 = np.ascontiguousarray(bz0, dtype=np.float64)
        bxa, bya = np.zeros((self.M, self.L + 1)), np.zeros((self.M + 1, self.L))
        self.L_.orc_upcont_plane(self.h, _dp(bz0), _dp(bxa), _dp(bya))
        return bxa, bya

    def upcont(self):
        self.L_.orc_upcont(self.h)

    # CPU twins of the air boundary (SplineGridConvIR2R_cpu.cpp, UpContinuation.cpp:70-97,
    # BilinearGridConvR2IR.cpp)
    def spline_ir2r_cpu(self, bz0):
        bz0 = np.ascontiguousarray(bz0, dtype=np.float64)
        out = np.zeros((self.M1, self.L1))
        self.L_.orc_spline_ir2r_cpu(self.h, _dp(bz0), _dp(out))
        return out

    def upcont_filter_cpu(self, bzair):
        bzair = np.ascontiguousarray(bzair, dtype=np.float64)
        bx, by = np.zeros((self.M1, self.L1)), np.zeros((self.M1, self.L1))
        self.L_.orc_upcont_filter_cpu(self.h, _dp(bzair), _dp(bx), _dp(by))
        return bx, by

    def bilinear_cpu(self, bxr, byr):
        bxr = np.ascontiguousarray(bxr, dtype=np.float64)
        byr = np.ascontiguousarray(byr, dtype=np.float64)
        bxa, bya = np.zeros((self.M, self.L + 1)), np.zeros((self.M + 1, self.L))
        self.L_.orc_bilinear_cpu(self.h, _dp(bxr), _dp(byr), _dp(bxa), _dp(bya))
        return bxa, bya

    def upcont_plane_cpu(self, bz0):
        bz0 = np.ascontiguousarray(bz0, dtype=np.float64)
        bxa, bya = np.zeros((self.M, self.L + 1)), np.zeros((self.M + 1, self.L))
        self.L_.orc_upcont_plane_cpu(self.h, _dp(bz0), _dp(bxa), _dp(bya))
        return bxa, bya

    def set_cpu_twins(self, on=True):
        """Steps (and run) take their air boundary from the CPU-twin chain."""
        self.L_.orc_set_cpu_twins(self.h, int(on))

    def update_e(self, lam, dt):
        self.L_.orc_update_e(self.h, lam, dt)

    def update_b(self, dtb):
        self.L_.orc_update_b(self.h, dtb)

    def update_bz(self):
        self.L_.orc_update_bz(self.h)

    def locate_tx(self, tx):
        p = TxPos()
        rc = self.L_.orc_locate_tx(self.h, C.byref(tx), C.byref(p))
        if rc != 0:
            raise ValueError("transmitter outside the model (%d)" % rc)
        return p

    def add_source(self, txpos, wave):
        self.L_.orc_add_source(self.h, C.byref(txpos), wave)

    def step(self, lam, dt1, wave, txpos, dtb):
        self.L_.orc_step(self.h, lam, dt1, wave, C.byref(txpos) if txpos is not None else None,
                         dtb)

    def rx_stencil(self, x, y, z, comp):
        """First node index per axis (x, y, z) and the 4 node coordinates per axis of the receiver
        stencil of component comp (EFInterp2Rx.cu:75-88)."""
        s, t = np.zeros(3, dtype=np.int32), np.zeros(12)
        self.L_.orc_rx_stencil(self.h, float(x), float(y), float(z), int(comp), _ip(s), _dp(t))
        return s, t.reshape(3, 4)

    def rx_values(self, rx_xyz, comps):
        """The current fields interpolated in space to the receivers: [nrx][len(comps)]."""
        rx = np.ascontiguousarray(rx_xyz, dtype=np.float64).reshape(-1, 3)
        cs = np.ascontiguousarray(comps, dtype=np.int32)
        out = np.zeros((rx.shape[0], cs.size))
        self.L_.orc_rx_values(self.h, rx.shape[0], _dp(rx), cs.size, _ip(cs), _dp(out))
        return out

    def run(self, rxtime, rx_xyz, tx, max_steps=0):
        rx = np.ascontiguousarray(rx_xyz, dtype=np.float64).reshape(-1, 3)
        ncomp = sum(1 for c in rxtime.is_comp if c)
        efs = np.zeros((tx.nr, ncomp, rxtime.Nt))
        n = self.L_.orc_run(self.h, C.byref(rxtime), rx.shape[0], _dp(rx), C.byref(tx),
                            max_steps, _dp(efs))
        if n < 0:
            raise ValueError("oracle run failed (%d)" % n)
        return efs, n

    def mindx(self):
        return self.L_.orc_mindx(self.h)


def waveform(ns, sigma_t):
    t, dt, w = np.zeros(2 * ns + 1), np.zeros(2 * ns + 2), np.zeros(2 * ns + 1)
    rc = lib().orc_waveform(ns, sigma_t, _dp(t), _dp(dt), _dp(w))
    if rc:
        raise ValueError("bad waveform parameters")
    return t, dt, w


def search_insert(arr, target):
    a = np.ascontiguousarray(arr, dtype=np.float64)
    return lib().orc_search_insert(_dp(a), len(a), float(target))


def times(rxtime):
    out = np.zeros(rxtime.Nt)
    lib().orc_times(C.byref(rxtime), _dp(out))
    return out


def fft2(re, im, sign):
    re = np.ascontiguousarray(re, dtype=np.float64).copy()
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    lib().orc_fft2(re.shape[0], re.shape[1], _dp(re), _dp(im), sign)
    return re, im


def spline4(x, y, t):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_spline4(_dp(x), _dp(y), float(t))


def schedule(setup, mindx, rxtime, cap):
    arrs = [np.zeros(cap) for _ in range(4)]
    n = lib().orc_schedule(C.byref(setup), mindx, C.byref(rxtime), cap, *[_dp(a) for a in arrs])
    return [a[:n] for a in arrs]


def threads():
    return lib().orc_threads()
