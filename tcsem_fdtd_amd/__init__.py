"""tcsem_fdtd_amd -- MI355X-native DuFort-Frankel time stepping for TCSEM_FDTD decks.

Python mirror of the C-ABI in include/tcsem.h (libtcsem.so, gfx950 HIP kernels).  It exposes the
reference's driver surface -- FDTD_setup.in, Cond.bin, Receiver.in, Transmitter.in and the
Tx<id>.bout traces -- and its step API (upward continuation, E update, B update, receiver
sampling) as one context object.  There is no CPU fallback: loading fails loudly when the library
is missing, and every compute call needs a gfx950 device.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtcsem.so")

FP64, FP32 = 64, 32
EX, EY, EZ, BX, BY, BZ, BXAIR, BYAIR = range(8)
COMP_NAMES = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")
ERRORS = {-1: "EINVAL", -2: "EDOMAIN", -3: "EIO", -4: "ENOMEM", -5: "EDEVICE", -6: "ESIZE",
          -7: "ESTATE"}


class TcsemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "E?"), code, msg))
        self.code = code


class Setup(C.Structure):
    """FDTD_setup.in in file order (ModelingSetup.cpp:51-59)."""
    _fields_ = [("sfa1", C.c_double), ("sfa2", C.c_double), ("n_src", C.c_int),
                ("sigma_t", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
                ("z0", C.c_double), ("dxmin", C.c_double), ("dymin", C.c_double),
                ("dzmin", C.c_double), ("arg1", C.c_double), ("arg2", C.c_double),
                ("ALx2", C.c_int), ("ALx1", C.c_int), ("Nx", C.c_int), ("ARx1", C.c_int),
                ("ARx2", C.c_int), ("Ay2", C.c_int), ("Ay1", C.c_int), ("Ny", C.c_int),
                ("Az2", C.c_int), ("Az1", C.c_int), ("Nz", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RxTime(C.Structure):
    """Receiver.in lines 1-3 (LoadRxTxPos.cpp:51-55)."""
    _fields_ = [("T0", C.c_double), ("Tn", C.c_double), ("Nt", C.c_int), ("is_log", C.c_int),
                ("is_comp", C.c_int * 6)]

    @property
    def ncomp(self):
        return sum(1 for c in self.is_comp if c)


class Tx(C.Structure):
    """One Transmitter.in row (LoadRxTxPos.cpp:151-178)."""
    _fields_ = [("id", C.c_int), ("ort", C.c_int), ("bipolen", C.c_double), ("x", C.c_double),
                ("y", C.c_double), ("z", C.c_double), ("azi_deg", C.c_double),
                ("dip_deg", C.c_double), ("bipolen1", C.c_double), ("x1", C.c_double),
                ("y1", C.c_double), ("z1", C.c_double), ("azi1_deg", C.c_double),
                ("dip1_deg", C.c_double), ("nr", C.c_int), ("rx_ids", C.POINTER(C.c_int))]


class GridInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("L", "M", "N", "AnLx_New", "AnRx_New", "Any_New", "L1",
                                       "M1")]


_lib = None


def lib():
    """The loaded libtcsem.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libtcsem.so is not built: run python -m tcsem_fdtd_amd.build "
                          "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "tcsem_version": (C.c_char_p, []),
        "tcsem_last_error": (C.c_char_p, []),
        "tcsem_read_setup": (C.c_int, [C.c_char_p, C.POINTER(Setup)]),
        "tcsem_read_cond": (C.c_int, [C.c_char_p, C.c_size_t, dp]),
        "tcsem_read_receivers": (C.c_int, [C.c_char_p, C.POINTER(RxTime), ip, dp, C.c_int]),
        "tcsem_read_transmitters": (C.c_int, [C.c_char_p, C.POINTER(Tx), C.c_int, ip, ip,
                                              C.c_int, ip]),
        "tcsem_write_bout": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, dp]),
        "tcsem_read_bout": (C.c_int, [C.c_char_p, ip, ip, ip, dp, C.c_size_t]),
        "tcsem_grid_info_of": (C.c_int, [C.POINTER(Setup), C.POINTER(GridInfo)]),
        "tcsem_grid_arrays": (C.c_int, [C.POINTER(Setup)] + [dp] * 8),
        "tcsem_grid_maps": (C.c_int, [C.POINTER(Setup)] + [ip] * 6),
        "tcsem_waveform": (C.c_int, [C.c_int, C.c_double, dp, dp, dp]),
        "tcsem_create": (C.c_int, [C.POINTER(Setup), dp, C.c_int, C.POINTER(P)]),
        "tcsem_destroy": (None, [P]),
        "tcsem_info": (C.c_int, [P, C.POINTER(GridInfo)]),
        "tcsem_set_receivers": (C.c_int, [P, C.POINTER(RxTime), C.c_int, dp]),
        "tcsem_set_transmitters": (C.c_int, [P, C.POINTER(Tx), C.c_int]),
        "tcsem_reserve_batch": (C.c_int, [P, C.c_int]),
        "tcsem_reset_fields": (C.c_int, [P, C.c_int, C.c_int]),
        "tcsem_autotune": (C.c_int, [P, C.c_int, C.c_int, ip, ip, ip, ip, dp]),
        "tcsem_begin": (C.c_int, [P, C.c_int, C.c_int]),
        "tcsem_step": (C.c_int, [P, C.c_int]),
        "tcsem_steps_remaining": (C.c_int, [P]),
        "tcsem_run": (C.c_int, [P, C.c_int, C.c_int, C.c_int, dp, ip]),
        "tcsem_get_traces": (C.c_int, [P, C.c_int, dp]),
        "tcsem_trace_shape": (C.c_int, [P, C.c_int, ip, ip, ip]),
        "tcsem_get_field": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "tcsem_set_field": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "tcsem_upcont_plane": (C.c_int, [P, dp, dp, dp]),
        "tcsem_upcont_plane_noalloc": (C.c_int, [P, P, P, P]),
        "tcsem_timing": (C.c_int, [P, dp, C.POINTER(C.c_long)]),
        "tcsem_enable_timing": (C.c_int, [P, C.c_int]),
        "tcsem_sync": (C.c_int, [P]),
        "tcsem_set_graph": (C.c_int, [P, C.c_int]),
        "tcsem_graph_replays": (C.c_long, [P]),
        "tcsem_sigma_table": (C.c_int, [P]),
        "tcsem_set_pairs": (C.c_int, [P, C.c_int, C.c_int]),
        "tcsem_pair_info": (C.c_int, [P, ip, ip, C.POINTER(C.c_long)]),
        "tcsem_pair_timing": (C.c_int, [P, dp, C.POINTER(C.c_long)]),
        "tcsem_batch_create": (C.c_int, [C.POINTER(P), C.c_int, C.POINTER(P)]),
        "tcsem_batch_autotune": (C.c_int, [P, C.c_int, ip, ip, dp]),
        "tcsem_batch_step": (C.c_int, [P, C.c_int]),
        "tcsem_batch_run": (C.c_int, [P, C.c_int, ip]),
        "tcsem_batch_sync": (C.c_int, [P]),
        "tcsem_batch_destroy": (None, [P]),
        "tcsem_batch_enable_timing": (C.c_int, [P, C.c_int]),
        "tcsem_batch_timing": (C.c_int, [P, dp, C.POINTER(C.c_long)]),
        "tcsem_batch_set_pairs": (C.c_int, [P, C.c_int]),
        "tcsem_batch_pair_info": (C.c_int, [P, ip, ip, C.POINTER(C.c_long)]),
        "tcsem_batch_pair_timing": (C.c_int, [P, dp, C.POINTER(C.c_long)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise TcsemError(rc, lib().tcsem_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


# ---- decks -------------------------------------------------------------------------------------
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
    t._ids = ids
    return t


def read_setup(path):
    s = Setup()
    _check(lib().tcsem_read_setup(os.fsencode(path), C.byref(s)))
    return s


def grid_info(setup):
    gi = GridInfo()
    _check(lib().tcsem_grid_info_of(C.byref(setup), C.byref(gi)))
    return gi


def grid_arrays(setup):
    gi = grid_info(setup)
    out = [np.zeros(n) for n in (gi.L, gi.M, gi.N, gi.L, gi.M, gi.N, gi.L1, gi.M1)]
    _check(lib().tcsem_grid_arrays(C.byref(setup), *[_dp(a) for a in out]))
    return dict(zip(("dx", "dy", "dz", "X", "Y", "Z", "Xn", "Yn"), out))


def grid_maps(setup):
    gi = grid_info(setup)
    out = [np.zeros(n, dtype=np.int32) for n in (gi.L1, gi.M1, gi.L, gi.M, gi.L + 1, gi.M + 1)]
    _check(lib().tcsem_grid_maps(C.byref(setup), *[_ip(a) for a in out]))
    return dict(zip(("xN2o", "yN2o", "xo2N", "yo2N", "xBx2N", "yBy2N"), out))


def waveform(ns, sigma_t):
    t, dt, w = np.zeros(2 * ns + 1), np.zeros(2 * ns + 2), np.zeros(2 * ns + 1)
    _check(lib().tcsem_waveform(ns, sigma_t, _dp(t), _dp(dt), _dp(w)))
    return t, dt, w


def read_cond(path, setup):
    gi = grid_info(setup)
    out = np.zeros((gi.N, gi.M, gi.L))
    _check(lib().tcsem_read_cond(os.fsencode(path), out.size, _dp(out)))
    return out


def read_receivers(path):
    rt, n = RxTime(), C.c_int()
    _check(lib().tcsem_read_receivers(os.fsencode(path), C.byref(rt), C.byref(n), None, 0))
    xyz = np.zeros((max(n.value, 1), 3))
    _check(lib().tcsem_read_receivers(os.fsencode(path), C.byref(rt), C.byref(n), _dp(xyz),
                                      n.value))
    return rt, xyz[:n.value]


def read_transmitters(path):
    ntx, nids = C.c_int(), C.c_int()
    _check(lib().tcsem_read_transmitters(os.fsencode(path), None, 0, C.byref(ntx), None, 0,
                                         C.byref(nids)))
    txs = (Tx * max(ntx.value, 1))()
    ids = np.zeros(max(nids.value, 1), dtype=np.int32)
    _check(lib().tcsem_read_transmitters(os.fsencode(path), txs, ntx.value, C.byref(ntx),
                                         _ip(ids), ids.size, C.byref(nids)))
    txs._ids = ids
    return txs, ntx.value


def write_bout(path, data):
    d = np.ascontiguousarray(data, dtype=np.float64)
    nr, nc, nt = d.shape
    _check(lib().tcsem_write_bout(os.fsencode(path), nr, nc, nt, _dp(d)))


def read_bout(path):
    nr, nc, nt = C.c_int(), C.c_int(), C.c_int()
    _check(lib().tcsem_read_bout(os.fsencode(path), C.byref(nr), C.byref(nc), C.byref(nt), None,
                                 0))
    d = np.zeros((nr.value, nc.value, nt.value))
    _check(lib().tcsem_read_bout(os.fsencode(path), C.byref(nr), C.byref(nc), C.byref(nt),
                                 _dp(d), d.size))
    return d


# ---- device context --------------------------------------------------------------------------
class Context:
    """One model on the GPU: fields of `nbatch` transmitters, the air-plane operators, receivers."""

    def __init__(self, setup, cond, precision=FP64):
        self.L_ = lib()
        self.setup = setup
        self.precision = precision
        c = np.ascontiguousarray(cond, dtype=np.float64)
        h = C.c_void_p()
        _check(self.L_.tcsem_create(C.byref(setup), _dp(c), precision, C.byref(h)))
        self.h = h
        gi = GridInfo()
        _check(self.L_.tcsem_info(self.h, C.byref(gi)))
        self.info = gi
        self.L, self.M, self.N = gi.L, gi.M, gi.N
        self.rt = None
        self.txs = []

    def close(self):
        if getattr(self, "h", None):
            self.L_.tcsem_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def shape(self, comp):
        L, M, N = self.L, self.M, self.N
        return {EX: (N + 1, M + 1, L), EY: (N + 1, M, L + 1), EZ: (N, M + 1, L + 1),
                BX: (N, M, L + 1), BY: (N, M + 1, L), BZ: (N + 1, M, L),
                BXAIR: (M, L + 1), BYAIR: (M + 1, L)}[comp]

    def set_receivers(self, rt, rx_xyz):
        xyz = np.ascontiguousarray(rx_xyz, dtype=np.float64).reshape(-1, 3)
        _check(self.L_.tcsem_set_receivers(self.h, C.byref(rt), xyz.shape[0], _dp(xyz)))
        self.rt = rt
        self.txs = []

    def set_transmitters(self, txs):
        arr = (Tx * max(len(txs), 1))(*txs)
        _check(self.L_.tcsem_set_transmitters(self.h, arr, len(txs)))
        self.txs = list(txs)

    def run(self, tx_first=0, ntx=1, max_steps=0):
        """Per-transmitter runs of main.cu:195-298, batched; returns a list of [nr][NComp][Nt]."""
        ncomp = self.rt.ncomp
        sizes = [self.txs[tx_first + b].nr for b in range(ntx)]
        out = np.zeros(sum(sizes) * ncomp * self.rt.Nt)
        steps = C.c_int()
        _check(self.L_.tcsem_run(self.h, tx_first, ntx, max_steps, _dp(out), C.byref(steps)))
        res, off = [], 0
        for nr in sizes:
            n = nr * ncomp * self.rt.Nt
            res.append(out[off:off + n].reshape(nr, ncomp, self.rt.Nt))
            off += n
        self.steps_done = steps.value
        return res

    def begin(self, tx_first=0, nbatch=1):
        _check(self.L_.tcsem_begin(self.h, tx_first, nbatch))

    def step(self, n=1):
        _check(self.L_.tcsem_step(self.h, n))

    def trace_shape(self, slot=0):
        """(nr, NComp, Nt) of slot `slot` of the begun run, from the library."""
        nr, nc, nt = C.c_int(), C.c_int(), C.c_int()
        _check(self.L_.tcsem_trace_shape(self.h, slot, C.byref(nr), C.byref(nc), C.byref(nt)))
        return nr.value, nc.value, nt.value

    def traces(self, slot=0):
        """Traces of batch slot `slot` of the begun run (its transmitter is tx_first + slot), in a
        buffer sized by the library's own record of that transmitter's receivers."""
        out = np.zeros(self.trace_shape(slot))
        _check(self.L_.tcsem_get_traces(self.h, slot, _dp(out)))
        return out

    def reserve(self, nbatch):
        _check(self.L_.tcsem_reserve_batch(self.h, nbatch))

    def reset(self, first=0, n=1):
        _check(self.L_.tcsem_reset_fields(self.h, first, n))

    def autotune(self, nbatch=1, nsteps=10):
        """Times every march tiling for nbatch-transmitter steps and keeps the fastest (zeroes the
        fields of slots [0, nbatch); a begun run is discarded)."""
        r, p, z, f, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _check(self.L_.tcsem_autotune(self.h, nbatch, nsteps, C.byref(r), C.byref(p), C.byref(z),
                                      C.byref(f), C.byref(ms)))
        pr, ppf, _ = self.pair_info()
        return {"rows": r.value, "pf": p.value, "zchunks": z.value,
                "zsum": "fused" if f.value else "pass", "sigma8": self.sigma_table() > 0,
                "pairs": [pr, ppf] if pr else None, "march_ms": ms.value}

    def get(self, comp, slot=0):
        a = np.zeros(self.shape(comp))
        _check(self.L_.tcsem_get_field(self.h, slot, comp, _dp(a)))
        return a

    def set(self, comp, arr, slot=0):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(comp))
        _check(self.L_.tcsem_set_field(self.h, slot, comp, _dp(a)))

    def upcont_plane(self, bz0):
        b = np.ascontiguousarray(bz0, dtype=np.float64).reshape(self.M, self.L)
        bx, by = np.zeros((self.M, self.L + 1)), np.zeros((self.M + 1, self.L))
        _check(self.L_.tcsem_upcont_plane(self.h, _dp(b), _dp(bx), _dp(by)))
        return bx, by

    def enable_timing(self, on=True):
        _check(self.L_.tcsem_enable_timing(self.h, int(on)))

    def timing(self):
        t, n = (C.c_double * 4)(), (C.c_long * 4)()
        _check(self.L_.tcsem_timing(self.h, t, n))
        return {"march_ms": t[0], "surface_ms": t[1], "upcont_ms": t[2], "gather_ms": t[3],
                "n": list(n)}

    def sync(self):
        _check(self.L_.tcsem_sync(self.h))

    def set_graph(self, nsteps):
        """Replay stretches of nsteps sample-free steady-state steps as HIP graphs (0: off; contexts of
        fewer than 2^21 cells start with 16)."""
        _check(self.L_.tcsem_set_graph(self.h, nsteps))

    def graph_replays(self):
        return self.L_.tcsem_graph_replays(self.h)

    def sigma_table(self):
        """Distinct conductivities (0 included) of the 8-bit index the march reads; 0: full values."""
        return self.L_.tcsem_sigma_table(self.h)

    def set_pairs(self, rows, pf=1):
        """Take two steady-state steps per lower pass (rows 8, 10, 12, 14 or 16 (FP64 8, 10, 11 or 12), prefetch pf 1 or 2; rows 0:
        off), over the library's own choice.  Results equal one step at a time (unchunked), bit for
        bit."""
        _check(self.L_.tcsem_set_pairs(self.h, rows, pf))

    def pair_info(self):
        """(rows, prefetch, pairs taken so far) of the two-step pass."""
        r, p, n = C.c_int(), C.c_int(), C.c_long()
        _check(self.L_.tcsem_pair_info(self.h, C.byref(r), C.byref(p), C.byref(n)))
        return r.value, p.value, n.value

    def pair_timing(self):
        """(average device time [ms] of the two-step pass alone, passes timed)."""
        ms, n = C.c_double(), C.c_long()
        _check(self.L_.tcsem_pair_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


# ---- model batch -----------------------------------------------------------------------------
class Batch:
    """Models of uneven grids advanced together, one launch per kernel for all of them.  Each
    member is a Context with receivers and transmitters set; begin(tx) starts every member's run
    with its transmitter tx[q] (default 0) and creates the batch."""

    def __init__(self, contexts, tx=None):
        self.L_ = lib()
        self.members = list(contexts)
        for q, c in enumerate(self.members):
            c.begin(0 if tx is None else tx[q], 1)
        hs = (C.c_void_p * len(self.members))(*[c.h.value for c in self.members])
        h = C.c_void_p()
        _check(self.L_.tcsem_batch_create(hs, len(self.members), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L_.tcsem_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def autotune(self, nsteps=10):
        r, p, ms = C.c_int(), C.c_int(), C.c_double()
        _check(self.L_.tcsem_batch_autotune(self.h, nsteps, C.byref(r), C.byref(p), C.byref(ms)))
        return {"rows": r.value, "pf": p.value, "pairs": self.pair_info()[1] or None, "step_ms": ms.value}

    def set_pairs(self, rows):
        """Two steps per pass for the batch (rows 8, 10, 12, 14 or 16, FP64 8, 10, 11 or 12; 0: one at a time)."""
        _check(self.L_.tcsem_batch_set_pairs(self.h, rows))

    def pair_info(self):
        """(rows of the last step call's passes, 0 for none; the autotuner's pick, -1 untuned;
        pairs taken so far)."""
        r, t, n = C.c_int(), C.c_int(), C.c_long()
        _check(self.L_.tcsem_batch_pair_info(self.h, C.byref(r), C.byref(t), C.byref(n)))
        return r.value, t.value, n.value

    def pair_timing(self):
        """(average device time [ms] of the two-step pass alone, passes timed)."""
        ms, n = C.c_double(), C.c_long()
        _check(self.L_.tcsem_batch_pair_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def step(self, n=1):
        _check(self.L_.tcsem_batch_step(self.h, n))

    def run(self, max_steps=0):
        """Steps until every member has all samples; returns each member's [nr][NComp][Nt]."""
        steps = C.c_int()
        _check(self.L_.tcsem_batch_run(self.h, max_steps, C.byref(steps)))
        self.steps_done = steps.value
        return [c.traces(0) for c in self.members]

    def sync(self):
        _check(self.L_.tcsem_batch_sync(self.h))

    def enable_timing(self, on=True):
        _check(self.L_.tcsem_batch_enable_timing(self.h, int(on)))

    def timing(self):
        t, n = (C.c_double * 3)(), (C.c_long * 3)()
        _check(self.L_.tcsem_batch_timing(self.h, t, n))
        return {"march_ms": t[0], "surface_ms": t[1], "upcont_ms": t[2], "n": list(n)}
