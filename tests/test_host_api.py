"""Host side of libtcsem.so without a GPU: exports, the reference file surface, the host setup
restatements (checked bit-for-bit against the independent oracle restatement), and error paths."""
import ctypes as C
import os
import re
import shutil
import tempfile

import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tcsem_fdtd_amd import workloads as W
from tests import decks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "tcsem.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?(tcsem_\w+)\(", hdr, re.M))
    assert len(declared) >= 30
    lib = C.CDLL(T.LIB_PATH)
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert T.lib().tcsem_version().startswith(b"tcsem")


def test_reference_example_deck_parses():
    s = T.read_setup(os.path.join(GOLDEN, "FDTD_setup.in"))
    assert (s.sfa1, s.sfa2, s.n_src, s.sigma_t) == (2.0, 2.0, 200, 0.001)
    assert (s.ALx2, s.ALx1, s.Nx, s.ARx1, s.ARx2) == (10, 40, 120, 40, 10)
    assert (s.Ay2, s.Ay1, s.Ny, s.Az2, s.Az1, s.Nz) == (10, 40, 100, 10, 40, 60)
    gi = T.grid_info(s)
    assert (gi.L, gi.M, gi.N, gi.L1, gi.M1) == (220, 200, 110, 3914, 3894)
    rt, xyz = T.read_receivers(os.path.join(GOLDEN, "Receiver.in"))
    assert (rt.T0, rt.Tn, rt.Nt, rt.is_log) == (0.01, 10.0, 100, 1)
    assert list(rt.is_comp) == [1, 1, 0, 1, 1, 1] and rt.ncomp == 5
    assert np.array_equal(xyz[:, 0], [1500.0, 2000.0, 2500.0, 3000.0, 3500.0])
    txs, n = T.read_transmitters(os.path.join(GOLDEN, "Transmitter.in"))
    # BipoLen is read before x, y, z (LoadRxTxPos.cpp:156-157), unlike the README's order
    t = txs[0]
    assert n == 1 and (t.id, t.ort, t.bipolen, t.x, t.y, t.z, t.nr) == (1, 1, 200.0, 0.0, 0.0, 600.0, 5)
    assert [t.rx_ids[q] for q in range(t.nr)] == [1, 2, 3, 4, 5]


def test_bout_round_trip_is_byte_exact():
    src = os.path.join(GOLDEN, "Tx1.bout")
    d = T.read_bout(src)
    assert d.shape == (5, 5, 100) and np.all(np.isfinite(d))
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "Tx1.bout")
        T.write_bout(out, d)
        assert open(out, "rb").read() == open(src, "rb").read()


def test_deck_files_round_trip():
    deck = W.marine_1d(T)
    with tempfile.TemporaryDirectory() as tmp:
        deck.write(tmp)
        s = T.read_setup(os.path.join(tmp, "FDTD_setup.in"))
        for k, v in deck.setup.items():
            assert getattr(s, k) == v, k
        assert np.array_equal(T.read_cond(os.path.join(tmp, "Cond.bin"), s), deck.cond)
        rt, xyz = T.read_receivers(os.path.join(tmp, "Receiver.in"))
        assert np.array_equal(xyz, deck.rx) and rt.Nt == deck.rxtime["Nt"]
        txs, n = T.read_transmitters(os.path.join(tmp, "Transmitter.in"))
        assert n == 1 and txs[0].z == deck.txs[0]["z"] and txs[0].nr == len(deck.rx)


@pytest.mark.parametrize("kind", ["small", "bluestein", "marine_1d", "marine_3d", "reference"])
def test_host_setup_equals_oracle(kind):
    """Two independent restatements of ModelingSetup/ModelGridMapping/SourceWave agree exactly."""
    if kind == "small":
        kw = dict()
    elif kind == "bluestein":
        kw = dict(nx=23, ny=17)
    if kind in ("small", "bluestein"):
        st, so = decks.small_setup(T, **kw), decks.small_setup(O, **kw)
    elif kind == "reference":
        st = T.read_setup(os.path.join(GOLDEN, "FDTD_setup.in"))
        so = O.make_setup(**st.as_dict())
    else:
        d = getattr(W, kind)(T)
        st, so = T.make_setup(**d.setup), O.make_setup(**d.setup)
    m = O.Model(so)
    go, gt = m.grid(), T.grid_arrays(st)
    for k in go:
        assert np.array_equal(go[k], gt[k]), k
    mo, mt = m.maps(), T.grid_maps(st)
    for k in mo:
        assert np.array_equal(mo[k], mt[k]), k
    ns = st.n_src // 2
    for a, b in zip(O.waveform(ns, st.sigma_t), T.waveform(ns, st.sigma_t)):
        assert np.array_equal(a, b)


def test_error_paths():
    s = decks.small_setup(T)
    s.ALx1 = s.ALx2 = 0          # no padding: the reference indexes X_Bzold[-1]
    with pytest.raises(T.TcsemError) as e:
        T.grid_info(s)
    assert e.value.code == -1
    with pytest.raises(T.TcsemError) as e:
        T.read_setup("/nonexistent/FDTD_setup.in")
    assert e.value.code == -3
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "Transmitter.in")
        open(p, "w").write("1\n1 7 100.0 0 0 100 1 1\n")
        with pytest.raises(T.TcsemError) as e:
            T.read_transmitters(p)
        assert e.value.code == -1   # illegal orientation (LoadRxTxPos.cpp:179-185)
        p = os.path.join(tmp, "Receiver.in")
        open(p, "w").write("0.01 1.0\n10 1\n1 1 1 1 1 1\n3\n1 0 0 0\n2 1 1 1\n")
        with pytest.raises(T.TcsemError) as e:
            T.read_receivers(p)
        assert e.value.code == -3   # fewer receivers than announced
    # the trace shape of a context that does not exist is an argument error, not a crash
    nr, nc, nt = C.c_int(), C.c_int(), C.c_int()
    assert T.lib().tcsem_trace_shape(None, 0, C.byref(nr), C.byref(nc), C.byref(nt)) == -1


def test_create_without_gpu_fails_loudly():
    """No CPU fallback: without a gfx950 device the context refuses to exist."""
    ndev = C.c_int(0)
    hip = C.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(C.byref(ndev)) == 0 and ndev.value > 0:
        pytest.skip("a GPU is present")
    s = decks.small_setup(T)
    gi = T.grid_info(s)
    with pytest.raises(T.TcsemError) as e:
        T.Context(s, np.ones((gi.N, gi.M, gi.L)))
    assert e.value.code == -5


def test_model_batch_argument_errors():
    """The model-batch entry points reject bad arguments before touching the device."""
    L = T.lib()
    h = C.c_void_p()
    assert L.tcsem_batch_create(None, 1, C.byref(h)) == -1
    arr = (C.c_void_p * 2)(None, None)
    assert L.tcsem_batch_create(arr, 0, C.byref(h)) == -1
    assert L.tcsem_batch_create(arr, 2, C.byref(h)) == -1  # null members
    assert b"null member" in L.tcsem_last_error()
    assert L.tcsem_batch_step(None, 1) == -1
    assert L.tcsem_batch_run(None, 0, None) == -1
    assert L.tcsem_batch_autotune(None, 5, None, None, None) == -1
    assert L.tcsem_batch_sync(None) == -1
    L.tcsem_batch_destroy(None)


def test_graph_argument_errors():
    L = T.lib()
    assert L.tcsem_set_graph(None, 8) == -1
    assert L.tcsem_graph_replays(None) == -1
