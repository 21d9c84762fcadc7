"""The drop-in driver tcsem_fdtd (main.cu behaviour): reads the four reference input files from a
directory and writes Tx<id>.bout per transmitter.  CPU tests cover its argument and input-error
handling; the GPU tests run it on a small seeded deck with two transmitters and compare every
Tx<id>.bout with the FP64 oracle's per-transmitter run (counters restarted per transmitter)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tcsem_fdtd_amd import workloads as W
from tests import decks

DRIVER = os.path.join(os.path.dirname(T.LIB_PATH), "tcsem_fdtd")


def _run(args, **kw):
    return subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=300, **kw)


def _small_deck():
    so = decks.small_setup(O)
    m = O.Model(so)
    g = m.grid()
    cond = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), seed=3,
                              blob=True)
    rx = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
    txs = [dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4]),
           dict(ort=2, bipolen=80.0, x=50.0, y=-25.0, z=320.0, rx_ids=[3, 1])]
    return W.Deck("small", decks.small_setup(T).as_dict(), cond,
                  dict(T0=0.02, Tn=0.25, Nt=16, is_log=1, comps=(1, 1, 1, 1, 1, 1)), rx, txs)


def test_driver_help():
    r = _run(["--help"])
    assert r.returncode == 0 and "Tx<id>.bout" in r.stdout


def test_driver_bad_arguments():
    assert _run(["--precision", "16"]).returncode == 1
    assert _run(["--bogus"]).returncode == 1


def test_driver_missing_input_is_an_io_error():
    with tempfile.TemporaryDirectory() as d:
        r = _run(["--dir", d])
    assert r.returncode == 3 and "FDTD_setup.in" in r.stderr


def test_driver_without_gpu_fails_loudly():
    import ctypes as C
    ndev = C.c_int(0)
    hip = C.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(C.byref(ndev)) == 0 and ndev.value > 0:
        pytest.skip("a GPU is present")
    with tempfile.TemporaryDirectory() as d:
        _small_deck().write(d)
        r = _run(["--dir", d])
        assert r.returncode == 5 and "device setup" in r.stderr
        assert not os.path.exists(os.path.join(d, "Tx1.bout"))


def _oracle_traces(deck):
    so = O.make_setup(**deck.setup)
    m = O.Model(so, deck.cond)
    r = deck.rxtime
    rt = O.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"])
    return [m.run(rt, deck.rx, O.make_tx(**t))[0] for t in deck.txs]


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(64, 1e-9), (32, 1e-4)])
def test_driver_bout_matches_oracle(precision, tol):
    deck = _small_deck()
    ref = _oracle_traces(deck)
    with tempfile.TemporaryDirectory() as d:
        deck.write(d)
        print("driver FP%d on %s" % (precision, d), flush=True)
        r = _run(["--dir", d, "--precision", str(precision), "--batch", "2"])
        assert r.returncode == 0, r.stderr
        for b, t in enumerate(deck.txs):
            got = T.read_bout(os.path.join(d, "Tx%d.bout" % (b + 1)))
            assert got.shape == ref[b].shape
            err = decks.trace_error(got, ref[b])
            assert np.all(err < tol), (b, err)


@pytest.mark.gpu
def test_driver_batch_size_does_not_change_results():
    deck = _small_deck()
    outs = []
    for batch in (1, 2):
        with tempfile.TemporaryDirectory() as d:
            deck.write(d)
            r = _run(["--dir", d, "--batch", str(batch), "--quiet"])
            assert r.returncode == 0, r.stderr
            outs.append([open(os.path.join(d, "Tx%d.bout" % (b + 1)), "rb").read() for b in range(2)])
    assert outs[0] == outs[1]


@pytest.mark.gpu
def test_driver_autotune_keeps_results():
    """--autotune may pick another march tiling (z-chunks reorder the FP64 Bz sums): the traces
    stay at round-off from the heuristic tiling's."""
    deck = _small_deck()
    outs = []
    for extra in ([], ["--autotune"]):
        with tempfile.TemporaryDirectory() as d:
            deck.write(d)
            r = _run(["--dir", d, "--batch", "2"] + extra)
            assert r.returncode == 0, r.stderr
            if extra:
                assert "autotuned march" in r.stdout, r.stdout
            outs.append([T.read_bout(os.path.join(d, "Tx%d.bout" % (b + 1))) for b in range(2)])
    for got, ref in zip(outs[1], outs[0]):
        err = decks.trace_error(got, ref)
        assert np.all(err < 1e-10), err


def _deck_of(kw, ntx, seed):
    so = decks.small_setup(O, **kw)
    m = O.Model(so)
    g = m.grid()
    cond = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), seed=seed,
                              blob=True)
    rx = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
    txs = [dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4]),
           dict(ort=2, bipolen=80.0, x=50.0, y=-25.0, z=320.0, rx_ids=[3, 1])][:ntx]
    return W.Deck("deck%d" % seed, decks.small_setup(T, **kw).as_dict(), cond,
                  dict(T0=0.02, Tn=0.25, Nt=16, is_log=1, comps=(1, 1, 1, 1, 1, 1)), rx, txs)


def test_driver_model_batch_missing_input_is_an_io_error():
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        _small_deck().write(d1)
        r = _run(["--dir", d1, "--dir", d2])
    assert r.returncode == 3 and "FDTD_setup.in" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 1e-5)])
def test_driver_model_batch_equals_separate_runs(precision, tol):
    """Three decks of different grids (2, 1 and 2 transmitters) as one model batch: every
    Tx<id>.bout equals the one the deck's own run writes."""
    dks = [_deck_of({}, 2, 3), _deck_of({"nx": 23, "ny": 17}, 1, 4), _deck_of({"nx": 30, "ny": 26, "nz": 10}, 2, 5)]
    with tempfile.TemporaryDirectory() as root:
        dirs = []
        for q, dk in enumerate(dks):
            for mode in ("batch", "single"):
                p = os.path.join(root, "%s%d" % (mode, q))
                os.makedirs(p)
                dk.write(p)
            dirs.append(q)
        args = []
        for q in dirs:
            args += ["--dir", os.path.join(root, "batch%d" % q)]
        print("driver model batch FP%d" % precision, flush=True)
        r = _run(args + ["--precision", str(precision)])
        assert r.returncode == 0, r.stderr
        for q, dk in enumerate(dks):
            r = _run(["--dir", os.path.join(root, "single%d" % q), "--precision", str(precision), "--quiet"])
            assert r.returncode == 0, r.stderr
            for b in range(len(dk.txs)):
                name = "Tx%d.bout" % (b + 1)
                got = T.read_bout(os.path.join(root, "batch%d" % q, name))
                ref = T.read_bout(os.path.join(root, "single%d" % q, name))
                assert got.shape == ref.shape and np.any(ref != 0)
                err = decks.trace_error(got, ref)
                assert np.all(err < tol), (q, b, err)
