"""Full-length accuracy of the FP32 path: the 256 x 256 x 160 marine model (seeded bathymetry and
resistive slab, 101 seafloor receivers, samples from 10 ms to 50 s, about 19 000 DuFort-Frankel
steps) run to its last sample with FP32 fields and with FP64 fields.  The FP64 path itself matches
the FP64 oracle at round-off (test_gpu_parity.py); here the FP32 traces are held to the 1e-4
relative trace error per receiver (components below 1e-6 of a receiver's largest norm, the
symmetry nulls of an inline x-dipole, excluded)."""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from tcsem_fdtd_amd import workloads as W
from tests import decks

pytestmark = pytest.mark.gpu


def _run(deck, precision, graph=None, replays=None):
    with T.Context(T.make_setup(**deck.setup), deck.cond, precision) as c:
        if graph is not None:
            c.set_graph(graph)
        r = deck.rxtime
        c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), deck.rx)
        c.set_transmitters([T.make_tx(**t) for t in deck.txs])
        tr = c.run(0, 1)
        if replays is not None:
            replays.append(c.graph_replays())
        return tr[0], c.steps_done


def test_marine_3d_full_run_fp32_vs_fp64():
    deck = W.marine_3d(T)
    print("marine_3d full run FP64", flush=True)
    ref, n64 = _run(deck, T.FP64)
    print("marine_3d full run FP32 (%d steps)" % n64, flush=True)
    got, n32 = _run(deck, T.FP32)
    assert n32 == n64 > 15000
    err = decks.trace_error(got, ref)
    print("steps %d, per-receiver FP32 trace error: max %.3e median %.3e" % (n32, err.max(), np.median(err)),
          flush=True)
    assert np.all(np.isfinite(got)) and np.all(err < 1e-4), err


@pytest.mark.parametrize("graph", [None, 0], ids=["graph-replays", "plain-steps"])
def test_marine_1d_full_run_vs_oracle(graph):
    """The 64 x 64 x 48 marine model (1000 m of seawater over sediment, 21 seafloor receivers) run
    to its last sample (3755 steps) on the HIP path in FP64 and FP32 and in the FP64 oracle: with
    the library's default for this grid (stretches of 16 sample-free steps after the source as
    replays of a captured HIP graph) and with plain steps (main.cu:219-273 launches every kernel
    of every step)."""
    from oracle import oracle as O
    deck = W.marine_1d(T)
    so = O.make_setup(**deck.setup)
    r = deck.rxtime
    m = O.Model(so, deck.cond)
    ref, nref = m.run(O.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), deck.rx,
                      O.make_tx(**deck.txs[0]))
    for precision, tol in ((T.FP64, 1e-9), (T.FP32, 1e-4)):
        replays = []
        got, n = _run(deck, precision, graph, replays)
        assert n == nref
        assert (replays[0] > 0) == (graph is None), replays
        err = decks.trace_error(got, ref)
        print("marine_1d FP%d vs oracle over %d steps: max %.3e median %.3e" % (precision, n, err.max(),
                                                                             np.median(err)), flush=True)
        assert np.all(np.isfinite(got)) and np.all(err < tol), err
