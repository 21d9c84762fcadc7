"""Graph replays (tcsem_set_graph): stretches of sample-free steady-state steps run as replays of
captured HIP graphs give the traces and fields of the plain steps, replay after replay, with the
plain steps of the sampling events and the host synchronisations of the run in between.  The
replays read the same step scalars from device memory that the plain launches take as
arguments, so the comparison is exact."""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tests import decks

pytestmark = pytest.mark.gpu

RX = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
TX = dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4])


def _run(precision, graph, **kw):
    so = decks.small_setup(O)
    m = O.Model(so)
    g = m.grid()
    cond = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), seed=5,
                              blob=True)
    with T.Context(decks.small_setup(T), cond, precision) as c:
        c.set_receivers(T.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1)), RX)
        c.set_transmitters([T.make_tx(**TX)])
        c.set_graph(graph)
        tr = c.run(0, 1, **kw)[0]
        fields = {comp: c.get(comp) for comp in (T.EX, T.EY, T.EZ, T.BX, T.BY, T.BXAIR, T.BYAIR)}
        return tr, fields, c.steps_done, c.graph_replays()


@pytest.mark.parametrize("precision,k", [(T.FP64, 8), (T.FP32, 4), (T.FP64, 32)])
def test_graph_run_equals_plain_run(precision, k):
    ref, fref, sref, _ = _run(precision, 0)
    got, fgot, sgot, nrep = _run(precision, k)
    print("steps", sgot, "replays", nrep)
    assert sgot == sref
    assert nrep > 0
    assert np.array_equal(got, ref)
    for comp in fref:
        assert np.array_equal(fgot[comp], fref[comp]), comp


@pytest.mark.parametrize("zsum", ["fused", "pass"])
def test_graph_run_with_z_chunks(zsum, monkeypatch):
    """Chunked marches inside the replays: the parities of the chunk sums and of the state."""
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", "2")
    monkeypatch.setenv("TCSEM_MARCH_ZSUM", zsum)
    ref, fref, _, _ = _run(T.FP64, 0, max_steps=700)
    got, fgot, _, nrep = _run(T.FP64, 6, max_steps=700)
    assert nrep > 0
    assert np.array_equal(got, ref)
    for comp in fref:
        assert np.array_equal(fgot[comp], fref[comp]), comp
