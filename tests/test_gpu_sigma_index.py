"""The 8-bit conductivity index of the volume march: a model of at most 255 distinct nonzero
conductivities is read as one byte per cell into a table of its values.  Over march tilings
(tile rows, prefetch depth, z-chunks with fused chunk sums or the k_zsum pass), FP64 and FP32, the
traces equal, bit for bit, those of the march that reads the full conductivity, and the FP64
traces equal the oracle (ElectricalFiled.cu:145-172 reads sigma per cell).  A model of more
distinct values keeps the full volume."""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tests import decks

pytestmark = pytest.mark.gpu

TX = dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4])
RX = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])


def _cond(distinct=None):
    """Seawater over sediment with a resistive blob (3 values), or `distinct` values."""
    m = O.Model(decks.small_setup(O))
    g = m.grid()
    c = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), blob=True)
    if distinct:
        vals = np.geomspace(0.05, 5.0, distinct)
        c = vals[np.random.default_rng(3).integers(0, distinct, c.shape)]
    return c


def _rt(mod):
    return mod.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1))


def _run(monkeypatch, precision, tiling, sigma8, cond, steps=300):
    rows, pf, z, zsum = tiling.split(",")
    monkeypatch.setenv("TCSEM_MARCH_ROWS", rows)
    monkeypatch.setenv("TCSEM_MARCH_PF", pf)
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", z)
    monkeypatch.setenv("TCSEM_MARCH_ZSUM", zsum)
    monkeypatch.setenv("TCSEM_SIGMA8", "1" if sigma8 else "0")
    with T.Context(decks.small_setup(T), cond, precision) as c:
        table = c.sigma_table()
        c.set_receivers(_rt(T), RX)
        c.set_transmitters([T.make_tx(**TX)])
        return c.run(0, 1, max_steps=steps)[0], table


@pytest.mark.parametrize("precision,tiling", [
    (T.FP32, "10,1,2,fused"), (T.FP32, "10,3,1,fused"), (T.FP32, "4,2,1,fused"), (T.FP32, "16,1,3,pass"),
    (T.FP64, "12,1,3,fused"), (T.FP64, "8,2,1,fused"), (T.FP64, "4,1,2,pass")])
def test_index_equals_the_full_conductivity(monkeypatch, precision, tiling):
    cond = _cond()
    on, n_on = _run(monkeypatch, precision, tiling, True, cond)
    off, n_off = _run(monkeypatch, precision, tiling, False, cond)
    assert n_on == 4 and n_off == 0  # 0 (the margins), sea, sediment, blob
    assert np.abs(on).max() > 0
    assert np.array_equal(on, off)


def test_index_fp64_equals_the_oracle(monkeypatch):
    cond = _cond()
    got, n = _run(monkeypatch, T.FP64, "12,1,3,fused", True, cond)
    assert n == 4
    ref, _ = O.Model(decks.small_setup(O), cond).run(_rt(O), RX, O.make_tx(**TX), max_steps=300)
    err = decks.trace_error(got, ref)
    assert np.all(err < 1e-9), err


@pytest.mark.parametrize("distinct,expect", [(255, 256), (256, 0)])
def test_index_holds_at_most_255_nonzero_values(monkeypatch, distinct, expect):
    cond = _cond(distinct)
    assert len(np.unique(cond)) == distinct
    got, n = _run(monkeypatch, T.FP64, "8,1,1,fused", True, cond)
    assert n == expect
    ref, _ = O.Model(decks.small_setup(O), cond).run(_rt(O), RX, O.make_tx(**TX), max_steps=300)
    err = decks.trace_error(got, ref)
    assert np.all(err < 1e-9), err


def _batch_run(monkeypatch, precision, sigma8):
    monkeypatch.setenv("TCSEM_SIGMA8", "1" if sigma8 else "0")
    ctxs = []
    try:
        for kw in ({}, {"nx": 23, "ny": 17}, {"nx": 30, "ny": 26, "nz": 10}):
            m = O.Model(decks.small_setup(O, **kw))
            g = m.grid()
            cond = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), blob=True)
            c = T.Context(decks.small_setup(T, **kw), cond, precision)
            assert c.sigma_table() == (4 if sigma8 else 0)
            c.set_receivers(_rt(T), RX)
            c.set_transmitters([T.make_tx(**TX)])
            ctxs.append(c)
        with T.Batch(ctxs) as b:
            return b.run(max_steps=300)
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_model_batch_index_equals_the_full_conductivity(monkeypatch, precision):
    """The model batch's march (one launch over every member's tiles) reads each member's index."""
    on = _batch_run(monkeypatch, precision, True)
    off = _batch_run(monkeypatch, precision, False)
    for q in range(len(on)):
        assert np.abs(on[q]).max() > 0
        assert np.array_equal(on[q], off[q]), q
