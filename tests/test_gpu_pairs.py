"""Two steps per lower pass (tcsem_set_pairs, k_march2): step B runs two levels behind step A in
one bottom-up pass and the intermediate state below level 2 never reaches memory.  Every update
is the one-step march's expression in the same operation order, so a run with pairs gives the
traces and the final fields of the run that takes one step at a time, bit for bit: over both
precisions, both workgroup heights, both prefetch depths, with and without the 8-bit
conductivity index, several transmitters at once, grids whose sides are not multiples of the
tile, and with the plain steps of the source wavelet and of the sampling events in between.
The one-step march it is compared with runs unchunked (one bottom-up Bz sum per column, as the
pass sums): a z-chunked march starts its upper chunks from partial sums and rounds differently.
The FP64 traces also equal the oracle (main.cu:219-273 steps one at a time)."""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tests import decks

pytestmark = pytest.mark.gpu

RX = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
TXS = [dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4]),
       dict(ort=2, bipolen=100.0, x=-150.0, y=50.0, z=300.0, rx_ids=[1, 2, 3, 4]),
       dict(ort=1, bipolen=100.0, x=200.0, y=-100.0, z=300.0, rx_ids=[1, 2, 3, 4])]
COMPS = (T.EX, T.EY, T.EZ, T.BX, T.BY, T.BXAIR, T.BYAIR)
BIG = dict(nx=100, ny=44, nz=16)  # 2 x 7 tiles of 12 rows, sides not multiples of the tile


@pytest.fixture(autouse=True)
def _unchunked(monkeypatch):
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", "1")


def _cond(kw):
    m = O.Model(decks.small_setup(O, **kw))
    g = m.grid()
    return decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), seed=5,
                              blob=True)


def _run(precision, pairs, kw=BIG, ntx=1, graph=0, steps=400):
    cond = _cond(kw)
    with T.Context(decks.small_setup(T, **kw), cond, precision) as c:
        c.set_receivers(T.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1)), RX)
        c.set_transmitters([T.make_tx(**t) for t in TXS[:ntx]])
        c.set_graph(graph)
        if pairs:
            c.set_pairs(*pairs)
        tr = c.run(0, ntx, max_steps=steps)
        fields = {(s, comp): c.get(comp, s) for s in range(ntx) for comp in COMPS}
        return tr, fields, c.steps_done, c.pair_info()[2]


def _same(a, b):
    ta, fa, sa, _ = a
    tb, fb, sb, _ = b
    assert sa == sb
    for q in range(len(ta)):
        assert np.abs(ta[q]).max() > 0
        assert np.array_equal(ta[q], tb[q]), q
    for key in fa:
        assert np.array_equal(fa[key], fb[key]), key


@pytest.mark.parametrize("precision,rows,pf,sigma8", [
    (T.FP32, 12, 1, True), (T.FP32, 16, 2, True), (T.FP32, 12, 2, False), (T.FP32, 16, 1, False),
    (T.FP32, 14, 2, True), (T.FP32, 14, 1, False),
    (T.FP64, 12, 1, True), (T.FP64, 8, 2, False), (T.FP64, 11, 1, True), (T.FP64, 11, 2, False),
    (T.FP32, 10, 2, False), (T.FP64, 10, 1, True)])
def test_pairs_equal_single_steps(monkeypatch, precision, rows, pf, sigma8):
    monkeypatch.setenv("TCSEM_SIGMA8", "1" if sigma8 else "0")
    ref = _run(precision, None)
    got = _run(precision, (rows, pf))
    print("steps", got[2], "pairs", got[3])
    assert got[3] > 50
    _same(got, ref)


def test_pairs_with_several_transmitters():
    ref = _run(T.FP32, None, ntx=3, steps=300)
    got = _run(T.FP32, (12, 1), ntx=3, steps=300)
    assert got[3] > 50
    _same(got, ref)


@pytest.mark.parametrize("kw", [dict(), dict(nx=23, ny=17, nz=10), dict(nx=130, ny=9, nz=8)])
def test_pairs_on_other_grids(kw):
    """One tile, ragged tiles, and a grid of few levels (N small) and rows."""
    ref = _run(T.FP64, None, kw=kw, steps=300)
    got = _run(T.FP64, (12, 2), kw=kw, steps=300)
    assert got[3] > 50
    _same(got, ref)


def test_graph_replays_take_precedence():
    ref = _run(T.FP32, None, graph=8)
    got = _run(T.FP32, (12, 1), graph=8)
    _same(got, ref)


def test_pairs_fp64_equal_the_oracle():
    kw = dict()
    so = decks.small_setup(O, **kw)
    cond = _cond(kw)
    ref, _ = O.Model(so, cond).run(O.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1)), RX,
                                   O.make_tx(**TXS[0]), max_steps=400)
    got = _run(T.FP64, (12, 1), kw=kw)
    assert got[3] > 50
    err = decks.trace_error(got[0][0], ref)
    assert np.all(err < 1e-9), err


def test_pair_settings_are_checked():
    with T.Context(decks.small_setup(T), _cond(dict()), T.FP32) as c:
        for rows, pf in ((9, 1), (12, 3), (6, 1), (4, 1), (11, 1)):
            with pytest.raises(T.TcsemError):
                c.set_pairs(rows, pf)
        c.set_pairs(16, 2)
        assert c.pair_info() == (16, 2, 0)
        c.set_pairs(0, 1)
        assert c.pair_info()[0] == 0
    with T.Context(decks.small_setup(T), _cond(dict()), T.FP64) as c:
        for rows in (14, 16):  # FP64 passes of 14 or 16 rows would spill their registers
            with pytest.raises(T.TcsemError):
                c.set_pairs(rows, 1)
        c.set_pairs(8, 2)
        assert c.pair_info()[:2] == (8, 2)


@pytest.mark.parametrize("precision,rows", [(T.FP32, 14), (T.FP64, 11)])
def test_default_pass_height_follows_wave_quantisation(monkeypatch, precision, rows):
    """Without a setting or the autotuner, a 512 x 512 plane takes 14-row FP32 (11-row FP64) passes:
    on MI355X's 256 CUs, at one workgroup per CU, its 468 (666) tiles run in 2 (3) rounds, where
    the 576 tiles of 12 rows need 3; a plane too small for two workgroups per CU takes none."""
    monkeypatch.delenv("TCSEM_MARCH_PAIRS", raising=False)
    for kw, want in ((dict(nx=500, ny=500, nz=8), rows), (dict(), 0)):
        with T.Context(decks.small_setup(T, **kw), _cond(kw), precision) as c:
            c.set_receivers(T.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1)), RX)
            c.set_transmitters([T.make_tx(**TXS[0])])
            c.begin(0, 1)
            assert c.pair_info()[0] == want, (c.L, c.M)
            if want:
                assert (c.L, c.M) == (512, 512)
                c.step(260)  # past the source wavelet: pairs are taken
                assert c.pair_info()[2] > 0
