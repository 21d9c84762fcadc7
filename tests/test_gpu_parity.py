"""Parity of the gfx950 HIP path (libtcsem.so, through the C-ABI) with the FP64 oracle.

FP64 contexts are compared at round-off level (rel L2 <= 1e-10 per field after a few steps, the
kernels reorder a few FP64 operations); FP32 contexts at the north-star tolerance of 1e-4
relative trace error per receiver (components below 1e-6 of the receiver's largest norm -- the
symmetry nulls -- excluded).
"""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tests import decks

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TRACE_TOL = 1e-4


def _pair(setup_kw=None, cond_seed=7, blob=True):
    kw = setup_kw or {}
    so = decks.small_setup(O, **kw)
    st = decks.small_setup(T, **kw)
    m = O.Model(so)
    g = m.grid()
    cond = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)),
                              seed=cond_seed, blob=blob)
    m = O.Model(so, cond)
    return so, st, m, cond


def _zero_walls(f, comp, L, M, N):
    """Entries the reference never updates (PEC-like walls) are kept at zero."""
    f = f.copy()
    if comp == O.EX:
        f[N:] = 0; f[:, 0] = 0; f[:, M] = 0
    elif comp == O.EY:
        f[N:] = 0; f[:, :, 0] = 0; f[:, :, L] = 0
    elif comp == O.EZ:
        f[:, 0] = 0; f[:, M] = 0; f[:, :, 0] = 0; f[:, :, L] = 0
    elif comp == O.BX:
        f[:, :, 0] = 0; f[:, :, L] = 0
    elif comp == O.BY:
        f[:, 0] = 0; f[:, M] = 0
    return f


def test_upcont_plane_fp64():
    so, st, m, cond = _pair()
    rng = np.random.default_rng(1)
    g = m.grid()
    X, Y = np.meshgrid(g["X"], g["Y"])
    bz0 = np.exp(-(X ** 2 + Y ** 2) / 4e5) * (1 + 0.1 * rng.standard_normal(X.shape))
    bxo, byo = m.upcont_plane(bz0)
    with T.Context(st, cond, T.FP64) as c:
        bx, by = c.upcont_plane(bz0)
    assert decks.rel_l2(bx, bxo) < 1e-12
    assert decks.rel_l2(by, byo) < 1e-12


def test_upcont_plane_bluestein_fp64():
    """A plane whose regular-grid sizes are not powers of two (Bluestein FFT lengths)."""
    so, st, m, cond = _pair({"nx": 23, "ny": 17})
    assert (m.L1 & (m.L1 - 1)) and (m.M1 & (m.M1 - 1))
    rng = np.random.default_rng(2)
    bz0 = rng.standard_normal((m.M, m.L))
    bxo, byo = m.upcont_plane(bz0)
    with T.Context(st, cond, T.FP64) as c:
        bx, by = c.upcont_plane(bz0)
    assert decks.rel_l2(bx, bxo) < 1e-11
    assert decks.rel_l2(by, byo) < 1e-11


# Decks whose regular air planes have power-of-two sides (the Stockham FFT path, not Bluestein):
# 128 x 64 (L = 82, M = 18) and 64 x 128 (L = 20, M = 84), so that a transposed x/y pass shows.
POW2_DECKS = {
    "128x64": {"nx": 72, "ny": 8, "pad": (2, 3), "arg": (1.5, 2.0)},
    "64x128": {"nx": 10, "ny": 74, "pad": (3, 2), "arg": (1.3, 2.0)},
}


@pytest.mark.parametrize("name", sorted(POW2_DECKS))
def test_upcont_plane_pow2_fp64(name):
    so, st, m, cond = _pair(POW2_DECKS[name])
    assert "%dx%d" % (m.L1, m.M1) == name
    rng = np.random.default_rng(3)
    bz0 = rng.standard_normal((m.M, m.L))
    bxo, byo = m.upcont_plane(bz0)
    with T.Context(st, cond, T.FP64) as c:
        bx, by = c.upcont_plane(bz0)
    assert decks.rel_l2(bx, bxo) < 1e-12
    assert decks.rel_l2(by, byo) < 1e-12


def _upcont_check(kw, seed, tol=1e-12):
    so, st, m, cond = _pair(kw)
    bz0 = np.random.default_rng(seed).standard_normal((m.M, m.L))
    bxo, byo = m.upcont_plane(bz0)
    with T.Context(st, cond, T.FP64) as c:
        bx, by = c.upcont_plane(bz0)
    assert decks.rel_l2(bx, bxo) < tol
    assert decks.rel_l2(by, byo) < tol
    return m


@pytest.mark.parametrize("kw", [{"nx": 23, "ny": 17}, POW2_DECKS["128x64"]], ids=["bluestein", "128x64"])
def test_upcont_plane_line_chain_fp64(kw, monkeypatch):
    """The six-kernel line chain, forced where the tiled chain would run."""
    monkeypatch.setenv("TCSEM_UPCONT_CHAIN", "lines")
    _upcont_check(kw, 5, 1e-11)


def test_upcont_plane_long_x_fp64():
    """L1 = 1156: Bluestein working length 4096 along x, beyond the tiled chain (line chain with
    in-place radix-2 lines of 256 threads)."""
    m = _upcont_check({"nx": 1100, "ny": 8, "pad": (2, 3), "arg": (1.5, 2.0)}, 6, 1e-11)
    assert m.L1 == 1156


@pytest.mark.parametrize("kw", [{}, {"nx": 23, "ny": 17}, POW2_DECKS["128x64"]],
                         ids=["stretched", "bluestein", "128x64"])
def test_upcont_plane_fp32(kw):
    """FP32 contexts run the tiled chain in FP32 (twiddles rounded once from FP64): within
    single-precision FFT error of the FP64 oracle."""
    so, st, m, cond = _pair(kw)
    bz0 = np.random.default_rng(8).standard_normal((m.M, m.L))
    bxo, byo = m.upcont_plane(bz0)
    with T.Context(st, cond, T.FP32) as c:
        bx, by = c.upcont_plane(bz0)
    ex, ey = decks.rel_l2(bx, bxo), decks.rel_l2(by, byo)
    print("FP32 air-plane chain vs FP64 oracle: Bx %.2e By %.2e" % (ex, ey))
    assert ex < 2e-5 and ey < 2e-5


def test_upcont_plane_marine_1d_fp64():
    """The 256 x 256 air plane of the 64 x 64 x 48 marine model (three radix-8/4 passes)."""
    from tcsem_fdtd_amd import workloads as W
    d = W.marine_1d(T)
    so = O.make_setup(**d.setup)
    m = O.Model(so, d.cond)
    assert (m.L1, m.M1) == (256, 256)
    g = m.grid()
    X, Y = np.meshgrid(g["X"], g["Y"])
    rng = np.random.default_rng(4)
    bz0 = np.exp(-(X ** 2 + Y ** 2) / 1e6) + 0.01 * rng.standard_normal(X.shape)
    bxo, byo = m.upcont_plane(bz0)
    with T.Context(T.make_setup(**d.setup), d.cond, T.FP64) as c:
        bx, by = c.upcont_plane(bz0)
    assert decks.rel_l2(bx, bxo) < 1e-12
    assert decks.rel_l2(by, byo) < 1e-12


def _random_state(m, seed):
    rng = np.random.default_rng(seed)
    st = {}
    for comp in (O.EX, O.EY, O.EZ, O.BX, O.BY):
        st[comp] = _zero_walls(rng.standard_normal(m.shape(comp)), comp, m.L, m.M, m.N)
    return st


@pytest.mark.parametrize("nsteps", [1, 4])
def test_step_fp64_random_state(nsteps):
    so, st, m, cond = _pair()
    state = _random_state(m, 11)
    for comp, f in state.items():
        m.set(comp, f)
    m.update_bz()
    lam, dte, dtb, amp = O.schedule(so, m.mindx(), O.make_rxtime(1e-3, 1.0, 10, 1, [1] * 6), 100)
    for n in range(nsteps):
        m.step(lam[n], dte[n], amp[n], None, dtb[n])
    with T.Context(st, cond, T.FP64) as c:
        c.begin(0, 1)
        for comp, f in state.items():
            c.set(comp, f)
        c.step(nsteps)
        for comp in (O.EX, O.EY, O.EZ, O.BX, O.BY, O.BZ, O.BXAIR, O.BYAIR):
            err = decks.rel_l2(c.get(comp), m.get(comp))
            assert err < FP64_TOL, (comp, err)


def _deck_run(precision, txs_kw, nsteps=0, setup_kw=None, comps=(1, 1, 1, 1, 1, 1), ntx=None):
    so, st, m, cond = _pair(setup_kw)
    g = m.grid()
    rt_o = O.make_rxtime(0.02, 0.25, 16, 1, comps)
    rt_t = T.make_rxtime(0.02, 0.25, 16, 1, comps)
    rx = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0],
                   [400.0, 0.0, 300.0]])
    ref = []
    for kw in txs_kw:
        efs, steps = m.run(rt_o, rx, O.make_tx(**kw), max_steps=nsteps)
        ref.append(efs)
    with T.Context(st, cond, precision) as c:
        c.set_receivers(rt_t, rx)
        c.set_transmitters([T.make_tx(**kw) for kw in txs_kw])
        got = c.run(0, len(txs_kw) if ntx is None else ntx, max_steps=nsteps)
    return ref, got


TX_X = dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4])


def test_traces_fp64():
    ref, got = _deck_run(T.FP64, [TX_X])
    err = decks.trace_error(got[0], ref[0])
    assert np.all(err < 1e-9), err


def test_traces_fp32():
    ref, got = _deck_run(T.FP32, [TX_X])
    err = decks.trace_error(got[0], ref[0])
    assert np.all(err < FP32_TRACE_TOL), err


@pytest.mark.parametrize("name", sorted(POW2_DECKS))
@pytest.mark.parametrize("precision,tol", [(T.FP64, 1e-9), (T.FP32, FP32_TRACE_TOL)])
def test_traces_pow2_decks(name, precision, tol):
    ref, got = _deck_run(precision, [TX_X], nsteps=400, setup_kw=POW2_DECKS[name])
    err = decks.trace_error(got[0], ref[0])
    assert np.any(np.abs(ref[0]) > 0)
    assert np.all(err < tol), err


@pytest.mark.parametrize("tx", [
    dict(ort=2, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4]),
    dict(ort=3, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[4, 3, 2, 1]),
    dict(ort=4, bipolen=150.0, x=25.0, y=-10.0, z=280.0, azi=30.0, dip=20.0, rx_ids=[1, 2, 3, 4]),
    dict(ort=5, bipolen=150.0, x=25.0, y=-10.0, z=280.0, azi=30.0, dip=20.0, bipolen1=100.0,
         x1=-40.0, y1=20.0, z1=320.0, azi1=-45.0, dip1=10.0, rx_ids=[2, 3]),
])
@pytest.mark.parametrize("precision,tol", [(T.FP64, 1e-9), (T.FP32, FP32_TRACE_TOL)])
def test_traces_orientations(tx, precision, tol):
    ref, got = _deck_run(precision, [tx], nsteps=150)
    err = decks.trace_error(got[0], ref[0])
    assert np.all(err < tol), err


@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_orientation_compositions_through_the_hip_path(precision):
    """The orientation pins of the oracle, on the HIP path: ort 4 at azimuth 0 and dip 0 is ort 1
    bit for bit (zero-current cells are dropped, the x weights are the same numbers); azimuth 90
    is ort 2 and dip 90 is ort 3 up to cos(pi/2) = 6e-17; ort 5 with an identical second bipole is
    ort 4 scaled by twice its length (ElectricalFiled.cu:532-641)."""
    pos = dict(x=10.0, y=10.0, z=310.0, rx_ids=[1, 2, 3, 4])
    txs = [dict(ort=1, bipolen=180.0, **pos), dict(ort=4, bipolen=180.0, azi=0.0, dip=0.0, **pos),
           dict(ort=2, bipolen=180.0, **pos), dict(ort=4, bipolen=180.0, azi=90.0, dip=0.0, **pos),
           dict(ort=3, bipolen=180.0, **pos), dict(ort=4, bipolen=180.0, azi=0.0, dip=90.0, **pos),
           dict(ort=4, bipolen=180.0, azi=30.0, dip=20.0, **pos),
           dict(ort=5, bipolen=180.0, azi=30.0, dip=20.0, bipolen1=180.0, x1=10.0, y1=10.0, z1=310.0,
                azi1=30.0, dip1=20.0, **pos)]
    so, st, m, cond = _pair()
    rt = T.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1))
    rx = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
    with T.Context(st, cond, precision) as c:
        c.set_receivers(rt, rx)
        c.set_transmitters([T.make_tx(txid=i + 1, **kw) for i, kw in enumerate(txs)])
        got = c.run(0, len(txs), max_steps=150)
    assert np.array_equal(got[0], got[1])
    tol = 1e-12 if precision == T.FP64 else 1e-5
    for a, b, f in ((2, 3, 1.0), (4, 5, 1.0), (6, 7, 360.0)):
        err = decks.trace_error(got[b], f * got[a])
        assert np.all(err < tol), (a, b, err)


def test_traces_of_a_run_begun_past_the_first_transmitter():
    """A run begun at tx_first = 1 reads the traces of transmitter 1, sized by its own receiver
    count (2 of the 4 receivers), equal to the oracle's run of that transmitter."""
    so, st, m, cond = _pair()
    rt_o, rt_t = O.make_rxtime(0.02, 0.25, 16, 1, (1,) * 6), T.make_rxtime(0.02, 0.25, 16, 1, (1,) * 6)
    rx = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
    txs = [dict(TX_X, txid=1), dict(TX_X, x=50.0, rx_ids=[3, 2], txid=2)]
    ref, _ = m.run(rt_o, rx, O.make_tx(**txs[1]), max_steps=120)
    with T.Context(st, cond, T.FP64) as c:
        c.set_receivers(rt_t, rx)
        c.set_transmitters([T.make_tx(**kw) for kw in txs])
        c.begin(1, 1)
        c.step(120)
        assert c.trace_shape(0) == (2, 6, 16)
        got = c.traces(0)
    assert np.all(decks.trace_error(got, ref) < 1e-9)


def test_batch_equals_independent_runs_fp64():
    txs = [dict(TX_X, x=x, txid=i + 1) for i, x in enumerate((-100.0, 0.0, 120.0))]
    ref, got = _deck_run(T.FP64, txs, nsteps=200)
    for b in range(3):
        err = decks.trace_error(got[b], ref[b])
        assert np.all(err < 1e-9), (b, err)


@pytest.mark.parametrize("zsum", ["fused", "pass"])
def test_batch_with_z_chunks_equals_independent_runs_fp64(zsum, monkeypatch):
    """Three transmitter slots with a z-chunked march: the chunk sums of B are planes per slot."""
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", "3")
    monkeypatch.setenv("TCSEM_MARCH_ZSUM", zsum)
    txs = [dict(TX_X, x=x, txid=i + 1) for i, x in enumerate((-100.0, 0.0, 120.0))]
    ref, got = _deck_run(T.FP64, txs, nsteps=300)
    for b in range(3):
        err = decks.trace_error(got[b], ref[b])
        assert np.all(err < 1e-9), (b, err)


@pytest.mark.parametrize("zsum", ["fused", "pass"])
@pytest.mark.parametrize("precision,tol", [(T.FP64, 1e-9), (T.FP32, FP32_TRACE_TOL)])
def test_source_pass_through_a_chunked_march(zsum, precision, tol, monkeypatch):
    """The bipole cells of every source step are patched by the pass after the march, here through
    a lower march cut into three z-chunks of six levels (19-14, 13-8, 7-2).  "fused": each step's
    chunks start from the chunk sums of B the previous step wrote, which the source pass moves
    along with the B cells it recomputes; "pass": from a k_zsum pass over the fields.  The sources
    sit in every chunk: an x-dipole at level 9-10, a vertical bipole across the boundary of the two
    deepest chunks, an ort-5 pair with one bipole in the middle chunk and one in the shallowest;
    the 3-transmitter batch equals the oracle's independent runs (ElectricalFiled.cu:347-641)."""
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", "3")
    monkeypatch.setenv("TCSEM_MARCH_ZSUM", zsum)
    txs = [dict(TX_X, x=-100.0, z=500.0, txid=1),
           dict(ort=3, bipolen=100.0, x=0.0, y=0.0, z=700.0, rx_ids=[4, 3, 2, 1], txid=2),
           dict(ort=5, bipolen=150.0, x=25.0, y=-10.0, z=560.0, azi=30.0, dip=20.0, bipolen1=100.0,
                x1=-40.0, y1=20.0, z1=320.0, azi1=-45.0, dip1=10.0, rx_ids=[2, 3], txid=3)]
    ref, got = _deck_run(precision, txs, nsteps=250)
    for b in range(3):
        err = decks.trace_error(got[b], ref[b])
        assert np.all(err < tol), (b, err)


def test_batch_graph_replays_equal_plain_steps():
    """Graph replays with three transmitter slots: the same traces as the plain steps."""
    txs = [dict(TX_X, x=x, txid=i + 1) for i, x in enumerate((-100.0, 0.0, 120.0))]
    so, st, m, cond = _pair()
    rt = T.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1))
    rx = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
    out = []
    for k in (0, 8):
        with T.Context(st, cond, T.FP32) as c:
            c.set_receivers(rt, rx)
            c.set_transmitters([T.make_tx(**kw) for kw in txs])
            c.set_graph(k)
            out.append(c.run(0, 3))
            if k:
                assert c.graph_replays() > 0
    for b in range(3):
        assert np.array_equal(out[0][b], out[1][b]), b


def test_upcont_plane_long_x_fp32():
    """The line chain in an FP32 context (FP64 arithmetic, FP32 planes out)."""
    so, st, m, cond = _pair({"nx": 1100, "ny": 8, "pad": (2, 3), "arg": (1.5, 2.0)})
    assert m.L1 == 1156
    bz0 = np.random.default_rng(9).standard_normal((m.M, m.L))
    bxo, byo = m.upcont_plane(bz0)
    with T.Context(st, cond, T.FP32) as c:
        bx, by = c.upcont_plane(bz0)
    assert decks.rel_l2(bx, bxo) < 2e-6 and decks.rel_l2(by, byo) < 2e-6
