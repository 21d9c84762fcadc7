"""The benchmark configurations at their real sizes on the HIP path (FP64 and FP32), each over a
window of steps that fits a test:

* large_5, 512 x 512 x 256: three FP64 steps from a random state equal the FP64 oracle in every
  field; FP32 against FP64 traces of all 401 receivers over a complete run.
* towline of 32 x-dipoles on the 256 x 256 x 160 grid, batched: every slot's FP64 traces equal the
  same transmitter run alone (main.cu:195-298 once per transmitter; the reference's counters are
  not reset between transmitters, so its multi-transmitter parity is N independent runs), and one
  slot equals the oracle.
* 64 conductivity models of uneven grids (96^3 .. 224 x 224 x 160) as one model batch: every
  member's FP64 traces equal the member run alone with the same march tiling, and two members
  equal the oracle.

The FP64 checks sample early windows (the first ms after the source) so that they close within
the steps run here; against the oracle they count the receivers whose traces stand above the
round-off floor of that window.  FP32 is compared with FP64 over a complete run, where every
receiver's trace is above it.
"""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tcsem_fdtd_amd import workloads as W
from tests import decks

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _zero_walls(f, comp, L, M, N):
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


def _rt(mod, r):
    return mod.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"])


def _above_floor(ref, decades=25):
    """Receivers whose trace norm is within `decades` of the strongest receiver's.  In an early
    window the far receivers' traces are 60 decades below the near ones, at the round-off floor of
    the run: two FP64 builds of the oracle (with and without fused multiply-add) differ there by
    20 %, while every receiver above this floor agrees to 1e-15."""
    norms = np.array([np.linalg.norm(ref[q]) for q in range(ref.shape[0])])
    keep = norms >= 10.0 ** -decades * norms.max()
    assert keep.sum() >= 3
    return keep


def test_large_5_fp64_steps_from_a_random_state_equal_the_oracle():
    d = W.large_5(T)
    so, st = O.make_setup(**d.setup), T.make_setup(**d.setup)
    m = O.Model(so, d.cond)
    _log("large_5 %dx%dx%d: oracle state" % (m.L, m.M, m.N))
    rng = np.random.default_rng(51)
    state = {comp: _zero_walls(rng.standard_normal(m.shape(comp)), comp, m.L, m.M, m.N)
             for comp in (O.EX, O.EY, O.EZ, O.BX, O.BY)}
    for comp, f in state.items():
        m.set(comp, f)
    m.update_bz()
    lam, dte, dtb, amp = O.schedule(so, m.mindx(), O.make_rxtime(1e-3, 1.0, 10, 1, [1] * 6), 10)
    for n in range(3):
        m.step(lam[n], dte[n], amp[n], None, dtb[n])
    _log("large_5: HIP steps")
    with T.Context(st, d.cond, T.FP64) as c:
        c.begin(0, 1)
        for comp, f in state.items():
            c.set(comp, f)
        del state
        c.step(3)
        for comp in (O.EX, O.EY, O.EZ, O.BX, O.BY, O.BZ, O.BXAIR, O.BYAIR):
            err = decks.rel_l2(c.get(comp), m.get(comp))
            _log("  field %d: rel L2 %.2e" % (comp, err))
            assert err < 1e-10, (comp, err)


def test_large_5_fp32_traces_equal_fp64():
    """Complete runs (to the last sample at 50 s, about 19 000 steps): the receivers up to 10 km from
    the source peak at 5-20 s, so a shorter window would compare traces that are still below the
    round-off floor of the near field (both precisions then record scheme noise, not signal)."""
    d = W.large_5(T)
    out = {}
    for prec in (T.FP64, T.FP32):
        with T.Context(T.make_setup(**d.setup), d.cond, prec) as c:
            c.set_receivers(_rt(T, d.rxtime), d.rx)
            c.set_transmitters([T.make_tx(**t) for t in d.txs])
            out[prec] = c.run(1, 1)[0]
            _log("large_5 FP%d: %d steps" % (prec, c.steps_done))
    err = decks.trace_error(out[T.FP32], out[T.FP64])
    _log("large_5 FP32 vs FP64, %d receivers: max %.2e median %.2e" % (err.size, err.max(), np.median(err)))
    assert np.abs(out[T.FP64]).max() > 0 and np.all(err < 1e-4), err.max()


def test_towline_32_batch_equals_independent_runs_and_the_oracle(monkeypatch):
    # one march tiling for the batch and the single runs (no z-chunks), so the FP64 results are
    # the same operations in the same order
    monkeypatch.setenv("TCSEM_MARCH_ROWS", "12")
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", "1")
    d = W.towline(T, ntx=32, nt=20, tn=0.008)
    d.rxtime = dict(d.rxtime, T0=0.002)
    txs = [T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(d.txs)]
    with T.Context(T.make_setup(**d.setup), d.cond, T.FP64) as c:
        c.set_receivers(_rt(T, d.rxtime), d.rx)
        c.set_transmitters(txs)
        batch = c.run(0, 32)
        nsteps = c.steps_done
        _log("towline: batch of 32, %d steps" % nsteps)
        for q in range(32):
            single = c.run(q, 1)[0]
            err = decks.trace_error(batch[q], single)
            assert np.abs(single).max() > 0
            assert np.array_equal(batch[q], single) or np.all(err < 1e-13), (q, err.max())
    m = O.Model(O.make_setup(**d.setup), d.cond)
    ref, n = m.run(_rt(O, d.rxtime), d.rx, O.make_tx(**d.txs[7]))
    assert n == nsteps
    keep = _above_floor(ref)
    err = decks.trace_error(batch[7][keep], ref[keep])
    _log("towline slot 7 vs oracle, %d receivers above the floor: max %.2e" % (keep.sum(), err.max()))
    assert np.all(err < 1e-9), err.max()


def _model_contexts(models, precision):
    ctxs = []
    for dm in models:
        c = T.Context(T.make_setup(**dm.setup), dm.cond, precision)
        c.set_receivers(_rt(T, dm.rxtime), dm.rx)
        c.set_transmitters([T.make_tx(**t) for t in dm.txs])
        ctxs.append(c)
    return ctxs


def test_uneven_64_model_batch_equals_single_runs_and_the_oracle(monkeypatch):
    # the batch's march tiling (8 rows, prefetch 1) for the single runs too, without z-chunks
    monkeypatch.setenv("TCSEM_MARCH_ROWS", "8")
    monkeypatch.setenv("TCSEM_MARCH_PF", "1")
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", "1")
    models = W.uneven_models(T, nt=12, tn=0.008)
    for dm in models:
        dm.rxtime = dict(dm.rxtime, T0=0.002)
    ctxs = _model_contexts(models, T.FP64)
    try:
        with T.Batch(ctxs) as b:
            res = b.run()
            _log("uneven models FP64: batch of %d, %d steps" % (len(ctxs), b.steps_done))
        for q, c in enumerate(ctxs):
            single = c.run(0, 1)[0]
            assert np.abs(single).max() > 0
            err = decks.trace_error(res[q], single)
            assert np.array_equal(res[q], single) or np.all(err < 1e-13), (q, err.max())
    finally:
        for c in ctxs:
            c.close()
    sizes = [dm.cond.size for dm in models]
    for q in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        dm = models[q]
        m = O.Model(O.make_setup(**dm.setup), dm.cond)
        ref, _ = m.run(_rt(O, dm.rxtime), dm.rx, O.make_tx(**dm.txs[0]))
        keep = _above_floor(ref)
        err = decks.trace_error(res[q][keep], ref[keep])
        _log("uneven model %d (%s) vs oracle, %d receivers above the floor: max %.2e"
             % (q, dm.name, keep.sum(), err.max()))
        assert np.all(err < 1e-9), (q, err.max())
