"""Model batch through the C ABI (tcsem_batch_*): conductivity models of uneven grids advanced
together, one launch per kernel for all of them, give each model's own traces.

* FP64: every member equals the same model run alone (tcsem_run) to round-off (the single runs
  may split the march into z-chunks, which reorders the FP64 Bz sums), and member 0 matches the
  FP64 oracle;
* FP32: every member within 1e-5 of the same model run alone in FP32;
* members whose sampling ends at different times (a shorter receiver window) are dropped from
  later launches and keep complete traces;
* the batch autotuner leaves the results unchanged.
"""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tests import decks

pytestmark = pytest.mark.gpu

# four grids: 36x32x20 (stretched), 35x29x20 (Bluestein air plane), 82x18x19 (128 x 64 air
# plane), 42x38x16
MODELS = [{}, {"nx": 23, "ny": 17}, {"nx": 72, "ny": 8, "pad": (2, 3), "arg": (1.5, 2.0)},
          {"nx": 30, "ny": 26, "nz": 10}]
RX = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
TX = dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4])


def _cond(kw, seed):
    m = O.Model(decks.small_setup(O, **kw))
    g = m.grid()
    return decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), seed=seed,
                              blob=True)


def _rt(mod, tn=0.25):
    return mod.make_rxtime(0.02, tn, 16, 1, (1, 1, 1, 1, 1, 1))


def _ctx(kw, cond, precision, tn=0.25):
    c = T.Context(decks.small_setup(T, **kw), cond, precision)
    c.set_receivers(_rt(T, tn), RX)
    c.set_transmitters([T.make_tx(**TX)])
    return c


def _singles(conds, precision, tns, max_steps=0):
    out = []
    for kw, cond, tn in zip(MODELS, conds, tns):
        with _ctx(kw, cond, precision, tn) as c:
            out.append(c.run(0, 1, max_steps=max_steps)[0])
    return out


def _batch(conds, precision, tns, max_steps=0, tune=False):
    ctxs = [_ctx(kw, cond, precision, tn) for kw, cond, tn in zip(MODELS, conds, tns)]
    try:
        with T.Batch(ctxs) as b:
            if tune:
                print("batch tiling:", b.autotune(nsteps=3))
            return b.run(max_steps=max_steps), b.steps_done
    finally:
        for c in ctxs:
            c.close()


CONDS = [_cond(kw, 11 + q) for q, kw in enumerate(MODELS)]


def test_batch_fp64_equals_single_models_and_the_oracle():
    tns = [0.25] * len(MODELS)
    ref = _singles(CONDS, T.FP64, tns, max_steps=400)
    got, steps = _batch(CONDS, T.FP64, tns, max_steps=400)
    assert steps == 400
    for q in range(len(MODELS)):
        assert np.any(np.abs(ref[q]) > 0)
        err = decks.trace_error(got[q], ref[q])
        print("member", q, "FP64 batch vs single:", err.max())
        assert np.all(err < 1e-11), (q, err)
    m = O.Model(decks.small_setup(O, **MODELS[0]), CONDS[0])
    efs, _ = m.run(_rt(O), RX, O.make_tx(**TX), max_steps=400)
    err = decks.trace_error(got[0], efs)
    assert np.all(err < 1e-9), err


def test_batch_fp32_equals_single_models():
    tns = [0.25] * len(MODELS)
    ref = _singles(CONDS, T.FP32, tns, max_steps=600)
    got, _ = _batch(CONDS, T.FP32, tns, max_steps=600)
    for q in range(len(MODELS)):
        err = decks.trace_error(got[q], ref[q])
        print("member", q, "FP32 batch vs single:", err.max())
        assert np.all(err < 1e-5), (q, err)


_ORACLE = {}


def _oracle(q, max_steps):
    if (q, max_steps) not in _ORACLE:
        m = O.Model(decks.small_setup(O, **MODELS[q]), CONDS[q])
        _ORACLE[q, max_steps] = m.run(_rt(O), RX, O.make_tx(**TX), max_steps=max_steps)[0]
    return _ORACLE[q, max_steps]


@pytest.mark.parametrize("precision,rows,pf,tol", [(T.FP64, 12, 1, 1e-9), (T.FP64, 4, 2, 1e-9),
                                                   (T.FP32, 10, 1, 1e-4), (T.FP32, 16, 2, 1e-4)])
def test_batch_tiling_matches_the_oracle(precision, rows, pf, tol, monkeypatch):
    """The model-batch march (one launch over every member's tiles) at tile heights other than its
    default, each member against the oracle (ElectricalFiled.cu:145-172, Bderivate.cu:79-255)."""
    monkeypatch.setenv("TCSEM_MARCH_ROWS", str(rows))
    monkeypatch.setenv("TCSEM_MARCH_PF", str(pf))
    got, steps = _batch(CONDS, precision, [0.25] * len(MODELS), max_steps=400)
    assert steps == 400
    for q in range(len(MODELS)):
        err = decks.trace_error(got[q], _oracle(q, 400))
        print("member", q, "rows", rows, "FP%d batch vs oracle:" % precision, err.max())
        assert np.all(err < tol), (q, err)


def test_batch_members_finishing_at_different_times():
    """Full runs; members 1 and 3 sample only to 0.08 s, so they finish first and are left out of
    the later launches (their records return at once)."""
    tns = [0.25, 0.08, 0.25, 0.08]
    ref = _singles(CONDS, T.FP64, tns)
    got, steps = _batch(CONDS, T.FP64, tns)
    for q in range(len(MODELS)):
        assert got[q].shape == ref[q].shape
        err = decks.trace_error(got[q], ref[q])
        assert np.all(err < 1e-11), (q, err)


def test_batch_autotune_keeps_results():
    tns = [0.25] * len(MODELS)
    ref, _ = _batch(CONDS, T.FP32, tns, max_steps=300)
    got, _ = _batch(CONDS, T.FP32, tns, max_steps=300, tune=True)
    for q in range(len(MODELS)):
        err = decks.trace_error(got[q], ref[q])
        assert np.all(err < 1e-5), (q, err)


def test_batch_refuses_inconsistent_members():
    """Members of different precisions, a member listed twice and a member that was not begun
    are refused; the members stay usable."""
    import ctypes as C
    a = _ctx(MODELS[0], CONDS[0], T.FP64)
    b = _ctx(MODELS[1], CONDS[1], T.FP32)
    c = _ctx(MODELS[2], CONDS[2], T.FP64)
    try:
        with pytest.raises(T.TcsemError, match="precision"):
            T.Batch([a, b])
        h = C.c_void_p()
        assert T.lib().tcsem_batch_create((C.c_void_p * 2)(a.h.value, a.h.value), 2, C.byref(h)) == -1
        assert b"twice" in T.lib().tcsem_last_error()
        assert T.lib().tcsem_batch_create((C.c_void_p * 2)(a.h.value, c.h.value), 2, C.byref(h)) == -7
        with T.Batch([a, c]) as bt:
            bt.run(max_steps=20)
            assert bt.steps_done == 20
    finally:
        for x in (a, b, c):
            x.close()


def test_batch_step_after_every_member_finished():
    ctxs = [_ctx(kw, cond, T.FP32, tn=0.03) for kw, cond in zip(MODELS[:2], CONDS[:2])]
    try:
        with T.Batch(ctxs) as bt:
            got = bt.run()
            steps = bt.steps_done
            bt.step(10)  # nothing left to step: a no-op
            assert all(np.array_equal(c.traces(0), g) for c, g in zip(ctxs, got))
            assert steps > 0
    finally:
        for c in ctxs:
            c.close()
