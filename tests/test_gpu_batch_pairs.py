"""Two steps per pass in a model batch (tcsem_batch_set_pairs, the model-batch form of k_march2):
one pass over every member's tiles takes two steps, then each step's top launches, chains and
surface levels run over all members.  The batch gives the traces and final fields of the same
batch stepped one step at a time, bit for bit -- FP64 and FP32, every pass height, members of
four uneven grids, members whose sampling ends early (left out of later passes) -- and its FP64
member 0 equals the oracle (main.cu:219-273, one model per run)."""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tests import decks

pytestmark = pytest.mark.gpu

# four grids: 36x32x20 (stretched), 35x29x20 (Bluestein air plane), 82x18x19, 42x38x16
MODELS = [{}, {"nx": 23, "ny": 17}, {"nx": 72, "ny": 8, "pad": (2, 3), "arg": (1.5, 2.0)},
          {"nx": 30, "ny": 26, "nz": 10}]
RX = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
TX = dict(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4])
COMPS = (T.EX, T.EY, T.EZ, T.BX, T.BY, T.BXAIR, T.BYAIR)


def _cond(kw, seed):
    m = O.Model(decks.small_setup(O, **kw))
    g = m.grid()
    return decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), seed=seed,
                              blob=True)


CONDS = [_cond(kw, 11 + q) for q, kw in enumerate(MODELS)]


def _rt(mod, tn=0.25):
    return mod.make_rxtime(0.02, tn, 16, 1, (1, 1, 1, 1, 1, 1))


def _batch(precision, rows, tns, max_steps=0):
    ctxs = []
    for kw, cond, tn in zip(MODELS, CONDS, tns):
        c = T.Context(decks.small_setup(T, **kw), cond, precision)
        c.set_receivers(_rt(T, tn), RX)
        c.set_transmitters([T.make_tx(**TX)])
        ctxs.append(c)
    try:
        with T.Batch(ctxs) as b:
            b.set_pairs(rows)
            tr = b.run(max_steps=max_steps)
            fields = {(q, comp): c.get(comp) for q, c in enumerate(ctxs) for comp in COMPS}
            return tr, fields, b.steps_done, b.pair_info()[2]
    finally:
        for c in ctxs:
            c.close()


def _same(got, ref):
    assert got[2] == ref[2]
    for q in range(len(MODELS)):
        assert np.abs(ref[0][q]).max() > 0
        assert got[0][q].shape == ref[0][q].shape
        assert np.array_equal(got[0][q], ref[0][q]), q
    for key in ref[1]:
        assert np.array_equal(got[1][key], ref[1][key]), key


@pytest.mark.parametrize("precision,rows", [(T.FP64, 12), (T.FP64, 8), (T.FP64, 11), (T.FP64, 10),
                                            (T.FP32, 12), (T.FP32, 16), (T.FP32, 14), (T.FP32, 10),
                                            (T.FP32, 8)])
def test_batch_pairs_equal_single_steps(precision, rows):
    tns = [0.25] * len(MODELS)
    ref = _batch(precision, 0, tns, max_steps=400)
    got = _batch(precision, rows, tns, max_steps=400)
    print("steps", got[2], "pairs", got[3])
    assert ref[3] == 0 and got[3] > 50
    _same(got, ref)


def test_batch_pairs_with_members_finishing_early():
    """Members 1 and 3 sample only to 0.08 s: they drop out of later passes and launches."""
    tns = [0.25, 0.08, 0.25, 0.08]
    ref = _batch(T.FP64, 0, tns)
    got = _batch(T.FP64, 12, tns)
    assert got[3] > 50
    _same(got, ref)


def test_batch_pairs_fp64_member_equals_the_oracle():
    got = _batch(T.FP64, 12, [0.25] * len(MODELS), max_steps=400)
    assert got[3] > 50
    m = O.Model(decks.small_setup(O, **MODELS[0]), CONDS[0])
    ref, _ = m.run(_rt(O), RX, O.make_tx(**TX), max_steps=400)
    err = decks.trace_error(got[0][0], ref)
    assert np.all(err < 1e-9), err


def test_batch_pair_settings_are_checked():
    ctxs = []
    for kw, cond in zip(MODELS[:2], CONDS[:2]):
        c = T.Context(decks.small_setup(T, **kw), cond, T.FP64)
        c.set_receivers(_rt(T), RX)
        c.set_transmitters([T.make_tx(**TX)])
        ctxs.append(c)
    try:
        with T.Batch(ctxs) as b:
            for rows in (16, 14, 9, 4):
                with pytest.raises(T.TcsemError):
                    b.set_pairs(rows)
            b.set_pairs(8)
            b.step(60)
            assert b.pair_info()[0] == 8 and b.pair_info()[2] > 0
    finally:
        for c in ctxs:
            c.close()
