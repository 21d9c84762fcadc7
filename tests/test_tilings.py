"""Every march tiling the library can select (rows per workgroup x prefetch depth x z-chunks x the
source of the chunks' Bz partial sums, set through the TCSEM_MARCH_* overrides that the heuristic and the autotuner also choose from) gives
the oracle's fields: a random state advanced four full steps, FP64 at round-off and FP32 within
single-precision accumulation error.  The deck is taller than one 16-row tile in y so that tile
seams, halo rows and chunk seams all fall inside the model."""
import numpy as np
import pytest

import tcsem_fdtd_amd as T
from oracle import oracle as O
from tests import decks

pytestmark = pytest.mark.gpu

SETUP = {"nx": 70, "ny": 36, "nz": 20}  # L = 82 (two 64-column blocks), M = 48, N = 26
COMPS = (O.EX, O.EY, O.EZ, O.BX, O.BY, O.BZ, O.BXAIR, O.BYAIR)


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


_REF = {}


def _reference(nsteps=4):
    if "ref" not in _REF:
        so = decks.small_setup(O, **SETUP)
        st = decks.small_setup(T, **SETUP)
        m = O.Model(so)
        g = m.grid()
        cond = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)),
                                  seed=3, blob=True)
        m = O.Model(so, cond)
        rng = np.random.default_rng(21)
        state = {c: _zero_walls(rng.standard_normal(m.shape(c)), c, m.L, m.M, m.N)
                 for c in (O.EX, O.EY, O.EZ, O.BX, O.BY)}
        for c, f in state.items():
            m.set(c, f)
        m.update_bz()
        lam, dte, dtb, amp = O.schedule(so, m.mindx(), O.make_rxtime(1e-3, 1.0, 10, 1, [1] * 6), 100)
        for n in range(nsteps):
            m.step(lam[n], dte[n], amp[n], None, dtb[n])
        _REF["ref"] = (st, cond, state, {c: m.get(c) for c in COMPS}, nsteps)
    return _REF["ref"]


# z-chunked tilings twice: chunks started from a k_zsum pass, and from the chunk sums of B that
# the previous step's march accumulated (steps 2-4; step 1 follows the field upload, a pass)
TILINGS = [(r, p, z, m) for r in (4, 8, 10, 12, 16) for p in (1, 2, 3, 4) for z, m in ((1, "fused"), (3, "pass"), (3, "fused"))]


@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
@pytest.mark.parametrize("rows,pf,zc,zsum", TILINGS, ids=["r%d-pf%d-z%d-%s" % t for t in TILINGS])
def test_tiling_matches_oracle(rows, pf, zc, zsum, precision, monkeypatch):
    if precision == T.FP64 and rows == 16:
        pytest.skip("FP64 tiles have 4, 8, 10 or 12 rows")
    if precision == T.FP32 and rows == 12:
        pytest.skip("FP32 tiles have 4, 8, 10 or 16 rows")
    monkeypatch.setenv("TCSEM_MARCH_ROWS", str(rows))
    monkeypatch.setenv("TCSEM_MARCH_PF", str(pf))
    monkeypatch.setenv("TCSEM_MARCH_ZCHUNKS", str(zc))
    monkeypatch.setenv("TCSEM_MARCH_ZSUM", zsum)
    st, cond, state, ref, nsteps = _reference()
    tol = 1e-10 if precision == T.FP64 else 2e-5
    with T.Context(st, cond, precision) as c:
        c.begin(0, 1)
        for comp, f in state.items():
            c.set(comp, f)
        c.step(nsteps)
        for comp in COMPS:
            err = decks.rel_l2(c.get(comp), ref[comp])
            assert err < tol, (comp, err)


@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_autotuned_tiling_matches_oracle(precision, monkeypatch):
    for v in ("TCSEM_MARCH_ROWS", "TCSEM_MARCH_PF", "TCSEM_MARCH_ZCHUNKS"):
        monkeypatch.delenv(v, raising=False)
    st, cond, state, ref, nsteps = _reference()
    tol = 1e-10 if precision == T.FP64 else 2e-5
    with T.Context(st, cond, precision) as c:
        tuned = c.autotune(1, 3)
        print("autotuned", precision, tuned, flush=True)
        assert tuned["rows"] in ((4, 8, 10, 12) if precision == T.FP64 else (4, 8, 10, 16))
        assert 1 <= tuned["pf"] <= 4 and tuned["zchunks"] >= 1 and tuned["march_ms"] > 0
        c.begin(0, 1)
        for comp, f in state.items():
            c.set(comp, f)
        c.step(nsteps)
        for comp in COMPS:
            err = decks.rel_l2(c.get(comp), ref[comp])
            assert err < tol, (comp, err)
