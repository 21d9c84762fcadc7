"""CPU twins of the air boundary: the reference's CPU formulation of the upward continuation -- SplineGridConvIR2R_cpu.cpp, the filter of UpContinuation.cpp:70-97 and
BilinearGridConvR2IR.cpp -- restated in the oracle and compared, function by function, with the
CUDA formulation that the HIP path follows (and matches at round-off).

Pins of the restatement:
  * spline: each 1-D pass of SplineGridConvIR2R_cpu is a natural cubic spline whose nodes sit at
    the cumulative cell widths of the window (the reference uses the cell width dx_i, not the
    centre distance, as node spacing), evaluated at the true offset from the bracketing node and
    continued linearly beyond the end nodes -- checked against scipy's CubicSpline on x-only and
    y-only fields, for which the composition of the passes reduces to one 1-D pass (except the
    right-hand corner blocks in the x-only case, whose y passes take their two core nodes two
    columns to the left, SplineGridConvIR2R_cpu.cpp:579-581,616-618);
  * filter: equal to the CUDA filter when dxmin = dymin; otherwise the CUDA Bx carries the extra
    factor dx/dy and By dy/dx (UpContinuation_gpu.cu:159-160);
  * bilinear: equal to the CUDA bilinear (its clamped indices give the same end-node pairs as the
    CPU extrapolation branches; the CPU code's one-past-the-row read at the II == L1, JJ == 0
    corner, BilinearGridConvR2IR.cpp:99,257, is restated with the neighbouring branches' nodes).
The two formulations differ in the spline: about 1 % of the plane's maximum in the padding zones,
and up to 3.5e-3 (median 1.8e-4) relative trace error per receiver over the full 64 x 64 x 48
marine run (measured with the oracle; the small-deck check below bounds it at 1e-2)."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O
from tests import decks

DECKS = {"stretched": {}, "bluestein": {"nx": 23, "ny": 17}}


def _model(kw, dy=None):
    s = decks.small_setup(O, **kw)
    if dy is not None:
        s.dymin = dy
    m = O.Model(s)
    # irregular padding counts and the core (ModelingSetup.cpp:79-84)
    m.AL, m.AR, m.AY = s.ALx1 + s.ALx2, s.ARx1 + s.ARx2, s.Ay1 + s.Ay2
    m.s = s
    return m


def _pass_1d(xs, d, ys, off, maps, xt_all, targets):
    """Expected values of one CPU-twin pass: window nodes xs[off:off+n] carrying ys, node spacing
    d[off:off+n-1] (cell widths), targets t with search index maps[t] and coordinate xt_all[t]."""
    n = len(ys)
    xp = np.concatenate([[0.0], np.cumsum(d[off:off + n - 1])])
    cs = CubicSpline(xp, ys, bc_type="natural")
    out = []
    for t in targets:
        h = xt_all[t] - xs[max(maps[t] - 1, 0)]
        idx = maps[t] - 1 - off
        if idx < 0:
            out.append(ys[0] + cs(xp[0], 1) * h)
        elif idx >= n - 1:
            out.append(ys[n - 1] + cs(xp[n - 1], 1) * h)
        else:
            c = cs.c[:, idx]
            out.append(((c[0] * h + c[1]) * h + c[2]) * h + c[3])
    return np.array(out)


@pytest.mark.parametrize("name", sorted(DECKS))
def test_spline_x_passes_are_natural_splines_on_cell_width_nodes(name):
    m = _model(DECKS[name])
    g, mp = m.grid(), m.maps()
    X, dx, Xn, xm = g["X"], g["dx"], g["Xn"], mp["xN2o"]
    f = lambda x: np.cos(x / 900.0) + 1e-4 * x
    bz = np.tile(f(X), (m.M, 1))
    out = m.spline_ir2r_cpu(bz)
    AL, AR, Nx, aL, aR, aY, Ny = m.AL, m.AR, m.s.Nx, m.AnLx, m.AnRx, m.Any, m.s.Ny
    left = _pass_1d(X, dx, f(X[:AL + 2]), 0, xm, Xn, range(aL))
    xb = Nx + AL - 2
    right = _pass_1d(X, dx, f(X[xb:xb + AR + 2]), xb, xm, Xn, range(aL + Nx, m.L1))
    scale = np.abs(out).max()
    # left block, every row (its corner y passes see constant columns)
    assert np.abs(out[:, :aL] - left[None, :]).max() < 1e-12 * scale
    # right block, rows of the core band
    assert np.abs(out[aY:aY + Ny, aL + Nx:] - right[None, :]).max() < 1e-12 * scale
    # the core is copied
    assert np.array_equal(out[aY:aY + Ny, aL:aL + Nx], bz[m.AY:m.AY + Ny, AL:AL + Nx])


@pytest.mark.parametrize("name", sorted(DECKS))
def test_spline_y_passes_are_natural_splines_on_cell_width_nodes(name):
    m = _model(DECKS[name])
    g, mp = m.grid(), m.maps()
    Y, dy, Yn, ym = g["Y"], g["dy"], g["Yn"], mp["yN2o"]
    f = lambda y: np.sin(y / 700.0) + 2e-4 * y
    bz = np.tile(f(Y)[:, None], (1, m.L))
    out = m.spline_ir2r_cpu(bz)
    AY, Ny, aY = m.AY, m.s.Ny, m.Any
    front = _pass_1d(Y, dy, f(Y[:AY + 2]), 0, ym, Yn, range(aY))
    yb = AY + Ny - 2
    back = _pass_1d(Y, dy, f(Y[yb:yb + AY + 2]), yb, ym, Yn, range(aY + Ny, m.M1))
    scale = np.abs(out).max()
    assert np.abs(out[:aY, :] - front[:, None]).max() < 1e-12 * scale
    assert np.abs(out[aY + Ny:, :] - back[:, None]).max() < 1e-12 * scale
    assert np.abs(out[aY:aY + Ny, :] - f(Y[AY:AY + Ny])[:, None]).max() < 1e-12 * scale


def test_spline_differs_from_the_cuda_form_in_the_padding_only():
    m = _model({})
    g = m.grid()
    X, Y = np.meshgrid(g["X"], g["Y"])
    bz = np.cos(X / 700.0) * np.exp(-Y ** 2 / 4e6) + 0.3
    a, b = m.spline_ir2r(bz), m.spline_ir2r_cpu(bz)
    aY, aL, Ny, Nx = m.Any, m.AnLx, m.s.Ny, m.s.Nx
    assert np.array_equal(a[aY:aY + Ny, aL:aL + Nx], b[aY:aY + Ny, aL:aL + Nx])
    d = np.abs(a - b).max() / np.abs(a).max()
    assert 1e-4 < d < 0.05, d


def test_filter_equals_the_cuda_filter_when_dx_equals_dy():
    m = _model({})
    bz = np.random.default_rng(5).standard_normal((m.M1, m.L1))
    bx1, by1 = m.upcont_filter(bz)
    bx2, by2 = m.upcont_filter_cpu(bz)
    assert decks.rel_l2(bx2, bx1) < 1e-12
    assert decks.rel_l2(by2, by1) < 1e-12


def test_filter_differs_from_the_cuda_filter_by_dx_over_dy():
    m = _model({}, dy=80.0)
    assert m.s.dymin != m.s.dxmin
    bz = np.random.default_rng(6).standard_normal((m.M1, m.L1))
    bx1, by1 = m.upcont_filter(bz)
    bx2, by2 = m.upcont_filter_cpu(bz)
    r = m.s.dxmin / m.s.dymin
    assert decks.rel_l2(bx1, r * bx2) < 1e-12
    assert decks.rel_l2(by1, by2 / r) < 1e-12


@pytest.mark.parametrize("name", sorted(DECKS))
def test_bilinear_equals_the_cuda_bilinear(name):
    m = _model(DECKS[name])
    rng = np.random.default_rng(7)
    bxr, byr = rng.standard_normal((m.M1, m.L1)), rng.standard_normal((m.M1, m.L1))
    a = m.bilinear(bxr, byr)
    b = m.bilinear_cpu(bxr, byr)
    assert decks.rel_l2(b[0], a[0]) < 1e-13
    assert decks.rel_l2(b[1], a[1]) < 1e-13


def test_traces_with_the_cpu_twin_air_boundary():
    """The same run with the CPU-twin air boundary: traces within 1e-2 of the CUDA form's (the
    spline difference reaches the receivers through the air term), and not identical."""
    so = decks.small_setup(O)
    m = O.Model(so)
    g = m.grid()
    cond = decks.layered_cond((m.N, m.M, m.L), g["Z"][:, None, None] * np.ones((1, m.M, m.L)), seed=7,
                              blob=True)
    rt = O.make_rxtime(0.02, 0.25, 16, 1, (1, 1, 1, 1, 1, 1))
    rx = np.array([[150.0, 0.0, 400.0], [300.0, 60.0, 400.0], [-200.0, -50.0, 350.0], [400.0, 0.0, 300.0]])
    tx = O.make_tx(ort=1, bipolen=100.0, x=0.0, y=0.0, z=300.0, rx_ids=[1, 2, 3, 4])
    res = []
    for twins in (False, True):
        mm = O.Model(so, cond)
        mm.set_cpu_twins(twins)
        res.append(mm.run(rt, rx, tx, max_steps=600)[0])
    err = decks.trace_error(res[1], res[0])
    print("CPU-twin vs CUDA-form air boundary, per-receiver trace difference:", err)
    assert np.all(err < 1e-2) and np.any(err > 0), err
