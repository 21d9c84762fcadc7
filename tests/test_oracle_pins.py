"""Pins of the FP64 oracle (CPU only).

Each test checks an oracle function against something it does not restate: the reference's own
golden file, the bracket invariants of the reference's DEBUG==1 build, a library routine (numpy FFT,
scipy natural cubic spline), a closed form (potential-field continuation of a cosine, the DF decay
factor), exactness on functions the operator reproduces, or a symmetry of the physics.  A planted
mistake in the oracle (wrong sign, dropped term, transposed index) fails at least one of them.
"""
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O
from tests import decks

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- a1: source schedule ------------------------------------------------------------------------
def test_waveform_matches_reference_golden_file():
    """testwaveform.dat is the reference's own dump of Waveform(100, 1e-3) (SourceWave.cpp:60-71):
    int32 N_steps, then t, dt, Wave (2N+1 doubles each).  Bit-exact."""
    raw = open(os.path.join(GOLDEN, "testwaveform.dat"), "rb").read()
    ns = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    t_ref, dt_ref, w_ref = np.frombuffer(raw[4:], dtype=np.float64).reshape(3, 2 * ns + 1)
    t, dt, w = O.waveform(ns, 1e-3)
    assert ns == 100
    assert np.array_equal(t, t_ref)
    assert np.array_equal(dt[:2 * ns + 1], dt_ref)
    assert np.array_equal(w, w_ref)
    # closed-form properties: symmetric time axis ending at 6 sigma, Gaussian peak 1/(sqrt(2 pi) s)
    assert t[-1] == pytest.approx(6e-3, rel=1e-15)
    assert w.max() == pytest.approx(1 / (np.sqrt(2 * np.pi) * 1e-3), rel=1e-12)


def test_schedule_step_recurrence():
    """After the source, dt2 = min(0.15 sfa1 sqrt(mu0 t sigma_min / 8) dmin, 1.2 dt1) with sigma_min
    = 0.1 and lambda = 7.5 sfa2 dt1 dt2 / (mu0 dmin^2) (main.cu:155-163,223,257-260)."""
    s = decks.small_setup(O)
    m = O.Model(s)
    lam, dte, dtb, amp = O.schedule(s, m.mindx(), O.make_rxtime(0.01, 0.3, 20, 1, [1] * 6), 5000)
    ns = s.n_src // 2
    mu0 = 4 * np.pi * 1e-7
    toT = np.concatenate([[0.0], np.cumsum(dte)])
    # the dt of step n (n >= 2 ns + 1) is chosen after step n-2, from the time reached then
    for n in range(2 * ns + 1, len(dte) - 1):
        want = min(0.15 * s.sfa1 * np.sqrt(mu0 * toT[n - 1] * 0.1 / 8) * m.mindx(), 1.2 * dte[n - 1])
        assert dte[n] == pytest.approx(want, rel=1e-12)
        assert lam[n] == pytest.approx(7.5 * s.sfa2 * dte[n] * dte[n + 1] / mu0 / m.mindx() ** 2, rel=1e-12)
        assert dtb[n] == pytest.approx(0.5 * (dte[n] + dte[n + 1]), rel=1e-14)
    _, _, w = O.waveform(ns, s.sigma_t)
    assert np.array_equal(amp[:2 * ns], w[:2 * ns]) and np.all(amp[2 * ns:] == 0)


# ---- a2: grid and mapping -------------------------------------------------------------------------
def _reference_deck_setup():
    vals = open(os.path.join(GOLDEN, "FDTD_setup.in")).read().split()
    f = [float(v) for v in vals]
    return O.make_setup(sfa1=f[0], sfa2=f[1], n_src=int(f[2]), sigma_t=f[3], x0=f[4], y0=f[5], z0=f[6],
                        dxmin=f[7], dymin=f[8], dzmin=f[9], arg1=f[10], arg2=f[11], ALx2=int(f[12]),
                        ALx1=int(f[13]), Nx=int(f[14]), ARx1=int(f[15]), ARx2=int(f[16]), Ay2=int(f[17]),
                        Ay1=int(f[18]), Ny=int(f[19]), Az2=int(f[20]), Az1=int(f[21]), Nz=int(f[22]))


@pytest.mark.parametrize("which", ["small", "reference_deck"])
def test_grid_closed_form(which):
    s = decks.small_setup(O) if which == "small" else _reference_deck_setup()
    m = O.Model(s)
    g = m.grid()
    # zone sizes: core dmin, geometric growth outward (ModelingSetup.cpp:100-152)
    AL = s.ALx1 + s.ALx2
    assert np.all(g["dx"][AL:AL + s.Nx] == s.dxmin)
    i = np.arange(s.ALx2, AL)
    assert np.allclose(g["dx"][i], s.dxmin * s.arg1 ** (AL - i), rtol=1e-14)
    assert np.allclose(g["dz"][s.Nz:s.Nz + s.Az1], s.dzmin * s.arg1 ** np.arange(1, s.Az1 + 1), rtol=1e-12)
    # centres: adjacent cells share a face; the core is centred on x0, y0; Z[0] = dz0/2 + z0
    assert np.allclose(np.diff(g["X"]), 0.5 * (g["dx"][1:] + g["dx"][:-1]), rtol=1e-13)
    assert np.allclose(np.diff(g["Z"]), 0.5 * (g["dz"][1:] + g["dz"][:-1]), rtol=1e-13)
    xc = g["X"][AL:AL + s.Nx]
    assert abs(0.5 * (xc[0] + xc[-1]) - s.x0) < 1e-9
    assert g["Z"][0] == 0.5 * g["dz"][0] + s.z0
    # regular plane: AnLx_New = ceil(left padding span / dxmin) (ModelingSetup.cpp:190)
    assert m.AnLx == int(np.ceil(g["dx"][:AL].sum() / s.dxmin - 1e-9))
    assert m.L1 == m.AnLx + m.AnRx + s.Nx and m.M1 == 2 * m.Any + s.Ny
    Xn = g["Xn"]
    assert np.allclose(np.diff(Xn), s.dxmin, rtol=1e-12)
    assert np.array_equal(Xn[m.AnLx:m.AnLx + s.Nx], g["X"][AL:AL + s.Nx])
    if which == "reference_deck":
        assert (m.L, m.M, m.N, m.L1, m.M1) == (220, 200, 110, 3914, 3894)


def _bracket(new, old, idx):
    """The DEBUG==1 checks of ModelGridMapping.cpp:104-390: idx==0 -> left of old[0]; idx==n ->
    right of old[n-1]; else old[idx-1] < new <= old[idx]."""
    n = len(old)
    for p, q in enumerate(idx):
        if q == 0:
            assert new[p] <= old[0]
        elif q == n:
            assert new[p] > old[n - 1]
        else:
            assert old[q - 1] < new[p] <= old[q]


@pytest.mark.parametrize("which", ["small", "reference_deck"])
def test_mapping_brackets(which):
    s = decks.small_setup(O) if which == "small" else _reference_deck_setup()
    m = O.Model(s)
    g, mp = m.grid(), m.maps()
    _bracket(g["Xn"], g["X"], mp["xN2o"])
    _bracket(g["Yn"], g["Y"], mp["yN2o"])
    _bracket(g["X"], g["Xn"], mp["xo2N"])
    _bracket(g["Y"], g["Yn"], mp["yo2N"])
    xb = np.concatenate([g["X"] - 0.5 * g["dx"], [g["X"][-1] + 0.5 * g["dx"][-1]]])
    yb = np.concatenate([g["Y"] - 0.5 * g["dy"], [g["Y"][-1] + 0.5 * g["dy"][-1]]])
    _bracket(xb, g["Xn"], mp["xBx2N"])
    _bracket(yb, g["Yn"], mp["yBy2N"])


def test_search_insert_is_binary_lower_bound():
    a = np.array([1.0, 3.0, 4.0, 6.0, 9.0])
    for t in (0.5, 1.0, 2.0, 3.0, 5.0, 6.0, 8.9, 9.0, 10.0):
        assert O.search_insert(a, t) == int(np.searchsorted(a, t, side="left"))


# ---- a10: natural-spline resampling -------------------------------------------------------------
def test_spline_ir2r_matches_scipy_natural_spline():
    """Each padding window is the textbook natural cubic spline through the window's nodes
    (scipy CubicSpline bc_type='natural'), extended linearly beyond the end nodes."""
    s = decks.small_setup(O)
    m = O.Model(s)
    g, mp = m.grid(), m.maps()
    rng = np.random.default_rng(5)
    bz0 = rng.standard_normal((m.M, m.L))
    out = m.spline_ir2r(bz0)
    AL, AY = s.ALx1 + s.ALx2, s.Ay1 + s.Ay2
    n1 = AL + 4

    def nat(xs, ys, xq):
        cs = CubicSpline(xs, ys, bc_type="natural")
        v = cs(xq)
        lo, hi = xq < xs[0], xq > xs[-1]
        v[lo] = ys[0] + cs(xs[0], 1) * (xq[lo] - xs[0])
        v[hi] = ys[-1] + cs(xs[-1], 1) * (xq[hi] - xs[-1])
        return v

    # x pass on the core rows (centre rows of the regular plane)
    for j in (AY, AY + 3, AY + s.Ny - 1):
        want = nat(g["X"][:n1], bz0[j, :n1], g["Xn"][:m.AnLx])
        got = out[j - AY + m.Any, :m.AnLx]
        assert np.allclose(got, want, rtol=1e-11, atol=1e-12)
    # y pass of a core column (x-pass is the identity there)
    for c in (m.AnLx, m.AnLx + 5):
        col = bz0[:, c - m.AnLx + AL]
        n3 = AY + 4
        want = nat(g["Y"][:n3], col[:n3], g["Yn"][:m.Any])
        assert np.allclose(out[:m.Any, c], want, rtol=1e-11, atol=1e-12)
    # the centre block is copied
    assert np.array_equal(out[m.Any:m.Any + s.Ny, m.AnLx:m.AnLx + s.Nx],
                          bz0[AY:AY + s.Ny, AL:AL + s.Nx])


def test_spline_reproduces_linear_fields():
    s = decks.small_setup(O)
    m = O.Model(s)
    g = m.grid()
    X, Y = np.meshgrid(g["X"], g["Y"])
    out = m.spline_ir2r(2.0 + 0.3 * X - 0.7 * Y)
    Xn, Yn = np.meshgrid(g["Xn"], g["Yn"])
    assert np.allclose(out, 2.0 + 0.3 * Xn - 0.7 * Yn, rtol=1e-11, atol=1e-8)


# ---- a11: FFT and the wavenumber filter --------------------------------------------------------
@pytest.mark.parametrize("shape", [(16, 32), (17, 23), (30, 7), (64, 48)])
def test_fft2_matches_numpy(shape):
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    re, im = rng.standard_normal(shape), rng.standard_normal(shape)
    fr, fi = O.fft2(re, im, -1)
    ref = np.fft.fft2(re + 1j * im)
    assert np.allclose(fr + 1j * fi, ref, rtol=1e-12, atol=1e-10 * np.abs(ref).max())
    br, bi = O.fft2(re, im, +1)
    ref = np.fft.ifft2(re + 1j * im) * re.size
    assert np.allclose(br + 1j * bi, ref, rtol=1e-12, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("dims", [(24, 20), (23, 17)])
def test_filter_is_potential_field_continuation(dims):
    """Closed form: Bz = cos(kx x) on the regular plane continues upward by dz/2 to
    Bx = -sin(kx x) exp(-|k| dz/2) (and By = 0); likewise cos(ky y) gives By = -sin(ky y) e^{-|k|dz/2}.
    The continuation is unnormalised (the bilinear step divides by L1*M1)."""
    s = decks.small_setup(O, nx=dims[0], ny=dims[1])
    m = O.Model(s)
    L1, M1 = m.L1, m.M1
    x = np.arange(L1) * s.dxmin
    y = np.arange(M1) * s.dymin
    for mode in (1, 3):
        kx = 2 * np.pi * mode / (L1 * s.dxmin)
        bz = np.tile(np.cos(kx * x), (M1, 1))
        bxr, byr = m.upcont_filter(bz)
        want = -np.sin(kx * x) * np.exp(-kx * s.dzmin / 2)
        assert np.allclose(bxr / (L1 * M1), np.tile(want, (M1, 1)), atol=1e-12)
        assert np.allclose(byr / (L1 * M1), 0.0, atol=1e-12)
        ky = 2 * np.pi * mode / (M1 * s.dymin)
        bz = np.tile(np.cos(ky * y)[:, None], (1, L1))
        bxr, byr = m.upcont_filter(bz)
        want = -np.sin(ky * y) * np.exp(-ky * s.dzmin / 2)
        assert np.allclose(byr / (L1 * M1), np.tile(want[:, None], (1, L1)), atol=1e-12)
        assert np.allclose(bxr / (L1 * M1), 0.0, atol=1e-12)


def test_filter_matches_numpy_rfft_semantics():
    """The half-spectrum filter followed by cuFFT/FFTW-style C2R equals numpy's irfft2 of the same
    half spectrum (which drops the imaginary parts of self-conjugate bins)."""
    s = decks.small_setup(O, nx=23, ny=20)
    m = O.Model(s)
    L1, M1 = m.L1, m.M1
    rng = np.random.default_rng(9)
    bz = rng.standard_normal((M1, L1))
    F = np.fft.rfft2(bz)
    iy = np.arange(M1)[:, None]
    ix = np.arange(L1 // 2 + 1)[None, :]
    v1 = np.where(iy < M1 // 2 + 1, np.pi * iy / M1, -np.pi * (M1 - iy) / M1) * (s.dzmin / s.dymin)
    u1 = np.pi * ix / L1 * (s.dzmin / s.dxmin)
    rr = u1 ** 2 + v1 ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        up = np.exp(-np.sqrt(rr)) / np.sqrt(rr)
        gx, gy = u1 * up, v1 * up
    gx[0, 0] = gy[0, 0] = 1.0
    bx_ref = np.fft.irfft2(1j * gx * F, s=(M1, L1)) * (L1 * M1)
    by_ref = np.fft.irfft2(1j * gy * F, s=(M1, L1)) * (L1 * M1)
    bxr, byr = m.upcont_filter(bz)
    assert np.allclose(bxr, bx_ref, rtol=1e-10, atol=1e-9 * np.abs(bx_ref).max())
    assert np.allclose(byr, by_ref, rtol=1e-10, atol=1e-9 * np.abs(by_ref).max())


# ---- a12: bilinear ----------------------------------------------------------------------------------
def test_bilinear_exact_on_bilinear_fields():
    s = decks.small_setup(O)
    m = O.Model(s)
    g = m.grid()
    Xn, Yn = np.meshgrid(g["Xn"], g["Yn"])
    f = lambda X, Y: 1.5 + 0.01 * X - 0.02 * Y + 1e-5 * X * Y
    norm = m.L1 * m.M1
    bxa, bya = m.bilinear(f(Xn, Yn) * norm, f(Xn, Yn) * norm)
    xb = np.concatenate([g["X"] - 0.5 * g["dx"], [g["X"][-1] + 0.5 * g["dx"][-1]]])
    yb = np.concatenate([g["Y"] - 0.5 * g["dy"], [g["Y"][-1] + 0.5 * g["dy"][-1]]])
    Xb, Yc = np.meshgrid(xb, g["Y"])
    assert np.allclose(bxa, f(Xb, Yc), rtol=1e-12, atol=1e-9)
    Xc, Yb = np.meshgrid(g["X"], yb)
    assert np.allclose(bya, f(Xc, Yb), rtol=1e-12, atol=1e-9)


# ---- a14: receiver interpolation --------------------------------------------------------------------
def test_spline4_is_natural_spline_on_middle_interval():
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = np.sort(rng.uniform(0, 10, 4))
        y = rng.standard_normal(4)
        cs = CubicSpline(x, y, bc_type="natural")
        for t in np.linspace(x[1], x[2], 5):
            assert O.spline4(x, y, t) == pytest.approx(float(cs(t)), rel=1e-10, abs=1e-12)
    x = np.array([0.0, 1.0, 2.5, 4.0])
    assert O.spline4(x, 3 * x - 1, 1.7) == pytest.approx(3 * 1.7 - 1, rel=1e-14)


# ---- a3-a9: the stencil --------------------------------------------------------------------------
def _homogeneous_model(sigma=2.0):
    s = decks.small_setup(O, nx=8, ny=8, nz=6, pad=(1, 1))
    m = O.Model(s)
    return s, O.Model(s, np.full((m.N, m.M, m.L), sigma))


def test_df_decay_factor_with_zero_curl():
    """With B = 0 (so curl B = 0 and the air fields vanish) the DF update scales E by exactly
    (2 lam - s dt)/(2 lam + s dt) with the 4-cell average s (ElectricalFiled.cu:138-139); at the
    surface the air cells count as s = 0."""
    s, m = _homogeneous_model(2.0)
    rng = np.random.default_rng(0)
    e0 = {c: rng.standard_normal(m.shape(c)) for c in (O.EX, O.EY, O.EZ)}
    for c, v in e0.items():
        m.set(c, v)
    lam, dt = 3e-4, 1e-4
    m.upcont()
    m.update_e(lam, dt)
    f = lambda sg: (2 * lam - sg * dt) / (2 * lam + sg * dt)
    ex, ey, ez = m.get(O.EX), m.get(O.EY), m.get(O.EZ)
    L, M, N = m.L, m.M, m.N
    assert np.allclose(ex[1:N, 1:M, :], f(2.0) * e0[O.EX][1:N, 1:M, :], rtol=1e-14)
    assert np.allclose(ex[0, 1:M, :], f(1.0) * e0[O.EX][0, 1:M, :], rtol=1e-14)
    assert np.allclose(ey[1:N, :, 1:L], f(2.0) * e0[O.EY][1:N, :, 1:L], rtol=1e-14)
    assert np.allclose(ey[0, :, 1:L], f(1.0) * e0[O.EY][0, :, 1:L], rtol=1e-14)
    assert np.allclose(ez[:, 1:M, 1:L], f(2.0) * e0[O.EZ][:, 1:M, 1:L], rtol=1e-14)
    # walls are never updated
    assert np.array_equal(ex[N], e0[O.EX][N]) and np.array_equal(ex[:, 0], e0[O.EX][:, 0])


def test_b_update_faraday_on_linear_fields_and_div_free():
    """Faraday dB/dt = -curl E is exact on linear E: with Ey = a z, Ez = b y, Ex = c z (at their
    own staggered positions) one step gives dBx = dt (a - b) and dBy = -dt c everywhere inside
    (Bderivate.cu:125-126,186-187).  Bz then closes the discrete divergence of every cell with
    Bz = 0 at the bottom (Bderivate.cu:227-253)."""
    s, m = _homogeneous_model()
    g = m.grid()
    L, M, N = m.L, m.M, m.N
    ze = np.concatenate([g["Z"] - 0.5 * g["dz"], [g["Z"][-1] + 0.5 * g["dz"][-1]]])   # z-edges, N+1
    ye = np.concatenate([g["Y"] - 0.5 * g["dy"], [g["Y"][-1] + 0.5 * g["dy"][-1]]])   # y-edges, M+1
    a, b, c, dt = 0.7, -0.3, 1.1, 0.25
    m.set(O.EY, a * ze[:, None, None] * np.ones(m.shape(O.EY)))
    m.set(O.EZ, b * ye[None, :, None] * np.ones(m.shape(O.EZ)))
    m.set(O.EX, c * ze[:, None, None] * np.ones(m.shape(O.EX)))
    m.update_b(dt)
    bx, by, bz = m.get(O.BX), m.get(O.BY), m.get(O.BZ)
    assert np.allclose(bx[:, :M - 1, 1:L], dt * (a - b), rtol=1e-11)
    assert np.allclose(by[:, 1:M, :], -dt * c, rtol=1e-11)
    rng = np.random.default_rng(1)
    for comp in (O.EX, O.EY, O.EZ, O.BX, O.BY):
        m.set(comp, rng.standard_normal(m.shape(comp)))
    m.update_b(dt)
    bx, by, bz = m.get(O.BX), m.get(O.BY), m.get(O.BZ)
    dx, dy, dz = g["dx"], g["dy"], g["dz"]
    div = ((bx[:, :, 1:] - bx[:, :, :-1]) / dx[None, None, :] + (by[:, 1:, :] - by[:, :-1, :]) / dy[None, :, None]
           + (bz[1:] - bz[:-1]) / dz[:, None, None])
    assert np.abs(div).max() < 1e-12 * np.abs(bx).max() / dx.min()
    assert np.all(bz[N] == 0)


def test_e_update_dufort_frankel_on_linear_fields():
    """The DuFort-Frankel E update (Wang & Hohmann 1993) E <- (2l - s dt)/(2l + s dt) E
    + 2/(2l/dt + s) (curl B)/mu0 is exact in space for linear B.  With Bz = a y, By = b z the
    x-curl is a - b; with Bx = c z, Bz = d x the y-curl is c - d; with By = e x, Bx = f y the z-curl
    is e - f (ElectricalFiled.cu:138-139,169-170,232-234,267-269,339-341)."""
    sig = 2.0
    s, m = _homogeneous_model(sig)
    g = m.grid()
    L, M, N = m.L, m.M, m.N
    mu0 = 4 * np.pi * 1e-7
    lam, dt = 3e-4, 1e-4
    ca = (2 * lam - sig * dt) / (2 * lam + sig * dt)
    gain = 2.0 / (2 * lam / dt + sig) / mu0
    X, Y, Z = g["X"], g["Y"], g["Z"]
    ones = lambda c: np.ones(m.shape(c))
    cases = [
        (O.EX, {O.BZ: ("y", 0.8), O.BY: ("z", -0.4)}, 0.8 - (-0.4)),
        (O.EY, {O.BX: ("z", 0.6), O.BZ: ("x", 0.25)}, 0.6 - 0.25),
        (O.EZ, {O.BY: ("x", -0.9), O.BX: ("y", 0.3)}, -0.9 - 0.3),
    ]
    for comp, bset, curl in cases:
        m.zero()
        e0 = np.random.default_rng(comp).standard_normal(m.shape(comp))
        m.set(comp, e0)
        for bc, (axis, slope) in bset.items():
            shape = m.shape(bc)
            coord = {"x": X[None, None, :], "y": Y[None, :, None], "z": Z[:, None, None]}[axis]
            v = slope * coord * np.ones(shape)
            m.set(bc, v)
        m.update_e(lam, dt)   # air planes stay zero; only interior levels are checked
        e = m.get(comp)
        want = ca * e0 + gain * curl
        if comp == O.EX:
            sl = (slice(1, N), slice(1, M), slice(0, L))
        elif comp == O.EY:
            sl = (slice(1, N), slice(0, M), slice(1, L))
        else:
            sl = (slice(0, N), slice(1, M), slice(1, L))
        assert np.allclose(e[sl], want[sl], rtol=1e-10, atol=1e-12 * abs(gain * curl)), comp


def test_mirror_symmetry_of_an_inline_x_dipole():
    """x-dipole on y = 0 in a model symmetric in y: Ex, By, Ez are even in y; Ey, Bx, Bz are odd.
    A transposed index or a wrong neighbour breaks this at O(1).  The reference filter treats the
    y-Nyquist bin of an even-sized air plane as a positive wavenumber (UpContinuation_gpu.cu:
    127-134), which is not mirror symmetric, so the symmetry holds to ~1e-7 rather than round-off."""
    s = decks.small_setup(O, nx=16, ny=16, nz=8, pad=(1, 2))
    m0 = O.Model(s)
    Z = m0.grid()["Z"][:, None, None] * np.ones((1, m0.M, m0.L))
    cond = decks.layered_cond((m0.N, m0.M, m0.L), Z, sea_depth=250.0)
    m = O.Model(s, cond)
    tx = O.make_tx(1, 100.0, 0.0, 0.0, 200.0, [1])
    tp = m.locate_tx(tx)
    lam, dte, dtb, amp = O.schedule(s, m.mindx(), O.make_rxtime(0.01, 0.3, 20, 1, [1] * 6), 200)
    for n in range(30):
        m.step(lam[n], dte[n], amp[n], tp, dtb[n])
    ex, ey, ez, bx, by, bz = (m.get(c) for c in (O.EX, O.EY, O.EZ, O.BX, O.BY, O.BZ))
    # y-staggering: Ex, Ez, By live on y-edges (index j <-> M - j); Ey, Bx, Bz on centres (j <-> M-1-j)
    sc, sb, tol = np.abs(ex).max(), np.abs(by).max(), 1e-5
    assert np.abs(ex - ex[:, ::-1, :]).max() < tol * sc
    assert np.abs(ez - ez[:, ::-1, :]).max() < tol * np.abs(ez).max()
    assert np.abs(by - by[:, ::-1, :]).max() < tol * sb
    assert np.abs(ey + ey[:, ::-1, :]).max() < tol * sc
    assert np.abs(bx + bx[:, ::-1, :]).max() < tol * sb
    assert np.abs(bz + bz[:, ::-1, :]).max() < tol * sb
    assert np.abs(ey).max() > 1e-8 * sc  # not trivially zero


def test_linearity_in_source_amplitude():
    s = decks.small_setup(O, nx=12, ny=12, nz=8, pad=(1, 2))
    outs = []
    for scale in (1.0, 2.5):
        m = O.Model(s)
        tp = m.locate_tx(O.make_tx(1, 100.0, 0.0, 0.0, 200.0, [1]))
        lam, dte, dtb, amp = O.schedule(s, m.mindx(), O.make_rxtime(0.01, 0.3, 20, 1, [1] * 6), 100)
        for n in range(15):
            m.step(lam[n], dte[n], scale * amp[n], tp, dtb[n])
        outs.append(m.get(O.EX))
    assert np.allclose(outs[1], 2.5 * outs[0], rtol=1e-10, atol=1e-14 * np.abs(outs[1]).max())


def test_source_injection_moment():
    """A one-cell x-bipole adds -2 I / (s dy dz) to that Ex with I = Wave / length and s the 4-cell
    average conductivity (ElectricalFiled.cu:347-396,538)."""
    s = decks.small_setup(O, nx=12, ny=12, nz=8, pad=(1, 2))
    m0 = O.Model(s)
    cond = np.full((m0.N, m0.M, m0.L), 1.7)
    m = O.Model(s, cond)
    g = m.grid()
    tp = m.locate_tx(O.make_tx(1, 50.0, 0.0, 0.0, 200.0, [1]))
    assert tp.ixa == tp.ixb
    m.add_source(tp, 3.0)
    ex = m.get(O.EX)
    length = g["X"][tp.ixb] - g["X"][tp.ixa] + g["dx"][tp.ixb] / 2 + g["dx"][tp.ixa] / 2
    want = -2 * (3.0 / length) / (1.7 * g["dy"][tp.icty] * g["dz"][tp.ictz])
    assert ex[tp.ictz, tp.icty, tp.ixa] == pytest.approx(want, rel=1e-14)
    assert np.count_nonzero(ex) == 1


# ---- a6: composition of the orientations (ElectricalFiled.cu:532-641, LoadRxTxPos.cpp:143-428) ----
# Bipoles span several cells and sit off the grid nodes, so that the cell search of ort 4/5
# (searchInsert at both ends, then the larger index decremented, LoadRxTxPos.cpp:297-302) selects
# the same cells as the ort 1-3 search (searchInsert(end) - 1).  For a bipole inside one cell the two
# searches differ: ort 1 then spans two cells, ort 4 one.
_ORT_RX = np.array([[160.0, 10.0, 400.0], [310.0, 70.0, 400.0], [-190.0, -40.0, 350.0], [410.0, 10.0, 290.0]])
_ORT_POS = dict(x=10.0, y=10.0, z=310.0)


def _ort_model():
    s = decks.small_setup(O)
    m0 = O.Model(s)
    Z = m0.grid()["Z"][:, None, None] * np.ones((1, m0.M, m0.L))
    return O.Model(s, decks.layered_cond((m0.N, m0.M, m0.L), Z, seed=11, blob=True))


def _ort_traces(m, **kw):
    tx = O.make_tx(rx_ids=[1, 2, 3, 4], **dict(_ORT_POS, **kw))
    efs, _ = m.run(O.make_rxtime(0.005, 0.2, 14, 1, [1] * 6), _ORT_RX, tx, max_steps=160)
    assert np.abs(efs).max() > 0
    return efs


def _rel_per_receiver(a, b):
    return np.array([np.linalg.norm(a[r] - b[r]) / np.linalg.norm(b[r]) for r in range(b.shape[0])])


def test_ort4_at_zero_angles_is_ort1_bit_exact():
    """azimuth 0, dip 0: the y and z currents are exactly zero (sin 0), the x cells and the x
    current equal those of an x-bipole (cos 0 = 1)."""
    m = _ort_model()
    a = _ort_traces(m, ort=1, bipolen=180.0)
    b = _ort_traces(m, ort=4, bipolen=180.0, azi=0.0, dip=0.0)
    assert np.array_equal(a, b)


def test_ort4_at_azimuth_90_is_ort2():
    """azimuth 90: a y-bipole.  cos(pi/2) = 6e-17 leaves an x current of that relative size."""
    m = _ort_model()
    a = _ort_traces(m, ort=2, bipolen=180.0)
    b = _ort_traces(m, ort=4, bipolen=180.0, azi=90.0, dip=0.0)
    assert np.all(_rel_per_receiver(b, a) < 1e-12)
    # and not an x-bipole
    assert np.all(_rel_per_receiver(_ort_traces(m, ort=1, bipolen=180.0), a) > 0.1)


def test_ort4_at_dip_90_is_ort3():
    m = _ort_model()
    a = _ort_traces(m, ort=3, bipolen=180.0)
    b = _ort_traces(m, ort=4, bipolen=180.0, azi=0.0, dip=90.0)
    assert np.all(_rel_per_receiver(b, a) < 1e-12)
    assert np.all(_rel_per_receiver(_ort_traces(m, ort=1, bipolen=180.0), a) > 0.1)


def test_ort4_splits_the_current_by_azimuth_and_dip():
    """A tilted bipole whose three projections span the same cells as the axis bipoles at the
    same centre carries the currents cos(dip)cos(azi), cos(dip)sin(azi), sin(dip) per unit moment
    of each axis bipole's own cell span, so by linearity it is that sum of ort 1-3 runs, each
    rescaled by the ratio of the span lengths (ElectricalFiled.cu:561-611)."""
    m = _ort_model()
    g = m.grid()
    azi, dip, bl = 35.0, 25.0, 300.0
    ca, sa, cd, sd = (np.cos(np.radians(azi)), np.sin(np.radians(azi)), np.cos(np.radians(dip)),
                      np.sin(np.radians(dip)))
    tilted = _ort_traces(m, ort=4, bipolen=bl, azi=azi, dip=dip)
    tp = m.locate_tx(O.make_tx(ort=4, bipolen=bl, azi=azi, dip=dip, rx_ids=[1], **_ORT_POS))
    # axis bipoles with the same cell spans: ort 1-3 lengths chosen so that their searches select
    # [ixa, ixb], [iya, iyb], [iza, izb] (each end inside its cell, centred on the tilted centre)
    X, Y, Z, dx, dy, dz = g["X"], g["Y"], g["Z"], g["dx"], g["dy"], g["dz"]
    span = {1: X[tp.ixb] - X[tp.ixa] + dx[tp.ixb] / 2 + dx[tp.ixa] / 2,
            2: Y[tp.iyb] - Y[tp.iya] + dy[tp.iyb] / 2 + dy[tp.iya] / 2,
            3: Z[tp.izb] - Z[tp.iza] + dz[tp.izb] / 2 + dz[tp.iza] / 2}
    centre = dict(x=0.5 * (X[tp.ixa] + X[tp.ixb]), y=0.5 * (Y[tp.iya] + Y[tp.iyb]), z=0.5 * (Z[tp.iza] + Z[tp.izb]))
    assert tp.ictx == tp.ixa or tp.ictx == tp.ixb or tp.ixa < tp.ictx < tp.ixb
    parts = []
    for ort, w in ((1, cd * ca), (2, cd * sa), (3, sd)):
        length = span[ort] - 1.0  # both ends 0.5 m inside the end cells
        tx = O.make_tx(ort=ort, bipolen=length, rx_ids=[1, 2, 3, 4], **centre)
        tq = m.locate_tx(tx)
        sel = {1: (tq.ixa, tq.ixb, tp.ixa, tp.ixb), 2: (tq.iya, tq.iyb, tp.iya, tp.iyb),
               3: (tq.iza, tq.izb, tp.iza, tp.izb)}[ort]
        assert (min(sel[:2]), max(sel[:2])) == (sel[2], sel[3]), (ort, sel)
        if (ort == 1 and (tq.icty, tq.ictz) != (tp.icty, tp.ictz)) or (ort == 2 and (tq.ictx, tq.ictz) != (tp.ictx, tp.ictz)) \
                or (ort == 3 and (tq.ictx, tq.icty) != (tp.ictx, tp.icty)):
            pytest.skip("the axis bipole's cross-section cell differs from the tilted one's")
        efs, _ = m.run(O.make_rxtime(0.005, 0.2, 14, 1, [1] * 6), _ORT_RX, tx, max_steps=160)
        parts.append(w * efs)
    assert np.all(_rel_per_receiver(tilted, sum(parts)) < 1e-11)


def test_ort5_without_second_bipole_is_ort4_times_its_length():
    """ort 5 multiplies the first bipole's currents by BipoLen (ElectricalFiled.cu:612-636); a
    second bipole of zero length adds nothing."""
    m = _ort_model()
    a = _ort_traces(m, ort=4, bipolen=180.0, azi=30.0, dip=20.0)
    b = _ort_traces(m, ort=5, bipolen=180.0, azi=30.0, dip=20.0, bipolen1=0.0, x1=-40.0, y1=20.0, z1=320.0,
                    azi1=-45.0, dip1=10.0)
    assert np.all(_rel_per_receiver(b, 180.0 * a) < 1e-12)


def test_ort5_second_bipole_is_injected_on_the_first_bipoles_cells():
    """The reference's AddTx kernels read the first bipole's cell indices for both bipoles (the
    second bipole's own indices only set its current's span length): a second bipole identical to
    the first doubles the response, and moving it leaves the cells where they are while its span
    length sets the current."""
    m = _ort_model()
    a = _ort_traces(m, ort=4, bipolen=180.0, azi=30.0, dip=20.0)
    same = _ort_traces(m, ort=5, bipolen=180.0, azi=30.0, dip=20.0, bipolen1=180.0, x1=10.0, y1=10.0,
                       z1=310.0, azi1=30.0, dip1=20.0)
    assert np.all(_rel_per_receiver(same, 360.0 * a) < 1e-12)
    # the second bipole moved by whole cells (same spans): the same currents, still on the first
    # bipole's cells, so the same traces
    moved = _ort_traces(m, ort=5, bipolen=180.0, azi=30.0, dip=20.0, bipolen1=180.0, x1=60.0, y1=60.0,
                        z1=360.0, azi1=30.0, dip1=20.0)
    assert np.all(_rel_per_receiver(moved, 360.0 * a) < 1e-12)


# ---- a14: receiver sampling (EFInterp2Rx.cu:46-585) ------------------------------------------------
def _face(c, d):
    """Lower faces of the cells and the upper face of the last one."""
    return np.concatenate([c - 0.5 * d, [c[-1] + 0.5 * d[-1]]])


# staggered positions of every component (x, y, z): c = cell centre, f = cell face
_STAGGER = {0: "cff", 1: "fcf", 2: "ffc", 3: "fcc", 4: "cfc", 5: "ccf"}


def test_gather_stencil_puts_the_receiver_in_its_middle_interval():
    """The 4-node stencil of every component brackets the receiver by its two middle nodes on
    each axis, and its nodes are the component's staggered positions (EFInterp2Rx.cu:75-88, the
    half-cell test on the face axes).  Above the second node in z (the surface cells) the start
    is clamped to level 0 and the receiver lies in the first interval."""
    s = decks.small_setup(O)
    m = O.Model(s)
    g = m.grid()
    pos = {"c": (g["X"], g["Y"], g["Z"]), "f": (_face(g["X"], g["dx"]), _face(g["Y"], g["dy"]), _face(g["Z"], g["dz"]))}
    rng = np.random.default_rng(21)
    for comp in range(6):
        for _ in range(300):
            x = rng.uniform(g["X"][2], g["X"][-3])
            y = rng.uniform(g["Y"][2], g["Y"][-3])
            z = rng.uniform(g["Z"][0], g["Z"][-4])
            st, t = m.rx_stencil(x, y, z, comp)
            for ax, v in enumerate((x, y, z)):
                nodes = pos[_STAGGER[comp][ax]][ax]
                assert np.array_equal(t[ax], nodes[st[ax]:st[ax] + 4]), (comp, ax)
                if ax < 2 or nodes[1] <= v:
                    assert t[ax][1] <= v <= t[ax][2], (comp, ax, v, t[ax])
                else:  # shallower than the second node: clamped to level 0
                    assert st[ax] == 0 and t[ax][0] <= v <= t[ax][1] or v < t[ax][0], (comp, v, t[ax])


def test_gather_is_the_tensor_natural_spline_of_its_stencil():
    """Random fields: the gathered value equals scipy's natural cubic spline through the 4 nodes,
    applied along x, then y, then z over the 4 x 4 x 4 block the stencil selects."""
    s = decks.small_setup(O)
    m = O.Model(s)
    rng = np.random.default_rng(22)
    for comp in range(6):
        m.set(comp, rng.standard_normal(m.shape(comp)))
    g = m.grid()
    rx = np.stack([rng.uniform(g["X"][2], g["X"][-3], 40), rng.uniform(g["Y"][2], g["Y"][-3], 40),
                   rng.uniform(g["Z"][1], g["Z"][-4], 40)], axis=1)
    got = m.rx_values(rx, list(range(6)))
    for c in range(6):
        F = m.get(c)
        for r, (x, y, z) in enumerate(rx):
            st, t = m.rx_stencil(x, y, z, c)
            blk = F[st[2]:st[2] + 4, st[1]:st[1] + 4, st[0]:st[0] + 4]
            vx = CubicSpline(t[0], blk, axis=2, bc_type="natural")(x)
            vy = CubicSpline(t[1], vx, axis=1, bc_type="natural")(y)
            want = float(CubicSpline(t[2], vy, bc_type="natural")(z))
            assert got[r, c] == pytest.approx(want, rel=1e-11, abs=1e-12), (c, r)


def test_gather_exact_on_linear_fields():
    s = decks.small_setup(O)
    m = O.Model(s)
    g = m.grid()
    pos = {"c": (g["X"], g["Y"], g["Z"]), "f": (_face(g["X"], g["dx"]), _face(g["Y"], g["dy"]), _face(g["Z"], g["dz"]))}
    a, b, c0, d = 0.3, 1.7e-3, -2.2e-3, 0.9e-3
    for comp in range(6):
        n = m.shape(comp)
        xs = pos[_STAGGER[comp][0]][0][:n[2]]
        ys = pos[_STAGGER[comp][1]][1][:n[1]]
        zs = pos[_STAGGER[comp][2]][2][:n[0]]
        m.set(comp, a + b * xs[None, None, :] + c0 * ys[None, :, None] + d * zs[:, None, None])
    rng = np.random.default_rng(23)
    rx = np.stack([rng.uniform(g["X"][2], g["X"][-3], 30), rng.uniform(g["Y"][2], g["Y"][-3], 30),
                   rng.uniform(g["Z"][1], g["Z"][-4], 30)], axis=1)
    got = m.rx_values(rx, list(range(6)))
    want = a + b * rx[:, 0] + c0 * rx[:, 1] + d * rx[:, 2]
    assert np.allclose(got, want[:, None], rtol=1e-12, atol=1e-12)


def _stepwise_states(m, s, rt_times, rx, tx, nsteps, ecomps, bcomps):
    """The field states of a manual run: after step n the E fields stand for time
    t_E(n) = toT + dt1 and the B fields for t_B(n) = t_E(n) + dt2 / 2, with (toT, dt1, dt2) the
    values of step n before the time advance (main.cu:227-248, the comments at the E and B
    calls)."""
    rt_long = O.make_rxtime(1e-3, 1e3, 4, 1, [1] * 6)
    lam, dte, dtb, amp = O.schedule(s, m.mindx(), rt_long, nsteps + 2)
    assert len(dte) >= nsteps + 1
    m.zero()
    tp = m.locate_tx(tx)
    vE, vB = [np.zeros((len(rx), len(ecomps)))], [np.zeros((len(rx), len(bcomps)))]  # the zero state at t = 0
    tE, tB = [0.0], [0.0]
    toT = 0.0
    for n in range(nsteps):
        m.step(lam[n], dte[n], amp[n], tp, dtb[n])
        tE.append(toT + dte[n])
        tB.append(toT + dte[n] + 0.5 * dte[n + 1])
        toT += dte[n]
        vE.append(m.rx_values(rx, ecomps))
        vB.append(m.rx_values(rx, bcomps))
    return (np.array(tE), np.array(vE)), (np.array(tB), np.array(vB))


def _linear_in_time(ts, states):
    t, v = states
    assert np.all(np.diff(t) > 0)
    n = np.searchsorted(t, ts, side="left")  # t[n-1] < ts <= t[n]
    assert np.all(n >= 1) and np.all(n < len(t))
    w = (ts - t[n - 1]) / (t[n] - t[n - 1])
    return (1.0 - w)[:, None, None] * v[n - 1] + w[:, None, None] * v[n], n


def _time_case(rt_args):
    # 200 source sub-steps (the reference's golden waveform size): monotone state times; with far
    # fewer sub-steps the reference's equal-area construction steps back in time
    s = decks.small_setup(O, n_src=200)
    m0 = O.Model(s)
    Z = m0.grid()["Z"][:, None, None] * np.ones((1, m0.M, m0.L))
    m = O.Model(s, decks.layered_cond((m0.N, m0.M, m0.L), Z, seed=4))
    rx = np.array([[160.0, 10.0, 400.0], [310.0, 70.0, 400.0], [-190.0, -40.0, 350.0]])
    tx = O.make_tx(1, 100.0, 0.0, 0.0, 300.0, [1, 2, 3])
    rt = O.make_rxtime(*rt_args, [1] * 6)
    ref, steps = m.run(rt, rx, tx)
    ts = O.times(rt)
    E, B = _stepwise_states(m, s, ts, rx, tx, steps, [0, 1, 2], [3, 4, 5])
    return ref, ts, E, B


def test_time_interpolation_is_linear_between_field_states():
    """Samples interpolated linearly in time between the two field states that bracket them,
    with the zero state at t = 0 before the first step (EFInterp2Rx.cu:472-585).  A swapped pair
    of weights or a state shifted by one step moves these by O(1) of the change per step."""
    ref, ts, E, B = _time_case((0.004, 0.05, 12, 1))
    for states, cs in ((E, slice(0, 3)), (B, slice(3, 6))):
        want, n = _linear_in_time(ts, states)
        gaps = np.diff(np.unique(n))
        assert np.all(gaps >= 2)  # no two samples in neighbouring step intervals (see the next test)
        got = np.moveaxis(ref[:, cs, :], 2, 0)  # [t][r][c]
        scale = np.abs(got).max(axis=0)
        assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300), np.abs(got - want).max()


def test_time_interpolation_of_samples_in_neighbouring_intervals():
    """The reference keeps the earlier state of an interval only when the sample it serves lies
    beyond the current state (slot 0 of ef0, EFInterp2Rx.cu:493-500); after interpolating (slot 1,
    :505-521) it does not move the current state into slot 0.  A sample whose predecessor fell in
    the immediately preceding step interval is therefore interpolated from a stale slot 0.  With
    dense samples, exactly the samples of such intervals differ from the linear interpolation;
    every other sample matches it."""
    ref, ts, E, B = _time_case((0.010, 0.016, 25, 0))
    for states, cs in ((E, slice(0, 3)), (B, slice(3, 6))):
        want, n = _linear_in_time(ts, states)
        got = np.moveaxis(ref[:, cs, :], 2, 0)
        scale = np.abs(got).max(axis=0)
        bad = np.abs(got - want) > 1e-12 * scale
        bad_t = bad.any(axis=(1, 2))
        prev = np.concatenate([[-10], n[:-1]])
        groups_first = n != prev  # first sample of each interval's group
        stale = np.zeros(len(ts), bool)
        for q in range(len(ts)):
            first = q if groups_first[q] else np.flatnonzero(groups_first[:q + 1])[-1]
            stale[q] = first > 0 and n[first] == n[first - 1] + 1
        assert stale.any() and (~stale).any()
        # every sample of a stale group differs where its interval's states differ
        moved = np.abs(states[1][n] - states[1][n - 1]).max(axis=(1, 2)) > 0
        assert np.array_equal(bad_t, stale & moved)
