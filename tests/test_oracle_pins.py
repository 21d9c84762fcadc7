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
