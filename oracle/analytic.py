"""Closed-form transient responses of a grounded electric source (TEST INFRASTRUCTURE).

The checker for the physics of the oracle and of the HIP path, used only by tests/.  Quasi-static
fields (no displacement current) of an x-directed electric dipole of unit moment driven by an
impulse current, in a homogeneous medium of conductivity sigma:

  whole space (Ward & Hohmann 1988, the frequency-domain dipole field of their whole-space section)
      E_x(s) = 1/(4 pi sigma r^3) e^{-a sqrt(s)} [ (x^2/r^2)(a^2 s + 3 a sqrt(s) + 3)
                                                 - (a^2 s + a sqrt(s) + 1) ],   a = r sqrt(mu0 sigma)
    with L^-1[e^{-a sqrt(s)}] = f(a, t) = a / (2 sqrt(pi) t^{3/2}) e^{-a^2/(4t)},
         L^-1[sqrt(s) e^{-a sqrt(s)}] = -df/da,  L^-1[s e^{-a sqrt(s)}] = d2f/da2;
  surface of a half space (source and receiver at z = 0)
      E_x(s) = 1/(2 pi sigma rho^3) [ 3x^2/rho^2 - 2 + (1 + a sqrt(s)) e^{-a sqrt(s)} ],
    whose response for t > 0 is the isotropic kernel
      e_x(t) = 1/(2 pi sigma rho^3) (2/sqrt(pi)) u^3 e^{-u^2} / t,   u = rho sqrt(mu0 sigma / (4t));
    its time integral gives the textbook turn-off response erf(u) - (2/sqrt(pi)) u e^{-u^2}
    (times 1/(2 pi sigma rho^3)), which pins the derivation.

The reference's own analytic module (AnalyticalSolution.cpp:240-300, tag 4, "Ex LOTEM impulse
response": a Hankel-transform quadrature with the 801-point tables of Hankel.h over the source
wire) cannot be built here (it needs alglib and fields its structs no longer hold); at z = 0 it
reduces to the line integral of the surface kernel above, which is what wire_response computes.

The simulated source is the reference's: a unit-area Gaussian current of width sigma_t centred at
3 sigma_t (SourceWave.cpp:8-58), nominally of moment 1 A m over a bipole of length ell
(ElectricalFiled.cu:532-641).  Its injection, however, carries more moment per step than the
Gaussian (step_source), so the model traces are the impulse responses convolved with the per-step
moments of step_source and averaged over the wire.
"""
import numpy as np

MU0 = 4e-7 * np.pi


def _f(a, t):
    return a / (2.0 * np.sqrt(np.pi) * t ** 1.5) * np.exp(-a * a / (4.0 * t))


def _df(a, t):
    return np.exp(-a * a / (4.0 * t)) / (2.0 * np.sqrt(np.pi) * t ** 1.5) * (1.0 - a * a / (2.0 * t))


def _d2f(a, t):
    return np.exp(-a * a / (4.0 * t)) / (2.0 * np.sqrt(np.pi) * t ** 1.5) * (a ** 3 / (4.0 * t * t) - 1.5 * a / t)


def wholespace_ex_impulse(x, y, z, sigma, t):
    """E_x at offset (x, y, z) from a unit x-dipole, impulse current, whole space; t > 0."""
    r = np.sqrt(x * x + y * y + z * z)
    a = r * np.sqrt(MU0 * sigma)
    f, df, d2f = _f(a, t), _df(a, t), _d2f(a, t)
    c = x * x / (r * r)
    return (c * (a * a * d2f - 3.0 * a * df + 3.0 * f) - (a * a * d2f - a * df + f)) / (4.0 * np.pi * sigma * r ** 3)


def halfspace_surface_ex_impulse(x, y, sigma, t):
    """E_x on the surface of a half space at offset (x, y) from a unit x-dipole on the surface,
    impulse current; t > 0 (the t = 0 impulse of the instantaneous term is left out)."""
    rho = np.sqrt(x * x + y * y)
    u = rho * np.sqrt(MU0 * sigma / (4.0 * t))
    return (2.0 / np.sqrt(np.pi)) * u ** 3 * np.exp(-u * u) / t / (2.0 * np.pi * sigma * rho ** 3)


def halfspace_surface_ex_turnoff(x, y, sigma, t):
    """Turn-off step response belonging to halfspace_surface_ex_impulse (its integral from t to
    infinity), the textbook form used to pin the kernel."""
    from scipy.special import erf
    rho = np.sqrt(x * x + y * y)
    u = rho * np.sqrt(MU0 * sigma / (4.0 * t))
    return (erf(u) - 2.0 / np.sqrt(np.pi) * u * np.exp(-u * u)) / (2.0 * np.pi * sigma * rho ** 3)


def gaussian_convolve(h, t, sigma_t, n=801):
    """(h * g)(t) for the unit-area Gaussian g centred at 3 sigma_t, truncated to [0, 6 sigma_t]
    (the extent of the reference's discretised source), by the trapezoid rule on n nodes."""
    tau = np.linspace(0.0, 6.0 * sigma_t, n)
    g = np.exp(-0.5 * ((tau - 3.0 * sigma_t) / sigma_t) ** 2) / (np.sqrt(2.0 * np.pi) * sigma_t)
    w = np.full(n, tau[1] - tau[0])
    w[0] *= 0.5
    w[-1] *= 0.5
    w *= g
    w /= w.sum()  # unit area after truncation, as the discretised source
    out = np.zeros_like(np.asarray(t, dtype=float))
    for k, tk in np.ndenumerate(np.asarray(t, dtype=float)):
        lag = tk - tau
        m = lag > 0
        out[k] = np.sum(w[m] * h(lag[m]))
    return out


def wire_response(kernel, rx, ry, x0, y0, ell, n=32):
    """Average of kernel(dx, dy) over an x-directed wire of length ell centred at (x0, y0)
    (Gauss-Legendre nodes): the field of a bipole of total moment 1 A m."""
    xg, wg = np.polynomial.legendre.leggauss(n)
    acc = 0.0
    for xi, wi in zip(xg, wg):
        acc = acc + 0.5 * wi * kernel(rx - (x0 + 0.5 * ell * xi), ry - y0)
    return acc


def step_source(lam, dte, amp, sigma_bar):
    """Times and moments of the current the reference's injection carries at each step.

    The bipole injection (ElectricalFiled.cu:347-475) subtracts 2 I / (sigma_bar d1 d2) from E per
    step, with I = W_n / l; the DuFort-Frankel E update of the same code (ElectricalFiled.cu:138-139,
    curl over spacing sums) takes a current density J as - 2 dt J / (2 lam + sigma_bar dt).  Equating
    the two, step n carries the moment W_n (dt_n + 2 lam_n / sigma_bar), i.e. more than the unit
    moment W_n dt_n of the Gaussian by the factor 1 + 2 lam_n / (sigma_bar dt_n).  sigma_bar is the
    conductivity average the injection uses at the source cells (half the earth value on the
    surface, where the air counts as zero).  Returns (T_n, s_n): E time after step n and moment."""
    lam, dte, amp = (np.asarray(v, dtype=float) for v in (lam, dte, amp))
    return np.cumsum(dte), amp * (dte + 2.0 * lam / sigma_bar)


def convolve_steps(h, t, T, s):
    """sum_n s_n h(t - T_n) over the steps that precede t."""
    keep = s != 0
    T, s = T[keep], s[keep]
    out = np.zeros(len(t))
    for k, tk in enumerate(t):
        lag = tk - T
        m = lag > 0
        out[k] = np.sum(s[m] * h(lag[m]))
    return out
