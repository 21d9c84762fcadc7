"""Physics pins against closed-form transient responses (oracle/analytic.py): the surface of a
homogeneous half space (source and receivers at z = 0, the geometry of the reference's LOTEM
analytic tags, AnalyticalSolution.cpp:240-300) and a homogeneous whole space (source and receivers
2 km deep, well before the diffusion front reaches the surface).

The closed forms themselves are pinned first by their time integrals (the textbook turn-off step
response and the DC dipole field).  The model traces are then compared with the impulse responses
convolved with the moment the reference's bipole injection actually carries per step
(analytic.step_source), from the peak of the analytic transient on: before it the DuFort-Frankel
fictitious-displacement wave dominates (a property of the scheme, not of this code).  The oracle
runs on the CPU; the HIP path (FP64 and FP32) on the GPU.  Tolerances are relative L2 over those
samples: the grid (25-50 m cells at 300-700 m offsets) limits agreement to a few percent.
"""
import numpy as np
import pytest
from scipy import integrate

import tcsem_fdtd_amd as T
from oracle import analytic as A
from oracle import oracle as O

SIG = 1.0
HALF_TOL = 0.08
WHOLE_TOL = 0.06

HALF = dict(setup=dict(sfa1=2.0, sfa2=2.0, n_src=200, sigma_t=1e-3, x0=0.0, y0=0.0, z0=0.0, dxmin=50.0,
                       dymin=50.0, dzmin=25.0, arg1=1.1, arg2=1.3, ALx2=8, ALx1=10, Nx=40, ARx1=10, ARx2=8,
                       Ay2=8, Ay1=10, Ny=40, Az2=8, Az1=10, Nz=20),
            rx=[[300.0, 0.0, 0.0], [500.0, 0.0, 0.0], [700.0, 0.0, 0.0], [0.0, 500.0, 0.0], [400.0, 300.0, 0.0]],
            times=(0.008, 0.2, 40), z=0.0, sigma_bar=0.5 * SIG)
WHOLE = dict(setup=dict(sfa1=2.0, sfa2=2.0, n_src=200, sigma_t=1e-3, x0=0.0, y0=0.0, z0=0.0, dxmin=50.0,
                        dymin=50.0, dzmin=50.0, arg1=1.1, arg2=1.3, ALx2=8, ALx1=10, Nx=40, ARx1=10, ARx2=8,
                        Ay2=8, Ay1=10, Ny=40, Az2=6, Az1=10, Nz=80),
             rx=[[300.0, 0.0, 2000.0], [400.0, 0.0, 2000.0], [500.0, 0.0, 2000.0], [0.0, 400.0, 2000.0],
                 [300.0, 300.0, 2000.0]],
             times=(0.004, 0.08, 30), z=2000.0, sigma_bar=SIG)
ELL = 100.0


def test_halfspace_kernel_integrates_to_textbook_turnoff():
    for rho, t in ((300.0, 0.01), (500.0, 0.03), (700.0, 0.2)):
        tail, _ = integrate.quad(lambda s: A.halfspace_surface_ex_impulse(rho, 0.0, SIG, s), t, np.inf,
                                 epsabs=0, epsrel=1e-10, limit=200)
        assert tail == pytest.approx(A.halfspace_surface_ex_turnoff(rho, 0.0, SIG, t), rel=1e-8)
    # the kernel is the same for any source-receiver azimuth (t > 0)
    assert A.halfspace_surface_ex_impulse(300.0, 400.0, SIG, 0.02) == pytest.approx(
        A.halfspace_surface_ex_impulse(500.0, 0.0, SIG, 0.02), rel=1e-14)


@pytest.mark.parametrize("xyz", [(300.0, 0.0, 0.0), (0.0, 300.0, 0.0), (200.0, 150.0, -100.0)])
def test_wholespace_kernel_integrates_to_dc_dipole_field(xyz):
    x, y, z = xyz
    r = np.sqrt(x * x + y * y + z * z)
    f = lambda s: A.wholespace_ex_impulse(x, y, z, SIG, s)
    head, _ = integrate.quad(f, 0.0, 1.0, epsabs=0, epsrel=1e-11, limit=400, points=[1e-4, 1e-3, 1e-2, 1e-1])
    tail, _ = integrate.quad(f, 1.0, np.inf, epsabs=0, epsrel=1e-11, limit=400)
    dc = head + tail
    assert dc == pytest.approx((3 * x * x / r ** 2 - 1) / (4 * np.pi * SIG * r ** 3), rel=1e-7)


def _analytic(case, rt_mod, rt, so, mindx):
    rx = np.array(case["rx"])
    lam, dte, dtb, amp = O.schedule(so, mindx, rt, 400)
    Tn, sn = A.step_source(lam, dte, amp, case["sigma_bar"])
    times = O.times(rt)
    out = []
    for r in range(len(rx)):
        if case["z"] == 0.0:
            ker = lambda dx, dy, tt: A.halfspace_surface_ex_impulse(dx, dy, SIG, tt)
        else:
            ker = lambda dx, dy, tt: A.wholespace_ex_impulse(dx, dy, 0.0, SIG, tt)
        h = lambda tt, r=r: A.wire_response(lambda dx, dy: ker(dx, dy, tt), rx[r, 0], rx[r, 1], 0.0, 0.0, ELL)
        out.append(A.convolve_steps(h, times, Tn, sn))
    return np.array(out)


def _late_error(got, ana):
    errs = []
    for g, a in zip(got, ana):
        k0 = int(np.argmax(np.abs(a)))
        sel = slice(k0, None)
        errs.append(np.linalg.norm(g[sel] - a[sel]) / np.linalg.norm(a[sel]))
    return np.array(errs)


def _case_inputs(case):
    so = O.make_setup(**case["setup"])
    t0, tn, nt = case["times"]
    rt = O.make_rxtime(t0, tn, nt, 1, [1, 0, 0, 0, 0, 0])
    tx = dict(ort=1, bipolen=ELL, x=0.0, y=0.0, z=case["z"], rx_ids=list(range(1, len(case["rx"]) + 1)))
    return so, rt, tx


@pytest.mark.parametrize("name", ["half", "whole"])
def test_oracle_matches_closed_form(name):
    case = HALF if name == "half" else WHOLE
    so, rt, tx = _case_inputs(case)
    m = O.Model(so)
    m = O.Model(so, np.full((m.N, m.M, m.L), SIG))
    efs, steps = m.run(rt, np.array(case["rx"]), O.make_tx(**tx))
    ana = _analytic(case, O, rt, so, m.mindx())
    err = _late_error(efs[:, 0], ana)
    assert np.all(err < (HALF_TOL if name == "half" else WHOLE_TOL)), err
    # the sign and the moment of the reference's source: a 30 % moment error would fail above,
    # an inverted source would give errors near 2
    assert np.all(np.sign(efs[:, 0][np.arange(len(ana)), np.argmax(np.abs(ana), axis=1)]) ==
                  np.sign(ana[np.arange(len(ana)), np.argmax(np.abs(ana), axis=1)]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["half", "whole"])
@pytest.mark.parametrize("precision", [64, 32])
def test_hip_path_matches_closed_form(name, precision):
    case = HALF if name == "half" else WHOLE
    so, rt_o, tx = _case_inputs(case)
    m = O.Model(so)
    st = T.make_setup(**case["setup"])
    t0, tn, nt = case["times"]
    rt = T.make_rxtime(t0, tn, nt, 1, [1, 0, 0, 0, 0, 0])
    print("closed form %s FP%d" % (name, precision), flush=True)
    with T.Context(st, np.full((m.N, m.M, m.L), SIG), precision) as c:
        c.set_receivers(rt, np.array(case["rx"]))
        c.set_transmitters([T.make_tx(**tx)])
        got = c.run(0, 1)[0]
    ana = _analytic(case, O, rt_o, so, m.mindx())
    err = _late_error(got[:, 0], ana)
    assert np.all(err < (HALF_TOL if name == "half" else WHOLE_TOL)), err
