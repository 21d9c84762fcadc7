"""Seeded synthetic decks for the benchmark configurations of this path.

Each builder returns a Deck whose FDTD_setup.in / Cond.bin / Receiver.in / Transmitter.in are
representable in the reference's own file formats (``Deck.write(dir)``), so every workload can
also be run through the drop-in driver.  Random draws use numpy's PCG64 with the seed recorded in
the deck.

  marine_1d   64 x 64 x 48 cells: 32 x 50 m horizontal core with 16 geometric padding cells per
              side (growth 1.0502 then 1.723, so the regular air plane is 256 x 256 and the 21
              seafloor receivers at 250 m inline fit); z log-stretched from 20 m to 200 m (one 20 m
              cell, 47 cells growing by 1.0502, 3.79 km total); 1000 m of 3.3 S/m seawater over
              1 S/m sediment; x-directed dipole 50 m above the seafloor.
  marine_3d   256 x 256 x 160 cells: 216 x 50 m horizontal core with 20 padding cells per side
              (growth 1.05 then 1.747: the air plane is 512 x 512, 7.4 km of padding); z: 110 x 20 m
              then 50 cells growing by 1.05 (to 229 m, 6.6 km total); the marine_1d background
              with seeded bathymetry (2-D Gaussian random field, std 50 m, covariance
              exp(-r^2/l^2), l = 2 km) and a 100 ohm-m slab of 4 x 3 km x 100 m whose top is 1 km
              below the nominal seafloor; 101 seafloor receivers from -5 to 5 km inline.
  towline     marine_3d background (no slab, flat seafloor) with a list of x-dipoles along a 6 km
              towline (spacing 6000/(n-1) m), all recording the same receivers.
  large_5     512 x 512 x 256 cells: 472 x 50 m horizontal core with 20 padding cells per side
              (growth 1.05 then 2.008: the air plane is 1024 x 1024); z: 200 x 20 m then 56 cells
              growing by 1.05; seawater over rough seeded bathymetry (std 100 m, l = 1 km, mean
              1000 m), ten sediment layers of log-uniform 0.5-5 ohm-m (seeded thicknesses of
              150-450 m), two 100 ohm-m reservoirs (4 x 3 km x 80 m and 3 x 2 km x 60 m); three
              x-dipoles 50 m above the local seafloor, 401 receivers 10 m above it from -10 to
              10 km inline.
  uneven_models  64 seeded models whose grids are drawn uniformly between 96 x 96 x 96 and
              224 x 224 x 160 cells (even L, M; 20 padding cells per side as marine_3d, 40 z cells
              growing by 1.05 under a 20 m core), 1 ohm-m sediment under 0.5-2 km of seawater
              with 5-10 axis-aligned ellipsoids of log-uniform 0.3-300 ohm-m; one x-dipole 50 m
              above the seafloor and 41 seafloor receivers from -1 to 1 km inline each.  All
              share dxmin, dymin, dzmin and the source, hence one time-step schedule.
"""
import os
from dataclasses import dataclass, field

import numpy as np

SEA, SEDIMENT = 3.3, 1.0


@dataclass
class Deck:
    name: str
    setup: dict
    cond: np.ndarray            # (N, M, L) FP64
    rxtime: dict                # T0 Tn Nt is_log comps
    rx: np.ndarray              # (nrx, 3)
    txs: list                   # dicts for make_tx
    seed: int = 0
    steps: int = 0              # nominal step count of the configuration
    notes: str = ""
    grid: dict = field(default_factory=dict)

    def write(self, d):
        """Write the four reference input files into directory d."""
        os.makedirs(d, exist_ok=True)
        s = self.setup
        with open(os.path.join(d, "FDTD_setup.in"), "w") as f:
            f.write("%r\t%r\n" % (s["sfa1"], s["sfa2"]))
            f.write("%d\t%r\n" % (s["n_src"], s["sigma_t"]))
            f.write("%r\t%r\t%r\n" % (s["x0"], s["y0"], s["z0"]))
            f.write("%r\t%r\t%r\n" % (s["dxmin"], s["dymin"], s["dzmin"]))
            f.write("%r\t%r\n" % (s["arg1"], s["arg2"]))
            f.write("%d %d %d %d %d\n" % (s["ALx2"], s["ALx1"], s["Nx"], s["ARx1"], s["ARx2"]))
            f.write("%d %d %d\n" % (s["Ay2"], s["Ay1"], s["Ny"]))
            f.write("%d %d %d\n" % (s["Az2"], s["Az1"], s["Nz"]))
        np.ascontiguousarray(self.cond, dtype="<f8").tofile(os.path.join(d, "Cond.bin"))
        r = self.rxtime
        with open(os.path.join(d, "Receiver.in"), "w") as f:
            f.write("%r %r\n%d %d\n" % (float(r["T0"]), float(r["Tn"]), r["Nt"], r["is_log"]))
            f.write(" ".join(str(int(c)) for c in r["comps"]) + "\n")
            f.write("%d\n" % len(self.rx))
            for i, (x, y, z) in enumerate(self.rx):
                f.write("%d %r %r %r\n" % (i + 1, float(x), float(y), float(z)))
        with open(os.path.join(d, "Transmitter.in"), "w") as f:
            f.write("%d\n" % len(self.txs))
            for i, t in enumerate(self.txs):
                ids = " ".join(str(v) for v in t["rx_ids"])
                if t["ort"] < 4:
                    f.write("%d %d %r %r %r %r %d %s\n" % (i + 1, t["ort"], float(t["bipolen"]),
                                                          float(t["x"]), float(t["y"]), float(t["z"]),
                                                          len(t["rx_ids"]), ids))
                else:
                    raise NotImplementedError("decks here use ort 1-3")


def _grid(setup_mod, setup):
    """Cell sizes and centres from the reference formulas (via the product's host restatement)."""
    s = setup_mod.make_setup(**setup)
    return setup_mod.grid_arrays(s)


def _grf(X, Y, std, corr, seed):
    """Seeded 2-D Gaussian random field with covariance std^2 exp(-r^2/corr^2), sampled at the cell
    centres X (L), Y (M); generated by spectral filtering on a 50 m lattice with 2 corr margin."""
    rng = np.random.default_rng(seed)
    h = 50.0
    x0, x1 = X.min() - 2 * corr, X.max() + 2 * corr
    y0, y1 = Y.min() - 2 * corr, Y.max() + 2 * corr
    nx, ny = int(np.ceil((x1 - x0) / h)) + 1, int(np.ceil((y1 - y0) / h)) + 1
    w = rng.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx, d=h) * 2 * np.pi
    ky = np.fft.fftfreq(ny, d=h) * 2 * np.pi
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    # spectral density of exp(-r^2/l^2) is proportional to exp(-k^2 l^2 / 4)
    amp = np.exp(-K2 * corr ** 2 / 8.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * amp))
    f *= std / f.std()
    ix = np.clip(np.round((X - x0) / h).astype(int), 0, nx - 1)
    iy = np.clip(np.round((Y - y0) / h).astype(int), 0, ny - 1)
    return f[np.ix_(iy, ix)]


def _setup(nx_core, pad_x, nz_core, az1, az2, dz=20.0, dx=50.0, arg1=1.05, arg2=1.15,
           n_src=200, sigma_t=1e-3):
    return dict(sfa1=2.0, sfa2=2.0, n_src=n_src, sigma_t=sigma_t, x0=0.0, y0=0.0, z0=0.0,
                dxmin=dx, dymin=dx, dzmin=dz, arg1=arg1, arg2=arg2, ALx2=pad_x[0],
                ALx1=pad_x[1], Nx=nx_core, ARx1=pad_x[1], ARx2=pad_x[0], Ay2=pad_x[0],
                Ay1=pad_x[1], Ny=nx_core, Az2=az2, Az1=az1, Nz=nz_core)


def marine_1d(setup_mod, nt=60, tn=2.0):
    setup = _setup(32, (6, 10), 1, 47, 0, arg1=1.0502, arg2=1.723)
    g = _grid(setup_mod, setup)
    N, M, L = len(g["dz"]), len(g["dy"]), len(g["dx"])
    cond = np.full((N, M, L), SEDIMENT)
    cond[g["Z"] < 1000.0] = SEA
    rx = np.array([[x, 0.0, 1000.0] for x in np.arange(-2500.0, 2501.0, 250.0)])
    txs = [dict(ort=1, bipolen=50.0, x=0.0, y=0.0, z=950.0, rx_ids=list(range(1, len(rx) + 1)))]
    return Deck("marine_1d 64x64x48", setup, cond,
                dict(T0=0.01, Tn=tn, Nt=nt, is_log=1, comps=(1, 1, 1, 1, 1, 1)), rx, txs,
                steps=2000, grid=g)


def marine_3d(setup_mod, seed=20261202, slab=True, bathymetry=True, nt=100, tn=50.0):
    setup = _setup(216, (6, 14), 110, 50, 0, arg1=1.05, arg2=1.747)
    g = _grid(setup_mod, setup)
    N, M, L = len(g["dz"]), len(g["dy"]), len(g["dx"])
    if bathymetry:
        floor = 1000.0 + _grf(g["X"], g["Y"], 50.0, 2000.0, seed)   # (M, L)
    else:
        floor = np.full((M, L), 1000.0)
    Z = g["Z"][:, None, None]
    cond = np.where(Z < floor[None, :, :], SEA, SEDIMENT).astype(np.float64)
    if slab:
        X, Y = g["X"][None, None, :], g["Y"][None, :, None]
        inside = (np.abs(X) <= 2000.0) & (np.abs(Y) <= 1500.0) & (Z >= 2000.0) & (Z < 2100.0)
        cond = np.where(inside, 0.01, cond)
    rx = np.array([[x, 0.0, 1000.0] for x in np.linspace(-5000.0, 5000.0, 101)])
    txs = [dict(ort=1, bipolen=50.0, x=0.0, y=0.0, z=950.0, rx_ids=list(range(1, len(rx) + 1)))]
    return Deck("marine_3d 256x256x160", setup, cond,
                dict(T0=0.01, Tn=tn, Nt=nt, is_log=1, comps=(1, 1, 1, 1, 1, 1)), rx, txs,
                seed=seed, steps=20000, grid=g)


def towline(setup_mod, ntx=32, nt=100, tn=50.0):
    d = marine_3d(setup_mod, slab=False, bathymetry=False, nt=nt, tn=tn)
    xs = np.linspace(-3000.0, 3000.0, ntx)
    d.txs = [dict(ort=1, bipolen=50.0, x=float(x), y=0.0, z=950.0,
                  rx_ids=list(range(1, len(d.rx) + 1))) for x in xs]
    d.name = "towline %d Tx x 256x256x160" % ntx
    return d


def large_5(setup_mod, seed=20261203, nt=100, tn=50.0):
    setup = _setup(472, (6, 14), 200, 56, 0, arg1=1.05, arg2=2.008)
    g = _grid(setup_mod, setup)
    N, M, L = len(g["dz"]), len(g["dy"]), len(g["dx"])
    rng = np.random.default_rng(seed)
    floor = 1000.0 + _grf(g["X"], g["Y"], 100.0, 1000.0, seed)   # (M, L)
    Z = g["Z"][:, None, None]
    depth = Z - floor[None, :, :]                                 # below the local seafloor
    rho = 10.0 ** rng.uniform(np.log10(0.5), np.log10(5.0), 10)
    tops = np.concatenate([[0.0], np.cumsum(rng.uniform(150.0, 450.0, 9))])
    layer = np.searchsorted(tops, depth, side="right") - 1          # -1 in the water
    cond = np.where(layer < 0, SEA, 1.0 / rho[np.clip(layer, 0, 9)])
    X, Y = g["X"][None, None, :], g["Y"][None, :, None]
    res1 = (np.abs(X - 1500.0) <= 2000.0) & (np.abs(Y) <= 1500.0) & (depth >= 1200.0) & (depth < 1280.0)
    res2 = (np.abs(X + 3000.0) <= 1500.0) & (np.abs(Y - 1000.0) <= 1000.0) & (depth >= 1800.0) & (depth < 1860.0)
    cond = np.where(res1 | res2, 0.01, cond).astype(np.float64)

    def floor_at(x, y):
        return float(floor[np.argmin(np.abs(g["Y"] - y)), np.argmin(np.abs(g["X"] - x))])

    rx = np.array([[x, 0.0, floor_at(x, 0.0) - 10.0] for x in np.linspace(-10000.0, 10000.0, 401)])
    ids = list(range(1, len(rx) + 1))
    txs = [dict(ort=1, bipolen=50.0, x=x, y=0.0, z=floor_at(x, 0.0) - 50.0, rx_ids=ids)
           for x in (-2000.0, 0.0, 2000.0)]
    return Deck("large_5 512x512x256", setup, cond,
                dict(T0=0.01, Tn=tn, Nt=nt, is_log=1, comps=(1, 1, 1, 1, 1, 1)), rx, txs, seed=seed,
                steps=50000, grid=g)


def uneven_models(setup_mod, n=64, seed=20261204, nt=40, tn=5.0):
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        L = 2 * int(rng.integers(48, 113))
        M = 2 * int(rng.integers(48, 113))
        N = int(rng.integers(96, 161))
        setup = _setup(L - 40, (6, 14), N - 40, 40, 0, arg1=1.05, arg2=1.747)
        setup["Ny"] = M - 40
        g = _grid(setup_mod, setup)
        depth = float(rng.uniform(500.0, 2000.0))
        Z = g["Z"][:, None, None]
        X, Y = g["X"][None, None, :], g["Y"][None, :, None]
        log_rho = np.where(Z < depth, np.log10(1.0 / SEA), 0.0) * np.ones((1, len(g["Y"]), len(g["X"])))
        for _ in range(int(rng.integers(5, 11))):
            cx, cy = rng.uniform(-1500.0, 1500.0, 2)
            cz = depth + rng.uniform(150.0, 1200.0)
            ax, ay = rng.uniform(200.0, 1500.0, 2)
            az = rng.uniform(30.0, 300.0)
            inside = ((X - cx) / ax) ** 2 + ((Y - cy) / ay) ** 2 + ((Z - cz) / az) ** 2 <= 1.0
            log_rho = np.where(inside, rng.uniform(np.log10(0.3), np.log10(300.0)), log_rho)
        cond = (10.0 ** -log_rho).astype(np.float64)
        rx = np.array([[x, 0.0, depth - 1.0] for x in np.linspace(-1000.0, 1000.0, 41)])
        txs = [dict(ort=1, bipolen=50.0, x=0.0, y=0.0, z=depth - 50.0, rx_ids=list(range(1, len(rx) + 1)))]
        out.append(Deck("uneven %d: %dx%dx%d" % (q, L, M, N), setup, cond,
                        dict(T0=0.01, Tn=tn, Nt=nt, is_log=1, comps=(1, 1, 1, 1, 1, 1)), rx, txs,
                        seed=seed, grid=g))
    return out

