"""Small synthetic TCSEM decks shared by the tests (seeded; the same inputs feed the HIP path
and the FP64 oracle)."""
import numpy as np

SEA = 3.3
SEDIMENT = 1.0


def small_setup(mod, nx=24, ny=20, nz=14, pad=(2, 4), dz=50.0, n_src=40, sigma_t=1e-3, arg=(1.1, 1.2),
                dx=50.0):
    """A setup with `pad` = (outer, inner) geometric cells on every side."""
    return mod.make_setup(sfa1=2.0, sfa2=2.0, n_src=n_src, sigma_t=sigma_t, x0=0.0, y0=0.0,
                          z0=0.0, dxmin=dx, dymin=dx, dzmin=dz, arg1=arg[0], arg2=arg[1],
                          ALx2=pad[0], ALx1=pad[1], Nx=nx, ARx1=pad[1], ARx2=pad[0],
                          Ay2=pad[0], Ay1=pad[1], Ny=ny, Az2=pad[0], Az1=pad[1], Nz=nz)


def layered_cond(shape, Z, sea_depth=400.0, seed=None, blob=False):
    """Seawater above `sea_depth`, sediment below; optional seeded resistive blob."""
    N, M, L = shape
    c = np.full(shape, SEDIMENT)
    c[np.asarray(Z) < sea_depth] = SEA
    if seed is not None:
        rng = np.random.default_rng(seed)
        c *= np.exp(0.3 * rng.standard_normal(shape))
    if blob:
        c[N // 2:N // 2 + 2, M // 3:2 * M // 3, L // 3:2 * L // 3] = 0.01
    return c


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def trace_error(got, ref, floor=1e-6):
    """Per receiver: ||a-b|| / ||b|| over all times and components, components whose norm is
    below floor x the receiver's largest component norm excluded (symmetry nulls)."""
    errs = []
    for r in range(ref.shape[0]):
        norms = np.linalg.norm(ref[r], axis=1)
        keep = norms >= floor * norms.max()
        errs.append(rel_l2(got[r][keep], ref[r][keep]))
    return np.array(errs)
