"""Named configurations shared by the tests and bench.py (sizes in 'N^3
interior' = N points per axis = grid_size N-1, dx = 1/(N-1))."""
import numpy as np

import opesci
from opesci.media import smooth_random_media, stable_dt


def eigen_model(N, order, steps, double=False, storage='native', legacy=False, lam=1.0, mu=1.0, rho=1.0,
                dt=None):
    """Eigenwave on the unit cube with N^3 interior points."""
    n = N - 1
    dx = 1.0 / n
    if dt is None:
        dt = 0.5 * dx / np.sqrt(3.0 * (lam + 2 * mu) / rho)
    vs = np.sqrt(mu / rho)
    vp = np.sqrt((lam + 2 * mu) / rho)
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5),
                           accuracy_order=[2, order, order, order], rho=rho, vp=vp, vs=vs,
                           double=double, storage=storage, legacy_levander=legacy)
    if lam is not None:
        g.set_variable('lambda', lam, 'float', True)
        g.set_variable('mu', mu, 'float', True)
        g._invalidate()
    return g, dict(order=order, n=(n, n, n), ntsteps=steps, dt=dt, rho=rho, lam=lam, mu=mu, double=double)


def hetero_model(N, order, steps, double=False, storage='native', seed=2016, sigma=8.0):
    n = N - 1
    dx = 1.0 / n
    dims = (N + order,) * 3
    rho, vp, vs = smooth_random_media(dims, seed=seed, sigma=sigma)
    dt = stable_dt(dx, float(vp.max()), order=4)
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5),
                           accuracy_order=[2, order, order, order], double=double, storage=storage,
                           media_arrays=(rho, vp, vs))
    return g, dict(order=order, n=(n, n, n), ntsteps=steps, dt=dt, double=double, media=(rho, vp, vs))


def box_model(sizes, order, steps, hetero=False, double=False, seed=2016, sigma=4.0, storage='native'):
    """Eigenwave (or the seeded smooth medium) on the unit cube with a
    different number of interior points per axis (dx_k = 1/(N_k - 1))."""
    n = tuple(int(v) - 1 for v in sizes)
    dmin = min(1.0 / v for v in n)
    if hetero:
        dims = tuple(int(v) + order for v in sizes)
        rho, vp, vs = smooth_random_media(dims, seed=seed, sigma=sigma)
        dt = stable_dt(dmin, float(vp.max()), order=4)
        g = opesci.eigenwave3d((1.0, 1.0, 1.0), n, dt, dt * (steps + 0.5), accuracy_order=[2, order, order, order],
                               double=double, storage=storage, media_arrays=(rho, vp, vs))
        return g, dict(order=order, n=n, ntsteps=steps, dt=dt, double=double, media=(rho, vp, vs))
    dt = 0.5 * dmin / np.sqrt(3.0 * 3.0)
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), n, dt, dt * (steps + 0.5), accuracy_order=[2, order, order, order],
                           rho=1.0, vp=np.sqrt(3.0), vs=1.0, double=double, storage=storage)
    g.set_variable('lambda', 1.0, 'float', True)
    g.set_variable('mu', 1.0, 'float', True)
    g._invalidate()
    return g, dict(order=order, n=n, ntsteps=steps, dt=dt, rho=1.0, lam=1.0, mu=1.0, double=double)


def batch_specs(count=64, seed=4, lo=48, hi=160, orders=(2, 4, 8)):
    """Seeded instance list: per-axis interior sizes uniform in [lo, hi],
    order drawn from `orders`, every other instance heterogeneous."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        sizes = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
        out.append(dict(sizes=sizes, order=int(rng.choice(orders)), hetero=bool(i % 2)))
    return out


def adjoint_model(N, order, steps, seed=5, double=True):
    """Small inversion case: seeded smooth medium, fields at rest, FP64 (or FP32),
    Ricker source (x100) in Txx/Tyy/Tzz at the centre, a receiver patch two
    cells below the top face plus three interior receivers.  Returns the grid
    and FP64 rho, lambda, mu."""
    from opesci.sources import ricker
    n = N - 1
    dims = (N + order,) * 3
    rho, vp, vs = smooth_random_media(dims, seed=seed, sigma=2.0)
    dx = 1.0 / n
    dt = 0.4 * 0.495 * dx / float(vp.max())
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5), accuracy_order=[2, order, order, order],
                           double=double, media_arrays=(rho, vp, vs), initial='zero', o_converge=False)
    m = order // 2
    c = m + n // 2
    g.set_source(ricker(20.0, dt, steps, t0=6 * dt) * 100.0, index=(c, c, c))
    rec = [(m + i, m + j, m + 2) for i in range(2, n - 1, 3) for j in range(2, n - 1, 3)]
    rec += [(c + 2, c - 1, c + 3), (m + 1, c, c), (c, m + 1, c)]
    g.set_receivers(index=rec)
    rho64, vp64, vs64 = (a.astype(np.float64) for a in (rho, vp, vs))
    lam = rho64 * (vp64 ** 2 - 2 * vs64 ** 2)
    mu = rho64 * vs64 ** 2
    return g, rho64, lam, mu


def lambda_blob(lam, amp=0.1, width=2.5):
    """lambda raised by `amp` in a Gaussian blob off the centre."""
    d = np.array(lam.shape)
    xx, yy, zz = np.meshgrid(*[np.arange(s) for s in d], indexing='ij')
    blob = np.exp(-((xx - d[0] * 0.4) ** 2 + (yy - d[1] * 0.6) ** 2 + (zz - d[2] * 0.5) ** 2) / (2 * width ** 2))
    return lam * (1 + amp * blob)
