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
