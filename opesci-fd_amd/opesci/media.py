"""Synthetic heterogeneous media for the read path (test and benchmark inputs).

Seeded uniform noise, Gaussian-smoothed and mapped affinely onto a velocity
and a density range; Vs = Vp/sqrt(3) (Poisson solid, lambda = mu).  Written
as the raw float32 rho/Vp/Vs arrays of dim1*dim2*dim3 values that the
reference read path consumes (opesci/staggeredgrid.py:522-551).
"""
import numpy as np

__all__ = ['smooth_random_media', 'write_media_files', 'stable_dt']


def _smooth(a, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(a, sigma=sigma, mode='nearest')


def _map(a, lo, hi):
    amin, amax = float(a.min()), float(a.max())
    if amax == amin:
        return np.full(a.shape, 0.5 * (lo + hi))
    return lo + (a - amin) * (hi - lo) / (amax - amin)


def smooth_random_media(dims, seed=2016, sigma=8.0, vp_range=(1.5, 4.5), rho_range=(1.0, 2.5)):
    """(rho, vp, vs) float32 arrays of shape dims."""
    rng = np.random.default_rng(seed)
    nv = rng.uniform(size=dims)
    nr = rng.uniform(size=dims)
    vp = _map(_smooth(nv, sigma), *vp_range).astype(np.float32)
    rho = _map(_smooth(nr, sigma), *rho_range).astype(np.float32)
    vs = (vp / np.float32(np.sqrt(3.0))).astype(np.float32)
    return rho, vp, vs


def write_media_files(prefix, rho, vp, vs):
    names = []
    for tag, a in (('RHO', rho), ('VP', vp), ('VS', vs)):
        fn = '%s%s.bin' % (prefix, tag)
        np.ascontiguousarray(a, dtype=np.float32).tofile(fn)
        names.append(fn)
    return names


def stable_dt(dx, vp_max, order=4, fraction=0.5):
    """fraction x the CFL bound of StaggeredGrid.get_time_step_limit
    (h/(sqrt(3) Vp) at 2nd order, 0.495 h/Vp at 4th; the 4th-order bound is
    used for higher orders too)."""
    lim = dx / vp_max / 3 ** 0.5 if order == 2 else 0.495 * dx / vp_max
    return fraction * lim
