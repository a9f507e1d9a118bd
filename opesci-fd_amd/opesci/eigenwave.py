"""The canonical elastic model description: the 3-D eigenwave of
tests/eigenwave3d.py:9-146, written against this package's API.

Nine fields, unit-cube staggered grid, constant or read media, analytic
eigenmode (U = cos(pi x)(sin(pi y) - sin(pi z)) cos(Omega t) and cyclic;
Txx = -A sin(pi x)(sin(pi y) - sin(pi z)) sin(Omega t) and cyclic; shear 0;
Omega = pi sqrt(2 mu beta), A = sqrt(2 mu / beta)), the nine velocity-stress
PDEs and free surfaces on all six faces.
"""
from sympy import Eq, Float, cos, pi, sin, sqrt, symbols

from .fields import SField, VField
from .staggeredgrid import StaggeredGrid

__all__ = ['eigenwave3d']


class _NoSolution(object):
    """Field proxy that ignores set_analytic_solution (fields start at rest)."""

    def __init__(self, field):
        self._f = field

    def set_analytic_solution(self, expr):
        pass

    def __getattr__(self, name):
        return getattr(self._f, name)


def eigenwave3d(domain_size, grid_size, dt, tmax, output_vts=False, o_converge=True,
                accuracy_order=(2, 4, 4, 4), omp=True, simd=False, ivdep=True, double=False,
                pluto=False, filename='test.cpp', read=False, expand=True, eval_const=True,
                rho_file='', vp_file='', vs_file='', fission=False, rho=1.0, vp=1.0, vs=0.5,
                lam=None, mu=None, media_arrays=None, storage='native', legacy_levander=False,
                verbose=False, initial='eigenmode'):
    """Build and return the eigenwave StaggeredGrid (same arguments as the
    reference, plus: rho/vp/vs or lam/mu for constant media, media_arrays =
    (rho, vp, vs) numpy arrays for the read path, storage = 'native'|'fp16'|'bf16',
    legacy_levander = the older generator's surfaces without the normal-stress
    recompute, as in benchmarks/test3d.cpp; initial = 'zero' starts from rest
    instead of the eigenmode, for source-driven runs)."""
    if verbose:
        print('domain size: ' + str(domain_size))
        print('grid size: ' + str(grid_size))
        print('approximation order: ' + str(list(accuracy_order)))
        print('dt: ' + str(dt))
        print('tmax: ' + str(tmax))
    Txx = SField('Txx', dimension=3, direction=(1, 1))
    Tyy = SField('Tyy', dimension=3, direction=(2, 2))
    Tzz = SField('Tzz', dimension=3, direction=(3, 3))
    Txy = SField('Txy', dimension=3, direction=(1, 2))
    Tyz = SField('Tyz', dimension=3, direction=(2, 3))
    Txz = SField('Txz', dimension=3, direction=(1, 3))
    U = VField('U', dimension=3, direction=1)
    V = VField('V', dimension=3, direction=2)
    W = VField('W', dimension=3, direction=3)
    grid = StaggeredGrid(dimension=3, domain_size=domain_size, grid_size=grid_size,
                         stress_fields=[Txx, Tyy, Tzz, Txy, Tyz, Txz],
                         velocity_fields=[U, V, W], pluto=pluto, fission=fission)
    grid.set_time_step(dt, tmax)
    grid.set_switches(omp=omp, simd=simd, ivdep=ivdep, double=double, expand=expand,
                      eval_const=eval_const, output_vts=output_vts, converge=o_converge,
                      storage=storage, legacy_levander=legacy_levander)
    rho_s, beta, lam_s, mu_s = symbols('rho beta lambda mu')
    t, x, y, z = symbols('_t x y z')
    grid.set_index([x, y, z])
    if read or media_arrays is not None:
        if media_arrays is not None:
            grid.set_media_params(read=True, rho_array=media_arrays[0], vp_array=media_arrays[1],
                                  vs_array=media_arrays[2])
        else:
            grid.set_media_params(read=True, rho_file=rho_file, vp_file=vp_file, vs_file=vs_file)
    else:
        grid.set_media_params(read=False, rho=rho, vp=vp, vs=vs)
        if lam is not None and mu is not None:
            grid.set_variable('lambda', lam, 'float', True)
            grid.set_variable('mu', mu, 'float', True)
    if verbose:
        print('require dt < ' + str(grid.get_time_step_limit()))
    Omega = pi * sqrt(2 * mu_s * beta)
    A = sqrt(2 * mu_s / beta)
    if initial not in ('eigenmode', 'zero'):
        raise ValueError("initial must be 'eigenmode' or 'zero'")
    if initial == 'zero':
        U = _NoSolution(U)
        V, W, Txx, Tyy, Tzz, Txy, Tyz, Txz = (_NoSolution(f) for f in (V, W, Txx, Tyy, Tzz, Txy, Tyz, Txz))
    U.set_analytic_solution(cos(pi * x) * (sin(pi * y) - sin(pi * z)) * cos(Omega * t))
    V.set_analytic_solution(cos(pi * y) * (sin(pi * z) - sin(pi * x)) * cos(Omega * t))
    W.set_analytic_solution(cos(pi * z) * (sin(pi * x) - sin(pi * y)) * cos(Omega * t))
    Txx.set_analytic_solution(-A * sin(pi * x) * (sin(pi * y) - sin(pi * z)) * sin(Omega * t))
    Tyy.set_analytic_solution(-A * sin(pi * y) * (sin(pi * z) - sin(pi * x)) * sin(Omega * t))
    Tzz.set_analytic_solution(-A * sin(pi * z) * (sin(pi * x) - sin(pi * y)) * sin(Omega * t))
    Txy.set_analytic_solution(Float(0))
    Tyz.set_analytic_solution(Float(0))
    Txz.set_analytic_solution(Float(0))
    grid.set_order(list(accuracy_order))
    grid.calc_derivatives()
    eqs = [
        Eq(U.d[0][1], beta * (Txx.d[1][1] + Txy.d[2][1] + Txz.d[3][1])),
        Eq(V.d[0][1], beta * (Txy.d[1][1] + Tyy.d[2][1] + Tyz.d[3][1])),
        Eq(W.d[0][1], beta * (Txz.d[1][1] + Tyz.d[2][1] + Tzz.d[3][1])),
        Eq(Txx.d[0][1], (lam_s + 2 * mu_s) * U.d[1][1] + lam_s * (V.d[2][1] + W.d[3][1])),
        Eq(Tyy.d[0][1], (lam_s + 2 * mu_s) * V.d[2][1] + lam_s * (U.d[1][1] + W.d[3][1])),
        Eq(Tzz.d[0][1], (lam_s + 2 * mu_s) * W.d[3][1] + lam_s * (U.d[1][1] + V.d[2][1])),
        Eq(Txy.d[0][1], mu_s * (U.d[2][1] + V.d[1][1])),
        Eq(Tyz.d[0][1], mu_s * (V.d[3][1] + W.d[2][1])),
        Eq(Txz.d[0][1], mu_s * (U.d[3][1] + W.d[1][1])),
    ]
    grid.solve_fd(eqs)
    for d in (1, 2, 3):
        for side in (0, 1):
            grid.set_free_surface_boundary(dimension=d, side=side)
    if verbose:
        print('stress kernel AI')
        print('%.2f, %.2f (weighted), %d ADD, %d MUL, %d LOAD, %d STORE' % grid.get_stress_kernel_ai())
        print('velocity kernel AI')
        print('%.2f, %.2f (weighted), %d ADD, %d MUL, %d LOAD, %d STORE' % grid.get_velocity_kernel_ai())
    return grid
