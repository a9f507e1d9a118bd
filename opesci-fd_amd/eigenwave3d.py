#!/usr/bin/env python3
"""Standalone eigenwave 3-D driver with the command line of the reference's
tests/eigenwave3d.py (:281-330): modes default / read / converge, and the
flags -so, -c, -x, -n, -o, -p, --papi-events.  `-c hip` (or no -c)
selects the gfx950 backend; -c with a CPU compiler name is refused, since
this package has no CPU execution path.

  python eigenwave3d.py default -c hip -x          # 100^3 cells, dt 0.002, tmax 1
  python eigenwave3d.py read -c hip -x --rho RHO --vp VP --vs VS
  python eigenwave3d.py converge -c hip -x

Pluto tiling and loop fission (--pluto, --tile, --fission) belong to the
reference's CPU code path and are not available here.
"""
import os
import sys
from argparse import ArgumentParser, RawTextHelpFormatter

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from opesci import eigenwave3d  # noqa: E402

_HERE = os.getcwd()


def _run(grid, filename, compiler, execute, nthreads):
    if compiler is None:
        grid.generate(filename)
        return
    grid.compile(filename, compiler=compiler, shared=True)
    if execute:
        grid.execute(filename, compiler=compiler, nthreads=nthreads)
        grid.convergence()
        grid.free()


def default(compiler=None, execute=False, nthreads=1, accuracy_order=(2, 4, 4, 4), output=False,
            profiling=False, papi_events=()):
    """Eigenwave test case on a unit cube grid (100 x 100 x 100 cells)."""
    grid = eigenwave3d((1.0, 1.0, 1.0), (100, 100, 100), 0.002, 1.0, accuracy_order=list(accuracy_order),
                       o_converge=True, verbose=True)
    grid.set_switches(output_vts=output, profiling=profiling)
    grid.set_papi_events(list(papi_events))
    _run(grid, os.path.join(_HERE, 'eigenwave3d.h'), compiler, execute, nthreads)
    return grid


def read_data(compiler=None, execute=False, nthreads=1, accuracy_order=(2, 4, 4, 4), output=False,
              profiling=False, papi_events=(), rho_file='RHOhomogx200', vp_file='VPhomogx200',
              vs_file='VShomogx200'):
    """Model from raw float32 rho/Vp/Vs files on a unit cube grid (195^3 cells)."""
    grid = eigenwave3d((1.0, 1.0, 1.0), (195, 195, 195), 0.002, 1.0, read=True,
                       accuracy_order=list(accuracy_order), o_converge=False, rho_file=rho_file,
                       vp_file=vp_file, vs_file=vs_file, verbose=True)
    grid.set_switches(output_vts=output, profiling=profiling)
    grid.set_papi_events(list(papi_events))
    _run(grid, os.path.join(_HERE, 'eigenwave3d_read.h'), compiler, execute, nthreads)
    return grid


def converge_test(compiler=None, execute=False, nthreads=1, sizes=(10, 20, 40, 80), tmax=5.0):
    """Convergence of the (2,4) scheme: spacing halves from 0.1, dt shrinks
    fourfold each time (dt = 0.4 s / s^2)."""
    grids = []
    for s in sizes:
        c = 0.4 * sizes[0]
        dt = c / (s ** 2)
        grid = eigenwave3d((1.0, 1.0, 1.0), (s, s, s), dt, tmax, o_converge=True, verbose=True)
        _run(grid, os.path.join(_HERE, 'test3d_%d.h' % s), compiler, execute, nthreads)
        grids.append(grid)
    return grids


def main(argv=None):
    mode_help = """Available testing modes:
default:   Eigenwave test case on a unit cube grid (100 x 100 x 100)

read:      Model initialisation from raw float32 rho/Vp/Vs files;
           eigenwave on a unit cube grid (195 x 195 x 195 cells)

converge:  Convergence test of the (2,4) scheme, which is 2nd order
           in time and 4th order in space: spacing halves from 0.1 and
           dt shrinks by a factor of 4 each step
"""
    p = ArgumentParser(description='Standalone testing script for the Eigenwave3D example',
                       formatter_class=RawTextHelpFormatter)
    p.add_argument('mode', choices=('default', 'read', 'converge'), nargs='?', default='default', help=mode_help)
    p.add_argument('-so', '--spatial_order', default=4, type=int, dest='so',
                   help='order of the spatial discretisation (2, 4, 6 or 8)')
    p.add_argument('-c', '--compiler', default=None, help="'hip' for the gfx950 backend; omit to only generate")
    p.add_argument('-x', '--execute', action='store_true', default=False, help='execute the generated model')
    p.add_argument('-n', '--nthreads', type=int, default=1, help='accepted for compatibility (GPU execution)')
    p.add_argument('-o', '--output', action='store_true', default=False,
                   help='write U after every time step as U_<step>.vts')
    p.add_argument('-p', '--profiling', action='store_true', default=False, help='print timing and MFlops/s')
    p.add_argument('--papi-events', dest='papi_events', nargs='+', default=[],
                   help='accepted for compatibility; rocprofv3 measures counters on this backend')
    p.add_argument('--rho', default='RHOhomogx200', help='read mode: raw float32 density file')
    p.add_argument('--vp', default='VPhomogx200', help='read mode: raw float32 Vp file')
    p.add_argument('--vs', default='VShomogx200', help='read mode: raw float32 Vs file')
    p.add_argument('--tile', default=None, help='not available on this backend')
    p.add_argument('--pluto', action='store_true', default=False, help='not available on this backend')
    p.add_argument('--fission', action='store_true', default=False, help='not available on this backend')
    args = p.parse_args(argv)
    if args.pluto or args.tile or args.fission:
        p.error('pluto tiling and loop fission belong to the CPU code path, not available on the gfx950 backend')
    if args.compiler not in (None, 'hip'):
        p.error('compiler %r: this backend executes on the GPU only (use -c hip)' % args.compiler)
    print('Eigenwave3D example (mode=%s)' % args.mode)
    order = [2, args.so, args.so, args.so]
    if args.mode == 'default':
        default(compiler=args.compiler, execute=args.execute, nthreads=args.nthreads, output=args.output,
                accuracy_order=order, profiling=args.profiling, papi_events=args.papi_events)
    elif args.mode == 'read':
        read_data(compiler=args.compiler, execute=args.execute, nthreads=args.nthreads, output=args.output,
                  accuracy_order=order, profiling=args.profiling, papi_events=args.papi_events,
                  rho_file=args.rho, vp_file=args.vp, vs_file=args.vs)
    else:
        converge_test(compiler=args.compiler, execute=args.execute, nthreads=args.nthreads)


if __name__ == '__main__':
    main()
