#!/usr/bin/env python3
"""Field-storage precision sweep: the same source-driven elastic run with
fields stored in HBM as FP32, FP16 and BF16 (FP32 arithmetic), each
compared with an FP64 run, per field and on a receiver plane.

Workload (default = the 8th-order 512^3 case): N^3 interior points on the
unit cube (dx = 1/(N-1) length units), seeded smooth heterogeneous medium
(Vp 1.5-4.5 length units/s, Vs = Vp/sqrt(3), rho 1.0-2.5, Gaussian smoothing
sigma = 8 cells), free surfaces on all faces, fields at rest at t = 0,
Ricker source f0 = 15 Hz (peak at t = 1/f0, wavelengths 0.1-0.3 units)
injected into Txx/Tyy/Tzz at the grid centre, dt = 0.8 x 0.495 dx / Vp_max
(s; the 8th-order stability bound is dx / (sqrt(3) Vp sum|w_k|) = 0.449 dx /
Vp; the fastest waves travel 396 cells in 1000 steps, so they reach the
plane), a receiver plane of up to 256 x 256 points at grid
index z = 8 recording U, V, W and pressure.  In these units stresses and
velocities are of the same magnitude (impedance rho Vp = 1.5-11), which
16-bit storage needs: with SI units (impedance ~1e6) FP16 would flush the
velocities to zero.

  python precision_sweep.py [--size 512] [--order 8] [--steps 1000] [--json out.json]

Errors reported: for every field ||F_s - F_64|| / ||F_64|| over the grid
after the last step, and the same over all receiver traces.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from opesci import eigenwave3d  # noqa: E402
from opesci.backend import Context  # noqa: E402
from opesci.media import smooth_random_media  # noqa: E402
from opesci.sources import ricker  # noqa: E402

FIELDS = ('U', 'V', 'W', 'Txx', 'Tyy', 'Tzz', 'Txy', 'Tyz', 'Txz')


def build(N, order, steps, media, double, storage, f0, plane):
    n = N - 1
    dx = 1.0 / n
    vpmax = float(media[1].max())
    dt = 0.8 * 0.495 * dx / vpmax
    g = eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5), accuracy_order=[2, order, order, order],
                    double=double, storage=storage, media_arrays=media, initial='zero', o_converge=False)
    m = order // 2
    c = m + n // 2
    g.set_source(ricker(f0, dt, steps), index=(c, c, c))
    g.set_receivers(index=plane)
    return g, dt


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--size', type=int, default=512)
    ap.add_argument('--order', type=int, default=8)
    ap.add_argument('--steps', type=int, default=1000)
    ap.add_argument('--f0', type=float, default=15.0)
    ap.add_argument('--plane', type=int, default=256, help='receiver plane edge (points)')
    ap.add_argument('--json', default=None)
    args = ap.parse_args(argv)
    N, order, steps = args.size, args.order, args.steps
    m = order // 2
    dims = (N + order,) * 3
    t0 = time.time()
    media = smooth_random_media(dims, seed=2016, sigma=8.0)
    e = min(args.plane, N)
    x0 = m + (N - e) // 2
    plane = [(x0 + i, x0 + j, 8) for j in range(e) for i in range(e)]
    print('media ready (%.1f s)' % (time.time() - t0), flush=True)

    g64, dt = build(N, order, steps, media, True, 'native', args.f0, plane)
    ref = Context(g64.lower())
    ref.init()
    t = time.time()
    ref.run()
    print('fp64 run: %.2f s' % (time.time() - t), flush=True)
    tr64 = ref.receivers().astype(np.float64)
    result = dict(workload='elastic %d^3 interior, order %d, %d steps, unit cube, dt %.6g s, Ricker %g Hz at the '
                           'centre, receiver plane %dx%d at z index 8' % (N, order, steps, dt, args.f0, e, e),
                  storages={})
    for storage in ('native', 'fp16', 'bf16'):
        g, _ = build(N, order, steps, media, False, storage, args.f0, plane)
        ctx = Context(g.lower())
        ctx.init()
        t = time.time()
        secs = ctx.run()
        wall = time.time() - t
        tr = ctx.receivers().astype(np.float64)
        per = {}
        for f, name in enumerate(FIELDS):
            a = ctx.download(f).astype(np.float64)
            b = ref.download(f)
            nb = float(np.sqrt(np.sum(b * b)))
            per[name] = float(np.sqrt(np.sum((a - b) ** 2)) / nb) if nb > 0 else 0.0
        ntr = float(np.sqrt(np.sum(tr64 ** 2)))
        rec = float(np.sqrt(np.sum((tr - tr64) ** 2)) / ntr)
        kind = 'fp32' if storage == 'native' else storage
        result['storages'][kind] = dict(field_rel_l2=per, receiver_rel_l2=rec,
                                        gpts_per_s=N ** 3 * steps / secs, wall_s=wall)
        print(kind, json.dumps(result['storages'][kind]), flush=True)
        ctx.close()
    ref.close()
    out = json.dumps(result)
    if args.json:
        with open(args.json, 'w') as f:
            f.write(out + '\n')
    print(out)


if __name__ == '__main__':
    main()
