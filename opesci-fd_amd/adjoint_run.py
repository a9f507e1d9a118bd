#!/usr/bin/env python3
"""Full-size misfit gradient: dJ/d(rho, lambda, mu) of a receiver misfit by
the hand-written adjoint, FP32 and FP64, with central finite differences.

Workload (defaults): 4th order, 256^3 interior points on the unit cube
(dx = 1/255 length units), 1000 steps, fields at rest at t = 0, Ricker
source f0 = 12 Hz (peak at t = 1/f0) in Txx/Tyy/Tzz at interior point
(128, 128, 16), a 128 x 128 receiver plane at interior z = 16 (x, y =
64..191) recording U, V, W.  Starting model: the seeded smooth medium (Vp
1.5-4.5 length units/s, Vs = Vp/sqrt(3), rho 1.0-2.5, smoothing sigma 8
cells); true model: the same with Vp raised by 10 % in a Gaussian of 20
cells full width at half maximum centred at interior (128, 128, 80).
dt = 0.8 x 0.495 dx / Vp_max.  Checkpoints every 32 steps.

Steps: observed traces from the true model (FP32); the FP32 gradient through
ElasticMisfit (timed; device tensors when a device is present); the FP64
gradient; the FP32 gradient's relative L2 difference from FP64 per
parameter, over the whole grid and over the cells more than two cells
(Chebyshev distance) from the source, plus the relative difference in the
source cell itself; central finite differences of the FP64
misfit (h = 1e-3 and 1e-5 of the value) at 16 distinct seeded points drawn
from the 20,000 largest gradient entries of each parameter (rho, lambda and
mu in turn).  Near the source the misfit is far from quadratic in one cell's
value, so h = 1e-3 carries truncation error there; h = 1e-5 does not.

The source cell's gradient is the time integral of the product of the
forward stresses there (driven directly by the source, the largest values on
the grid) and the adjoint strain rates, which nearly cancel over the
source's duration: its FP32 value carries the rounding of the largest
forward values against a small net sum, so its FP32-vs-FP64 difference is
reported on its own, apart from the rest of the grid.

  python adjoint_run.py [--size 256] [--steps 1000] [--every 32] [--fd 16] [--no-fp64] [--json out.json]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from opesci import eigenwave3d  # noqa: E402
from opesci.autograd import ElasticMisfit, misfit  # noqa: E402
from opesci.backend import Context  # noqa: E402
from opesci.media import smooth_random_media  # noqa: E402
from opesci.sources import ricker  # noqa: E402

PARAMS = ('rho', 'lambda', 'mu')


def log(msg):
    print(msg, flush=True)


def build(N, order, steps, media, double, f0, src, plane):
    n = N - 1
    dx = 1.0 / n
    dt = 0.8 * 0.495 * dx / float(media[1].max())
    g = eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5), accuracy_order=[2, order, order, order],
                    double=double, media_arrays=media, initial='zero', o_converge=False)
    g.set_source(ricker(f0, dt, steps), index=src)
    g.set_receivers(index=plane)
    return g, dt


def elastic(rho, vp, vs):
    rho, vp, vs = (np.asarray(a, np.float64) for a in (rho, vp, vs))
    return rho, rho * (vp * vp - 2.0 * vs * vs), rho * vs * vs


def gradient(ctx, model, obs, every):
    dev = 'cuda' if torch.cuda.is_available() else 'cpu'
    t = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in model]
    o = torch.from_numpy(np.ascontiguousarray(obs)).to(dev)
    if dev == 'cuda':
        torch.cuda.synchronize()
    t0 = time.time()
    J = ElasticMisfit.apply(*t, ctx, o, every)
    J.backward()
    if dev == 'cuda':
        torch.cuda.synchronize()
    secs = time.time() - t0
    return float(J.detach().cpu()), {k: v.grad.detach().cpu().numpy() for k, v in zip(PARAMS, t)}, secs


def split_rel_l2(a, b, src, r=2):
    """(relative L2 of a - b beyond Chebyshev distance r of src, relative
    difference at src)."""
    mask = np.ones(a.shape, bool)
    mask[tuple(slice(max(0, c - r), c + r + 1) for c in src)] = False
    far = float(np.linalg.norm((a - b)[mask]) / np.linalg.norm(b[mask]))
    at = float(abs(a[src] - b[src]) / abs(b[src])) if b[src] != 0 else float(abs(a[src]))
    return far, at


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--size', type=int, default=256)
    ap.add_argument('--order', type=int, default=4)
    ap.add_argument('--steps', type=int, default=1000)
    ap.add_argument('--every', type=int, default=32)
    ap.add_argument('--f0', type=float, default=12.0)
    ap.add_argument('--fd', type=int, default=16, help='finite-difference points')
    ap.add_argument('--fd-steps', type=float, nargs='+', default=[1e-3, 1e-5],
                    help='relative finite-difference steps')
    ap.add_argument('--no-fp64', action='store_true', help='FP32 gradient only (no FP64 run, no differences)')
    ap.add_argument('--json', default=None)
    args = ap.parse_args(argv)
    N, order, steps = args.size, args.order, args.steps
    m = order // 2
    dims = (N + order,) * 3
    # interior coordinates scaled from the 256^3 layout
    sc = N / 256.0
    at = lambda v: m + int(round(v * sc))
    src = (at(128), at(128), at(16))
    lo, e = at(64), max(1, int(round(128 * sc)))
    plane = [(lo + i, lo + j, at(16)) for j in range(e) for i in range(e)]
    t = time.time()
    rho, vp, vs = smooth_random_media(dims, seed=2016, sigma=8.0)
    xx, yy, zz = np.meshgrid(*[np.arange(d, dtype=np.float32) for d in dims], indexing='ij')
    sig = 20.0 * sc / 2.3548
    blob = np.exp(-((xx - at(128)) ** 2 + (yy - at(128)) ** 2 + (zz - at(80)) ** 2) / (2 * sig * sig))
    del xx, yy, zz
    vp_true = (vp * (1.0 + 0.1 * blob)).astype(np.float32)
    start, true = elastic(rho, vp, vs), elastic(rho, vp_true, vs)
    log('media ready (%.1f s)' % (time.time() - t))
    result = dict(workload='elastic %d^3 interior, order %d, %d steps, Ricker %g Hz at interior %s, receiver plane '
                           '%dx%d at interior z %d, Vp +10%% Gaussian (FWHM %.0f cells), checkpoints every %d'
                           % (N, order, steps, args.f0, tuple(v - m for v in src), e, e, at(16) - m, 20 * sc,
                              args.every))

    # FP32: observed data, gradient
    g32, dt = build(N, order, steps, (rho, vp, vs), False, args.f0, src, plane)
    ctx = Context(g32.lower())
    t = time.time()
    misfit(ctx, *(torch.from_numpy(a) for a in true), np.zeros((steps, len(plane), 3), np.float32))
    obs = ctx.receivers()[:, :, :3].copy()
    log('observed traces (%.1f s)' % (time.time() - t))
    ctx.init()
    fwd = ctx.run()
    J32, g32r, secs32 = gradient(ctx, start, obs, args.every)
    t = time.time()
    ctx.gradient(obs, args.every)
    dev32 = time.time() - t
    log('fp32 gradient: J %.6e, %.2f s through autograd, %.2f s on the device (forward alone %.3f s)'
        % (J32, secs32, dev32, fwd))
    ctx.close()
    result['fp32'] = dict(J=J32, gradient_s=secs32, device_gradient_s=dev32, forward_s=fwd,
                          gpts_per_s_forward=N ** 3 * steps / fwd)

    if args.no_fp64:
        print(json.dumps(result))
        return
    # FP64 gradient and finite differences
    g64, _ = build(N, order, steps, (rho, vp, vs), True, args.f0, src, plane)
    ctx = Context(g64.lower())
    J64, g64r, secs64 = gradient(ctx, start, obs, args.every)
    log('fp64 gradient: J %.6e, %.2f s' % (J64, secs64))
    result['fp64'] = dict(J=J64, gradient_s=secs64)
    result['fp32_vs_fp64_rel_l2'] = {
        k: float(np.linalg.norm(g32r[k] - g64r[k]) / np.linalg.norm(g64r[k])) for k in PARAMS}
    split = {k: split_rel_l2(g32r[k], g64r[k], src) for k in PARAMS}
    result['fp32_vs_fp64_rel_l2_beyond_2_cells_of_source'] = {k: v[0] for k, v in split.items()}
    result['fp32_vs_fp64_rel_at_source'] = {k: v[1] for k, v in split.items()}
    log('fp32 vs fp64 gradient: %s; beyond 2 cells of the source %s; at the source %s'
        % (result['fp32_vs_fp64_rel_l2'], result['fp32_vs_fp64_rel_l2_beyond_2_cells_of_source'],
           result['fp32_vs_fp64_rel_at_source']))
    # the bare device gradient (media already resident), without the host-side
    # effective-media transform and its backward
    t = time.time()
    ctx.gradient(obs, args.every)
    result['fp64']['device_gradient_s'] = time.time() - t
    rng = np.random.default_rng(16)
    picks = []
    per = [args.fd // 3 + (1 if i < args.fd % 3 else 0) for i in range(3)]
    for k, cnt in zip(PARAMS, per):
        a = np.abs(g64r[k]).ravel()
        top = np.argpartition(a, -20000)[-20000:] if a.size > 20000 else np.arange(a.size)
        for j in rng.choice(top, size=cnt, replace=False):
            picks.append((k, tuple(int(v) for v in np.unravel_index(int(j), g64r[k].shape))))
    fd_rows = []
    for k, idx in picks:
        row = dict(param=k, index=idx, adjoint=float(g64r[k][idx]), fp32=float(g32r[k][idx]))
        for rel in args.fd_steps:
            vals = []
            h = rel * abs(start[PARAMS.index(k)][idx])
            for sgn in (1, -1):
                mod = [a.copy() if p == k else a for p, a in zip(PARAMS, start)]
                mod[PARAMS.index(k)][idx] += sgn * h
                vals.append(misfit(ctx, *(torch.from_numpy(a) for a in mod), obs))
            fd = (vals[0] - vals[1]) / (2 * h)
            row['fd_h%g' % rel] = fd
            row['rel_h%g' % rel] = abs(row['adjoint'] - fd) / abs(fd)
        fd_rows.append(row)
        log('fd %s' % json.dumps(row))
    ctx.close()
    result['finite_differences'] = fd_rows
    for rel in args.fd_steps:
        result['fd_max_rel_h%g' % rel] = max(r['rel_h%g' % rel] for r in fd_rows) if fd_rows else None
    out = json.dumps(result)
    if args.json:
        with open(args.json, 'w') as f:
            f.write(out + '\n')
    print(out)


if __name__ == '__main__':
    main()
