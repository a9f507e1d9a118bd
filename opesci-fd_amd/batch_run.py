#!/usr/bin/env python3
"""Batched propagators: many independent eigenwave / heterogeneous models of
uneven shape and spatial order stepped together on one GPU (one graph launch
per two steps, autotuned interior launch shapes), against the same models
run one after another.

Workload (defaults): 64 instances, interior points per axis drawn uniformly
from 48..160 (not cubic), order from {2, 4, 8}, every other instance in a
seeded smooth heterogeneous medium, FP32, 200 steps.

  python batch_run.py [--count 64] [--steps 200] [--json out.json]

Reports aggregate grid-point updates/s of the batch and of the one-by-one
runs, and the largest per-instance L2 difference between the two (the
kernels are identical, so it is expected to be exactly 0).
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), 'tests'))

from opesci.backend import Context  # noqa: E402
from opesci.batch import Batch  # noqa: E402
from workloads import batch_specs, box_model  # noqa: E402


def _lower(args):
    spec, steps = args
    g, _ = box_model(spec['sizes'], spec['order'], steps, hetero=spec['hetero'])
    return g.lower()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--count', type=int, default=64)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--workers', type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument('--json', default=None)
    args = ap.parse_args(argv)
    specs = batch_specs(args.count)
    t = time.time()
    with ProcessPoolExecutor(args.workers) as ex:
        plans = list(ex.map(_lower, [(s, args.steps) for s in specs]))
    pts = [float(np.prod(s['sizes'])) for s in specs]
    print('%d models lowered (%.1f s), %.1f M interior points' % (len(plans), time.time() - t, sum(pts) / 1e6),
          flush=True)
    t = time.time()
    with Batch(plans) as b:
        print('contexts ready and tuned (%.1f s): %s' % (time.time() - t, sorted(set(b.configs))), flush=True)
        b.run(2)                     # warm-up: builds the batch graph
        for c in b.contexts:
            c.init()
        secs = b.run(args.steps)
        l2_batch = [c.l2() for c in b.contexts]
        configs = list(b.configs)
    batch_rate = sum(pts) * args.steps / secs
    print('batch: %.3f s, %.2f G pt/s' % (secs, batch_rate / 1e9), flush=True)
    seq = 0.0
    diff = 0.0
    for p, cfg, l2b in zip(plans, configs, l2_batch):
        with Context(p) as c:
            if cfg is not None:
                c.configure(*cfg)
            c.init()
            c.run(2)
            c.init()
            seq += c.run(args.steps)
            l2 = c.l2()
        diff = max(diff, max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(l2b, l2)))
    seq_rate = sum(pts) * args.steps / seq
    res = dict(workload='%d instances, per-axis interior sizes 48..160, orders {2,4,8}, half heterogeneous, FP32, '
                        '%d steps' % (args.count, args.steps),
               interior_points=sum(pts), batch_seconds=secs, batch_gpts_per_s=batch_rate,
               one_by_one_seconds=seq, one_by_one_gpts_per_s=seq_rate, speedup=seq / secs,
               max_rel_l2_difference=diff, configs={str(k): configs.count(k) for k in set(configs)})
    out = json.dumps(res)
    if args.json:
        with open(args.json, 'w') as f:
            f.write(out + '\n')
    print(out)


if __name__ == '__main__':
    main()
