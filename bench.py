"""Benchmark: elastic grid-point updates/s on one MI355X.

Workload: the eigenwave of tests/eigenwave3d.py at 256^3 interior points
(grid_size 255, dims 260^3), 4th order in space / 2nd in time, FP32,
constant media rho = lambda = mu = 1, dt = 0.5 dx / sqrt(3 (lambda+2mu)/rho),
Levander free surfaces on all faces.  A step is one full leapfrog step
(stress update, stress boundary launches, velocity update, velocity boundary
launches) over the whole grid, run by libopesci_hip through its C ABI with
the fields resident in HBM.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--copy-probe MIB]

--copy-probe runs three device-to-device copies of MIB MiB before the
workload: a plain copy kernel of known size against which rocprofv3's
FETCH_SIZE / WRITE_SIZE counters are scaled.  The HBM traffic of the stress
kernel measured that way is read from profiles/*_pmc.md (lines starting with
"traffic:"); it is null for a workload without such a line.

Prints one JSON line.  The roofline entry is for the stress kernel (the
dominant launch: 9 field loads + 6 stores per point); its duration is taken
with HIP events on the context stream.  cpu_baseline times the oracle (the
CPU/OpenMP restatement of the reference's generated code, reference GNU
flags) on a bounded number of steps of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, 'opesci-fd_amd'), os.path.join(ROOT, 'tests'), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PROFILES = os.path.join(ROOT, 'profiles')
METRIC = 'elastic grid-pt updates/s and % of HBM roofline, 1 MI355X, at stated L2 error'
HBM_PEAK = 8.0e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--size', type=int, default=256, help='interior points per axis')
    ap.add_argument('--order', type=int, default=4)
    ap.add_argument('--storage', default='native')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-steps', type=int, default=100)
    ap.add_argument('--variant', type=int, default=None, help='interior kernel variant (0 sweep, 1 gather, 2 one-point-per-lane sweep)')
    ap.add_argument('--x-chunk', type=int, default=0)
    ap.add_argument('--tile-y', type=int, default=0)
    ap.add_argument('--copy-probe', type=int, default=0, help='MiB copied device-to-device 3x first')
    ap.add_argument('--hetero', action='store_true',
                    help='seeded smooth heterogeneous medium through the read path (8 more media streams)')
    ap.add_argument('--fused-step', action='store_true',
                    help='both half steps in one sweep over the rows clear of the x and y faces (opt-in, '
                         'slower than the separate sweeps on MI355X)')
    ap.add_argument('--no-split-z', action='store_true',
                    help='keep the deep rows of the stress z-face loops in the boundary launches '
                         '(default: split off and applied by the stress sweep)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on exactly one GPU: this backend has no multi-GPU path '
              '(the slab decomposition is a design note only); use --gpus 1', file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from opesci.backend import Context
    from workloads import eigen_model, hetero_model

    hip = _hip()
    if args.copy_probe:
        copy_probe(hip, args.copy_probe << 20)
    N, order, K, W = args.size, args.order, args.steps, args.warmup
    if args.hetero:
        g, kw = hetero_model(N, order, K + W, storage=args.storage)
    else:
        g, kw = eigen_model(N, order, K + W, storage=args.storage)
    if args.no_split_z:
        g.fuse_boundary = False
    plan = g.lower()
    plan.fuse_step = args.fused_step
    if args.variant is not None:
        plan.kernel_variant = args.variant
    plan.x_chunk = args.x_chunk
    plan.tile_y = args.tile_y
    pts = float(N) ** 3
    elem = 2 if args.storage in ('fp16', 'bf16') else 4
    with Context(plan) as ctx:
        ctx.init()
        ctx.run(W)
        _check(hip.hipDeviceSynchronize(), 'hipDeviceSynchronize')
        t0 = time.perf_counter()
        secs = ctx.run(K)   # HIP events on the context stream, synchronised on return
        _check(hip.hipDeviceSynchronize(), 'hipDeviceSynchronize')
        wall = time.perf_counter() - t0
        l2 = ctx.l2()
        kt = ctx.kernel_times(10)   # per-launch averages, after the timed region
        bc_stats = ctx.stats(full=True)
        fused_F = ctx.fused_margin()
    # the traffic record of the stress launch that ran: the pair sweep with
    # or without the deep z-face statements
    workload_key = 'stress_sweep %d^3 order %d %s%s%s' % (N, order, args.storage, ' hetero' if args.hetero else '',
                                                         ' z-faces in sweep' if bc_stats[3] else '')
    # flops per point-step: ADD + MUL of the stress and velocity kernels in the
    # reference's own operation count (get_*_kernel_ai)
    flops_pt = sum(ai[2] + ai[3] for ai in (g.get_stress_kernel_ai(), g.get_velocity_kernel_ai()))
    ms_step = secs * 1e3 / K
    value = pts * K / secs
    media_streams = (5, 3) if args.hetero else (0, 0)   # lambda, mu, mu12/13/23 | beta1..3 (FP32)
    stress_bytes = pts * (15 * elem + 4 * media_streams[0])   # 9 loads + 6 stores (+ media) per point
    achieved = stress_bytes / (kt[0] * 1e-3)
    step_bytes = pts * (27 * elem + 4 * sum(media_streams))   # reference LOAD/STORE model, both kernels
    out = {
        'metric': METRIC,
        'value': value,
        'unit': 'grid-pt updates/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': W,
        'ms_per_step': ms_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32' if elem == 4 else 'f32 compute, %s storage' % args.storage,
        'data': 'synthetic',
        'config': {
            'workload': ('eigenwave3d %d^3 interior (dims %d^3), order %d space / 2 time, %s, ' % (N, N + order, order,
                                                                                         args.storage)) +
                        ('seeded smooth medium (Vp 1.5-4.5, rho 1-2.5, sigma 8), read path, dt = 0.5 x CFL'
                         if args.hetero else 'rho=lambda=mu=1, dt=dx/6') + ', Levander surfaces',
            'grid_points': int(pts),
            'l2_error_vs_analytic': dict(zip(('U', 'V', 'W', 'Txx', 'Tyy', 'Tzz', 'Txy', 'Tyz', 'Txz'),
                                             [float('%.6e' % v) for v in l2])),
            'step_hbm_roofline_frac': value * step_bytes / pts / HBM_PEAK,
            'wall_s': wall,
            'kernel_ms': {'stress': kt[0], 'velocity': kt[1], 'boundary': kt[2]},
            'flops_per_point_step': flops_pt,
            'mflops': flops_pt * value / 1e6,
            'boundary_launches_per_step': bc_stats[0],
            'deep_z_face_statements': bc_stats[1],
            'deep_z_faces_in_stress_sweep': bool(bc_stats[3]),
            'fused_step_margin': fused_F,
        },
        'roofline': {
            'bound': 'hbm',
            'achieved': achieved / 1e9,
            'peak': HBM_PEAK / 1e9,
            'unit': 'GB/s',
            'frac': achieved / HBM_PEAK,
            'traffic': committed_traffic(workload_key),
            'kernel': 'stress update (6 stresses at t1 from U,V,W and stresses at t0%s)'
                      % (', deep z-face statements in place' if bc_stats[3] else ''),
            'algorithmic_bytes_per_launch': stress_bytes,
        },
    }
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(N, order, args.cpu_steps, args.hetero)
    print(json.dumps(out))


def _hip():
    import ctypes
    hip = ctypes.CDLL('libamdhip64.so')
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError('%s failed with HIP error %d' % (what, rc))


def copy_probe(hip, nbytes):
    """Three hipMemcpyDeviceToDevice of nbytes (the runtime's copy kernel)."""
    import ctypes
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _check(hip.hipMalloc(ctypes.byref(a), nbytes), 'hipMalloc')
    _check(hip.hipMalloc(ctypes.byref(b), nbytes), 'hipMalloc')
    for _ in range(3):
        _check(hip.hipMemcpy(b, a, nbytes, 3), 'hipMemcpy')   # 3 = hipMemcpyDeviceToDevice
    _check(hip.hipDeviceSynchronize(), 'hipDeviceSynchronize')
    hip.hipFree(a)
    hip.hipFree(b)


def committed_traffic(key):
    """HBM bytes per launch from a committed PMC summary, or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(PROFILES, '*_pmc.md'))):
        with open(path) as f:
            for line in f:
                if line.startswith('traffic:'):
                    rec = json.loads(line[len('traffic:'):])
                    if rec.get('workload') == key:
                        return float(rec['bytes_per_launch'])
    return None


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(N, order, steps, hetero=False, runs=3):
    """Median of `runs` timed runs of the oracle's time loop, built with the
    reference GNU flags, and of its -march=native build beside it, on the
    cores this process may run on (its affinity mask, at most 16: the host
    share of one GPU).  The whole call (allocation, initial state, the time
    loop, the norms) is timed beside the loop."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    from workloads import eigen_model, hetero_model
    _, kw = (hetero_model if hetero else eigen_model)(N, order, steps)
    cores = min(16, len(os.sched_getaffinity(0)))
    kw['nthreads'] = cores
    kw['want_fields'] = False
    pts = float(N) ** 3 * steps
    rates, whole = {}, {}
    for native in (False, True):
        loop, call = [], []
        for _ in range(runs):
            r = oracle.run(native=native, **kw)
            loop.append(r['seconds'])
            call.append(r['seconds_total'])
        rates[native] = pts / sorted(loop)[len(loop) // 2]
        whole[native] = sorted(call)[len(call) // 2]
    return {'value': rates[False], 'unit': 'grid-pt updates/s', 'cores': cores, 'kind': 'port',
            'cpu': cpu_model(),
            'sample': 'median of %d runs of %d steps of the same %d^3 order-%d FP32 %s, oracle/elastic_oracle.cpp '
                      'built with the reference GNU flags (-O3 -fno-tree-vectorize -fopenmp), time loop only, '
                      'OpenMP on %d threads' % (runs, steps, N, order,
                                                'heterogeneous model' if hetero else 'eigenwave', cores),
            'whole_run_s': whole[False],
            'whole_run_sample': 'the same runs timed around the whole call: allocation, initial state and '
                                'boundary values, the time loop, the L2 norms (execute() and convergence())',
            'native_value': rates[True],
            'native_whole_run_s': whole[True],
            'native_sample': 'the same built with -O3 -march=native -fopenmp (the stronger CPU baseline)'}


if __name__ == '__main__':
    main()
