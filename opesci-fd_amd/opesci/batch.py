"""Many independent propagators stepped together on one GPU.

Each model keeps its own device context (fields, media, boundary program);
a batch captures two time steps of every context as parallel branches of
one hipGraph, so one graph launch advances all of them and the kernels of
small grids of different shapes and orders run side by side.

The autotuner replaces the reference's Pluto tile-size sweep
(opesci/regulargrid.py:21-23, tests/eigenwave3d.py --tile): for every
distinct (grid shape, order, precision, media) it times the interior launch
configurations of CONFIGS -- kernel variant, rows per tile, x chunk -- each
after one warm-up step, as the median of five single steps, and keeps the
fastest.  The kernels use the same arithmetic in every configuration, so
tuning never changes the results.
"""
import ctypes

from . import backend

__all__ = ['Batch', 'autotune', 'CONFIGS']

# (kernel_variant, tile_y, x_chunk): the default sweep (two z points per lane
# for FP32 arithmetic, 128-point z tiles) with its default rows, 8 or 4 rows
# per tile, and x chunks of 8, 16 or 32 planes beside the automatic choice;
# the one-point-per-lane sweep (64-point z tiles) with 16 or 8 rows,
# automatic or 16-plane chunks
CONFIGS = ((0, 0, 0), (0, 8, 0), (0, 4, 0), (0, 0, 8), (0, 0, 16), (0, 0, 32), (0, 4, 16),
           (2, 16, 0), (2, 8, 0), (2, 8, 16))


def _key(plan):
    return (tuple(plan.dims), plan.order, plan.real_bytes, plan.storage, bool(plan.media_arrays))


def autotune(ctx, configs=CONFIGS, steps=5, cache=None):
    """Pick the fastest interior launch configuration for a context (the
    context is re-initialised afterwards): per configuration one warm-up
    step, then the median of `steps` single steps of the stress and velocity
    launches.  Returns (config, times_ms)."""
    key = _key(ctx.plan)
    if cache is not None and key in cache:
        best = cache[key]
        ctx.configure(*best)
        return best, None
    times = {}
    for cfg in configs:
        ctx.configure(*cfg)
        ctx.init()
        ctx.kernel_times(1)
        samples = sorted(sum(ctx.kernel_times(1)[:2]) for _ in range(steps))
        times[cfg] = samples[len(samples) // 2]
    best = min(times, key=times.get)
    ctx.configure(*best)
    ctx.init()
    if cache is not None:
        cache[key] = best
    return best, times


class Batch(object):
    """Contexts for `plans`, initialised (and autotuned), stepped together."""

    def __init__(self, plans, tune=True):
        self.contexts = []
        self.configs = []
        cache = {}
        try:
            for p in plans:
                c = backend.Context(p)
                c.init()
                if tune:
                    cfg, _ = autotune(c, cache=cache)
                    c.init()
                else:
                    cfg = None
                self.contexts.append(c)
                self.configs.append(cfg)
            self.lib = self.contexts[0].lib
            arr = (ctypes.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
            h = ctypes.c_void_p()
            backend._check(self.lib.opesci_batch_create(arr, len(self.contexts), ctypes.byref(h)),
                           'opesci_batch_create')
            self.h = h
        except Exception:
            self.close()
            raise

    def run(self, nsteps):
        """Advance every context by nsteps; returns the GPU seconds."""
        s = ctypes.c_double()
        backend._check(self.lib.opesci_batch_run(self.h, int(nsteps), ctypes.byref(s)), 'opesci_batch_run')
        return s.value

    def close(self):
        if getattr(self, 'h', None):
            self.lib.opesci_batch_destroy(self.h)
            self.h = None
        for c in self.contexts:
            c.close()
        self.contexts = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.contexts)

    def __getitem__(self, i):
        return self.contexts[i]
