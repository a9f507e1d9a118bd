"""Batched propagators: several models of different shapes, orders, media and
storage stepped together by one graph launch of batched kernels (autotuned
launch shapes; one launch per kind of model and per boundary level index)
give exactly the fields and L2 norms of running each model alone."""
import numpy as np
import pytest

import oracle
from opesci.backend import Context
from opesci.batch import Batch
from workloads import box_model

pytestmark = pytest.mark.gpu

# (sizes, order, heterogeneous, FP64, storage): kinds that share a batched
# launch (the first two, and the two FP16 ones) and 16-bit storage
SPECS = [((30, 24, 20), 2, False, False, 'native'), ((19, 27, 23), 2, False, False, 'native'),
         ((22, 28, 26), 4, True, False, 'native'), ((26, 20, 30), 8, False, False, 'native'),
         ((20, 22, 24), 4, False, True, 'native'), ((24, 26, 22), 8, True, False, 'native'),
         ((23, 25, 27), 4, False, False, 'fp16'), ((27, 21, 25), 4, False, False, 'fp16'),
         ((21, 23, 29), 2, True, False, 'bf16')]


def single(plan, steps):
    with Context(plan) as c:
        c.init()
        c.run(steps)
        return c.fields(), c.l2()


def test_batch_equals_single_runs():
    models = [box_model(sizes, order, 12, hetero=het, double=dbl, sigma=2.0, storage=st)
              for sizes, order, het, dbl, st in SPECS]
    plans = [g.lower() for g, _ in models]
    with Batch(plans) as b:
        assert len(b) == len(SPECS) and all(c is not None for c in b.configs)
        b.run(5)          # odd count: graph pairs plus single steps
        b.run(7)
        got = [(c.fields(), c.l2(), c.steps()) for c in b.contexts]
    for (fields, l2, steps), plan in zip(got, plans):
        assert steps == 12
        ref_fields, ref_l2 = single(plan, 12)
        for f in range(9):
            assert np.array_equal(fields[f], ref_fields[f]), oracle.FIELDS[f]
        assert l2 == ref_l2


def test_batch_l2_matches_oracle():
    """One batched non-cubic FP64 instance per order against the oracle
    (1e-10 relative; dx differs per axis)."""
    models = [box_model(s, o, 10, double=True) for s, o in (((28, 22, 25), 2), ((21, 27, 24), 4),
                                                            ((25, 23, 29), 8))]
    with Batch([g.lower() for g, _ in models]) as b:
        b.run(10)
        l2s = [c.l2() for c in b.contexts]
    for (g, kw), l2 in zip(models, l2s):
        ref = oracle.run(**kw)
        for f in range(9):
            assert abs(l2[f] - ref['l2'][f]) <= 1e-10 * ref['l2'][f] + 1e-15, (kw['order'], oracle.FIELDS[f])
