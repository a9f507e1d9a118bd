"""The closed-form pins of tests/test_oracle_pins.py applied to the device:
the Levander normal-stress recompute from a linear velocity state (one step
through the C ABI) and the read path's effective media on a two-layer
medium (the arrays the device's media kernels build)."""
import numpy as np
import pytest

from opesci.backend import Context
from test_oracle_pins import (DIMS, DT, LAM, MU, N, ORDER, TOL, check_recompute, expected_media, linear_state,
                              two_layer_media)
from workloads import eigen_model

pytestmark = pytest.mark.gpu


def device_step_from(plan, state, dtype):
    with Context(plan) as ctx:
        for f in range(9):
            ctx.upload(f, 0, np.asarray(state[f], dtype))
        ctx.run(1)
        assert ctx.steps() == 1
        return [ctx.download(f, 1) for f in range(9)]


@pytest.mark.parametrize("double,tol", [(True, TOL), (False, 2e-5)])
def test_levander_recompute_closed_form_device(double, tol):
    """FP64: float-literal coefficients bound the deviation (1e-6 of the
    increment); FP32 adds the rounding of T(t0) + increment (2e-5)."""
    g, _ = eigen_model(N + 1, ORDER, 1, double=double, lam=LAM, mu=MU, rho=1.0, dt=DT)
    plan = g.lower()
    assert plan.dims == DIMS
    state = linear_state()
    dtype = np.float64 if double else np.float32
    state = [np.asarray(a, dtype) for a in state]
    out = device_step_from(plan, state, dtype)
    assert check_recompute([np.asarray(a, np.float64) for a in state], out, LAM, MU) < tol


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_effective_media_two_layer_closed_form_device(axis):
    import opesci
    cut = DIMS[axis] // 2
    rho, vp, vs = two_layer_media(axis, cut)
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (N, N, N), DT, DT * 1.5, accuracy_order=[2, ORDER, ORDER, ORDER],
                           double=True, media_arrays=(rho, vp, vs))
    plan = g.lower()
    prof = expected_media(axis, cut)
    with Context(plan) as ctx:
        for name in ('beta1', 'beta2', 'beta3', 'lambda', 'mu', 'mu12', 'mu13', 'mu23'):
            got = ctx.download_media(name)
            want = np.broadcast_to(prof[name].reshape([-1 if a == axis else 1 for a in range(3)]), DIMS)
            assert np.allclose(got, want, rtol=1e-14, atol=0), name
