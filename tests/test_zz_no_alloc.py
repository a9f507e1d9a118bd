"""opesci_ctx_l2_noalloc (the allocation-free L2 entry point) against
opesci_ctx_l2 through the C ABI: two calls in a row on two contexts, the
second grid smaller than the first, each result equal to the allocating
entry point's norms (same kernel, block count and summation order; the final
square root runs on the device, so the bar is 1e-14 relative)."""
import ctypes

import numpy as np
import pytest

from opesci.backend import Context
from workloads import eigen_model, hetero_model

pytestmark = pytest.mark.gpu


def noalloc_l2(ctx, torch):
    # a stream of torch's own (default priority), so the handle is not the
    # null stream and the copy below is ordered after the entry point's work
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        out = torch.full((9,), float('nan'), dtype=torch.float64, device='cuda')
    stream.synchronize()
    lib = ctx.lib
    lib.opesci_ctx_l2_noalloc.restype = ctypes.c_int
    lib.opesci_ctx_l2_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rc = lib.opesci_ctx_l2_noalloc(ctx.h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, lib.opesci_last_error()
    stream.synchronize()
    return out.cpu().numpy()


def test_l2_noalloc_equals_l2_two_calls_shrinking():
    torch = pytest.importorskip('torch')
    assert torch.cuda.is_available()
    models = [eigen_model(49, 4, 10)[0], hetero_model(21, 8, 6, double=True, sigma=2.0)[0]]
    for g in models:
        plan = g.lower()
        with Context(plan) as ctx:
            ctx.init()
            ctx.run()
            want = np.asarray(ctx.l2())
            got = noalloc_l2(ctx, torch)
        assert np.all(np.isfinite(got)), got
        assert np.allclose(got, want, rtol=1e-14, atol=0), (got, want)


def test_l2_noalloc_rejects_uninitialised_context():
    torch = pytest.importorskip('torch')
    g, _ = eigen_model(17, 4, 2)
    with Context(g.lower()) as ctx:
        out = torch.zeros(9, dtype=torch.float64, device='cuda')
        lib = ctx.lib
        lib.opesci_ctx_l2_noalloc.restype = ctypes.c_int
        lib.opesci_ctx_l2_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        assert lib.opesci_ctx_l2_noalloc(ctx.h, ctypes.c_void_p(out.data_ptr()), None) != 0
        assert b'opesci_ctx_init' in lib.opesci_last_error()
