"""The FP64 CPU adjoint (tests/planadj.py) against central finite
differences of the FP64 CPU forward it differentiates.

Directional derivatives: dJ/dm . d for a seeded random direction d over every
cell of the eight effective media arrays, and for directions supported only
on the boundary slabs (the free-surface statements, whose coefficients
depend on lambda and mu) and only inside them.  Step h = 1e-6 of each
value; tolerance 1e-8 relative (FP64 roundoff at this h is ~1e-10 and the
truncation O(h^2) smaller still).
"""
import numpy as np
import pytest

from planadj import AdjointVM
from workloads import adjoint_model, lambda_blob

torch = pytest.importorskip('torch')


def _case(N, order, steps):
    from opesci.autograd import effective_media
    g, rho, lam, mu = adjoint_model(N, order, steps)
    plan = g.lower()
    as_np = lambda e: {k: v.numpy() for k, v in e.items()}
    true = as_np(effective_media(torch.from_numpy(rho), torch.from_numpy(lambda_blob(lam)), torch.from_numpy(mu)))
    eff = as_np(effective_media(torch.from_numpy(rho), torch.from_numpy(lam), torch.from_numpy(mu)))
    obs = AdjointVM(plan, true).forward().astype(np.float32)
    return plan, eff, obs


@pytest.mark.parametrize("order", [4, 8])
def test_cpu_adjoint_directional_finite_differences(order):
    plan, eff, obs = _case(12, order, 24)
    J, grad, other = AdjointVM(plan, eff).gradient(obs)
    assert other == 0.0          # every coefficient reads effective arrays only
    assert J == pytest.approx(AdjointVM(plan, eff).misfit(obs), rel=1e-14)
    m = order // 2
    dims = plan.dims
    slab = np.zeros(dims, bool)
    for ax in range(3):
        sl = [slice(None)] * 3
        sl[ax] = slice(0, m + 2)
        slab[tuple(sl)] = True
        sl[ax] = slice(dims[ax] - m - 2, dims[ax])
        slab[tuple(sl)] = True
    rng = np.random.default_rng(order)
    for support in (None, slab, ~slab):
        d = {}
        for k, v in eff.items():
            r = rng.standard_normal(v.shape) * v
            d[k] = r if support is None else np.where(support, r, 0.0)
        h = 1e-6
        jp = AdjointVM(plan, {k: eff[k] + h * d[k] for k in eff}).misfit(obs)
        jm = AdjointVM(plan, {k: eff[k] - h * d[k] for k in eff}).misfit(obs)
        fd = (jp - jm) / (2 * h)
        an = sum(float(np.sum(grad[k] * d[k])) for k in eff)
        assert abs(an - fd) <= 1e-8 * abs(fd), (support is None, an, fd)


def test_dual_numbers_differentiate_a_free_surface_coefficient():
    """eval_dual on lambda/(lambda + 2 mu): the textbook quotient rule."""
    from opesci import exprvm
    from planadj import eval_dual
    code = [(exprvm.OP_MEDIA, (4, 0, 0, 0), 0.0), (exprvm.OP_MEDIA, (4, 0, 0, 0), 0.0),
            (exprvm.OP_CONST, (0, 0, 0, 0), 2.0), (exprvm.OP_MEDIA, (5, 0, 0, 0), 0.0),
            (exprvm.OP_MUL, (0, 0, 0, 0), 0.0), (exprvm.OP_ADD, (0, 0, 0, 0), 0.0),
            (exprvm.OP_DIV, (0, 0, 0, 0), 0.0)]
    lam, mu = np.array([1.0, 3.0]), np.array([2.0, 0.5])
    v, d = eval_dual(code, lambda mid, offs: lam if mid == 4 else mu)
    assert np.allclose(v, lam / (lam + 2 * mu))
    assert np.allclose(d[(4, (0, 0, 0))], 2 * mu / (lam + 2 * mu) ** 2)
    assert np.allclose(d[(5, (0, 0, 0))], -2 * lam / (lam + 2 * mu) ** 2)


def test_torch_effective_media_match_the_plan_interpreter():
    """opesci.autograd.effective_media (torch, differentiable) forms the same
    effective arrays as the plan interpreter builds from rho/Vp/Vs (which
    tests/test_host.py pins against the oracle's read path), FP64."""
    from opesci.autograd import effective_media
    from planvm import PlanVM
    g, rho, lam, mu = adjoint_model(9, 4, 2)
    vm = PlanVM(g.lower())
    eff = effective_media(torch.from_numpy(rho), torch.from_numpy(lam), torch.from_numpy(mu))
    for k, mid in (('lambda', 4), ('mu', 5), ('beta1', 6), ('beta2', 7), ('beta3', 8), ('mu12', 9), ('mu13', 10),
                   ('mu23', 11)):
        a, b = eff[k].numpy(), vm.media[mid]
        assert np.allclose(a, b, rtol=1e-6, atol=0), k   # the interpreter's media start from float32 rho/Vp/Vs
