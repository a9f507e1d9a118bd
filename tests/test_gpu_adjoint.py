"""Adjoint: gradients of a receiver misfit by the hand-written adjoint
against central finite differences (FP64).

Setup (small C5-like case): heterogeneous seeded medium, fields at rest,
Ricker source in Txx/Tyy/Tzz, a receiver patch below the top face plus a few
interior receivers, 'observed' data from the same model with lambda raised
by 10 % in a Gaussian blob.  Checks:
  * dJ/d(effective media arrays) at 16 seeded cells per order (interior
    update and free-surface coefficients), relative step 1e-5 on FP64
    traces, 1e-5 relative tolerance;
  * the checkpoint interval does not change the gradient;
  * every cell of the gradient against the FP64 CPU adjoint (tests/planadj.py);
  * the sweep-based interior transposes against the per-cell gather kernels;
  * torch autograd through ElasticMisfit: dJ/d(rho, lambda, mu) at 16
    seeded points, relative step 1e-3 (tolerance 1e-4 of the local value
    plus 1e-6 of the largest gradient entry), three of them again with
    1e-5 steps at 1e-5 tolerance;
  * FP32 contexts: the pair-sweep transposes against the gather kernels,
    FP32 against FP64, and two FP32 runs bit for bit;
  * the gradient in two halves (forward, then reverse) and ElasticMisfit on
    device tensors against the one-call gradient.
"""
import numpy as np
import pytest

from opesci.backend import Context
from workloads import adjoint_model, lambda_blob

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

EFFECTIVE = ('lambda', 'mu', 'beta1', 'beta2', 'beta3', 'mu12', 'mu13', 'mu23')


def build(N, order, steps):
    return adjoint_model(N, order, steps)


def observed_from(ctx, rho, lam, mu):
    from opesci.autograd import effective_media
    eff = effective_media(torch.from_numpy(rho), torch.from_numpy(lambda_blob(lam)), torch.from_numpy(mu))
    for k, v in eff.items():
        ctx.upload_media(k, v.numpy())
    ctx.init()
    ctx.run()
    return ctx.receivers()[:, :, :3].copy()


def J_of(ctx, eff, obs):
    for k, v in eff.items():
        ctx.upload_media(k, v)
    ctx.init()
    ctx.run()
    tr = ctx.receivers(np.float64)[:, :, :3]
    return 0.5 * float(np.sum((tr - obs) ** 2))


@pytest.mark.parametrize("order", [4, 8])
def test_effective_media_gradient_finite_differences(order):
    from opesci.autograd import effective_media
    N, steps = 14, 36
    g, rho, lam, mu = build(N, order, steps)
    plan = g.lower()
    with Context(plan) as ctx:
        obs = observed_from(ctx, rho, lam, mu)
        eff = {k: v.numpy().copy() for k, v in effective_media(torch.from_numpy(rho), torch.from_numpy(lam),
                                                               torch.from_numpy(mu)).items()}
        for k, v in eff.items():
            ctx.upload_media(k, v)
        J = ctx.gradient(obs, checkpoint_every=8)
        assert J == pytest.approx(J_of(ctx, eff, obs), rel=1e-12)
        grads = {k: ctx.download_gradient(k) for k in eff}
        # media cells with sizeable gradient, and cells on the top surface plane
        rng = np.random.default_rng(order)
        cands = []
        for k in eff:
            a = np.abs(grads[k])
            if a.max() == 0:
                continue
            big = np.argwhere(a > 0.05 * a.max())
            for idx in big[rng.choice(len(big), size=min(3, len(big)), replace=False)]:
                cands.append((k, tuple(int(v) for v in idx)))
        m = order // 2
        for k in ('lambda', 'mu'):
            a = np.abs(grads[k][:, :, m])
            idx = np.unravel_index(int(np.argmax(a)), a.shape)
            cands.append((k, (int(idx[0]), int(idx[1]), m)))
        picks = [cands[i] for i in rng.choice(len(cands), size=min(16, len(cands)), replace=False)]
        gmax = max(float(np.abs(v).max()) for v in grads.values())
        for k, idx in picks:
            base = eff[k][idx]
            h = 1e-5 * abs(base)
            e2 = {kk: vv.copy() for kk, vv in eff.items()}
            e2[k][idx] = base + h
            jp = J_of(ctx, e2, obs)
            e2[k][idx] = base - h
            jm = J_of(ctx, e2, obs)
            fd = (jp - jm) / (2 * h)
            an = grads[k][idx]
            assert abs(an - fd) <= 1e-5 * abs(fd) + 1e-8 * gmax, (k, idx, an, fd)


def test_checkpoint_interval_does_not_change_gradient():
    from opesci.autograd import effective_media
    g, rho, lam, mu = build(12, 4, 20)
    with Context(g.lower()) as ctx:
        obs = observed_from(ctx, rho, lam, mu)
        eff = effective_media(torch.from_numpy(rho), torch.from_numpy(lam), torch.from_numpy(mu))
        for k, v in eff.items():
            ctx.upload_media(k, v.numpy())
        out = []
        for every in (1, 7, 20):
            J = ctx.gradient(obs, checkpoint_every=every)
            out.append((J, {k: ctx.download_gradient(k) for k in eff}))
    for J, gr in out[1:]:
        assert J == out[0][0]
        for k in gr:
            assert np.allclose(gr[k], out[0][1][k], rtol=1e-12, atol=1e-14 * np.abs(out[0][1][k]).max())


def test_autograd_rho_lambda_mu_finite_differences():
    from opesci.autograd import ElasticMisfit, misfit
    g, rho, lam, mu = build(14, 4, 36)
    with Context(g.lower()) as ctx:
        obs = observed_from(ctx, rho, lam, mu)
        r = torch.from_numpy(rho).requires_grad_(True)
        l = torch.from_numpy(lam).requires_grad_(True)
        m = torch.from_numpy(mu).requires_grad_(True)
        J = ElasticMisfit.apply(r, l, m, ctx, obs, 8)
        J.backward()
        grads = {'rho': r.grad.numpy(), 'lambda': l.grad.numpy(), 'mu': m.grad.numpy()}
        base = {'rho': rho, 'lambda': lam, 'mu': mu}
        assert float(J.detach()) == pytest.approx(misfit(ctx, torch.from_numpy(rho), torch.from_numpy(lam),
                                                 torch.from_numpy(mu), obs), rel=1e-12)
        rng = np.random.default_rng(16)
        gmax = max(float(np.abs(v).max()) for v in grads.values())
        picks = []
        for k in ('rho', 'lambda', 'mu'):
            a = np.abs(grads[k])
            big = np.argwhere(a > 0.05 * a.max())
            for idx in big[rng.choice(len(big), size=5, replace=False)]:
                picks.append((k, tuple(int(v) for v in idx)))
        picks.append(('mu', tuple(int(v) for v in np.unravel_index(int(np.argmax(np.abs(grads['mu']))),
                                                                     grads['mu'].shape))))
        assert len(picks) == 16
        def fd_at(k, idx, rel):
            h = rel * abs(base[k][idx])
            vals = []
            for sgn in (1, -1):
                b2 = {kk: vv.copy() for kk, vv in base.items()}
                b2[k][idx] += sgn * h
                vals.append(misfit(ctx, *(torch.from_numpy(b2[kk]) for kk in ('rho', 'lambda', 'mu')), obs))
            return (vals[0] - vals[1]) / (2 * h)

        # relative step 1e-3 ...
        for k, idx in picks:
            fd = fd_at(k, idx, 1e-3)
            an = grads[k][idx]
            assert abs(an - fd) <= 1e-4 * abs(fd) + 1e-6 * gmax, (k, idx, an, fd)
        # ... and a tight check with 1e-5 steps on the first point of each parameter
        for k, idx in (picks[0], picks[5], picks[10]):
            fd = fd_at(k, idx, 1e-5)
            an = grads[k][idx]
            assert abs(an - fd) <= 1e-5 * abs(fd) + 1e-8 * gmax, (k, idx, an, fd)


def test_split_gradient_and_device_tensors_match_gradient():
    """misfit_forward + gradient_reverse equal gradient(): J and the
    gradients bit for bit (every accumulation runs in a fixed order), and
    ElasticMisfit on device tensors (media uploaded and gradients downloaded
    device to device, observed traces on the device) equals it on host
    tensors to 1e-12 of each array's largest entry (torch's device and host
    kernels for the media averaging may contract or order FP64 operations
    differently)."""
    from opesci.autograd import ElasticMisfit, effective_media
    g, rho, lam, mu = build(12, 4, 20)
    with Context(g.lower()) as ctx:
        obs = observed_from(ctx, rho, lam, mu)
        eff = {k: v.numpy().copy() for k, v in effective_media(torch.from_numpy(rho), torch.from_numpy(lam),
                                                               torch.from_numpy(mu)).items()}
        for k, v in eff.items():
            ctx.upload_media(k, v)
        J0 = ctx.gradient(obs, checkpoint_every=6)
        g0 = {k: ctx.download_gradient(k) for k in EFFECTIVE}
        J1 = ctx.misfit_forward(obs, checkpoint_every=6)
        ctx.gradient_reverse()
        g1 = {k: ctx.download_gradient(k) for k in EFFECTIVE}
        assert J1 == J0
        for k in EFFECTIVE:
            assert np.array_equal(g1[k], g0[k]), k
        grads = {}
        for dev in ('cpu', 'cuda'):
            ts = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in (rho, lam, mu)]
            o = torch.from_numpy(np.ascontiguousarray(obs)).to(dev)
            J = ElasticMisfit.apply(*ts, ctx, o, 6)
            assert J.device.type == dev
            assert float(J.detach().cpu()) == pytest.approx(J0, rel=1e-12)
            J.backward()
            grads[dev] = [t.grad.detach().cpu().numpy() for t in ts]
        for a, b in zip(grads['cpu'], grads['cuda']):
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
        # device-side download of the raw gradient equals the host one
        out = torch.empty(tuple(g.lower().dims), dtype=torch.float64, device='cuda')
        ctx.misfit_forward(obs, checkpoint_every=6)
        ctx.gradient_reverse()
        ctx.download_gradient('mu', out)
        assert np.array_equal(out.cpu().numpy(), g0['mu'])


@pytest.mark.parametrize("order", [4, 8])
def test_gradient_matches_fp64_cpu_adjoint(order):
    """Every cell of the eight effective-media gradients against the FP64 CPU
    adjoint (tests/planadj.py): max |g_gpu - g_cpu| <= 1e-9 of the array's
    largest entry, and J to 1e-12 relative."""
    from opesci.autograd import effective_media
    from planadj import AdjointVM
    g, rho, lam, mu = build(12, order, 24)
    plan = g.lower()
    as_np = lambda e: {k: v.numpy().copy() for k, v in e.items()}
    eff = as_np(effective_media(torch.from_numpy(rho), torch.from_numpy(lam), torch.from_numpy(mu)))
    with Context(plan) as ctx:
        obs = observed_from(ctx, rho, lam, mu)
        for k, v in eff.items():
            ctx.upload_media(k, v)
        J = ctx.gradient(obs, checkpoint_every=5)
        grads = {k: ctx.download_gradient(k) for k in eff}
    Jc, gc, other = AdjointVM(plan, eff).gradient(obs)
    assert other == 0.0
    assert J == pytest.approx(Jc, rel=1e-12)
    for k in eff:
        scale = float(np.abs(gc[k]).max())
        assert scale > 0, k
        err = float(np.abs(grads[k] - gc[k]).max())
        assert err <= 1e-9 * scale, (k, err, scale)


@pytest.mark.parametrize("order,initial,every", [(2, 'zero', 4), (4, 'zero', 4), (8, 'zero', 4),
                                                  (4, 'eigenmode', 5), (2, 'eigenmode', 4)])
def test_fast_interior_transposes_equal_simple_kernels(order, initial, every):
    """The sweep-based interior transposes, the level swaps and the
    slot-based segment scheme (default) against the per-cell gather kernels,
    explicit level moves and full-state segments (launch variant 1), FP64,
    on a box of uneven sides, from rest and from the eigenmode (whose
    initial boundary values sit in cells the steps never write), with even
    and odd checkpoint intervals: every gradient entry to 1e-12 of the
    array's largest."""
    import opesci
    from opesci.media import smooth_random_media
    from opesci.sources import ricker
    sizes, steps = (10, 13, 17), 18
    n = tuple(s - 1 for s in sizes)
    dims = tuple(s + order for s in sizes)
    rho, vp, vs = smooth_random_media(dims, seed=9, sigma=2.0)
    dt = 0.4 * 0.495 / max(n) / float(vp.max())
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), n, dt, dt * (steps + 0.5), accuracy_order=[2, order, order, order],
                           double=True, media_arrays=(rho, vp, vs), initial=initial, o_converge=False)
    m = order // 2
    g.set_source(ricker(20.0, dt, steps, t0=6 * dt) * 100.0, index=(m + 4, m + 6, m + 8))
    g.set_receivers(index=[(m + i, m + j, m + 1) for i in range(0, sizes[0], 3) for j in range(0, sizes[1], 3)])
    with Context(g.lower()) as ctx:
        ctx.init()
        ctx.run()
        obs = (ctx.receivers()[:, :, :3] * 0.9).astype(np.float32)
        out = []
        for variant in (0, 1):
            ctx.configure(variant)
            J = ctx.gradient(obs, checkpoint_every=every)
            out.append((J, {k: ctx.download_gradient(k) for k in EFFECTIVE}))
    assert out[0][0] == pytest.approx(out[1][0], rel=1e-13)
    for k in EFFECTIVE:
        ref = out[1][1][k]
        scale = float(np.abs(ref).max())
        assert scale > 0, k
        assert float(np.abs(out[0][1][k] - ref).max()) <= 1e-12 * scale, k


def test_gradient_errors_and_state():
    """Refusals leave the context usable; a gradient leaves it needing init."""
    from workloads import eigen_model
    g, rho, lam, mu = build(10, 4, 6)
    with Context(g.lower()) as ctx:
        ctx.init()
        ctx.run()
        obs = ctx.receivers()[:, :, :3].copy()
        with pytest.raises(RuntimeError, match='observed'):
            ctx.gradient(obs[:-1])
        with pytest.raises(RuntimeError):
            ctx.gradient(obs, checkpoint_every=0)
        ctx.init()
        ctx.run()                                  # still steps after the refusals
        ctx.gradient(obs, checkpoint_every=4)
        with pytest.raises(RuntimeError, match='init'):
            ctx.run()
        ctx.init()
        ctx.run()
        assert np.array_equal(ctx.receivers()[:, :, :3], obs)   # the model is unchanged
        # the two halves: the reverse sweep needs its own forward pass
        with pytest.raises(RuntimeError, match='forward'):
            ctx.gradient_reverse()
        J = ctx.misfit_forward(obs, checkpoint_every=4)
        ctx.init()                                 # drops the forward pass
        with pytest.raises(RuntimeError, match='forward'):
            ctx.gradient_reverse()
        assert ctx.misfit_forward(obs, checkpoint_every=4) == J
        ctx.upload_media('lambda', lam)            # so does new media
        with pytest.raises(RuntimeError, match='forward'):
            ctx.gradient_reverse()
        assert ctx.misfit_forward(obs, checkpoint_every=4) == J
        ctx.gradient_reverse()
        with pytest.raises(RuntimeError, match='forward'):
            ctx.gradient_reverse()                 # once per forward pass
    plain, _ = eigen_model(10, 4, 4, double=True)    # constant media: no media arrays
    with Context(plain.lower()) as ctx:
        ctx.init()
        with pytest.raises(RuntimeError, match='read path'):
            ctx.gradient(np.zeros((4, 0, 3), np.float32))


@pytest.mark.parametrize("order", [4, 8])
def test_fp32_gradient_against_fp64_and_simple_kernels(order):
    """FP32 contexts, bounds stated before the run:
      * kernel variant 0 (the pair sweeps as interior transposes, in place)
        against variant 1 (per-cell gather kernels): every effective-media
        gradient within 1e-5 of its largest entry (the same operations,
        FP32 sums in a different order);
      * against the FP64 gradient of the same model: every array within
        1e-3 relative L2 (FP32 rounding of the forward and adjoint fields
        and of the gradient sums);
      * two FP32 gradients in a row, in a fresh context each: bit for bit
        equal (the boundary transposes, the media maps of their
        coefficients and the receiver residuals are gathered in a fixed
        order, no atomics)."""
    from opesci.autograd import effective_media
    N, steps = 14, 36
    g64, rho, lam, mu = adjoint_model(N, order, steps)
    with Context(g64.lower()) as ctx:
        obs = observed_from(ctx, rho, lam, mu)
        eff = {k: v.numpy().copy() for k, v in effective_media(torch.from_numpy(rho), torch.from_numpy(lam),
                                                               torch.from_numpy(mu)).items()}
        for k, v in eff.items():
            ctx.upload_media(k, v)
        ctx.gradient(obs, checkpoint_every=8)
        ref = {k: ctx.download_gradient(k) for k in EFFECTIVE}
    res = {}
    for variant, rep in ((0, 0), (0, 1), (1, 0)):
        g32, _, _, _ = adjoint_model(N, order, steps, double=False)
        plan = g32.lower()
        plan.kernel_variant = variant
        with Context(plan) as ctx:
            for k, v in eff.items():
                ctx.upload_media(k, v.astype(np.float32))
            ctx.gradient(obs, checkpoint_every=8)
            res[(variant, rep)] = {k: ctx.download_gradient(k).astype(np.float64) for k in EFFECTIVE}
    for k in EFFECTIVE:
        scale = np.abs(ref[k]).max()
        if scale == 0:
            continue
        a, b, c = res[(0, 0)][k], res[(0, 1)][k], res[(1, 0)][k]
        assert np.abs(a - c).max() <= 1e-5 * scale, (k, 'variant', np.abs(a - c).max() / scale)
        assert np.array_equal(a, b), (k, 'repeat', np.abs(a - b).max() / scale)
        rel = np.linalg.norm(a - ref[k]) / np.linalg.norm(ref[k])
        assert rel <= 1e-3, (k, 'fp32 vs fp64', rel)
