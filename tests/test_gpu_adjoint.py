"""Adjoint: gradients of a receiver misfit by the hand-written adjoint
against central finite differences (FP64).

Setup (small C5-like case): heterogeneous seeded medium, fields at rest,
Ricker source in Txx/Tyy/Tzz, a receiver patch below the top face plus a few
interior receivers, 'observed' data from the same model with lambda raised
by 10 % in a Gaussian blob.  Checks:
  * dJ/d(effective media arrays) at 16 seeded cells per order (interior
    update and free-surface coefficients), relative step 1e-5, 1e-5
    relative tolerance (1e-3 steps leave ~1e-4 truncation error near the
    source);
  * the checkpoint interval does not change the gradient;
  * torch autograd through ElasticMisfit: dJ/d(rho, lambda, mu) at 16
    seeded points, relative step 1e-3 (tolerance 1e-3 of the local value
    plus 1e-6 of the largest gradient entry), three of them again with
    1e-5 steps at 1e-5 tolerance.
"""
import numpy as np
import pytest

import opesci
from opesci.backend import Context
from opesci.media import smooth_random_media
from opesci.sources import ricker

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def build(N, order, steps):
    n = N - 1
    dims = (N + order,) * 3
    rho, vp, vs = smooth_random_media(dims, seed=5, sigma=2.0)
    dx = 1.0 / n
    dt = 0.4 * 0.495 * dx / float(vp.max())
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5), accuracy_order=[2, order, order, order],
                           double=True, media_arrays=(rho, vp, vs), initial='zero', o_converge=False)
    m = order // 2
    c = m + n // 2
    g.set_source(ricker(20.0, dt, steps, t0=6 * dt) * 100.0, index=(c, c, c))
    rec = [(m + i, m + j, m + 2) for i in range(2, n - 1, 3) for j in range(2, n - 1, 3)]
    rec += [(c + 2, c - 1, c + 3), (m + 1, c, c), (c, m + 1, c)]
    g.set_receivers(index=rec)
    rho64, vp64, vs64 = (a.astype(np.float64) for a in (rho, vp, vs))
    lam = rho64 * (vp64 ** 2 - 2 * vs64 ** 2)
    mu = rho64 * vs64 ** 2
    return g, rho64, lam, mu


def observed_from(ctx, rho, lam, mu):
    from opesci.autograd import effective_media
    d = np.array(lam.shape)
    xx, yy, zz = np.meshgrid(*[np.arange(s) for s in d], indexing='ij')
    blob = np.exp(-((xx - d[0] * 0.4) ** 2 + (yy - d[1] * 0.6) ** 2 + (zz - d[2] * 0.5) ** 2) / (2 * 2.5 ** 2))
    eff = effective_media(torch.from_numpy(rho), torch.from_numpy(lam * (1 + 0.1 * blob)), torch.from_numpy(mu))
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
            h = 1e-5 * abs(base)   # FD truncation of 1e-3 steps reaches 1e-4 near the source
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
        assert float(J) == pytest.approx(misfit(ctx, torch.from_numpy(rho), torch.from_numpy(lam),
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

        # relative step 1e-3 (truncation error up to ~1e-3 of the value) ...
        for k, idx in picks:
            fd = fd_at(k, idx, 1e-3)
            an = grads[k][idx]
            assert abs(an - fd) <= 1e-3 * abs(fd) + 1e-6 * gmax, (k, idx, an, fd)
        # ... and a tight check with 1e-5 steps on the first point of each parameter
        for k, idx in (picks[0], picks[5], picks[10]):
            fd = fd_at(k, idx, 1e-5)
            an = grads[k][idx]
            assert abs(an - fd) <= 1e-5 * abs(fd) + 1e-8 * gmax, (k, idx, an, fd)
