"""Receiver-misfit gradients with respect to rho, lambda and mu through
torch.autograd.

The propagator's own adjoint (libopesci_hip: opesci_ctx_misfit_forward in
the autograd forward, opesci_ctx_gradient_reverse in the backward) gives
dJ/d(effective media arrays): lambda, mu, the shear moduli mu12/mu13/mu23
and the buoyancies beta1/beta2/beta3 the kernels read, including the media
dependence of the free-surface coefficients.  The map from (rho, lambda, mu)
to those arrays -- the reference's read-path averaging
(opesci/staggeredgrid.py:572-597): beta = 1/rho, beta_i the arithmetic mean
of beta and its +i neighbour, mu_ij the harmonic mean of mu over the 4 cells
of the (i, j) face, the last plane of each axis taking the clamped
neighbour -- is written here in torch, so autograd carries the gradient
back to rho, lambda and mu.
"""
import numpy as np
import torch

__all__ = ['effective_media', 'ElasticMisfit', 'misfit']

EFFECTIVE = ('lambda', 'mu', 'beta1', 'beta2', 'beta3', 'mu12', 'mu13', 'mu23')


def _shift(a, axis):
    """a at +1 along axis, the last plane repeated (clamped neighbour)."""
    idx = torch.arange(a.shape[axis], device=a.device)
    idx = torch.clamp(idx + 1, max=a.shape[axis] - 1)
    return torch.index_select(a, axis, idx)


def effective_media(rho, lam, mu):
    """Dict of the effective media arrays from (rho, lambda, mu)."""
    beta = 1.0 / rho
    im = 1.0 / mu
    ix, iy, iz = _shift(im, 0), _shift(im, 1), _shift(im, 2)
    return {
        'lambda': lam,
        'mu': mu,
        'beta1': 0.5 * (beta + _shift(beta, 0)),
        'beta2': 0.5 * (beta + _shift(beta, 1)),
        'beta3': 0.5 * (beta + _shift(beta, 2)),
        'mu12': 1.0 / (0.25 * (im + ix + iy + _shift(ix, 1))),
        'mu13': 1.0 / (0.25 * (im + ix + iz + _shift(ix, 2))),
        'mu23': 1.0 / (0.25 * (im + iy + iz + _shift(iy, 2))),
    }


def _upload(context, eff):
    """Device tensors go to the context device to device; others through
    host memory."""
    for k in EFFECTIVE:
        t = eff[k].detach()
        context.upload_media(k, t if t.is_cuda else t.cpu().numpy())


def _download_gradients(context, like):
    """dJ/d(effective media) of the last reverse sweep, as tensors on
    `like`'s device."""
    out = {}
    for k in EFFECTIVE:
        if like.is_cuda:
            dt = torch.float64 if context.dtype == np.float64 else torch.float32
            out[k] = context.download_gradient(k, torch.empty(tuple(context.plan.dims), dtype=dt, device=like.device))
        else:
            out[k] = torch.from_numpy(context.download_gradient(k))
    return out


def misfit(context, rho, lam, mu, observed):
    """Forward only: J for the given media (no gradient)."""
    with torch.no_grad():
        _upload(context, effective_media(rho, lam, mu))
    context.init()
    context.run()
    tr = context.receivers(np.float64)[:, :, :3]
    return 0.5 * float(np.sum((tr - np.asarray(observed, dtype=np.float64)) ** 2))


class ElasticMisfit(torch.autograd.Function):
    """J(rho, lambda, mu) = 1/2 sum_t,r,c (receiver U,V,W - observed)^2 on a
    device context of the read path.  forward runs the propagator once with
    checkpoints every `checkpoint_every` steps; backward runs the
    hand-written adjoint from those checkpoints.  The context must not step,
    init or take new media between the two (the reverse sweep then fails
    with OPESCI_ESTATE); J's backward runs at most once."""

    @staticmethod
    def forward(fctx, rho, lam, mu, context, observed, checkpoint_every=32):
        with torch.no_grad():
            _upload(context, effective_media(rho, lam, mu))
        J = context.misfit_forward(observed, checkpoint_every)
        fctx.context = context
        fctx.save_for_backward(rho, lam, mu)
        return torch.tensor(J, dtype=rho.dtype, device=rho.device)

    @staticmethod
    def backward(fctx, gJ):
        rho, lam, mu = fctx.saved_tensors
        fctx.context.gradient_reverse()
        grads = _download_gradients(fctx.context, rho)
        with torch.enable_grad():
            r, l, m = (t.detach().requires_grad_(True) for t in (rho, lam, mu))
            eff = effective_media(r, l, m)
            outs = [eff[k] for k in EFFECTIVE]
            gouts = [grads[k].to(dtype=r.dtype, device=r.device) * gJ for k in EFFECTIVE]
            gr, gl, gm = torch.autograd.grad(outs, (r, l, m), gouts, allow_unused=True)
        return gr, gl, gm, None, None, None
