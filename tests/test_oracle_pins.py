"""Closed-form pins of the two oracle functions no reference vector covers.

* Levander normal-stress recompute (oracle ``stress_tnn``, restating
  opesci/fields.py:318-353): from a state whose velocities are linear in
  x, y, z every difference stencil is exact, so one stress phase must give,
  on a free-surface plane normal to axis d and for a normal stress T_jj
  (j != d, k the third axis),
      T_jj(t1) = T_jj(t0) + dt [(l2m - lam^2/l2m) dV_j/dx_j + (lam - lam^2/l2m) dV_k/dx_k]
  (l2m = lam + 2 mu), where an interior point gets
      T_jj(t1) = T_jj(t0) + dt [l2m dV_j/dx_j + lam (dV_k/dx_k + dV_d/dx_d)].
  The gradients are chosen distinct, so a swapped, dropped or mis-signed
  term moves the result by O(dt); the oracle with the recompute left out
  (bc_algo=2, the older generator) must fail the check.
* Read-path effective media (oracle ``build_media``, restating
  opesci/staggeredgrid.py:572-597): on a two-layer medium beta_i is the
  arithmetic mean of beta over the cell and its +i neighbour and mu_ij the
  harmonic mean of mu over the 4 cells of the (i, j) face, so at a layer
  interface they take the closed values (bL + bR)/2 and 2 muL muR/(muL + muR)
  and inside a layer the layer's own values; uniform media through the
  read path must reproduce the constant-media run.

The same closed forms pin the product's lowering on the CPU (the plan
interpreter executes exactly the coefficients and boundary program the
device runs) and, in tests/test_gpu_pins.py, the device itself.
"""
import numpy as np
import pytest

import oracle
from workloads import eigen_model

N = 12          # cells per axis
ORDER = 4
M = ORDER // 2
DIMS = (N + 1 + 2 * M,) * 3
DX = 1.0 / N
DT = DX / 6.0
# velocity gradients (row = U, V, W; column = d/dx, d/dy, d/dz), all distinct
GRAD = np.array([[0.30, -0.20, 0.10],
                 [0.15, 0.45, -0.25],
                 [-0.35, 0.05, 0.60]])


def linear_state(seed=3):
    """Velocities linear in the grid coordinates (every cell, ghosts
    included); stresses at t0 seeded noise of the increments' size."""
    ii = np.meshgrid(*[np.arange(d, dtype=np.float64) * DX for d in DIMS], indexing='ij')
    vel = [sum(GRAD[f, a] * ii[a] for a in range(3)) + 0.1 * (f + 1) for f in range(3)]
    rng = np.random.default_rng(seed)
    st = [0.01 * rng.standard_normal(DIMS) for _ in range(6)]
    return vel + st


def surface_cells(d, side):
    """Index tuple of the recompute cells on the surface normal to d: the
    plane b (m or dim-m-1), tangential indices in [m+1, dim-m-1)."""
    b = M if side == 0 else DIMS[d] - M - 1
    sl = [slice(M + 1, DIMS[a] - M - 1) for a in range(3)]
    sl[d] = b
    return tuple(sl)


def closed_form_recompute(t0, j, d, lam, mu):
    k = 3 - j - d
    l2m = lam + 2 * mu
    return t0 + DT * ((l2m - lam * lam / l2m) * GRAD[j, j] + (lam - lam * lam / l2m) * GRAD[k, k])


def closed_form_interior(t0, j, lam, mu):
    l2m = lam + 2 * mu
    others = [a for a in range(3) if a != j]
    return t0 + DT * (l2m * GRAD[j, j] + lam * sum(GRAD[a, a] for a in others))


def check_recompute(state, out, lam, mu):
    """Max relative deviation from the closed forms over every recompute
    cell (all six faces, both tangential normal stresses) and the deep
    interior, relative to the increment dt * (...)."""
    worst = 0.0
    for d in range(3):
        for side in range(2):
            sl = surface_cells(d, side)
            for j in range(3):
                if j == d:
                    continue
                want = closed_form_recompute(state[3 + j][sl], j, d, lam, mu)
                inc = np.abs(want - state[3 + j][sl]).max()
                worst = max(worst, np.abs(out[3 + j][sl] - want).max() / inc)
    deep = tuple(slice(2 * M + 1, DIMS[a] - 2 * M - 1) for a in range(3))
    for j in range(3):
        want = closed_form_interior(state[3 + j][deep], j, lam, mu)
        inc = np.abs(want - state[3 + j][deep]).max()
        worst = max(worst, np.abs(out[3 + j][deep] - want).max() / inc)
    return worst


# The coefficients are float literals (every folded constant is printed
# with an F suffix, opesci/codeprinter.py:42-63), so even an FP64 run meets
# the closed form only to float rounding of about 8 coefficients: 1e-6 of
# the increment.  A wrong term is off by O(1) of it.
TOL = 1e-6
LAM, MU = 0.75, 0.5


def test_levander_recompute_closed_form_oracle():
    state = linear_state()
    out = oracle.run_from(state, order=ORDER, n=(N, N, N), ntsteps=1, dt=DT, double=True, rho=1.0, lam=LAM, mu=MU)
    assert check_recompute(state, out, LAM, MU) < TOL


def test_levander_recompute_planted_mistakes_fail():
    """The check discriminates: the oracle without the recompute fails it,
    and so does the closed form with the two coefficients swapped."""
    state = linear_state()
    out = oracle.run_from(state, order=ORDER, n=(N, N, N), ntsteps=1, dt=DT, double=True, rho=1.0, lam=LAM, mu=MU,
                          bc_algo=2)
    assert check_recompute(state, out, LAM, MU) > 1e-2
    out = oracle.run_from(state, order=ORDER, n=(N, N, N), ntsteps=1, dt=DT, double=True, rho=1.0, lam=LAM, mu=MU)
    sl = surface_cells(2, 0)
    l2m = LAM + 2 * MU
    swapped = state[3][sl] + DT * ((LAM - LAM ** 2 / l2m) * GRAD[0, 0] + (l2m - LAM ** 2 / l2m) * GRAD[1, 1])
    inc = np.abs(swapped - state[3][sl]).max()
    assert np.abs(out[3][sl] - swapped).max() / inc > 1e-2


def uniform_media(rho=1.5, vp=2.0, vs=1.0):
    return tuple(np.full(DIMS, v, dtype=np.float32) for v in (rho, vp, vs))


def test_levander_recompute_closed_form_oracle_read_path():
    """The same on the read path with uniform arrays (rho 1.5, Vp 2, Vs 1:
    lambda = 3, mu = 1.5 at every cell)."""
    state = linear_state()
    out = oracle.run_from(state, order=ORDER, n=(N, N, N), ntsteps=1, dt=DT, double=True, media=uniform_media())
    assert check_recompute(state, out, 3.0, 1.5) < TOL


def plan_step_from(plan, state):
    from planvm import PlanVM
    vm = PlanVM(plan)
    for f in range(9):
        vm.F[f, 0] = state[f]
    vm.step()
    return vm.latest()


def test_levander_recompute_closed_form_product_lowering():
    """The product's lowering (the coefficients and boundary program the
    device executes, run by the plan interpreter) meets the closed form."""
    g, _ = eigen_model(N + 1, ORDER, 1, double=True, lam=LAM, mu=MU, rho=1.0, dt=DT)
    plan = g.lower()
    assert plan.dims == DIMS
    state = linear_state()
    assert check_recompute(state, plan_step_from(plan, state), LAM, MU) < TOL


def test_read_path_uniform_media_equals_constant_run():
    """Uniform rho/Vp/Vs through the read path (arithmetic and harmonic means
    of equal values) reproduce the constant-media eigenmode run.  The read
    path multiplies float literals by media values where the constant path
    folds them into one literal, so the two differ by float rounding of the
    coefficients only: 2e-6 of the field maximum after 20 FP64 steps."""
    rho, vp, vs = 1.5, 2.0, 1.0
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    kw = dict(order=ORDER, n=(N, N, N), ntsteps=20, dt=DT, double=True)
    a = oracle.run(rho=rho, lam=lam, mu=mu, **kw)
    b = oracle.run(media=uniform_media(rho, vp, vs), **kw)
    for f in range(9):
        scale = np.abs(a['fields'][f]).max() + 1e-30
        assert np.abs(a['fields'][f] - b['fields'][f]).max() / max(scale, np.abs(a['fields'][0]).max()) < 2e-6, \
            oracle.FIELDS[f]


def two_layer_media(axis, cut, left=(1.0, 2.0, 1.0), right=(2.5, 4.0, 1.6)):
    """rho/Vp/Vs taking the `left` values at indices < cut along `axis` and
    `right` from cut on."""
    idx = np.arange(DIMS[axis]).reshape([-1 if a == axis else 1 for a in range(3)])
    out = []
    for k in range(3):
        a = np.where(idx < cut, left[k], right[k]) * np.ones(DIMS)
        out.append(a.astype(np.float32))
    return out


def layer_values(rho, vp, vs):
    rho, vp, vs = (float(np.float32(v)) for v in (rho, vp, vs))
    return 1.0 / rho, rho * (vp * vp - 2 * vs * vs), rho * vs * vs


def expected_media(axis, cut, left=(1.0, 2.0, 1.0), right=(2.5, 4.0, 1.6)):
    """Closed-form effective media of a two-layer medium, per cell index
    along `axis`: beta_i, lambda, mu and mu_ij as 1-D profiles."""
    bL, lL, mL = layer_values(*left)
    bR, lR, mR = layer_values(*right)
    n = DIMS[axis]
    prof = {}
    side = lambda i: (bL, lL, mL) if i < cut else (bR, lR, mR)
    # the last plane's +1 neighbour is clamped to the plane itself
    nxt = lambda i: min(i + 1, n - 1)
    b = np.array([side(i)[0] for i in range(n)])
    prof['beta'] = b
    prof['lambda'] = np.array([side(i)[1] for i in range(n)])
    mu = np.array([side(i)[2] for i in range(n)])
    prof['mu'] = mu
    hm2 = lambda a, c: 1.0 / (0.25 * (2.0 / a + 2.0 / c))
    names_i = {0: 'beta1', 1: 'beta2', 2: 'beta3'}
    for i in range(3):
        if i == axis:
            prof[names_i[i]] = np.array([(b[k] + b[nxt(k)]) / 2.0 for k in range(n)])
        else:
            prof[names_i[i]] = b.copy()
    for name, (i, j) in (('mu12', (0, 1)), ('mu13', (0, 2)), ('mu23', (1, 2))):
        if axis in (i, j):
            prof[name] = np.array([hm2(mu[k], mu[nxt(k)]) for k in range(n)])
        else:
            prof[name] = mu.copy()
    return prof


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_effective_media_two_layer_closed_form_oracle(axis):
    cut = DIMS[axis] // 2
    med = oracle.media(*two_layer_media(axis, cut), order=ORDER, double=True)
    prof = expected_media(axis, cut)
    for name, p in prof.items():
        want = np.broadcast_to(p.reshape([-1 if a == axis else 1 for a in range(3)]), DIMS)
        assert np.allclose(med[name], want, rtol=1e-14, atol=0), name
    # the interface values really are the means (not one side's value)
    bL, _, mL = layer_values(1.0, 2.0, 1.0)
    bR, _, mR = layer_values(2.5, 4.0, 1.6)
    at = [slice(None)] * 3
    at[axis] = cut - 1
    bname = ('beta1', 'beta2', 'beta3')[axis]
    assert np.allclose(med[bname][tuple(at)], (bL + bR) / 2)
    shear = {0: 'mu12', 1: 'mu23', 2: 'mu13'}[axis]
    assert np.allclose(med[shear][tuple(at)], 2 * mL * mR / (mL + mR))


@pytest.mark.parametrize("axis", [0, 2])
def test_effective_media_two_layer_closed_form_product(axis):
    """The torch effective media of the autograd path (the product's
    differentiable media transform) meet the same closed forms."""
    torch = pytest.importorskip("torch")
    from opesci.autograd import effective_media
    cut = DIMS[axis] // 2
    rho, vp, vs = two_layer_media(axis, cut)
    r = torch.tensor(rho, dtype=torch.float64)
    lam = r * (torch.tensor(vp, dtype=torch.float64) ** 2 - 2 * torch.tensor(vs, dtype=torch.float64) ** 2)
    mu = r * torch.tensor(vs, dtype=torch.float64) ** 2
    med = effective_media(r, lam, mu)
    prof = expected_media(axis, cut)
    for name in ('beta1', 'beta2', 'beta3', 'mu12', 'mu13', 'mu23'):
        want = np.broadcast_to(prof[name].reshape([-1 if a == axis else 1 for a in range(3)]), DIMS)
        assert np.allclose(med[name].numpy(), want, rtol=1e-12, atol=0), name
