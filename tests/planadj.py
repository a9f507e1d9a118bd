"""FP64 CPU adjoint of a lowered model plan -- test infrastructure.

The reverse-mode derivative of PlanVM's time step (interior stress and
velocity updates, the boundary program in its launch order, receivers and
the point source), written independently of the device kernels: interior
transposes scatter (the device kernels gather), and the media-dependent
boundary coefficients are differentiated by forward-mode dual numbers over
their bytecode (the device uses the planner's symbolic derivative
programs).  Misfit J = 1/2 sum_{step, receiver, U/V/W} (trace - observed)^2.

The forward it differentiates is the one PlanVM executes, which is pinned
against the oracle (tests/test_host.py); the adjoint itself is pinned by
central finite differences of that FP64 forward (tests/test_adjoint_cpu.py).
"""
import numpy as np

from opesci import exprvm
from opesci.plan import SHEAR_PAIRS, shear_slot

from planvm import PlanVM, _code

EFFECTIVE = {'lambda': 4, 'mu': 5, 'beta1': 6, 'beta2': 7, 'beta3': 8, 'mu12': 9, 'mu13': 10, 'mu23': 11}
SHEAR_MEDIA = (9, 11, 10)     # media id of the shear stress in slot s of SHEAR_PAIRS


def eval_dual(code, media_at):
    """Value and {(media id, offsets): derivative} of a coefficient program."""
    st = []
    for op, a, val in code:
        if op == exprvm.OP_CONST:
            st.append((val, {}))
        elif op == exprvm.OP_MEDIA:
            v = media_at(a[0], a[1:4])
            st.append((v, {(a[0], tuple(a[1:4])): np.ones_like(v)}))
        elif op in (exprvm.OP_ADD, exprvm.OP_SUB, exprvm.OP_MUL, exprvm.OP_DIV):
            (y, dy), (x, dx) = st.pop(), st.pop()
            if op == exprvm.OP_ADD:
                v, fx, fy = x + y, 1.0, 1.0
            elif op == exprvm.OP_SUB:
                v, fx, fy = x - y, 1.0, -1.0
            elif op == exprvm.OP_MUL:
                v, fx, fy = x * y, y, x
            else:
                v, fx, fy = x / y, 1.0 / y, -x / (y * y)
            d = {}
            for k, g in dx.items():
                d[k] = d.get(k, 0.0) + fx * g
            for k, g in dy.items():
                d[k] = d.get(k, 0.0) + fy * g
            st.append((v, d))
        else:
            x, dx = st.pop()
            if op == exprvm.OP_NEG:
                v, f = -x, -1.0
            elif op == exprvm.OP_POWC:
                v, f = np.power(x, val), val * np.power(x, val - 1.0)
            elif op == exprvm.OP_SQRT:
                v = np.sqrt(x)
                f = 0.5 / v
            elif op == exprvm.OP_SIN:
                v, f = np.sin(x), np.cos(x)
            elif op == exprvm.OP_COS:
                v, f = np.cos(x), -np.sin(x)
            elif op == exprvm.OP_EXP:
                v = np.exp(x)
                f = v
            elif op == exprvm.OP_LOG:
                v, f = np.log(x), 1.0 / x
            elif op == exprvm.OP_ROUNDF:
                if dx:
                    raise ValueError('rounded literal depends on the media')
                v, f = np.asarray(x, np.float32).astype(np.float64), 1.0
            else:
                raise ValueError('opcode %d in a boundary coefficient' % op)
            st.append((v, {k: f * g for k, g in dx.items()}))
    return st[0]


class AdjointVM(PlanVM):
    """FP64 forward with source and receivers, and its adjoint."""

    def __init__(self, plan, effective):
        assert plan.real_bytes == 8 and plan.media_arrays
        super().__init__(plan)
        for k, mid in EFFECTIVE.items():
            self.media[mid] = np.array(effective[k], dtype=np.float64)
        self.rec = np.asarray(plan.rec_idx).reshape(-1, 3)
        self.src = None if plan.src_wavelet is None else np.asarray(plan.src_wavelet, np.float64)
        self.src_idx = tuple(plan.src_idx)

    # ---------- forward ----------
    def _record_source(self, t1, s):
        F = self.F
        r = self.rec
        tr = np.stack([F[f, t1][r[:, 0], r[:, 1], r[:, 2]] for f in range(3)], axis=1)
        if self.src is not None and s < self.src.shape[1]:
            for f in range(3):
                w = self.src[f if self.src.shape[0] == 3 else 0, s]
                F[3 + f, t1][self.src_idx] -= w / 3.0
        return tr

    def forward(self, keep_states=False):
        """Traces (steps, receivers, 3) from fields at rest."""
        self.F[...] = 0
        self.states = []
        out = []
        for s in range(self.p.ntsteps):
            if keep_states:
                self.states.append(self.F.copy())
            t0 = s % 2
            t1 = 1 - t0
            self.stress(t0, t1)
            self.bc((4, 2), t0, t1)
            self.velocity(t0, t1)
            self.bc(3, t0, t1)
            out.append(self._record_source(t1, s))
        self.steps = self.p.ntsteps
        return np.array(out)

    def misfit(self, observed):
        tr = self.forward()
        return 0.5 * float(np.sum((tr - np.asarray(observed, np.float64)) ** 2))

    # ---------- adjoint pieces ----------
    def _stmt_record(self, i, lv):
        """Forward statement i, returning what its transpose needs."""
        p = self.p
        terms = []
        for t in p.terms[p.stmts[i]['term0']:p.stmts[i]['term0'] + p.stmts[i]['nterms']]:
            sl, media_slice = self.term_slices(i, t)
            src = self.F[int(t['field']), lv[int(t['level'])]][sl].copy()
            if t['prog'] >= 0:
                code = _code(p.bc_code, int(t['prog']), int(t['prog_len']))
                v, d = eval_dual(code, lambda mid, offs: self.media[mid][media_slice(offs)])
                scale = float(t['coef'])
                terms.append((t, sl, src, np.asarray(v) * scale, {k: g * scale for k, g in d.items()}, media_slice))
            else:
                terms.append((t, sl, src, float(t['coef']), {}, media_slice))
        self.apply_stmt(i, lv)
        return i, terms

    def _stmt_adjoint(self, rec, lv, G, gm):
        i, terms = rec
        s = self.p.stmts[i]
        dsl = self.dst_slice(i)
        a = G[int(s['field']), lv[int(s['level'])]][dsl].copy()
        G[int(s['field']), lv[int(s['level'])]][dsl] = 0
        for t, sl, src, c, dc, media_slice in terms:
            G[int(t['field']), lv[int(t['level'])]][sl] += c * a
            for (mid, offs), g in dc.items():
                gm[mid][media_slice(offs)] += a * src * g

    def _stress_adjoint(self, t0, t1, F, G, gm):
        p, M = self.p, self.p.M
        inner = tuple(slice(M, d - M) for d in p.dims)
        for i in range(3):
            a = G[3 + i, t1][inner].copy()
            G[3 + i, t1][inner] = 0
            G[3 + i, t0][inner] += a
            lam, mu = self.media[4][inner], self.media[5][inner]
            for j in range(3):
                m = lam + 2.0 * mu if j == i else lam
                for k in range(M):
                    c = float(p.cn[i, j, k])
                    diff = self._sh(F[j, t0], j, k) - self._sh(F[j, t0], j, -1 - k)
                    self._sh(G[j, t0], j, k)[...] += c * m * a
                    self._sh(G[j, t0], j, -1 - k)[...] -= c * m * a
                    gm[4][inner] += a * c * diff
                    if j == i:
                        gm[5][inner] += 2.0 * a * c * diff
        for s, (ia, ja) in enumerate(SHEAR_PAIRS):
            a = G[6 + s, t1][inner].copy()
            G[6 + s, t1][inner] = 0
            G[6 + s, t0][inner] += a
            mid = SHEAR_MEDIA[s]
            mq = self.media[mid][inner]
            for g, (src, ax) in enumerate(((ia, ja), (ja, ia))):
                for k in range(M):
                    c = float(p.cs[s, g, k])
                    diff = self._sh(F[src, t0], ax, 1 + k) - self._sh(F[src, t0], ax, -k)
                    self._sh(G[src, t0], ax, 1 + k)[...] += c * mq * a
                    self._sh(G[src, t0], ax, -k)[...] -= c * mq * a
                    gm[mid][inner] += a * c * diff

    def _velocity_adjoint(self, t0, t1, F, G, gm):
        p, M = self.p, self.p.M
        inner = tuple(slice(M, d - M) for d in p.dims)
        for i in range(3):
            a = G[i, t1][inner].copy()
            G[i, t1][inner] = 0
            G[i, t0][inner] += a
            b = self.media[6 + i][inner]
            for j in range(3):
                src = 3 + i if j == i else shear_slot(i, j)
                o = 1 if j == i else 0
                for k in range(M):
                    c = float(p.cv[i, j, k])
                    diff = self._sh(F[src, t1], j, o + k) - self._sh(F[src, t1], j, o - 1 - k)
                    self._sh(G[src, t1], j, o + k)[...] += c * b * a
                    self._sh(G[src, t1], j, o - 1 - k)[...] -= c * b * a
                    gm[6 + i][inner] += a * c * diff

    def gradient(self, observed):
        """(J, {effective name: dJ/d array}, largest |dJ/d| of rho/vp/vs/beta)."""
        obs = np.asarray(observed, np.float64)
        tr = self.forward(keep_states=True)
        res = tr - obs
        J = 0.5 * float(np.sum(res ** 2))
        G = np.zeros_like(self.F)
        gm = [np.zeros(self.p.dims) for _ in range(12)]
        r = self.rec
        for s in range(self.p.ntsteps - 1, -1, -1):
            t0 = s % 2
            t1 = 1 - t0
            lv = (t0, t1)
            # replay step s, keeping what the transposes read
            self.F[...] = self.states[s]
            F_stress = self.F.copy()
            self.stress(t0, t1)
            recs42 = [self._stmt_record(i, lv) for i in self.stmt_order((4, 2))]
            F_vel = self.F.copy()
            self.velocity(t0, t1)
            recs3 = [self._stmt_record(i, lv) for i in self.stmt_order(3)]
            # reverse
            for f in range(3):
                np.add.at(G[f, t1], (r[:, 0], r[:, 1], r[:, 2]), res[s, :, f])
            for rec in reversed(recs3):
                self._stmt_adjoint(rec, lv, G, gm)
            self._velocity_adjoint(t0, t1, F_vel, G, gm)
            for rec in reversed(recs42):
                self._stmt_adjoint(rec, lv, G, gm)
            self._stress_adjoint(t0, t1, F_stress, G, gm)
        other = max(float(np.abs(gm[i]).max()) for i in range(4))
        return J, {k: gm[mid] for k, mid in EFFECTIVE.items()}, other
