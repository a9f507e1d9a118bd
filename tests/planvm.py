"""NumPy interpreter of a lowered model plan -- test infrastructure.

Executes exactly what libopesci_hip is asked to execute (folded interior
coefficients, the boundary program, the analytic-solution bytecode), on the
CPU, so the lowering in opesci/plan.py can be checked against the oracle
independently of any GPU, and the GPU can be checked against this.
"""
import numpy as np

from opesci import exprvm
from opesci.plan import SHEAR_PAIRS, STAG, shear_slot

TIME_STAG = (1, 1, 1, 0, 0, 0, 0, 0, 0)


def eval_code(code, coords, t, media_at):
    st = []
    for op, a, val in code:
        if op == exprvm.OP_CONST:
            st.append(val)
        elif op == exprvm.OP_COORD:
            st.append(coords[a[0]])
        elif op == exprvm.OP_TIME:
            st.append(t)
        elif op == exprvm.OP_MEDIA:
            st.append(media_at(a[0], a[1:4]))
        elif op == exprvm.OP_ADD:
            y = st.pop(); st.append(st.pop() + y)
        elif op == exprvm.OP_SUB:
            y = st.pop(); st.append(st.pop() - y)
        elif op == exprvm.OP_MUL:
            y = st.pop(); st.append(st.pop() * y)
        elif op == exprvm.OP_DIV:
            y = st.pop(); st.append(st.pop() / y)
        elif op == exprvm.OP_NEG:
            st.append(-st.pop())
        elif op == exprvm.OP_POWC:
            st.append(np.power(st.pop(), val))
        elif op == exprvm.OP_SQRT:
            st.append(np.sqrt(st.pop()))
        elif op == exprvm.OP_SIN:
            st.append(np.sin(st.pop()))
        elif op == exprvm.OP_COS:
            st.append(np.cos(st.pop()))
        elif op == exprvm.OP_EXP:
            st.append(np.exp(st.pop()))
        elif op == exprvm.OP_LOG:
            st.append(np.log(st.pop()))
        elif op == exprvm.OP_ROUNDF:
            st.append(np.asarray(st.pop(), np.float32).astype(np.float64))
    return st[0]


def _code(arr, off, n):
    return [(int(r['op']), tuple(int(v) for v in r['a']), float(r['val'])) for r in arr[off:off + n]]


class PlanVM(object):
    def __init__(self, plan, order_by_sched=True):
        self.p = plan
        self.dt_real = np.float64 if plan.real_bytes == 8 else np.float32
        self.F = np.zeros((9, 2) + tuple(plan.dims), dtype=self.dt_real)
        self.order_by_sched = order_by_sched
        self.steps = 0
        self.media = None
        if plan.media_arrays:
            self._build_media()

    def _build_media(self):
        R = self.dt_real
        rho, vp, vs = [np.asarray(a, np.float64) for a in self.p.media]
        d0, d1, d2 = self.p.dims
        beta = (1.0 / rho).astype(R)
        lam = (rho * (vp * vp - 2.0 * vs * vs)).astype(R)
        mu = (rho * (vs * vs)).astype(R)
        cl = lambda a, ax: np.concatenate([np.take(a, range(1, a.shape[ax]), axis=ax),
                                           np.take(a, [a.shape[ax] - 1], axis=ax)], axis=ax)
        b = beta.astype(np.float64)
        m = mu.astype(np.float64)
        beta1 = ((b + cl(b, 0)) / 2.0).astype(R)
        beta2 = ((b + cl(b, 1)) / 2.0).astype(R)
        beta3 = ((b + cl(b, 2)) / 2.0).astype(R)
        with np.errstate(divide='ignore'):
            inv = 1.0 / m
        hm = lambda a, c, e: (1.0 / (0.25 * (inv + a + c + e))).astype(R)
        mu12 = hm(cl(inv, 0), cl(inv, 1), cl(cl(inv, 0), 1))
        mu13 = hm(cl(inv, 0), cl(inv, 2), cl(cl(inv, 0), 2))
        mu23 = hm(cl(inv, 1), cl(inv, 2), cl(cl(inv, 1), 2))
        self.media = [rho.astype(np.float32), vp.astype(np.float32), vs.astype(np.float32),
                      beta, lam, mu, beta1, beta2, beta3, mu12, mu13, mu23]

    # ---------- analytic solution ----------
    def coord(self, axis, i, stag):
        R = self.dt_real
        dx = R(self.p.spacing[axis])
        return np.asarray(np.float64(dx) * ((i - self.p.M) + (0.5 if stag else 0.0)), dtype=R).astype(np.float64)

    def solution(self, f, t, ranges):
        p = self.p
        code = _code(p.sol_code, p.sol_off[f], p.sol_len[f])
        ix = [np.arange(lo, hi) for lo, hi in ranges]
        X, Y, Z = np.meshgrid(*[self.coord(a, ix[a], STAG[f][a]) for a in range(3)], indexing='ij')

        def media_at(mid, offs):
            sl = tuple(slice(lo + o, hi + o) for (lo, hi), o in zip(ranges, offs))
            return self.media[mid][sl].astype(np.float64)
        v = eval_code(code, (X, Y, Z), t, media_at)
        return np.broadcast_to(v, X.shape)

    def ranges(self, f):
        M = self.p.M
        return [(M, d - M - (1 if STAG[f][a] else 0)) for a, d in enumerate(self.p.dims)]

    def init(self):
        p = self.p
        for f in range(9):
            t0 = p.dt / 2.0 if TIME_STAG[f] else 0.0
            rg = self.ranges(f)
            sl = tuple(slice(lo, hi) for lo, hi in rg)
            self.F[f, 0][sl] = self.solution(f, t0, rg).astype(self.dt_real)
        self.bc(0, 0, 0)
        self.bc(1, 0, 0)
        self.steps = 0

    # ---------- interior ----------
    def _sh(self, a, axis, off):
        """view of a shifted by off along axis over the interior box"""
        M = self.p.M
        sl = []
        for ax, d in enumerate(self.p.dims):
            o = off if ax == axis else 0
            sl.append(slice(M + o, d - M + o))
        return a[tuple(sl)]

    def stress(self, t0, t1):
        p, F, R = self.p, self.F, self.dt_real
        M = p.M
        inner = tuple(slice(M, d - M) for d in p.dims)
        med = self.media
        for i in range(3):
            acc = F[3 + i, t0][inner].copy()
            for j in range(3):
                for k in range(M):
                    c = R(p.cn[i, j, k])
                    fp = self._sh(F[j, t0], j, k)
                    fm = self._sh(F[j, t0], j, -1 - k)
                    if not p.media_arrays:
                        acc = acc + c * fp
                        acc = acc - c * fm
                    else:
                        lam = med[4][inner]
                        m = lam + R(2) * med[5][inner] if j == i else lam
                        acc = acc + (c * m) * (fp - fm)
            F[3 + i, t1][inner] = acc
        for s, (ia, ja) in enumerate(SHEAR_PAIRS):
            acc = F[6 + s, t0][inner].copy()
            mq = med[(9, 11, 10)[s]][inner] if p.media_arrays else None
            for g, (src, ax) in enumerate(((ia, ja), (ja, ia))):
                for k in range(M):
                    c = R(p.cs[s, g, k])
                    fp = self._sh(F[src, t0], ax, 1 + k)
                    fm = self._sh(F[src, t0], ax, -k)
                    if not p.media_arrays:
                        acc = acc + c * fp
                        acc = acc - c * fm
                    else:
                        acc = acc + (c * mq) * (fp - fm)
            F[6 + s, t1][inner] = acc

    def velocity(self, t0, t1):
        p, F, R = self.p, self.F, self.dt_real
        M = p.M
        inner = tuple(slice(M, d - M) for d in p.dims)
        for i in range(3):
            acc = F[i, t0][inner].copy()
            b = self.media[6 + i][inner] if p.media_arrays else None
            for j in range(3):
                src = 3 + i if j == i else shear_slot(i, j)
                a = 1 if j == i else 0
                for k in range(M):
                    c = R(p.cv[i, j, k])
                    fp = self._sh(F[src, t1], j, a + k)
                    fm = self._sh(F[src, t1], j, a - 1 - k)
                    if not p.media_arrays:
                        acc = acc + c * fp
                        acc = acc - c * fm
                    else:
                        acc = acc + (c * b) * (fp - fm)
            F[i, t1][inner] = acc

    # ---------- boundary program ----------
    def bc(self, phases, t0, t1):
        """Run the statements of the given phases.  Scheduled order: phase by
        phase as listed (statements folded into the interior kernel first),
        launch level, then position.  Sequential order: the reference emission
        order (pieces of one split statement together)."""
        for i in self.stmt_order(phases):
            self.apply_stmt(i, (t0, t1))

    def stmt_order(self, phases):
        p = self.p
        st = p.stmts
        if isinstance(phases, int):
            phases = (phases,)
        idx = [i for i in range(len(st)) if int(st['phase'][i]) in phases]
        if self.order_by_sched:
            rank = {ph: r for r, ph in enumerate(phases)}
            return sorted(idx, key=lambda i: (rank[int(st['phase'][i])], st['sched'][i], i))
        seq = getattr(p, 'stmt_seq', np.arange(len(st)))
        return sorted(idx, key=lambda i: (seq[i], i))

    def stmt_geometry(self, i):
        """(axis, in-plane axes, lo, hi) of statement i, or None when empty."""
        s = self.p.stmts[i]
        ax = int(s['axis'])
        oth = [a for a in range(3) if a != ax]
        lo, hi = [int(v) for v in s['lo']], [int(v) for v in s['hi']]
        if hi[0] - lo[0] <= 0 or hi[1] - lo[1] <= 0:
            return None
        return ax, oth, lo, hi

    def term_slices(self, i, t):
        """Source slice of term t of statement i, and the media slicer for an
        offset vector of its coefficient program."""
        ax, oth, lo, hi = self.stmt_geometry(i)
        sl = [None] * 3
        sl[ax] = int(t['plane'])
        for k, a in enumerate(oth):
            sl[a] = slice(lo[k] + int(t['off'][k]), hi[k] + int(t['off'][k]))

        def media_slice(offs):
            msl = [None] * 3
            msl[ax] = int(offs[ax])
            for k, a in enumerate(oth):
                msl[a] = slice(lo[k] + int(offs[a]), hi[k] + int(offs[a]))
            return tuple(msl)
        return tuple(sl), media_slice

    def dst_slice(self, i):
        ax, oth, lo, hi = self.stmt_geometry(i)
        dsl = [None] * 3
        dsl[ax] = int(self.p.stmts[i]['plane'])
        for k, a in enumerate(oth):
            dsl[a] = slice(lo[k], hi[k])
        return tuple(dsl)

    def apply_stmt(self, i, lv):
        p, F, R = self.p, self.F, self.dt_real
        if self.stmt_geometry(i) is None:
            return
        s = p.stmts[i]
        ax, oth, lo, hi = self.stmt_geometry(i)
        acc = np.zeros((hi[0] - lo[0], hi[1] - lo[1]), dtype=R)
        for t in p.terms[s['term0']:s['term0'] + s['nterms']]:
            sl, media_slice = self.term_slices(i, t)
            src = F[int(t['field']), lv[int(t['level'])]][sl]
            if t['prog'] >= 0:
                code = _code(p.bc_code, int(t['prog']), int(t['prog_len']))
                c = np.asarray(eval_code(code, None, 0.0, lambda mid, offs: self.media[mid][media_slice(offs)]
                                         .astype(np.float64))) * float(t['coef'])
                acc = acc + c.astype(R) * src
            else:
                acc = acc + R(t['coef']) * src
        F[int(s['field']), lv[int(s['level'])]][self.dst_slice(i)] = acc

    def step(self):
        t0 = self.steps % 2
        t1 = 1 - t0
        self.stress(t0, t1)
        self.bc((4, 2), t0, t1)
        self.velocity(t0, t1)
        self.bc(3, t0, t1)
        self.steps += 1

    def run(self, n=None):
        for _ in range(self.p.ntsteps if n is None else n):
            self.step()

    def latest(self):
        return self.F[:, self.steps % 2]

    def l2(self):
        p = self.p
        out = []
        vol = p.spacing[0] * p.spacing[1] * p.spacing[2]
        for f in range(9):
            tn = p.dt * self.steps + (p.dt / 2.0 if TIME_STAG[f] else 0.0)
            rg = self.ranges(f)
            sl = tuple(slice(lo, hi) for lo, hi in rg)
            e = self.latest()[f][sl].astype(np.float64) - self.solution(f, tn, rg)
            out.append(float(np.sqrt(np.sum(e * e) * vol)))
        return out
