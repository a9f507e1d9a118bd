"""Lowering of a StaggeredGrid model to the data the gfx950 library runs.

The reference prints C++ loop nests (opesci/regulargrid.py:566-619,
opesci/staggeredgrid.py:728-864); here the same derived expressions are
lowered to:

* folded stencil coefficients for the hand-written interior kernels, after
  checking that every derived kernel has exactly the velocity-stress stencil
  structure those kernels implement (anything else raises);
* a boundary program: one statement per ghost-cell equation, in the
  reference's emission order (stress_bc_getter / velocity_bc /
  initialise_bc), each tagged with a launch index from a read/write hazard
  analysis so that statements touching disjoint cells share a launch while
  the result stays identical to the sequential loops;
* device bytecode for the analytic solutions (initial state, L2 norms).

Coefficients follow the reference's printed literals: every folded Float is
printed with 15 significant digits and an F suffix (codeprinter.py:42-63),
so the values are float-rounded even in the double build.
"""
import os

import numpy as np
from sympy import Add, Float, Indexed, Integer, Mul, Symbol, expand

from . import exprvm

SLOT_NAMES = ('U', 'V', 'W', 'Txx', 'Tyy', 'Tzz', 'Txy', 'Tyz', 'Txz')
STAG = ((1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, 0), (0, 0, 0), (0, 0, 0),
        (1, 1, 0), (0, 1, 1), (1, 0, 1))
SHEAR_PAIRS = ((0, 1), (1, 2), (0, 2))
SHEAR_MEDIA = ('mu12', 'mu23', 'mu13')
STORAGE_CODE = {'native': 0, 'fp32': 0, 'fp64': 0, 'fp16': 1, 'bf16': 2}

STMT_DTYPE = np.dtype([('field', '<i4'), ('level', '<i4'), ('axis', '<i4'), ('plane', '<i4'),
                       ('lo', '<i4', (2,)), ('hi', '<i4', (2,)), ('term0', '<i4'),
                       ('nterms', '<i4'), ('phase', '<i4'), ('sched', '<i4')])
TERM_DTYPE = np.dtype([('field', '<i4'), ('level', '<i4'), ('plane', '<i4'), ('off', '<i4', (2,)),
                       ('prog', '<i4'), ('prog_len', '<i4'), ('pad', '<i4'), ('coef', '<f8')])
DTERM_DTYPE = np.dtype([('term', '<i4'), ('media', '<i4'), ('off', '<i4', (3,)), ('prog', '<i4'),
                        ('prog_len', '<i4'), ('pad', '<i4')])


class PlanError(NotImplementedError):
    pass


def f_lit(v):
    """Value of a folded constant as the reference prints it: %.15g + F suffix."""
    return float(np.float32(float('%.15g' % float(v))))


def shear_slot(i, j):
    i, j = min(i, j), max(i, j)
    return {(0, 1): 6, (1, 2): 7, (0, 2): 8}[(i, j)]


class ModelPlan(object):
    """Everything libopesci_hip needs for one model (see OpesciModelDesc)."""

    def __repr__(self):
        return '<ModelPlan order=%d dims=%s steps=%d real=%d storage=%s media=%s bc=%d stmts/%d launches>' % (
            self.order, self.dims, self.ntsteps, self.real_bytes, self.storage,
            'arrays' if self.media_arrays else 'const', len(self.stmts), self.nlaunches_step)


def _canonical_slots(grid):
    """Map the grid's fields onto U,V,W,Txx,Tyy,Tzz,Txy,Tyz,Txz by direction."""
    slots = [None] * 9
    for v in grid.vfields:
        slots[v.direction - 1] = v
    for s in grid.sfields:
        a, b = sorted(s.direction)
        slots[3 + a - 1 if a == b else shear_slot(a - 1, b - 1)] = s
    if any(s is None for s in slots):
        raise PlanError('the elastic backend needs 3 velocity and 6 stress fields')
    if [f for f in slots] != list(grid.fields):
        raise PlanError('fields must be given as velocity (x,y,z) then stress '
                        '(xx,yy,zz,xy,yz,xz) components, the order of StaggeredGrid.fields')
    return slots


def _time_level(ind, grid):
    s = str(ind)
    if s in ('_t0', str(grid.t)):
        return 0
    if s in ('_t1', str(grid.t + 1)):
        return 1
    raise PlanError('unexpected time index %s' % ind)


def _split(term, field_names):
    """term = number * factors * field -> (number, [other factors], field Indexed)."""
    num = Integer(1)
    other = []
    fld = None
    for f in Mul.make_args(term):
        if f.is_Number:
            num *= f
        elif isinstance(f, Indexed) and str(f.base.label) in field_names:
            if fld is not None:
                raise PlanError('non-linear term %s' % term)
            fld = f
        else:
            other.append(f)
    return num, other, fld


# ---------------------------------------------------------------- interior
def _extract_interior(grid, slots):
    names = {str(f.label): k for k, f in enumerate(slots)}
    M = grid.order[1] // 2
    idx = grid.index
    cn = np.zeros((3, 3, 4))
    cs = np.zeros((3, 2, 4))
    cv = np.zeros((3, 3, 4))
    for slot, field in enumerate(slots):
        if field.kernel_aligned is None:
            raise PlanError('solve_fd() has not been called')
        k = grid.kernel_sympy(field)
        table = {}
        for term in Add.make_args(expand(k)):
            num, other, fld = _split(term, names)
            if fld is None:
                raise PlanError('%s kernel has a term without a field: %s' % (field, term))
            src = names[str(fld.base.label)]
            lvl = _time_level(fld.indices[0], grid)
            off = tuple(int(fld.indices[1 + a] - idx[a]) for a in range(3))
            media = None
            if other:
                if len(other) != 1 or not isinstance(other[0], Indexed):
                    raise PlanError('unsupported coefficient %s in %s kernel' % (other, field))
                m = other[0]
                if any(m.indices[a] != idx[a] for a in range(3)):
                    raise PlanError('media must be read at the update point: %s' % m)
                media = str(m.base.label)
            table.setdefault((src, lvl, off), {})
            table[(src, lvl, off)][media] = table[(src, lvl, off)].get(media, 0) + num
        if table.pop((slot, 0, (0, 0, 0)), None) != {None: 1}:
            raise PlanError('%s kernel must be F[t1] = F[t0] + ...' % field)

        def take(src, lvl, axis, a, kk, want):
            op, om = [0, 0, 0], [0, 0, 0]
            op[axis] = a + kk
            om[axis] = a - 1 - kk
            cp = table.pop((src, lvl, tuple(op)), None)
            cm = table.pop((src, lvl, tuple(om)), None)
            if cp is None or cm is None:
                raise PlanError('%s kernel misses the %s stencil pair %d along axis %d'
                                % (field, SLOT_NAMES[src], kk, axis))
            if set(cp) != set(want) or set(cm) != set(want):
                raise PlanError('%s kernel: media factors %s, expected %s' % (field, list(cp), want))
            for w in want:
                if f_lit(cp[w]) != -f_lit(cm[w]):
                    raise PlanError('%s kernel stencil is not antisymmetric' % field)
            return cp

        if slot < 3:
            i = slot
            for j in range(3):
                src = 3 + i if j == i else shear_slot(i, j)
                a = 1 if j == i else 0
                for kk in range(M):
                    want = [None] if not grid.read else ['beta%d' % (i + 1)]
                    c = take(src, 1, j, a, kk, want)
                    cv[i, j, kk] = f_lit(c[want[0]])
        elif slot < 6:
            i = slot - 3
            for j in range(3):
                for kk in range(M):
                    if not grid.read:
                        c = take(j, 0, j, 0, kk, [None])
                        cn[i, j, kk] = f_lit(c[None])
                    else:
                        want = ['lambda', 'mu'] if j == i else ['lambda']
                        c = take(j, 0, j, 0, kk, want)
                        if j == i and f_lit(c['mu']) != 2 * f_lit(c['lambda']):
                            raise PlanError('%s kernel: expected (lambda + 2 mu) on its own axis' % field)
                        cn[i, j, kk] = f_lit(c['lambda'])
        else:
            s = slot - 6
            ia, ja = SHEAR_PAIRS[s]
            for g, (src, axis) in enumerate(((ia, ja), (ja, ia))):
                for kk in range(M):
                    want = [None] if not grid.read else [SHEAR_MEDIA[s]]
                    c = take(src, 0, axis, 1, kk, want)
                    cs[s, g, kk] = f_lit(c[want[0]])
        if table:
            raise PlanError('%s kernel has terms outside the elastic stencil: %s'
                            % (field, sorted(table)[:4]))
    return cn, cs, cv


# ---------------------------------------------------------------- boundaries
def _bc_ops(grid, slots, init):
    """(phase, field, axis, side, ranges) in the reference's emission order."""
    m = grid.margin.value
    dims = [d.value for d in grid.dim]
    ops = []
    stress = [s for s in grid.sfields]
    for f in stress:
        normal = f.direction[0] == f.direction[1]
        for d in range(3):
            if normal and f.direction[0] != d + 1 and (init or grid.legacy_levander):
                continue
            for side in range(2):
                bc = f.bc[d + 1][side]
                if not bc:
                    continue
                rng = []
                for d2 in range(3):
                    if d2 == d:
                        continue
                    if normal and f.direction[0] != d + 1:
                        rng.append((m + 1, dims[d2] - m - 1))
                    else:
                        rng.append((0, dims[d2]))
                ops.append((0 if init else 2, f, d, side, rng))
    for d in range(3):
        seq = [f for f in grid.vfields if f.staggered[d + 1]] + \
              [f for f in grid.vfields if not f.staggered[d + 1]]
        for f in seq:
            for side in range(2):
                if not f.bc[d + 1][side]:
                    continue
                rng = [(1, dims[d2] - 1) for d2 in range(3) if d2 != d]
                ops.append((1 if init else 3, f, d, side, rng))
    return ops


class _BcBuilder(object):
    def __init__(self, grid, slots):
        self.grid = grid
        self.slots = slots
        self.names = {str(f.label): k for k, f in enumerate(slots)}
        self.stmts = []
        self.terms = []
        self.code = []
        self.dterms = []
        self.op_of_stmt = []

    def media_anchor_for(self, axis):
        idx = self.grid.index

        def anchor(e):
            name = str(e.base.label)
            if name not in exprvm.MEDIA_IDS:
                raise PlanError('unknown media array %s' % name)
            offs = []
            for a in range(3):
                v = e.indices[a]
                offs.append(int(v) if a == axis else int(v - idx[a]))
            return exprvm.MEDIA_IDS[name], offs
        return anchor

    def add_op(self, op_id, phase, field, axis, side, ranges):
        g = self.grid
        idx = g.index
        others = [a for a in range(3) if a != axis]
        eqs = g.transform_bc(field, axis + 1, side)
        if g.read:
            eqs = [g.resolve_media_params(e) for e in eqs]
        for eq in eqs:
            lhs = eq.lhs
            if not isinstance(lhs, Indexed) or str(lhs.base.label) not in self.names:
                raise PlanError('boundary equation with lhs %s' % lhs)
            dst = self.names[str(lhs.base.label)]
            for a in others:
                if lhs.indices[1 + a] != idx[a]:
                    raise PlanError('boundary lhs must loop over the tangential axes: %s' % lhs)
            plane = int(lhs.indices[1 + axis])
            t0 = len(self.terms)
            rhs = expand(eq.rhs)
            # group the terms reading the same cell: constant coefficients add,
            # media-dependent ones become one summed coefficient program
            groups = {}
            order = []
            if rhs != 0:
                for term in Add.make_args(rhs):
                    num, other, fld = _split(term, self.names)
                    if fld is None:
                        raise PlanError('boundary term without a field: %s' % term)
                    src = self.names[str(fld.base.label)]
                    lvl = _time_level(fld.indices[0], g)
                    p = int(fld.indices[1 + axis])
                    offs = tuple(int(fld.indices[1 + a] - idx[a]) for a in others)
                    key = (src, lvl, p, offs)
                    if key not in groups:
                        groups[key] = []
                        order.append(key)
                    groups[key].append((num, other))
            # one canonical term order (field, level, plane, offsets): the
            # boundary launches and the sweep kernels that apply a statement
            # in place sum its terms in the same order
            for key in sorted(order):
                src, lvl, p, offs = key
                parts = groups[key]
                if all(not o for _, o in parts):
                    prog, plen, coef = -1, 0, f_lit(sum(f_lit(n) for n, _ in parts))
                elif len(parts) == 1:
                    num, other = parts[0]
                    expr = Mul(*other)
                    code = exprvm.compile_expr(expr, coords={}, consts={},
                                               media_anchor=self.media_anchor_for(axis))
                    code = [(op, a, f_lit(v) if op == exprvm.OP_CONST else v) for op, a, v in code]
                    prog, plen, coef = len(self.code), len(code), f_lit(num)
                    self.code += code
                    self._derivatives(expr, axis)
                else:
                    expr = Add(*[Float(f_lit(n)) * Mul(*o) for n, o in parts], evaluate=False)
                    code = exprvm.compile_expr(expr, coords={}, consts={},
                                               media_anchor=self.media_anchor_for(axis))
                    prog, plen, coef = len(self.code), len(code), 1.0
                    self.code += code
                    self._derivatives(expr, axis)
                self.terms.append((src, lvl, p, list(offs), prog, plen, coef))
            self.stmts.append(dict(field=dst, level=_time_level(lhs.indices[0], g), axis=axis,
                                   plane=plane, lo=[r[0] for r in ranges], hi=[r[1] for r in ranges],
                                   term0=t0, nterms=len(self.terms) - t0, phase=phase, sched=0))
            self.op_of_stmt.append(op_id)

    def _derivatives(self, expr, axis):
        """Partial derivatives of a media-dependent coefficient program with
        respect to each media cell it reads (for the adjoint's media
        gradient): one (term, media id, offsets, program) entry per cell; the
        term is the one about to be appended."""
        term = len(self.terms)
        anchor = self.media_anchor_for(axis)
        cells = sorted((e for e in expr.atoms(Indexed) if str(e.base.label) in exprvm.MEDIA_IDS), key=str)
        for cell in cells:
            d = expr.diff(cell)
            if d == 0:
                continue
            mid, offs = anchor(cell)
            code = exprvm.compile_expr(d, coords={}, consts={}, media_anchor=anchor)
            self.dterms.append((term, mid, list(offs), len(self.code), len(code)))
            self.code += code

    # -------- hazard analysis
    def _boxes(self, s, lo=None, hi=None):
        """write box and read boxes of a statement over its loop ranges (or over
        [lo, hi) instead): (field, level, ((lo,hi) x3))."""
        ax = s['axis']
        lo = s['lo'] if lo is None else lo
        hi = s['hi'] if hi is None else hi
        others = [a for a in range(3) if a != ax]
        w = [None] * 3
        w[ax] = (s['plane'], s['plane'] + 1)
        for k, a in enumerate(others):
            w[a] = (lo[k], hi[k])
        wbox = (s['field'], s['level'], tuple(w))
        rboxes = []
        for t in self.terms[s['term0']:s['term0'] + s['nterms']]:
            src, lvl, p, offs = t[0], t[1], t[2], t[3]
            r = [None] * 3
            r[ax] = (p, p + 1)
            for k, a in enumerate(others):
                r[a] = (lo[k] + offs[k], hi[k] + offs[k])
            rboxes.append((src, lvl, tuple(r)))
        return wbox, rboxes

    def _conflict(self, a, b):
        (wa, ra), (wb, rb) = a, b
        return self._hit(wa, wb) or any(self._hit(wa, r) for r in rb) or any(self._hit(r, wb) for r in ra)

    # -------- z-face statements folded into the interior kernel
    @staticmethod
    def _deep_range(s, F, dims):
        lo = [max(s['lo'][0], F), max(s['lo'][1], F)]
        hi = [min(s['hi'][0], dims[0] - F), min(s['hi'][1], dims[1] - F)]
        if lo[0] >= hi[0] or lo[1] >= hi[1]:
            return None
        return lo, hi

    @staticmethod
    def _frame(s, lo, hi):
        """Loop box of s minus [lo, hi): up to four rectangles (x slabs, then y slabs)."""
        L, H = s['lo'], s['hi']
        out = []
        for r in (((L[0], lo[0]), (L[1], H[1])), ((hi[0], H[0]), (L[1], H[1])),
                  ((lo[0], hi[0]), (L[1], lo[1])), ((lo[0], hi[0]), (hi[1], H[1]))):
            if r[0][0] < r[0][1] and r[1][0] < r[1][1]:
                out.append(([r[0][0], r[1][0]], [r[0][1], r[1][1]]))
        return out

    def _try_fuse(self, idx, F, dims):
        """Rows (x, y) in [F, dim-F)^2 of the z-face statements idx (reference
        order within one phase), moved to run right after the interior update
        and before every other boundary statement of the phase.  Valid when
          * a moved part reads fields written in the phase only in its own row
            (offsets 0, 0) and only cells no moved part writes: it then sees the
            interior kernel's value, as in the reference order;
          * moved parts never write the same cell twice;
          * no statement left in place that precedes a moved part in the
            reference order conflicts with it (read-after-write,
            write-after-read, write-after-write).
        Returns {statement index: (lo, hi)} or None."""
        st = self.stmts
        phase = st[idx[0]]['phase']
        members = [i for i, s in enumerate(st) if s['phase'] == phase]
        written = set((st[i]['field'], st[i]['level']) for i in members)
        deep = {}
        for j in idx:
            d = self._deep_range(st[j], F, dims)
            if d is not None:
                deep[j] = d
        if not deep:
            return None
        wcells = set()
        for j, (lo, hi) in deep.items():
            s = st[j]
            key = (s['field'], s['level'], s['plane'])
            if key in wcells:
                return None
            wcells.add(key)
        for j in deep:
            for t in self.terms[st[j]['term0']:st[j]['term0'] + st[j]['nterms']]:
                if (t[0], t[1]) in written:
                    if tuple(t[3]) != (0, 0) or (t[0], t[1], t[2]) in wcells:
                        return None
        for j, (lo, hi) in deep.items():
            bj = self._boxes(st[j], lo, hi)
            for i in members:
                if i >= j:
                    break
                if i in deep:
                    pieces = [self._boxes(st[i], a, b) for a, b in self._frame(st[i], *deep[i])]
                else:
                    pieces = [self._boxes(st[i])]
                if any(self._conflict(bi, bj) for bi in pieces):
                    return None
        return deep

    def fuse_z(self, phase, fused_phase, dims, margin):
        """Split the z-face statements of a phase into a part folded into the
        interior kernel (fused_phase) and the frame rows that stay in place,
        for the smallest frame width F that keeps the reference semantics."""
        idx = [i for i, s in enumerate(self.stmts) if s['phase'] == phase and s['axis'] == 2]
        if not idx:
            return 0
        for F in range(margin + 1, min(dims[0], dims[1]) // 2):
            deep = self._try_fuse(idx, F, dims)
            if deep is not None:
                break
        else:
            return 0
        stmts, ops, seq = [], [], []
        for i, s in enumerate(self.stmts):
            if i not in deep:
                stmts.append(s)
                ops.append(self.op_of_stmt[i])
                seq.append(self.seq[i])
                continue
            lo, hi = deep[i]
            stmts.append(dict(s, lo=list(lo), hi=list(hi), phase=fused_phase))
            ops.append(self.op_of_stmt[i])
            seq.append(self.seq[i])
            for a, b in self._frame(s, lo, hi):
                stmts.append(dict(s, lo=list(a), hi=list(b)))
                ops.append(self.op_of_stmt[i])
                seq.append(self.seq[i])
        self.stmts, self.op_of_stmt, self.seq = stmts, ops, seq
        return F

    @staticmethod
    def _hit(a, b):
        if a[0] != b[0] or a[1] != b[1]:
            return False
        return all(x[0] < y[1] and y[0] < x[1] for x, y in zip(a[2], b[2]))

    def schedule(self):
        boxes = [self._boxes(s) for s in self.stmts]
        for j, sj in enumerate(self.stmts):
            wj, rj = boxes[j]
            lvl = 0
            for i in range(j):
                si = self.stmts[i]
                if si['phase'] != sj['phase']:
                    continue
                wi, ri = boxes[i]
                conflict = self._hit(wi, wj) or any(self._hit(wi, r) for r in rj) or \
                    any(self._hit(r, wj) for r in ri)
                if conflict:
                    if self.op_of_stmt[i] == self.op_of_stmt[j]:
                        raise PlanError('boundary statements of one loop depend on each other; '
                                        'the sequential per-point semantics cannot be scheduled')
                    lvl = max(lvl, si['sched'] + 1)
            sj['sched'] = lvl

    def arrays(self):
        st = np.zeros(len(self.stmts), dtype=STMT_DTYPE)
        for i, s in enumerate(self.stmts):
            for k in ('field', 'level', 'axis', 'plane', 'term0', 'nterms', 'phase', 'sched'):
                st[i][k] = s[k]
            st[i]['lo'] = s['lo']
            st[i]['hi'] = s['hi']
        te = np.zeros(len(self.terms), dtype=TERM_DTYPE)
        for i, (src, lvl, p, offs, prog, plen, coef) in enumerate(self.terms):
            te[i]['field'] = src
            te[i]['level'] = lvl
            te[i]['plane'] = p
            te[i]['off'] = offs
            te[i]['prog'] = prog
            te[i]['prog_len'] = plen
            te[i]['coef'] = coef
        dt = np.zeros(len(self.dterms), dtype=DTERM_DTYPE)
        for i, (term, mid, offs, prog, plen) in enumerate(self.dterms):
            dt[i]['term'] = term
            dt[i]['media'] = mid
            dt[i]['off'] = offs
            dt[i]['prog'] = prog
            dt[i]['prog_len'] = plen
        return st, te, exprvm.to_array(self.code), dt


def _build_bc(grid, slots):
    b = _BcBuilder(grid, slots)
    op_id = 0
    for init in (True, False):
        for phase, f, d, side, rng in _bc_ops(grid, slots, init):
            b.add_op(op_id, phase, f, d, side, rng)
            op_id += 1
    # stable order by phase keeps the reference order inside each phase
    order = sorted(range(len(b.stmts)), key=lambda i: b.stmts[i]['phase'])
    b.stmts = [b.stmts[i] for i in order]
    b.op_of_stmt = [b.op_of_stmt[i] for i in order]
    b.seq = list(range(len(b.stmts)))
    b.fused_frame = 0
    # split by default where the FP32 pair stress sweep applies the deep rows
    # in place (constant media); elsewhere they would only add a launch
    if getattr(grid, 'fuse_boundary', not grid.read and not grid.double):
        dims = [d.value for d in grid.dim]
        b.fused_frame = b.fuse_z(2, 4, dims, grid.margin.value)
    b.schedule()
    return b


# ---------------------------------------------------------------- analytic
def _media_consts(grid):
    out = {}
    if not grid.read:
        for n in ('rho', 'beta', 'lambda', 'mu'):
            v = grid.defined_variable[n]
            out[Symbol(n)] = float(np.float32(v.value)) if v.type == 'float' else float(v.value)
    grid.create_const_dict()
    for s, v in grid.const_dict.items():
        if s not in out and isinstance(v, (int, float)) and s.name not in ('dim1', 'dim2', 'dim3'):
            out[s] = float(v)
    return out


def _solutions(grid, slots):
    consts = _media_consts(grid)
    coords = {s: a for a, s in enumerate(grid.index)}
    code, off, ln = [], [], []
    for f in slots:
        sol = f.sol
        if sol is None:
            sol = Integer(0)
        if grid.read:
            sol = grid.resolve_media_params(sol.subs(grid.media_dict))

        def anchor(e):
            name = str(e.base.label)
            offs = [int(e.indices[a] - grid.index[a]) for a in range(3)]
            return exprvm.MEDIA_IDS[name], offs
        c = exprvm.compile_expr(sol, coords=coords, time=grid.t, consts=consts, media_anchor=anchor)
        off.append(len(code))
        ln.append(len(c))
        code += c
    return exprvm.to_array(code), off, ln


def _media_arrays(grid, dims):
    shape = tuple(dims)
    n = dims[0] * dims[1] * dims[2]
    if grid.media_arrays is not None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(shape) for a in grid.media_arrays]
    else:
        arrs = []
        for fn in (grid.rho_file, grid.vp_file, grid.vs_file):
            if not fn or not os.path.exists(fn):
                raise IOError('media file %r not found' % fn)
            a = np.fromfile(fn, dtype=np.float32)
            if a.size < n:
                raise IOError('media file %s holds %d values, %d needed' % (fn, a.size, n))
            arrs.append(a[:n].reshape(shape))
    return arrs


# ---------------------------------------------------------------- sources
def _sources(grid, p):
    """Source traces as (1 or 3, ntsteps) float32, zero after their end (the
    reference injects while ti < snt, tests/src/test_ref_iso_elastic.cpp:
    275-290), the source index and the receiver indices (n, 3), all checked
    against the grid."""
    src, idx, rec = getattr(grid, 'src_wavelet', None), getattr(grid, 'src_idx', None), \
        getattr(grid, 'rec_idx', None)
    nt = max(p.ntsteps, 1)
    if src is not None:
        w = np.atleast_2d(np.asarray(src, dtype=np.float32))
        if w.ndim != 2 or w.shape[0] not in (1, 3):
            raise PlanError('a source is one trace or three (x, y, z)')
        out = np.zeros((w.shape[0], nt), dtype=np.float32)
        n = min(nt, w.shape[1])
        out[:, :n] = w[:, :n]
        src = out
        idx = tuple(int(v) for v in idx)
        if any(not (0 <= i < d) for i, d in zip(idx, p.dims)):
            raise PlanError('source index %s outside the grid %s' % (idx, p.dims))
    if rec is not None:
        rec = np.ascontiguousarray(np.asarray(rec, dtype=np.int32).reshape(-1, 3))
        if rec.size and ((rec < 0).any() or (rec >= np.asarray(p.dims)).any()):
            raise PlanError('receiver outside the grid %s' % (p.dims,))
    return src, (idx if idx is not None else (0, 0, 0)), rec


# ---------------------------------------------------------------- entry
def build_plan(grid):
    if grid.dimension != 3:
        raise PlanError('the elastic backend is 3-D')
    if grid.order[0] != 2:
        raise PlanError('leapfrog: time order must be 2')
    orders = set(grid.order[1:])
    if len(orders) != 1 or grid.order[1] not in (2, 4, 6, 8):
        raise PlanError('equal spatial orders 2, 4, 6 or 8 are supported')
    slots = _canonical_slots(grid)
    p = ModelPlan()
    p.order = grid.order[1]
    p.M = p.order // 2
    p.dims = tuple(d.value for d in grid.dim)
    p.grid_size = tuple(grid.grid_size)
    p.ntsteps = int(grid.ntsteps.value)
    p.dt = float(grid.dt.value)
    p.spacing = tuple(float(s.value) for s in grid.spacing)
    p.real_bytes = 8 if grid.double else 4
    p.storage = grid.storage
    st = STORAGE_CODE.get(grid.storage)
    if st is None:
        raise PlanError('unknown storage %s' % grid.storage)
    if st and grid.double:
        raise PlanError('16-bit storage needs the float (FP32 compute) build')
    p.storage_code = st
    p.media_arrays = bool(grid.read)
    p.media_const = np.zeros(12)
    if not grid.read:
        dv = grid.defined_variable
        for n, i in (('rho', 0), ('beta', 3), ('lambda', 4), ('mu', 5)):
            p.media_const[i] = float(np.float32(dv[n].value))
        rho, lam, mu = dv['rho'].value, dv['lambda'].value, dv['mu'].value
        p.media_const[1] = float(np.float32(((lam + 2 * mu) / rho) ** 0.5))
        p.media_const[2] = float(np.float32((mu / rho) ** 0.5))
        p.media = None
    else:
        p.media = _media_arrays(grid, p.dims)
    p.cn, p.cs, p.cv = _extract_interior(grid, slots)
    b = _build_bc(grid, slots)
    p.stmts, p.terms, p.bc_code, p.dterms = b.arrays()
    p.stmt_seq = np.asarray(b.seq, dtype=np.int64)
    p.fused_frame = b.fused_frame
    p.nlaunches_step = int(sum(len(set(p.stmts['sched'][p.stmts['phase'] == ph])) for ph in (2, 3)))
    p.sol_code, p.sol_off, p.sol_len = _solutions(grid, slots)
    p.src_wavelet, p.src_idx, p.rec_idx = _sources(grid, p)
    p.tile_y = getattr(grid, 'tile_y', 0)
    p.x_chunk = getattr(grid, 'x_chunk', 0)
    p.kernel_variant = getattr(grid, 'kernel_variant', 0)
    # both half steps in one sweep where it applies (off by default: slower
    # than the separate launches on MI355X)
    p.fuse_step = bool(getattr(grid, 'fuse_step', False))
    # the analytic solution evaluated per point as given, without the
    # per-axis tables of its one-coordinate parts (same values, slower)
    p.pointwise_solution = bool(getattr(grid, 'pointwise_solution', False))
    p.use_graph = 1 if grid.use_graph else 0
    p.flops_per_point = None
    p.field_names = tuple(str(f.label) for f in slots)
    return p
