"""SymPy expression -> device bytecode (stack machine evaluated in double).

Used for the analytic solutions (initial state and L2 norms,
opesci/staggeredgrid.py:612-659, :892-945) and for media-dependent boundary
coefficients of the read path.  Opcodes match OPESCI_OP_* in
include/opesci_hip.h.
"""
import numpy as np
from sympy import Float, Indexed, Integer, Rational, Symbol, cos, exp, log, pi, sin, sqrt
from sympy.core.numbers import Number, NumberSymbol

OP_CONST, OP_COORD, OP_TIME, OP_MEDIA, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_NEG, OP_POWC, \
    OP_SQRT, OP_SIN, OP_COS, OP_EXP, OP_LOG, OP_ROUNDF = range(16)

MEDIA_IDS = {'rho': 0, 'vp': 1, 'vs': 2, 'beta': 3, 'lambda': 4, 'mu': 5, 'beta1': 6,
             'beta2': 7, 'beta3': 8, 'mu12': 9, 'mu13': 10, 'mu23': 11}

OP_DTYPE = np.dtype([('op', '<i4'), ('a', '<i4', (4,)), ('pad', '<i4'), ('val', '<f8')])


class ExprCompiler(object):
    """Compiles an expression into a list of (op, a0..a3, val).

    coords: {Symbol: axis}; time: Symbol or None; consts: {Symbol: value};
    media_anchor: function(Indexed) -> (media_id, (o0, o1, o2)) for media loads.
    """

    def __init__(self, coords, time=None, consts=None, media_anchor=None):
        self.coords = coords
        self.time = time
        self.consts = consts or {}
        self.media_anchor = media_anchor
        self.code = []
        self.depth = 0
        self.max_depth = 0

    def _push(self, op, a=(0, 0, 0, 0), val=0.0, delta=1):
        self.code.append((op, tuple(a) + (0,) * (4 - len(a)), float(val)))
        self.depth += delta
        self.max_depth = max(self.max_depth, self.depth)

    def compile(self, e):
        if isinstance(e, (Number, NumberSymbol)) or e in (pi,):
            self._push(OP_CONST, val=float(e))
            return
        if e.is_Symbol:
            if e in self.coords:
                self._push(OP_COORD, (self.coords[e],))
            elif self.time is not None and e == self.time:
                self._push(OP_TIME)
            elif e in self.consts:
                self._push(OP_CONST, val=float(self.consts[e]))
            else:
                raise ValueError('unbound symbol %s in analytic expression' % e)
            return
        if isinstance(e, Indexed):
            if self.media_anchor is None:
                raise ValueError('array reference %s not allowed here' % e)
            mid, offs = self.media_anchor(e)
            self._push(OP_MEDIA, (mid,) + tuple(offs))
            return
        if e.is_Add:
            args = e.args
            self.compile(args[0])
            for a in args[1:]:
                self.compile(a)
                self._push(OP_ADD, delta=-1)
            return
        if e.is_Mul:
            args = e.args
            self.compile(args[0])
            for a in args[1:]:
                self.compile(a)
                self._push(OP_MUL, delta=-1)
            return
        if e.is_Pow:
            b, ex = e.args
            if ex == Rational(1, 2):
                self.compile(b)
                self._push(OP_SQRT, delta=0)
            elif ex == -1:
                self._push(OP_CONST, val=1.0)
                self.compile(b)
                self._push(OP_DIV, delta=-1)
            elif ex == Rational(-1, 2):
                self._push(OP_CONST, val=1.0)
                self.compile(b)
                self._push(OP_SQRT, delta=0)
                self._push(OP_DIV, delta=-1)
            elif ex.is_Number:
                self.compile(b)
                self._push(OP_POWC, val=float(ex), delta=0)
            else:
                self.compile(log(b) * ex)
                self._push(OP_EXP, delta=0)
            return
        fn = {sin: OP_SIN, cos: OP_COS, exp: OP_EXP, log: OP_LOG}
        for k, op in fn.items():
            if isinstance(e, k):
                self.compile(e.args[0])
                self._push(op, delta=0)
                return
        if e.func == sqrt:
            self.compile(e.args[0])
            self._push(OP_SQRT, delta=0)
            return
        raise NotImplementedError('cannot compile %s (%s) for the device' % (e, type(e).__name__))


def compile_expr(e, **kw):
    c = ExprCompiler(**kw)
    c.compile(e)
    if c.max_depth > 16:
        raise ValueError('expression too deep for the device evaluator (%d > 16)' % c.max_depth)
    return c.code


def to_array(code):
    arr = np.zeros(len(code), dtype=OP_DTYPE)
    for i, (op, a, val) in enumerate(code):
        arr[i]['op'] = op
        arr[i]['a'] = a
        arr[i]['val'] = val
    return arr


def evaluate(code, coords=(0.0, 0.0, 0.0), t=0.0, media=None):
    """Host evaluator with the device's semantics (used by the tests)."""
    st = []
    for op, a, val in code:
        if op == OP_CONST:
            st.append(val)
        elif op == OP_COORD:
            st.append(coords[a[0]])
        elif op == OP_TIME:
            st.append(t)
        elif op == OP_MEDIA:
            st.append(media(a[0], a[1:]))
        elif op in (OP_ADD, OP_SUB, OP_MUL, OP_DIV):
            y = st.pop()
            x = st.pop()
            st.append({OP_ADD: x + y, OP_SUB: x - y, OP_MUL: x * y, OP_DIV: x / y if y != 0 else float('inf') * np.sign(x)}[op])
        elif op == OP_NEG:
            st.append(-st.pop())
        elif op == OP_POWC:
            st.append(float(np.power(st.pop(), val)))
        elif op == OP_SQRT:
            st.append(float(np.sqrt(st.pop())))
        elif op == OP_SIN:
            st.append(float(np.sin(st.pop())))
        elif op == OP_COS:
            st.append(float(np.cos(st.pop())))
        elif op == OP_EXP:
            st.append(float(np.exp(st.pop())))
        elif op == OP_LOG:
            st.append(float(np.log(st.pop())))
        elif op == OP_ROUNDF:
            st.append(float(np.float32(st.pop())))
    assert len(st) == 1
    return st[0]
