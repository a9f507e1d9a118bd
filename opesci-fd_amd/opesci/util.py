"""Finite-difference weights and expression helpers.

Mirrors the public helpers of opesci/util.py: staggered first-derivative
stencils (Taylor_half/Deriv_half, util.py:103-118,195-236), their generic
one-sided form (Taylor_generic/Deriv_generic, :121-133,176-192), the op
counter behind the arithmetic-intensity model (get_ops_expr, :30-64) and the
expression walkers.  Weights are exact rationals from the moment equations.
"""
from fractions import Fraction

from sympy import Basic, Indexed, Matrix, Rational, Symbol, factorial

__all__ = ['hf', 'get_all_objects', 'variable_to_symbol', 'get_ops_expr', 'IndexedBases',
           'tc', 'Taylor', 'Taylor_half', 'Taylor_generic', 'Deriv', 'Deriv_half',
           'Deriv_generic', 'is_half', 'shift_grid', 'shift_index', 'staggered_weights']

hf = Rational(1, 2)


def get_all_objects(expr, typ):
    """All sub-expressions of ``expr`` that are instances of ``typ`` (depth first)."""
    if isinstance(expr, typ):
        return [expr]
    if not isinstance(expr, Basic):
        return []
    out = []
    for a in expr.args:
        out += get_all_objects(a, typ)
    return out


def variable_to_symbol(variables):
    return [Symbol(v.name) for v in variables]


def get_ops_expr(expr, arrays):
    """(#ADD, #MUL, arrays): an n-ary Add counts n-1 additions, an n-ary Mul
    n-1 multiplications; every distinct IndexedBase label is one load."""
    if isinstance(expr, Indexed):
        name = str(expr.base.label)
        if name not in arrays:
            arrays = arrays + [name]
        return 0, 0, arrays
    add = mul = 0
    if expr.is_Add or expr.is_Mul:
        n = len(expr.args) - 1
        if expr.is_Add:
            add += n
        else:
            mul += n
        for a in expr.args:
            a2, m2, arrays = get_ops_expr(a, arrays)
            add += a2
            mul += m2
    return add, mul, arrays


def IndexedBases(s):
    from sympy import IndexedBase
    return tuple(IndexedBase(n) for n in s.split())


def tc(dx, n):
    """Taylor coefficient dx^n / n!"""
    return dx ** n / factorial(n)


def Taylor(dx, n):
    """Rows [(i dx)^j / j!] for i = -n/2..n/2 (centred points)."""
    return Matrix([[tc(i * dx, j) for j in range(n + 1)] for i in range(-(n // 2), n // 2 + 1)])


def Taylor_half(dx, n):
    """Rows [(i dx/2)^j / j!] for the 2n half-integer points i = -(2n-1)..(2n-1) step 2."""
    return Matrix([[tc(i * dx / 2, j) for j in range(2 * n)] for i in range(-2 * n + 1, 2 * n, 2)])


def Taylor_generic(delta, left, right, half=False):
    shift = hf if half else 0
    n = left + right + 1
    return Matrix([[tc((i + shift) * delta, j) for j in range(n)] for i in range(-left, right + 1)])


def _stencil(U, index, dimension, offsets):
    out = []
    for o in offsets:
        idx = list(index)
        idx[dimension] = index[dimension] + o
        out.append(U[idx] if len(idx) > 1 else U[idx[0]])
    return Matrix(out)


def Deriv(U, index, k, d, n):
    """Centred FD approximations of all derivatives up to n (D = M^-1 R)."""
    offs = list(range(-(n // 2), n // 2 + 1))
    return Taylor(d, n).inv() * _stencil(U, index, k, offs)


def Deriv_generic(U, index, dimension, delta, left, right, half=False):
    shift = hf if half else 0
    offs = [i + shift for i in range(-left, right + 1)]
    return Taylor_generic(delta, left, right, half).inv() * _stencil(U, index, dimension, offs)


def Deriv_half(U, index, dimension, delta, order):
    """Staggered FD approximations on the half-integer stencil of 2*order points."""
    offs = [Rational(i, 2) for i in range(-2 * order + 1, 2 * order, 2)]
    return Taylor_half(delta, order).inv() * _stencil(U, index, dimension, offs)


_WEIGHT_CACHE = {}


def staggered_weights(M):
    """Exact w_k, k < M, with d/dx f = sum_k w_k (f(x+(k+1/2)h) - f(x-(k+1/2)h))/h + O(h^2M).

    Solves the odd moment conditions sum_k w_k 2 p_k^(2r+1) = [r == 0]
    for p_k = k + 1/2 (the antisymmetric half of Deriv_half's system)."""
    if M in _WEIGHT_CACHE:
        return _WEIGHT_CACHE[M]
    p = [Fraction(2 * k + 1, 2) for k in range(M)]
    A = [[2 * pk ** (2 * r + 1) for pk in p] for r in range(M)]
    b = [Fraction(1 if r == 0 else 0) for r in range(M)]
    for c in range(M):
        piv = next(r for r in range(c, M) if A[r][c] != 0)
        A[c], A[piv] = A[piv], A[c]
        b[c], b[piv] = b[piv], b[c]
        for r in range(M):
            if r != c and A[r][c] != 0:
                f = A[r][c] / A[c][c]
                A[r] = [x - f * y for x, y in zip(A[r], A[c])]
                b[r] -= f * b[c]
    w = [Rational(b[i].numerator, b[i].denominator) / Rational(A[i][i].numerator, A[i][i].denominator)
         for i in range(M)]
    _WEIGHT_CACHE[M] = w
    return w


def is_half(expr):
    """True when the constant part of an index is integer + 1/2."""
    d = {s: 0 for s in expr.free_symbols}
    return not expr.subs(d).is_Integer


def shift_grid(expr):
    if expr.is_Symbol or expr.is_Number:
        return expr
    if isinstance(expr, Indexed):
        return Indexed(expr.base, *[i - hf if is_half(i) else i for i in expr.indices])
    return expr.func(*[shift_grid(a) for a in expr.args])


def shift_index(expr, k, s):
    if expr.is_Symbol or expr.is_Number:
        return expr
    if isinstance(expr, Indexed):
        idx = list(expr.indices)
        idx[k] += s
        return Indexed(expr.base, *idx)
    return expr.func(*[shift_index(a, k, s) for a in expr.args])
