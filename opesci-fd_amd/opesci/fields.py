"""Staggered fields of the velocity-stress model.

Same roles as opesci/fields.py: ``VField`` (velocity, staggered in time and
along its own direction), ``SField`` (stress: normal components unstaggered,
shear components staggered along both of their directions), ``Media`` (media
parameter arrays of the read path).  A field wraps a SymPy ``IndexedBase``
rather than subclassing it, so SymPy's rebuilding of expressions can never
drop the staggering metadata; ``field[t, x, y, z]`` still returns an
``Indexed`` that prints as the reference's array reference.
"""
from sympy import Eq, Indexed, IndexedBase, Rational, expand

from .derivative import DDerivative
from .util import Deriv_half, get_all_objects, hf, staggered_weights

__all__ = ['Field', 'SField', 'VField', 'Media', 'solve_linear_for']

_REGISTRY = {}


def solve_linear_for(expr, target):
    """Solve expr == 0 for ``target`` (an Indexed or Symbol) when expr is linear in it."""
    from sympy import Dummy
    if hasattr(expr, 'lhs') and hasattr(expr, 'rhs'):
        expr = expr.lhs - expr.rhs
    s = Dummy('s')
    e = expand(expr.xreplace({target: s}))
    a = e.coeff(s)
    if a == 0 or e.diff(s, 2) != 0:
        raise ValueError('expression is not linear in %s' % target)
    b = e.subs(s, 0)
    return expand(-b / a)


class _Base(object):
    """Common SymPy plumbing: an IndexedBase plus metadata."""

    def __init__(self, name):
        self.base = IndexedBase(name)
        self.label = self.base.label
        self.name = name
        self._peers = {}
        _REGISTRY[name] = self

    def __getitem__(self, idx):
        if not isinstance(idx, (tuple, list)):
            idx = (idx,)
        return Indexed(self.base, *idx)

    def __repr__(self):
        return self.name

    __str__ = __repr__

    def _lookup(self, base):
        name = str(base.label)
        return self._peers.get(name) or _REGISTRY.get(name)


class Field(_Base):
    """A discretised field: staggering, FD derivatives, update kernel, ghost-cell equations."""

    def __init__(self, name, **kwargs):
        super(Field, self).__init__(name)
        self.sol = None
        self.d = None
        self.kernel = None
        self.kernel_aligned = None
        self.dt = None
        self.indices = None
        self.spacing = None
        self.order = None
        self.sfields = None
        self.dimension = None
        self.staggered = None
        self.bc = None
        if kwargs:
            self.set(**kwargs)

    def set(self, dimension, staggered):
        self.dimension = dimension
        self.staggered = list(staggered)
        self.bc = [[None] * 2 for _ in range(dimension + 1)]

    def set_analytic_solution(self, function):
        """Exact solution used for the initial state and the L2 norms."""
        self.sol = function

    def set_indices(self, indices):
        self.indices = list(indices)

    def set_spacing(self, spacing):
        self.spacing = list(spacing)

    def set_order(self, order):
        self.order = list(order)

    def calc_derivative(self, l, k, d, n, order_of_derivative):
        """FD approximation of d^order F/dx_k^order of accuracy n on the
        half-integer stencil (util.Deriv_half semantics)."""
        if order_of_derivative == 1:
            M = n // 2
            w = staggered_weights(M)
            expr = 0
            for j in range(M):
                off = Rational(2 * j + 1, 2)
                ip = list(l)
                im = list(l)
                ip[k] = l[k] + off
                im[k] = l[k] - off
                expr += w[j] * self[ip] / d - w[j] * self[im] / d
            return expr
        return Deriv_half(self, l, k, d, n // 2)[order_of_derivative]

    def populate_derivatives(self, max_order=1):
        """self.d[axis][order] = DDerivative with fd[accuracy] for accuracy 2..order[axis]."""
        self.d = [[None] * (max_order + 1) for _ in range(self.dimension + 1)]
        for ax in range(self.dimension + 1):
            index = self.indices[ax]
            for order in range(1, max_order + 1):
                name = '\\partial %s/\\partial %s' % (self.name, index)
                D = DDerivative(name, index, order, self.order[ax])
                for acc in range(2, self.order[ax] + 2, 2):
                    D.fd[acc] = self.calc_derivative(self.indices, ax, self.spacing[ax], acc, order)
                self.d[ax][order] = D

    def align(self, expr):
        """Rewrite half-integer references to other fields as integer array
        offsets in this field's frame: +1/2 where this field is staggered and
        the other is not, -1/2 in the opposite case; media get +1/2 along every
        axis this field is staggered on."""
        if expr.is_Symbol or expr.is_Number:
            return expr
        if isinstance(expr, Indexed):
            f = self._lookup(expr.base)
            if f is None:
                return expr
            if isinstance(f, Media):
                idx = list(expr.indices)
                for ax in range(self.dimension):
                    if self.staggered[ax + 1]:
                        idx[ax] += hf
                return f[idx]
            if not isinstance(f, Field) or f.staggered is None:
                return expr
            idx = []
            for k, ind in enumerate(expr.indices):
                if self.staggered[k] == f.staggered[k]:
                    idx.append(ind)
                elif self.staggered[k]:
                    idx.append(ind + hf)
                else:
                    idx.append(ind - hf)
            return f[idx]
        return expr.func(*[self.align(a) for a in expr.args])

    def set_fd_kernel(self, kernel):
        self.kernel = kernel
        self.kernel_aligned = self.align(kernel)

    def set_dt(self, dt):
        """Right-hand side of this field's PDE (d F/dt = dt), with DDerivative symbols."""
        self.dt = dt

    def associate_stress_fields(self, sfields):
        self.sfields = sfields


class VField(Field):
    """Velocity component: staggered in time and along ``direction``."""

    def set(self, dimension, direction):
        self.direction = direction
        stag = [False] * (dimension + 1)
        stag[0] = True
        stag[direction] = True
        Field.set(self, dimension, stag)

    def set_free_surface(self, d, b, side, algo='robertsson'):
        """Ghost-cell equations of a free surface normal to axis d at index b
        (fields.py:192-261).  'levander' (4th order): the velocity staggered
        along d from T_dd' = 0, the others from T_id'(b-1/2) = T_id'(b+1/2),
        both with 2nd-order differences.  'robertsson': ghosts set to zero."""
        if algo == 'levander':
            field = self.sfields[d]
            idx = list(self.indices)
            expr = field.dt
            expr = expr.subs({D: D.fd[2] for D in get_all_objects(expr, DDerivative)})
            if self.staggered[d]:
                eq = expr
                shift = hf
                tb = b - hf
            else:
                eq = expr.subs(idx[d], idx[d] - hf) - expr.subs(idx[d], idx[d] + hf)
                shift = 1
                tb = b
            idx[d] -= ((-1) ** side) * shift
            lhs = self[idx]
            rhs = solve_linear_for(eq, lhs)
            lhs = lhs.subs(self.indices[d], tb)
            rhs = self.align(rhs.subs(self.indices[d], tb))
            lhs = lhs.subs(idx[0], idx[0] + 1)
            rhs = rhs.subs(idx[0], idx[0] + 1)
            self.bc[d][side] = [Eq(lhs, rhs, evaluate=False)]
        elif algo == 'robertsson':
            idx = list(self.indices)
            if self.staggered[d]:
                idx[d] = b - (1 - side)
            else:
                idx[d] = b - (-1) ** side
            out = [Eq(self[idx], 0, evaluate=False).subs(idx[0], idx[0] + 1)]
            for _ in range(self.order[d] // 2 - 1):
                idx[d] -= (-1) ** side
                out.append(Eq(self[idx], 0, evaluate=False).subs(idx[0], idx[0] + 1))
            self.bc[d][side] = out
        else:
            raise ValueError('Unknown boundary condition algorithm')


class SField(Field):
    """Stress component: normal (i,i) unstaggered, shear (i,j) staggered along i and j."""

    def set(self, dimension, direction):
        self.direction = tuple(direction)
        stag = [False] * (dimension + 1)
        if direction[0] != direction[1]:
            for a in direction:
                stag[a] = True
        Field.set(self, dimension, stag)

    def set_free_surface(self, d, b, side, algo='robertsson'):
        """Ghost/boundary equations of a free surface normal to axis d at index
        b (fields.py:294-381): antisymmetry about the surface for components
        involving d; with 'levander', normal stresses T_jj (j != d) on the
        surface plane are recomputed with d(V_d)/d(x_d) eliminated via
        T_dd' = 0 and 4th-order tangential differences."""
        idx = list(self.indices)
        if d not in self.direction:
            if algo != 'levander' or self.direction[0] != self.direction[1]:
                self.bc[d][side] = []
                return
            expr = self.dt
            target = sub = None
            for D in get_all_objects(expr, DDerivative):
                if D.var == idx[d]:
                    target = D
                    sub = solve_linear_for(self.sfields[d].dt, D)
                    break
            expr = expr.subs(target, sub)
            expr = expr.subs({D: D.fd[4] for D in get_all_objects(expr, DDerivative)})
            eq = Eq(self.d[0][1].fd[2], expr).subs(idx[d], b)
            t = idx[0]
            idx[0] = t + hf
            idx[d] = b
            lhs = self[idx]
            rhs = self.align(solve_linear_for(eq, lhs))
            lhs = lhs.subs(t, t + hf)
            rhs = rhs.subs(t, t + hf)
            self.bc[d][side] = [Eq(lhs, rhs, evaluate=False)]
            return
        idx = list(self.indices)
        idx2 = list(self.indices)
        if not self.staggered[d]:
            idx[d] = b
            idx2[d] = b
            eq1 = Eq(self[idx], 0, evaluate=False)
        else:
            idx[d] = b - (1 - side)
            idx2[d] = idx[d] + (-1) ** side
            eq1 = Eq(self[idx], -self[idx2], evaluate=False)
        out = [eq1.subs(idx[0], idx[0] + 1)]
        for _ in range(self.order[d] // 2 - 1):
            idx[d] -= (-1) ** side
            idx2[d] += (-1) ** side
            out.append(Eq(self[idx], -self[idx2], evaluate=False).subs(idx[0], idx[0] + 1))
        self.bc[d][side] = out


class Media(_Base):
    """A media parameter array (rho, vp, vs, beta*, lambda, mu*) of the read path."""

    def __init__(self, name, dimension=3, staggered=None, index=None):
        super(Media, self).__init__(name)
        self.dimension = dimension
        self.staggered = list(staggered) if staggered is not None else [False] * dimension
        self.index = index

    def set(self, dimension, staggered, index):
        self.dimension = dimension
        self.staggered = list(staggered)
        self.index = index
