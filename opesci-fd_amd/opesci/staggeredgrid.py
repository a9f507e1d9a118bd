"""Velocity-stress elastic model on a staggered grid.

Same public API and semantics as opesci/staggeredgrid.py:
 * fields / media set-up (:55-133, :234-352)
 * PDE -> leapfrog kernels (solve_fd, :135-170)
 * free-surface equations (set_free_surface_boundary, :214-232)
 * AI model (:361-510)
 * emission order of the boundary loops (stress_bc_getter, velocity_bc,
   initialise_bc, :754-879), here turned into a boundary program by
   opesci.plan instead of C loops.
"""
import numpy as np
from sympy import Eq, Indexed, Rational, Symbol, expand

from .derivative import DDerivative
from .fields import Media, solve_linear_for
from .regulargrid import RegularGrid
from .util import get_all_objects, get_ops_expr, hf, is_half

__all__ = ['StaggeredGrid']


class StaggeredGrid(RegularGrid):
    """3-D isotropic elastic velocity-stress scheme (Virieux/Graves) on a
    staggered grid; 2nd-order leapfrog in time, 2nd..8th order in space,
    free surfaces by Levander (4th order) or Robertsson (other orders)."""

    _switches = ['omp', 'ivdep', 'simd', 'double', 'expand', 'eval_const',
                 'output_vts', 'converge', 'profiling', 'pluto', 'fission']

    def __init__(self, stress_fields=None, velocity_fields=None, converge=False, **kwargs):
        self.sfields = []
        self.vfields = []
        kwargs.pop('fields', None)
        super(StaggeredGrid, self).__init__(**kwargs)
        self.converge = converge
        self.read = False
        self.media_arrays = None
        if stress_fields:
            self.set_stress_fields(stress_fields)
        if velocity_fields:
            self.set_velocity_fields(velocity_fields)

    @property
    def fields(self):
        return self.vfields + self.sfields

    @property
    def io(self):
        return self.read or self.output_vts

    def set_alignment(self, alignment):
        self.alignment = alignment

    def _register(self, fields):
        peers = {f.name: f for f in self.fields}
        for m in self._media_objects():
            peers[m.name] = m
        for f in self.fields:
            f._peers = peers
        if fields:
            for f in fields:
                if getattr(self, 'index', None) is not None:
                    f.set_indices([self.t] + self.index)
                f.set_order(self.order)
        self.set_field_spacing()

    def _media_objects(self):
        if not self.read:
            return []
        return [self.rho, self.vp, self.vs] + self.beta + [self.lam] + self.mu

    def get_time_step_limit(self):
        """CFL bound: h/(sqrt(3) Vp) at 2nd order, 0.495 h/Vp at 4th."""
        if self.read:
            return 'physical parameters to be read from file, please compile and run the executable'
        l = self.defined_variable['lambda'].value
        m = self.defined_variable['mu'].value
        r = self.defined_variable['rho'].value
        Vp = ((l + 2 * m) / r) ** 0.5
        h = min(sp.value for sp in self.spacing)
        if self.order[1] == 2:
            return h / Vp / (3 ** 0.5)
        elif self.order[1] == 4:
            return 0.495 * h / Vp
        return 'not implemented yet'

    def set_stress_fields(self, sfields):
        num = self.dimension + self.dimension * (self.dimension - 1) // 2
        if len(sfields) != num:
            raise Exception('wrong number of stress fields: ' + str(num) + ' fields required.')
        self.sfields = list(sfields)
        self._register(sfields)

    def set_velocity_fields(self, vfields):
        if len(vfields) != self.dimension:
            raise Exception('wrong number of velocity fields: ' + str(self.dimension) + ' fields required.')
        self.vfields = list(vfields)
        self._register(vfields)

    def set_index(self, index):
        super(StaggeredGrid, self).set_index(index)
        if self.read:
            for m in self._media_objects():
                m.index = self.index
            self.create_media_dict()

    def calc_derivatives(self):
        for f in self.sfields + self.vfields:
            f.populate_derivatives(max_order=1)

    def solve_fd(self, equations):
        """Substitute the highest-accuracy FD stencils into the PDEs and solve
        each for its field at t+1/2; shifting t by 1/2 gives the leapfrog
        F[t+1] = F[t] + dt*RHS(t+1/2), then align() maps half indices to array
        offsets in the field's own frame."""
        if len(equations) != len(self.fields):
            raise KeyError('Number of equations must be the same as number of fields. '
                           'Number of fields is ', len(self.fields))
        self.eq = []
        for field, eq in zip(self.fields, equations):
            if self.read:
                eq = eq.subs(self.media_dict)
            field.set_dt(eq.rhs)
            self.eq.append(eq)
        eqs = []
        for eq in self.eq:
            for D in get_all_objects(eq, DDerivative):
                eq = eq.subs(D, D.fd[D.max_accuracy])
            eqs.append(eq)
        t = self.t
        shift = self.order[0] // 2 - 1
        t1 = t + hf + shift
        for field, eq in zip(self.fields, eqs):
            kernel = solve_linear_for(eq, field[[t1] + self.index])
            kernel = kernel.subs({t: t + hf - shift})
            field.set_fd_kernel(kernel)
        self._invalidate()

    def transform_bc(self, field, dimension, side):
        out = []
        for bc in field.bc[dimension][side]:
            b = bc
            if self.expand:
                b = Eq(expand(b.lhs), expand(b.rhs), evaluate=False)
            if self.eval_const:
                self.create_const_dict()
                b = Eq(b.lhs.subs(self.const_dict), b.rhs.subs(self.const_dict), evaluate=False)
            out.append(b)
        return out

    def associate_fields(self):
        """Velocity V_i <-> stresses (T_i1, T_i2, T_i3); normal stresses <-> (Txx, Tyy, Tzz)."""
        for v in self.vfields:
            fl = [None]
            for d in range(1, self.dimension + 1):
                look = tuple(sorted([v.direction, d]))
                match = [f for f in self.sfields if tuple(f.direction) == look]
                if not match:
                    raise Exception('error in field directions')
                fl.append(match[0])
            v.associate_stress_fields(fl)
        normals = [None] + [f for f in self.sfields if f.direction[0] == f.direction[1]]
        for s in normals[1:]:
            s.associate_stress_fields(normals)

    def set_free_surface_boundary(self, dimension, side):
        """Free surface normal to ``dimension`` (1..3) on the low (0) or high
        (1) side; Levander at 4th order, Robertsson otherwise."""
        algo = 'levander' if self.order[dimension] == 4 else 'robertsson'
        self.associate_fields()
        m = self.margin.value
        for f in self.sfields + self.vfields:
            b = m if side == 0 else self.dim[dimension - 1] - m - 1
            f.set_free_surface(dimension, b, side, algo=algo)
        self._invalidate()

    # ---------------- media ----------------
    def set_media_params(self, read=False, rho=1.0, vp=1.0, vs=0.5, rho_file='', vp_file='',
                         vs_file='', rho_array=None, vp_array=None, vs_array=None):
        """Constant media (rho, vp, vs) or the read path: rho/Vp/Vs from raw
        float32 files (or arrays) of dim1*dim2*dim3 values, with buoyancy and
        shear modulus averaged onto the staggered points."""
        self.read = read
        if read:
            self.rho_file, self.vp_file, self.vs_file = rho_file, vp_file, vs_file
            self.media_arrays = (rho_array, vp_array, vs_array) if rho_array is not None else None
            st = [False, False, False]
            mk = lambda n: Media(n, dimension=3, staggered=st, index=self.index)
            self.rho, self.vp, self.vs = mk('rho'), mk('vp'), mk('vs')
            self.beta = [mk('beta'), mk('beta1'), mk('beta2'), mk('beta3')]
            self.lam = mk('lambda')
            self.mu = [mk('mu'), mk('mu12'), mk('mu13'), mk('mu23')]
            self.create_media_dict()
            self._register(None)
        else:
            self.set_variable('rho', rho, 'float', True)
            self.set_variable('beta', 1.0 / rho, 'float', True)
            self.set_variable('lambda', rho * (vp ** 2 - 2 * vs ** 2), 'float', True)
            self.set_variable('mu', rho * (vs ** 2), 'float', True)
        self._invalidate()

    # ---------------- point source and receivers ----------------
    def set_source(self, wavelet, location=None, index=None, sdt=None):
        """Point stress source (tests/src/test_ref_iso_elastic.cpp:274-290):
        after every velocity update and receiver sampling, Txx, Tyy and Tzz at
        the source point decrease by w[t]/3 while t is inside the trace.
        wavelet: one trace (the same for the three) or three (x, y, z);
        location: model coordinates (nearest unstaggered grid point) or index:
        grid indices; sdt: sampling interval of the traces when it differs
        from the time step (spectral resampling, src/opesciHandy.cpp:125-195)."""
        from . import sources
        w = np.atleast_2d(np.asarray(wavelet, dtype=np.float32))
        if w.ndim != 2 or w.shape[0] not in (1, 3):
            raise ValueError('a source is one trace or three (x, y, z)')
        if sdt is not None:
            w = np.stack([sources.resample_timeseries(t, self.dt.value, sdt) for t in w])
        if (location is None) == (index is None):
            raise ValueError('give the source location or its grid index')
        if index is None:
            index = sources.nearest_index(location, [s.value for s in self.spacing], self.margin.value)
        self.src_wavelet = w
        self.src_idx = tuple(int(i) for i in index)
        self._invalidate()

    def set_receivers(self, locations=None, index=None):
        """Receivers (test_ref_iso_elastic.cpp:227-272): U, V, W and the
        pressure (Txx + Tyy + Tzz)/3 at each receiver's nearest grid point
        after every velocity update, before the source injection.  Read the
        traces with Context.receivers(): [step][receiver][U, V, W, p]."""
        from . import sources
        if (locations is None) == (index is None):
            raise ValueError('give receiver locations or grid indices')
        if index is None:
            h = [s.value for s in self.spacing]
            index = [sources.nearest_index(c, h, self.margin.value) for c in np.asarray(locations).reshape(-1, 3)]
        self.rec_idx = np.asarray(index, dtype=np.int32).reshape(-1, 3)
        self._invalidate()

    def create_media_dict(self):
        self.media_dict = {Symbol('beta'): self.beta[0][self.index],
                           Symbol('lambda'): self.lam[self.index],
                           Symbol('mu'): self.mu[0][self.index]}

    def resolve_individual_media_param(self, expr):
        """beta[x+1/2,y,z] -> beta1[x,y,z], mu[x+1/2,y+1/2,z] -> mu12[x,y,z], ..."""
        idx = list(expr.indices)
        name = str(expr.base.label)
        if name == 'beta':
            for k in range(3):
                if is_half(idx[k]):
                    idx[k] -= hf
                    return self.beta[k + 1][idx]
            return expr
        if name == 'mu':
            if is_half(idx[0]):
                idx[0] -= hf
                if is_half(idx[1]):
                    idx[1] -= hf
                    return self.mu[1][idx]
                idx[2] -= hf
                return self.mu[2][idx]
            if is_half(idx[1]):
                idx[1] -= hf
                idx[2] -= hf
                return self.mu[3][idx]
            return expr
        return expr

    def resolve_media_params(self, expr):
        if isinstance(expr, Eq):
            return Eq(expr.lhs, self.resolve_media_params(expr.rhs), evaluate=False)
        if expr.is_Symbol or expr.is_Number:
            return expr
        if isinstance(expr, Indexed):
            name = str(expr.base.label)
            if self.read and name in ('beta', 'mu'):
                return self.resolve_individual_media_param(expr)
            return expr
        return expr.func(*[self.resolve_media_params(a) for a in expr.args])

    def kernel_sympy(self, field):
        k = self.transform_kernel(field)
        if self.read:
            k = self.resolve_media_params(k)
        return k.xreplace({self.t + 1: self.time[1], self.t: self.time[0]})

    # ---------------- AI ----------------
    def get_velocity_kernel_ai(self):
        return self.get_kernel_ai(self.vfields)

    def get_stress_kernel_ai(self):
        return self.get_kernel_ai(self.sfields)

    def _bc_ai(self, fields):
        out = []
        word = 8 if self.double else 4
        for f in fields:
            for dim in range(1, 4):
                for side in range(2):
                    add = mul = store = 0
                    arrays = []
                    for bc in self.transform_bc(f, dim, side):
                        store += 1
                        a, m, arrays = get_ops_expr(bc.rhs, arrays)
                        add += a
                        mul += m
                    load = len(arrays)
                    if store == 0:
                        ai = ai_w = 0
                        weight = 0
                    else:
                        ai = float(add + mul) / (load + store) / word
                        weight = 1.0 / (self.dim[dim - 1].value - 2 * self.margin.value)
                        ai_w = ai if (add == 0 and mul == 0) else ai * (add + mul) / max(add, mul) / 2.0
                    out.append({'weight': weight, 'ai': ai, 'ai_w': ai_w, 'add': add, 'mul': mul,
                                'load': load, 'store': store})
        return out

    def get_velocity_bc_ai(self):
        return self._bc_ai(self.vfields)

    def get_stress_bc_ai(self):
        return self._bc_ai(self.sfields)

    def get_overall_kernel_ai(self):
        v = self.get_velocity_kernel_ai()
        s = self.get_stress_kernel_ai()
        total = 2.0
        ai = v[0] + s[0]
        ai_w = v[1] + s[1]
        for b in self.get_velocity_bc_ai() + self.get_stress_bc_ai():
            total += b['weight']
            ai += b['ai'] * b['weight']
            ai_w += b['ai_w'] * b['weight']
        ai /= total
        ai_w /= total
        adj = 1.0
        for d in self.dim[1:]:
            adj *= 1 - float(self.margin.value) / d.value
        return ai * adj, ai_w * adj

    def algorithmic_flops_per_point(self):
        """ADD+MUL of the stress and velocity kernels (the reference op model)."""
        v = self.get_velocity_kernel_ai()
        s = self.get_stress_kernel_ai()
        return v[2] + v[3] + s[2] + s[3]

    # ---------------- listings (the reference's code fragments) ----------------
    @property
    def stress_loop(self):
        from .emit import kernel_listing
        return kernel_listing(self, self.sfields)

    @property
    def velocity_loop(self):
        from .emit import kernel_listing
        return kernel_listing(self, self.vfields)

    @property
    def stress_bc(self):
        from .emit import bc_listing
        return bc_listing(self, phase=2)

    @property
    def velocity_bc(self):
        from .emit import bc_listing
        return bc_listing(self, phase=3)

    @property
    def initialise_bc(self):
        from .emit import bc_listing
        return bc_listing(self, phase=0) + '\n' + bc_listing(self, phase=1)
