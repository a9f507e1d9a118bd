"""Grid geometry, switches, constants and the arithmetic-intensity model.

Mirrors the parts of opesci/regulargrid.py the staggered elastic path uses:
geometry (__init__/set_order/set_grid_size/_update_spacing, :22-80,
:106-131, :145-155, :187-201), time stepping (:165-172), switches
(:174-185), constant folding (create_const_dict/transform_kernel,
:281-291, :329-342) and the op-count AI model (get_kernel_ai, :293-327).
The acoustic RegularField path of the reference is not part of this backend.
"""
from sympy import Symbol, expand

from .grid import Grid
from .util import get_ops_expr, variable_to_symbol
from .variable import Variable

__all__ = ['RegularGrid']

STORAGE_KINDS = ('native', 'fp32', 'fp64', 'fp16', 'bf16')


class RegularGrid(Grid):
    _papi_events = []
    _switches = ['omp', 'ivdep', 'simd', 'double', 'expand', 'eval_const',
                 'output_vts', 'converge', 'profiling', 'pluto', 'fission']
    # switches added by this backend (not bool-only)
    _extra_switches = ['storage', 'autotune', 'use_graph', 'legacy_levander']

    def __init__(self, dimension, index=None, fields=None, double=False, profiling=False,
                 pluto=False, fission=False, omp=True, ivdep=True, simd=False, io=False,
                 expand=True, eval_const=True, grid_size=(10, 10, 10), domain_size=None,
                 output_vts=False, storage='native', autotune=False, use_graph=True,
                 legacy_levander=False):
        super(RegularGrid, self).__init__()
        self.dimension = dimension
        self.double = double
        self.real_t = 'double' if double else 'float'
        self.dt = Variable('dt', 0.01, self.real_t, True)
        self.ntsteps = Variable('ntsteps', 100, 'int', True)
        self.alignment = 4096
        self.margin = Variable('margin', 2, 'int', True)
        self.dim = [Variable('dim%d' % (k + 1), 10 + 1 + 2 * self.margin.value, 'int', True)
                    for k in range(dimension)]
        self.size = [1.0] * dimension
        self.t = Symbol('_t')
        self.grid_size = tuple(grid_size)
        self.max_derivative_order = 1
        self.read = False
        self.set_order([2] + [4] * dimension)
        self.set_grid_size(grid_size)
        self.set_index(index)
        self.defined_variable = {}
        self.pluto = pluto
        self.omp = omp
        self.ivdep = ivdep
        self.simd = simd
        self.output_vts = output_vts
        self.expand = expand
        self.eval_const = eval_const
        self.profiling = profiling
        self.fission = fission
        self.converge = False
        self.storage = storage
        self.autotune = autotune
        self.use_graph = use_graph
        self.legacy_levander = legacy_levander
        if domain_size:
            self.set_domain_size(domain_size)

    # ---------------- model variables ----------------
    def set_variable(self, var, value=0, type='int', constant=False):
        if isinstance(var, Symbol):
            var = var.name
        self.defined_variable[var] = Variable(var, value, type, constant)

    def set_order(self, order):
        """[time order, space orders...]; every order must be even."""
        for x in order:
            if x % 2 == 1:
                raise ValueError(str(x) + ' is not a valid order (require even integer)')
        self.order = list(order)
        ntime = max(self.order[0], self.max_derivative_order + 1)
        self.tp = Variable('tp', ntime, 'int', True)
        self.time = [Variable(str(self.t) + str(k), k, 'int', False) for k in range(ntime)]
        self.margin.value = self.order[1] // 2
        self.set_grid_size(self.grid_size)
        self.update_field_order()
        self._update_spacing()

    def set_domain_size(self, size):
        self.size = list(size)
        self._update_spacing()

    def set_grid_size(self, size):
        """Number of cells per axis; dim_k = size_k + 1 + 2*margin."""
        self.grid_size = tuple(size)
        self.dim = [Variable('dim%d' % (k + 1), size[k] + 1 + 2 * self.margin.value, 'int', True)
                    for k in range(self.dimension)]
        self._update_spacing()

    def update_field_order(self):
        for f in getattr(self, 'fields', []) or []:
            f.set_order(self.order)

    def set_time_step(self, dt, tmax):
        self.dt.value = dt
        self.ntsteps.value = int(tmax / dt)

    def set_papi_events(self, events=()):
        """Hardware events to report with profiling (opesci/regulargrid.py:
        346-371).  The names are kept; on this backend counters come from
        rocprofv3 (e.g. --pmc FETCH_SIZE), not from PAPI."""
        self._papi_events = list(events)

    @property
    def numevents_papi(self):
        return len(self._papi_events)

    def set_switches(self, **kwargs):
        for switch, value in kwargs.items():
            if switch in self._extra_switches:
                if switch == 'storage' and value not in STORAGE_KINDS:
                    raise ValueError('storage must be one of %s' % (STORAGE_KINDS,))
                setattr(self, switch, value)
                continue
            if switch not in self._switches:
                raise KeyError('Unsupported switch: ', switch)
            if not isinstance(value, bool):
                raise ValueError('Only boolean values allowed for switches')
            setattr(self, switch, value)
            if switch == 'double':
                self.real_t = 'double' if self.double else 'float'
                self._update_spacing()
        self._invalidate()

    def _update_spacing(self):
        m = self.margin.value
        self.spacing = [Variable('dx%d' % (k + 1), self.size[k] / (self.dim[k].value - 1 - 2 * m),
                                 getattr(self, 'real_t', 'float'), True)
                        for k in range(self.dimension)]
        n = self.order[0]
        for d in self.dim:
            n *= d.value
        self.vec_size = Variable('vec_size', n, 'int', True)
        if hasattr(self, 'dt'):
            self.dt.type = getattr(self, 'real_t', 'float')
        self.set_field_spacing()

    def set_index(self, index):
        self.index = [Symbol('x%d' % (k + 1)) for k in range(self.dimension)] if index is None else list(index)
        for f in getattr(self, 'fields', []) or []:
            f.set_indices([self.t] + self.index)

    def set_field_spacing(self):
        for f in getattr(self, 'fields', []) or []:
            f.set_spacing(variable_to_symbol([self.dt] + self.spacing))

    def get_all_variables(self):
        return (self.dim + self.spacing + self.time + [self.tp, self.dt, self.margin, self.ntsteps]
                + list(self.defined_variable.values()))

    def create_const_dict(self):
        """{symbol: value} of every constant Variable; the Variable objects
        themselves are keys too, since they appear in boundary indices
        (dim1 - m - 1) and are not equal to plain Symbols of the same name."""
        d = {}
        for v in self.get_all_variables():
            if v.constant:
                d[Symbol(v.name)] = v.value
                d[v] = v.value
        self.const_dict = d

    def transform_kernel(self, field):
        """expand (de-factorise) then substitute constant values, as the
        expand / eval_const switches ask."""
        k = field.kernel_aligned
        if self.expand:
            k = expand(k)
        if self.eval_const:
            self.create_const_dict()
            k = k.subs(self.const_dict)
        return k

    # ---------------- arithmetic intensity ----------------
    def get_kernel_ai(self, fields=None):
        """(AI, AI_w, #ADD, #MUL, #LOAD, #STORE); AI = (ADD+MUL)/((LOAD+STORE)*word),
        AI_w = AI*(ADD+MUL)/(2*max(ADD, MUL))."""
        if fields is None:
            fields = self.fields
        if self.eval_const:
            self.create_const_dict()
        add = mul = store = 0
        arrays = []
        for f in fields:
            store += 1
            a, m, arrays = get_ops_expr(self.transform_kernel(f), arrays)
            add += a
            mul += m
        word = 8 if self.double else 4
        load = len(arrays)
        ai = float(add + mul) / (load + store) / word
        ai_w = ai * (add + mul) / max(add, mul) / 2.0
        return (ai, ai_w, add, mul, load, store)
