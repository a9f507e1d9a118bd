"""The drop-in boundary: generate -> compile -> execute -> convergence.

Same entry points and error behaviour as opesci/grid.py:59-156, with the
generated C++/OpenMP translation unit replaced by a model description
(opesci.plan.ModelPlan) executed by the prebuilt gfx950 library
libopesci_hip.so through its C ABI (include/opesci_hip.h).  There is no CPU
execution path: when the HIP library or a GPU is missing, execute() raises.
"""
import ctypes
import os

from .compilation import HIPCompiler, get_compiler

__all__ = ['Grid']


class Grid(object):
    """Code generation, compilation and execution infrastructure of a grid."""

    _compiler = None
    src_code = None
    src_file = None
    src_lib = None
    _library = None
    _arg_grid = None
    _arg_profiling = None
    _plan = None
    _ctx = None
    _ctx_output = None
    _host_fields = None

    def __init__(self):
        self._compiler = HIPCompiler()

    def _invalidate(self):
        self._plan = None
        self.src_code = None

    @property
    def compiler(self):
        return self._compiler

    @compiler.setter
    def compiler(self, compiler):
        self._compiler = get_compiler(compiler)

    def _load_library(self, src_lib=None):
        from . import backend
        try:
            self._library = backend.load(src_lib or self.src_lib)
        except OSError as e:
            raise Exception('Failed to load %s: %s' % (src_lib or self.src_lib, e))

    # ---------------- generate / compile ----------------
    def lower(self):
        """Build (once) the model plan: folded coefficients, boundary program,
        analytic-solution bytecode."""
        if self._plan is None:
            from .plan import build_plan
            self._plan = build_plan(self)
        return self._plan

    def generate(self, filename, compiler=None):
        """Write the generated model description (a C header with the tables the
        library consumes and the derived kernels as comments)."""
        if compiler:
            self.compiler = compiler
        from .emit import emit_header
        self.src_code = emit_header(self, self.lower())
        self.src_file = filename
        d = os.path.dirname(os.path.abspath(filename))
        if not os.path.isdir(d):
            os.makedirs(d)
        with open(filename, 'w') as f:
            f.write(self.src_code)
        print('Generated:', self.src_file)

    def compile(self, filename, compiler=None, shared=True):
        """Generate if needed and make sure the gfx950 library exists."""
        if compiler:
            self.compiler = compiler
        if self.src_file is None:
            self.generate(filename)
        out = self.compiler.compile(self.src_file, shared=shared)
        if shared:
            self.src_lib = out
        return out

    # ---------------- execute / convergence ----------------
    def execute(self, filename, compiler=None, nthreads=1, affinity='close'):
        """Run the model on the GPU; grid.<field> arrays come back through the
        OpesciGrid struct exactly like the reference's ctypes call."""
        if affinity not in ('close', 'spread', 'compact', 'scatter'):
            print("""ERROR: Only the following affinity settings are supported:
 * OMP_PROC_BIND: 'close', 'spread'
 * KMP_AFFINITY: 'compact', 'scatter'""")
            raise ValueError('Unknown thread affinity setting: %s' % affinity)
        if self.src_lib is None:
            self.compile(filename, compiler=compiler, shared=True)
        self._load_library(self.src_lib)
        from . import backend
        plan = self.lower()
        real = ctypes.c_double if plan.real_bytes == 8 else ctypes.c_float

        class OpesciGrid(ctypes.Structure):
            _fields_ = [(str(f.label), ctypes.POINTER(real)) for f in self.fields]

        class OpesciProfiling(ctypes.Structure):
            _fields_ = [(v, ctypes.c_float) for v in ('rtime', 'ptime', 'mflops')]

        self._arg_grid = OpesciGrid()
        self._arg_profiling = OpesciProfiling()
        print('Executing on %s' % backend.device_string())
        if getattr(self, 'output_vts', False):
            self._execute_with_output(plan, real)
        else:
            self._ctx_output = None
            backend.execute(plan, self._arg_grid, self._arg_profiling)
        self._executed_plan = plan
        if self.profiling:
            print('PAPI:: Max real_time: %f (sec)' % self._arg_profiling.rtime)
            print('PAPI:: Max proc_time: %f (sec)' % self._arg_profiling.ptime)
            print('PAPI:: Total MFlops/s: %f' % self._arg_profiling.mflops)

    def _execute_with_output(self, plan, real):
        """Time loop with the first field written after every step as
        <name>_<step>.vts (RegularGrid.output_step, opesci/regulargrid.py:
        702-719; the template calls it at the end of each step).  Runs one
        step per launch on a context and keeps the context for
        convergence(); the final fields come back through the OpesciGrid
        struct as with opesci_execute."""
        import numpy as np
        from . import backend
        from .output import write_vts
        if self._ctx_output is not None:
            self._ctx_output.close()
        ctx = backend.Context(plan)
        ctx.init()
        name = str(self.fields[0].label)
        spacing = [float(s.value) for s in self.spacing]
        secs = 0.0
        for ti in range(plan.ntsteps):
            secs += ctx.run(1)
            write_vts('%s_%d' % (name, ti), ctx.download(0, -1), spacing, plan.M)
        self._host_fields = []
        for f in range(len(self.fields)):
            both = np.stack([ctx.download(f, 0), ctx.download(f, 1)]).astype(np.float64 if real is ctypes.c_double
                                                                               else np.float32)
            self._host_fields.append(np.ascontiguousarray(both))
            setattr(self._arg_grid, str(self.fields[f].label), self._host_fields[-1].ctypes.data_as(ctypes.POINTER(real)))
        self._arg_profiling.rtime = self._arg_profiling.ptime = secs
        self._ctx_output = ctx

    def convergence(self):
        """L2 norms of every field against its analytic solution."""
        if self._library is None:
            self._load_library()
        if self._arg_grid is None:
            raise RuntimeError("""Convergence could not find grid argument!
You need to you run grid.execute() first!""")
        from . import backend
        plan = self._executed_plan
        if getattr(self, '_ctx_output', None) is not None:
            real = float if plan.real_bytes == 8 else (lambda v: float(__import__('numpy').float32(v)))
            vals = [real(v) for v in self._ctx_output.l2()]
        else:
            vals = backend.convergence(plan, self._arg_grid)
        print('Convergence:')
        out = {}
        for f, v in zip(self.fields, vals):
            name = '%s_l2' % f.label
            out[name] = v
            print('%s: %.10f' % (name, v))
        return out

    def field_array(self, name, level=None):
        """NumPy view of a field returned by execute(): [tp][dim1][dim2][dim3],
        or one time level."""
        import numpy as np
        if self._arg_grid is None:
            raise RuntimeError('execute() first')
        plan = self._executed_plan
        ptr = getattr(self._arg_grid, name)
        n = 2 * plan.dims[0] * plan.dims[1] * plan.dims[2]
        arr = np.ctypeslib.as_array(ptr, shape=(n,)).reshape((2,) + tuple(plan.dims))
        return arr if level is None else arr[level]

    def free(self):
        """Release the host arrays handed out by execute() (opesci_free)."""
        from . import backend
        if getattr(self, '_ctx_output', None) is not None:
            self._ctx_output.close()
            self._ctx_output = None
            self._host_fields = None
            self._arg_grid = None
        if self._arg_grid is not None:
            backend.free(self._arg_grid)
            self._arg_grid = None
