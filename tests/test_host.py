"""CPU tests of the host layer: the SymPy derivation and its lowering against
the oracle (through the plan interpreter), reference fixtures, error
behaviour, and the C ABI library surface (no GPU calls)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import opesci
import oracle
from opesci import backend
from opesci.compilation import library_path
from opesci.plan import PlanError
from planvm import PlanVM
from workloads import eigen_model, hetero_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'opesci_hip.h')


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_plan_matches_oracle_fp64(order):
    """Two independent derivations: SymPy lowering (executed by the plan
    interpreter) vs the hand-written oracle."""
    g, kw = eigen_model(11, order, 12, double=True)
    plan = g.lower()
    vm = PlanVM(plan)
    vm.init()
    vm.run()
    ref = oracle.run(**kw)
    scale = np.abs(ref['fields'][0]).max()
    for f in range(9):
        assert np.abs(vm.latest()[f] - ref['fields'][f]).max() / scale < 1e-12, oracle.FIELDS[f]
    assert np.allclose(vm.l2(), ref['l2'], rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("order", [4, 8])
def test_plan_matches_oracle_heterogeneous(order):
    g, kw = hetero_model(11, order, 10, double=True, sigma=1.5)
    vm = PlanVM(g.lower())
    vm.init()
    vm.run()
    ref = oracle.run(**kw)
    scale = np.abs(ref['fields'][0]).max()
    for f in range(9):
        assert np.abs(vm.latest()[f] - ref['fields'][f]).max() / scale < 1e-8, oracle.FIELDS[f]


@pytest.mark.parametrize("split", [False, True])
def test_schedule_equals_sequential_order(split):
    """Hazard-scheduled boundary launches reproduce the reference's sequential
    loop order bit for bit, also with the deep z-face rows of the stress
    phase split off and moved ahead of every other stress statement."""
    g, _ = eigen_model(13, 4, 6)
    g.fuse_boundary = split
    plan = g.lower()
    assert (plan.stmts['phase'] == 4).any() == split
    a, b = PlanVM(plan, True), PlanVM(plan, False)
    for vm in (a, b):
        vm.init()
        vm.run()
    assert np.array_equal(a.F, b.F)
    assert plan.nlaunches_step < sum(1 for s in plan.stmts if s['phase'] >= 2)


def test_fd_weights_fixture(fixtures):
    w = opesci.staggered_weights(2)
    assert [float(v) for v in w] == [9 / 8, -1 / 24]
    for M in (1, 2, 3, 4):
        assert np.allclose([float(v) for v in opesci.staggered_weights(M)], oracle.fd_weights(M), rtol=0, atol=1e-16)


def test_8th_order_velocity_coefficients_fixture(fixtures):
    """tests/eigenwave3d-higher_order.ipynb cell 31: dt/dx = 0.2, beta = 1."""
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (100, 100, 100), 0.002, 0.01, accuracy_order=[2, 8, 8, 8])
    plan = g.lower()
    txx = {t[2]: t[0] for t in fixtures['velocity_8th_U_terms']['terms'] if t[1] == 'Txx'}
    for k in range(4):
        assert plan.cv[0, 0, k] == np.float32(txx['x + %d' % (k + 1)])
        lo = 'x' if k == 0 else 'x - %d' % k
        assert -plan.cv[0, 0, k] == np.float32(txx[lo])


def test_ai_fixtures(fixtures):
    """Op counts (ADD, MUL, LOAD, STORE) of the folded kernels against the
    reference notebooks' printed get_*_kernel_ai() outputs.  4th order:
    velocity (36, 36, 9, 3), stress (60, 60, 9, 6), as printed by
    eigenwave3d-higher_order.ipynb cells 27-28 and eigenwave3d-read.ipynb
    cells 14-15 (both set accuracy [1,2,2,2] of the older API: 4th order).
    8th order: (72, 72, 9, 3) / (120, 120, 9, 6), as printed by
    eigenwave3d.ipynb cells 18-19, whose saved outputs are 8th-order counts
    (3 axes x 8 stencil points per update) although the notebook's source now
    sets [2,4,4,4].  The per-term count fixes which order each belongs to:
    an update of order 2M has 3 x 2M stencil terms (velocity, normal stress)
    or 2 x 2M (shear), one ADD and one MUL each."""
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (20, 20, 20), 0.002, 0.01, accuracy_order=[2, 4, 4, 4])
    g8 = opesci.eigenwave3d((1.0, 1.0, 1.0), (20, 20, 20), 0.002, 0.01, accuracy_order=[2, 8, 8, 8])
    for order, grid in ((4, g), (8, g8)):
        M = order // 2
        assert grid.get_velocity_kernel_ai()[2:] == (3 * 3 * 2 * M, 3 * 3 * 2 * M, 9, 3)
        assert grid.get_stress_kernel_ai()[2:] == (3 * 3 * 2 * M + 3 * 2 * 2 * M,) * 2 + (9, 6)
    for key in ('ai_nb_higher_order', 'ai_nb_read'):
        assert g.get_velocity_kernel_ai()[2:] == tuple(fixtures[key]['velocity'][2:]), key
        assert g.get_stress_kernel_ai()[2:] == tuple(fixtures[key]['stress'][2:]), key
    assert g8.get_velocity_kernel_ai()[2:] == tuple(fixtures['ai_nb_eigenwave3d']['velocity'][2:])
    assert g8.get_stress_kernel_ai()[2:] == tuple(fixtures['ai_nb_eigenwave3d']['stress'][2:])


def test_robertsson_fixture(fixtures):
    """8th-order ghost planes: U at x: 3,2,1,0 and dim1-5..dim1-2 (zero);
    Txx at x: 4 = 0, 3..1 = -Txx[5..7] (higher_order notebook cells 34-35)."""
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (20, 20, 20), 0.002, 0.01, accuracy_order=[2, 8, 8, 8])
    plan = g.lower()
    dim1 = plan.dims[0]
    st, te = plan.stmts, plan.terms
    sel = st[(st['phase'] == 3) & (st['field'] == 0) & (st['axis'] == 0)]
    assert sorted(sel['plane']) == [0, 1, 2, 3, dim1 - 5, dim1 - 4, dim1 - 3, dim1 - 2]
    assert all(s['nterms'] == 0 for s in sel)
    sel = st[(st['phase'] == 2) & (st['field'] == 3) & (st['axis'] == 0)]
    got = {int(s['plane']): [(int(t['plane']), float(t['coef'])) for t in te[s['term0']:s['term0'] + s['nterms']]]
           for s in sel}
    assert got[4] == [] and got[3] == [(5, -1.0)] and got[2] == [(6, -1.0)] and got[1] == [(7, -1.0)]
    assert got[dim1 - 5] == [] and got[dim1 - 2] == [(dim1 - 8, -1.0)]
    assert '[Txx[t + 1, 4, y, z] == 0' in fixtures['robertsson_8th_Txx_bc']['text'].replace('\n', '')


def test_cfl_fixture(fixtures):
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (100, 100, 100), 0.002, 1.0)
    assert g.get_time_step_limit() == pytest.approx(fixtures['cfl_4th_h001']['value'], rel=1e-12)


def test_error_behaviour(tmp_path):
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (10, 10, 10), 0.01, 0.05)
    with pytest.raises(ValueError):
        g.set_order([2, 3, 3, 3])
    with pytest.raises(KeyError):
        g.set_switches(nonsense=True)
    with pytest.raises(ValueError):
        g.set_switches(omp=1)
    with pytest.raises(Exception):
        g.set_stress_fields([])
    with pytest.raises(RuntimeError):
        g.convergence()
    with pytest.raises(ValueError):
        g.compiler = 'fortran'
    with pytest.raises(IOError):
        opesci.eigenwave3d((1.0, 1.0, 1.0), (10, 10, 10), 0.01, 0.05, read=True,
                           rho_file=str(tmp_path / 'missing'), vp_file='x', vs_file='y').lower()
    with pytest.raises(PlanError):
        opesci.eigenwave3d((1.0, 1.0, 1.0), (10, 10, 10), 0.01, 0.05, double=True, storage='bf16').lower()


def test_generate_writes_model_header(tmp_path):
    g = opesci.eigenwave3d((1.0, 1.0, 1.0), (10, 10, 10), 0.01, 0.05)
    fn = str(tmp_path / 'eigenwave3d.h')
    g.generate(fn)
    text = open(fn).read()
    assert 'opesci_cn' in text and 'Txx[_t1][x][y][z]=' in text and 'launch' in text


def _declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(opesci_\w+)\s*\(', src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(library_path())
    names = _declared_functions()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(backend.EXPORTS)
    assert lib.opesci_abi_version() == backend.ABI_VERSION


def test_one_hip_runtime_with_torch():
    """torch imported before the library (tests/conftest.py): the library
    binds to torch's HIP runtime, so the process maps exactly one; a fresh
    process that loads the library first and torch after maps two, and the
    device-tensor paths refuse to run there."""
    pytest.importorskip('torch')
    backend.load()
    assert len(backend._hip_runtimes()) == 1, backend._hip_runtimes()
    code = ('import sys; sys.path.insert(0, %r)\n'
            'from opesci import backend\n'
            'backend.load()\n'
            'import torch\n'
            'assert len(backend._hip_runtimes()) == 2, backend._hip_runtimes()\n'
            'class T:\n'
            '    is_cuda = True\n'
            '    def data_ptr(self): return 0\n'
            'try:\n'
            '    backend._is_device_tensor(T())\n'
            'except RuntimeError as e:\n'
            '    assert "import torch before" in str(e)\n'
            'else:\n'
            '    raise AssertionError("no refusal")\n' % os.path.dirname(os.path.dirname(backend.__file__)))
    subprocess.run([sys.executable, '-c', code], check=True, timeout=300)


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirror of OpesciModelDesc/OpesciBcStmt/OpesciBcTerm/OpesciExprOp == the C layout."""
    from opesci import exprvm, plan as planmod
    fields = [f[0] for f in backend.OpesciModelDesc._fields_]
    prog = ['#include <stdio.h>', '#include <stddef.h>', '#include "opesci_hip.h"', 'int main(void){',
            'printf("%zu %zu %zu %zu\\n", sizeof(OpesciModelDesc), sizeof(OpesciBcStmt), sizeof(OpesciBcTerm), '
            'sizeof(OpesciExprOp));']
    for f in fields:
        prog.append('printf("%%zu\\n", offsetof(OpesciModelDesc, %s));' % f)
    prog.append('return 0;}')
    c = tmp_path / 'layout.c'
    c.write_text('\n'.join(prog))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-I', os.path.dirname(HEADER), str(c), '-o', str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    sizes = [int(v) for v in out[:4]]
    assert sizes == [ctypes.sizeof(backend.OpesciModelDesc), planmod.STMT_DTYPE.itemsize,
                     planmod.TERM_DTYPE.itemsize, exprvm.OP_DTYPE.itemsize]
    for f, off in zip(fields, out[4:]):
        assert getattr(backend.OpesciModelDesc, f).offset == int(off), f


def test_malformed_description_rejected_before_device():
    """opesci_ctx_create validates the description before touching a GPU."""
    lib = backend.load()
    g, _ = eigen_model(11, 4, 2)
    plan = g.lower()
    desc, keep = backend.make_desc(plan)
    desc.abi_version = 99
    h = ctypes.c_void_p()
    assert lib.opesci_ctx_create(ctypes.byref(desc), ctypes.byref(h)) == 1
    assert b'ABI' in lib.opesci_last_error()
    desc, keep = backend.make_desc(plan)
    bad = plan.terms.copy()
    bad[0]['plane'] = 10 ** 6
    desc.terms = bad.ctypes.data
    assert lib.opesci_ctx_create(ctypes.byref(desc), ctypes.byref(h)) == 1
    desc, keep = backend.make_desc(plan)
    desc.space_order = 10
    assert lib.opesci_ctx_create(ctypes.byref(desc), ctypes.byref(h)) == 6


def test_boundary_coefficient_derivatives():
    """The adjoint's derivative programs (plan.dterms) equal central finite
    differences of each media-dependent boundary coefficient with respect to
    the media cell they name (FP64, relative step 1e-6), and every media
    cell a coefficient reads has one."""
    from opesci.plan import DTERM_DTYPE
    from planvm import _code, eval_code
    g, _ = hetero_model(14, 4, 2, double=True, sigma=2.0)
    p = g.lower()
    assert p.dterms.dtype == DTERM_DTYPE and len(p.dterms) > 0
    rng = np.random.default_rng(1)
    cells = {}
    checked = 0
    for i in rng.choice(len(p.dterms), size=min(40, len(p.dterms)), replace=False):
        d = p.dterms[i]
        t = p.terms[d['term']]
        prog = _code(p.bc_code, int(t['prog']), int(t['prog_len']))
        dprog = _code(p.bc_code, int(d['prog']), int(d['prog_len']))
        vals = {}

        def media_at(mid, offs):
            key = (int(mid), tuple(int(v) for v in offs))
            if key not in vals:
                vals[key] = 1.0 + rng.uniform(0.2, 2.0)
            return vals[key]
        eval_code(prog, None, 0.0, media_at)
        key = (int(d['media']), tuple(int(v) for v in d['off']))
        assert key in vals
        base = vals[key]
        h = 1e-6 * base

        def f(v):
            vals[key] = v
            return eval_code(prog, None, 0.0, media_at)
        fd = (f(base + h) - f(base - h)) / (2 * h)
        vals[key] = base
        an = eval_code(dprog, None, 0.0, media_at)
        assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (i, an, fd)
        checked += 1
        cells.setdefault(int(d['term']), set()).add(key)
    assert checked > 10
    # completeness: every media cell of a sampled term has a derivative entry
    for term, keys in list(cells.items())[:10]:
        t = p.terms[term]
        prog = _code(p.bc_code, int(t['prog']), int(t['prog_len']))
        reads = set((a[0], tuple(a[1:4])) for op, a, v in prog if op == 3)
        have = set((int(d['media']), tuple(int(v) for v in d['off'])) for d in p.dterms if d['term'] == term)
        assert reads == have
