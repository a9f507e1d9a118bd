"""GPU parity of the point source and the receivers against the oracle
(tests/src/test_ref_iso_elastic.cpp:227-290 semantics: sample U, V, W and
(Txx+Tyy+Tzz)/3 after the velocity update, then Tnn -= w[t]/3).

Tolerances: FP64 fields 1e-9 and traces 1e-9 of their maximum; FP32 traces
1e-5 of their maximum (float32 trace storage, different summation order).
"""
import numpy as np
import pytest

import oracle
from opesci import sources
from opesci.backend import Context
from workloads import eigen_model, hetero_model

pytestmark = pytest.mark.gpu


def run_gpu(plan):
    with Context(plan) as ctx:
        ctx.init()
        ctx.run()
        return ctx.fields(), ctx.receivers()


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


def test_source_receivers_heterogeneous_fp64():
    g, kw = hetero_model(24, 4, 30, double=True, sigma=2.0)
    w = sources.ricker(10.0, g.dt.value, 30, t0=8 * g.dt.value) * 50.0
    src_idx = (13, 12, 14)
    rec = [src_idx, (4, 5, 6), (2, 2, 2), (25, 20, 3), (13, 12, 25)]
    g.set_source(w, index=src_idx)
    g.set_receivers(index=rec)
    fields, tr = run_gpu(g.lower())
    ref = oracle.run(src=w, src_idx=src_idx, rec_idx=rec, **kw)
    assert tr.shape == (30, len(rec), 4)
    assert rel(tr, ref['rec']) < 1e-9
    scale = np.abs(ref['fields'][0]).max()
    for f in range(9):
        assert np.abs(fields[f] - ref['fields'][f]).max() <= 1e-9 * scale, oracle.FIELDS[f]
    # the source is felt: traces differ from a run without it
    g.set_source(0 * w, index=src_idx)
    _, tr0 = run_gpu(g.lower())
    assert np.abs(tr[:, 0, 3] - tr0[:, 0, 3]).max() > 1e-3 * np.abs(tr[:, 0, 3]).max()


def test_three_traces_fp32_8th_order():
    g, kw = eigen_model(30, 8, 25)
    dt = g.dt.value
    w = np.stack([sources.ricker(f, dt, 25, t0=6 * dt) for f in (8.0, 11.0, 14.0)]) * 10.0
    src_idx = (15, 17, 16)
    rec = [src_idx, (10, 10, 10), (30, 31, 5)]
    g.set_source(w, index=src_idx)
    g.set_receivers(index=rec)
    _, tr = run_gpu(g.lower())
    ref = oracle.run(src=w, src_idx=src_idx, rec_idx=rec, **kw)
    assert rel(tr, ref['rec']) < 1e-5


def test_receiver_plane_many_blocks():
    """A 40 x 40 receiver plane at z = 8 (several receiver blocks), FP32."""
    g, kw = eigen_model(44, 4, 12)
    m = 2
    rec = [(m + i, m + j, 8) for j in range(40) for i in range(40)]
    w = sources.ricker(10.0, g.dt.value, 12, t0=3 * g.dt.value) * 10.0
    src_idx = (22, 22, 8)
    g.set_source(w, index=src_idx)
    g.set_receivers(index=rec)
    _, tr = run_gpu(g.lower())
    ref = oracle.run(src=w, src_idx=src_idx, rec_idx=rec, **kw)
    assert tr.shape == (12, 1600, 4)
    assert rel(tr, ref['rec']) < 1e-5


def test_media_files_path_equals_arrays(tmp_path):
    """Heterogeneous run from raw rho/Vp/Vs files == the same media as arrays,
    bit for bit (effective media computed on the device either way)."""
    import opesci
    from opesci.media import smooth_random_media, write_media_files
    N, order, steps = 22, 4, 10
    n = N - 1
    dims = (N + order,) * 3
    rho, vp, vs = smooth_random_media(dims, seed=11, sigma=2.0)
    names = write_media_files(str(tmp_path / 'm'), rho, vp, vs)
    dt = 0.2 / n / float(vp.max())
    kw = dict(accuracy_order=[2, order, order, order], double=False)
    a = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5), media_arrays=(rho, vp, vs), **kw)
    b = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), dt, dt * (steps + 0.5), read=True, rho_file=names[0],
                           vp_file=names[1], vs_file=names[2], **kw)
    fa, _ = run_gpu(a.lower())
    fb, _ = run_gpu(b.lower())
    for f in range(9):
        assert np.array_equal(fa[f], fb[f]), oracle.FIELDS[f]


def test_execute_with_vts_output(tmp_path, monkeypatch, capsys):
    """output_vts: U written after every step as U_<step>.vts; the final file
    holds the returned field and the L2 norms match a run without output."""
    from opesci.output import read_vts
    monkeypatch.chdir(tmp_path)
    g, kw = eigen_model(17, 4, 6, double=True)
    g.execute(str(tmp_path / 'model.h'), compiler='hip')
    conv_plain = g.convergence()
    U_plain = g.field_array('U', 6 % 2).copy()
    g.free()
    g.set_switches(output_vts=True)
    g.execute(str(tmp_path / 'model.h'), compiler='hip')
    conv = g.convergence()
    for ti in range(6):
        assert (tmp_path / ('U_%d.vts' % ti)).exists()
    last, _ = read_vts(str(tmp_path / 'U_5.vts'))
    assert np.allclose(last, g.field_array('U', 6 % 2), rtol=0, atol=1e-6 * np.abs(U_plain).max())
    assert np.array_equal(g.field_array('U', 6 % 2), U_plain)
    for k in conv:
        assert conv[k] == pytest.approx(conv_plain[k], rel=1e-12, abs=1e-15)
    g.free()


def test_precision_sweep_small(tmp_path):
    """16-bit field storage against FP64 on a short 8th-order run (64^3,
    300 steps, Ricker source, 48 x 48 receiver plane): FP32 storage within
    1e-5, FP16 within 2e-2 and BF16 within 1e-1 relative L2 (per field and
    on the receivers).  Longer, finer runs stagnate in 16 bits (increments
    below half an ulp are lost)."""
    import json
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'opesci-fd_amd'))
    import precision_sweep
    out = tmp_path / 'p.json'
    precision_sweep.main(['--size', '64', '--order', '8', '--steps', '300', '--plane', '48', '--json', str(out)])
    res = json.loads(out.read_text())['storages']
    for kind, bound in (('fp32', 1e-5), ('fp16', 2e-2), ('bf16', 1e-1)):
        errs = list(res[kind]['field_rel_l2'].values()) + [res[kind]['receiver_rel_l2']]
        assert max(errs) < bound, (kind, max(errs))
        assert min(errs) > 0.0
