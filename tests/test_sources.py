"""Source, receiver and media-input helpers (host side), against closed forms
and the oracle's restatement of the reference's resampling.

* Ricker: value 1 at its peak, zeros at tau = +-1/(pi f0 sqrt(2)).
* nearest_index: C round() (halves away from zero), coordinate 0 at margin.
* resample_timeseries: the reference's unitary transforms scale a pure tone
  by sqrt(n/n2) and keep its frequency; up then down returns the trace;
  identity when dt == sdt; oracle (float DFT as src/opesciHandy.cpp) within
  2e-6 of the trace maximum.
* calculate_dt / calculate_lame_constants: closed forms.
* geometry / trace files round trip; short media files raise.
"""
import numpy as np
import pytest

import oracle
from opesci import sources
from workloads import eigen_model


def test_ricker_closed_form():
    f0, dt = 10.0, 1e-4
    w = sources.ricker(f0, dt, 6000, t0=3.0 / f0)
    peak = 3000
    assert w[peak] == pytest.approx(1.0, abs=1e-6)
    assert int(np.argmax(w)) == peak
    zc = 1.0 / (np.pi * f0 * np.sqrt(2.0))
    for t in (3.0 / f0 - zc, 3.0 / f0 + zc):
        i = int(np.floor(t / dt))
        assert w[i] * w[i + 1] <= 0.0
    # zero mean (its spectrum vanishes at zero frequency); tails are below 1e-30
    assert abs(float(np.sum(w.astype(np.float64))) * dt) < 1e-6
    # default delay 1/f0
    assert int(np.argmax(sources.ricker(f0, dt, 3000))) == 1000


def test_nearest_index_rounds_like_c():
    assert sources.nearest_index((0.0, 0.0, 0.0), (0.1, 0.1, 0.1), 2) == (2, 2, 2)
    # exact halves (binary fractions) round away from zero
    assert sources.nearest_index((0.25, 0.375, 0.125), (0.25, 0.25, 0.25), 2) == (3, 4, 3)
    assert sources.nearest_index((-0.375, 0.3, 1.0), (0.25, 0.25, 0.5), 4) == (2, 5, 6)


@pytest.mark.parametrize("n,k,factor", [(128, 5, 2), (200, 7, 3), (96, 3, 0.5)])
def test_resample_pure_tone(n, k, factor):
    """cos(2 pi k t/n) resampled to n2 samples is sqrt(n/n2) cos(2 pi k t'/n2)."""
    sdt = 1.0
    dt = sdt / factor
    t = np.arange(n)
    x = np.cos(2 * np.pi * k * t / n).astype(np.float32)
    y = sources.resample_timeseries(x, dt, sdt)
    n2 = int(round(n * sdt / float(np.float32(dt))))
    assert y.size == n2
    t2 = np.arange(n2)
    expect = np.sqrt(n / n2) * np.cos(2 * np.pi * k * t2 / n2)
    assert np.abs(y - expect).max() < 1e-5


def test_resample_round_trip_and_identity():
    w = sources.ricker(15.0, 0.002, 160)
    up = sources.resample_timeseries(w, 0.001, 0.002)
    back = sources.resample_timeseries(up, 0.002, 0.001)
    assert up.size == 320 and back.size == 160
    assert np.abs(back - w).max() < 1e-5 * np.abs(w).max() + 1e-6
    same = sources.resample_timeseries(w, 0.002, 0.002)
    assert np.array_equal(same, w)


@pytest.mark.parametrize("snt,dt,sdt", [(200, 0.001, 0.002), (300, 0.002, 0.001), (128, 0.0007, 0.001),
                                        (101, 0.003, 0.001), (77, 0.001, 0.0013)])
def test_resample_matches_oracle(snt, dt, sdt):
    rng = np.random.default_rng(snt)
    w = (sources.ricker(12.0, sdt, snt) + 0.1 * rng.standard_normal(snt)).astype(np.float32)
    a = sources.resample_timeseries(w, dt, sdt)
    b = oracle.resample(w, dt, sdt)
    assert a.shape == b.shape
    assert np.abs(a - b).max() <= 2e-6 * np.abs(b).max()


def test_calculate_dt_and_lame():
    rng = np.random.default_rng(3)
    vp = rng.uniform(1.5, 4.5, 1000).astype(np.float32)
    vs = (vp / np.float32(np.sqrt(3.0))).astype(np.float32)
    rho = rng.uniform(1.0, 2.5, 1000).astype(np.float32)
    dt = sources.calculate_dt(vp, 0.01)
    assert dt == pytest.approx((6 / 7) / np.sqrt(3) * 0.01 / vp.max(), rel=1e-6)
    mu, lam = sources.calculate_lame_constants(vp, vs, rho)
    assert np.allclose(mu, rho.astype(np.float64) * vs.astype(np.float64) ** 2, rtol=1e-6)
    assert np.allclose(lam, rho.astype(np.float64) * (vp.astype(np.float64) ** 2 - 2 * vs.astype(np.float64) ** 2),
                       rtol=1e-5)
    # Poisson solid: lambda == mu up to float rounding
    assert np.allclose(lam, mu, rtol=1e-5)


def test_geometry_and_trace_files(tmp_path):
    xyz = tmp_path / 'src.txt'
    xyz.write_text('x y z\n0.5 0.25 0.125\n\n1 2 3\n')
    tr = [sources.ricker(f, 0.001, 50) for f in (10.0, 12.0, 14.0)]
    names = []
    for tag, t in zip('xyz', tr):
        fn = tmp_path / ('%ssrc.bin' % tag)
        t.tofile(str(fn))
        names.append(str(fn))
    pos, traces = sources.read_sources(str(xyz), *names)
    assert np.allclose(pos, [[0.5, 0.25, 0.125], [1, 2, 3]])
    assert traces.shape == (3, 50) and np.array_equal(traces[1], tr[1])
    rec = sources.read_receivers(str(xyz))
    assert rec.shape == (2, 3)
    short = tmp_path / 'short.bin'
    np.zeros(10, np.float32).tofile(str(short))
    with pytest.raises(IOError):
        sources.read_simple_binary(str(short), 11)


def test_plan_source_and_receivers():
    g, _ = eigen_model(20, 4, 30)
    w = sources.ricker(15.0, g.dt.value, 12)
    g.set_source(w, location=(0.5, 0.5, 0.25))
    g.set_receivers(locations=[(0.5, 0.5, 0.25), (0.1, 0.2, 0.3)])
    p = g.lower()
    m = p.M
    h = p.spacing[0]
    assert p.src_idx == sources.nearest_index((0.5, 0.5, 0.25), p.spacing, m)
    assert p.src_wavelet.shape == (1, 30)
    assert np.array_equal(p.src_wavelet[0, :12], w) and not p.src_wavelet[0, 12:].any()
    assert p.rec_idx.shape == (2, 3) and tuple(p.rec_idx[1]) == (int(round(0.1 / h)) + m, int(round(0.2 / h)) + m,
                                                                int(round(0.3 / h)) + m)
    g.set_source(np.stack([w, 2 * w, 3 * w]), index=(40, 2, 2))
    with pytest.raises(Exception):
        g.lower()


def test_oracle_three_traces_linearity():
    """Three equal traces inject exactly what one trace does."""
    kw = dict(order=4, n=(16, 16, 16), ntsteps=25, dt=0.01, rho=1.0, lam=1.0, mu=1.0, double=True,
              src_idx=(9, 8, 10), rec_idx=[(9, 8, 10), (4, 5, 6)])
    w = sources.ricker(8.0, 0.01, 25)
    a = oracle.run(src=w, **kw)
    b = oracle.run(src=np.stack([w, w, w]), **kw)
    assert np.array_equal(a['rec'], b['rec'])
    c = oracle.run(src=np.stack([w, 0 * w, 0 * w]), **kw)
    d = oracle.run(src=np.stack([0 * w, w, w]), **kw)
    base = oracle.run(src=0 * w, **kw)
    # linear in the injected traces: (x only) + (y, z only) - nothing == all
    # three, to the float32 rounding of the recorded traces
    lhs = c['rec'].astype(np.float64) + d['rec'] - base['rec']
    assert np.abs(lhs - a['rec']).max() <= 4e-7 * np.abs(a['rec']).max()
    assert np.abs(a['rec'] - base['rec']).max() > 1e-3 * np.abs(a['rec']).max()


def test_media_files_equal_arrays(tmp_path):
    """The read path from raw float32 rho/Vp/Vs files (opesciIO.cpp:319-342,
    staggeredgrid.py:522-551) hands the device the same media as arrays."""
    import opesci
    from opesci.media import smooth_random_media, write_media_files
    N, order = 14, 4
    n = N - 1
    dims = (N + order,) * 3
    rho, vp, vs = smooth_random_media(dims, seed=7, sigma=2.0)
    names = write_media_files(str(tmp_path / 'm'), rho, vp, vs)
    kw = dict(accuracy_order=[2, order, order, order], double=False)
    a = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), 0.002, 0.011, media_arrays=(rho, vp, vs), **kw)
    b = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), 0.002, 0.011, read=True, rho_file=names[0],
                           vp_file=names[1], vs_file=names[2], **kw)
    pa, pb = a.lower(), b.lower()
    for x, y in zip(pa.media, pb.media):
        assert np.array_equal(x, y)
    assert np.array_equal(pa.stmts, pb.stmts) and np.array_equal(pa.bc_code, pb.bc_code)
    # a short file is an error, not a read past its end
    np.zeros(10, np.float32).tofile(names[1])
    c = opesci.eigenwave3d((1.0, 1.0, 1.0), (n, n, n), 0.002, 0.011, read=True, rho_file=names[0],
                           vp_file=names[1], vs_file=names[2], **kw)
    with pytest.raises(IOError):
        c.lower()


def test_vts_writer_round_trip(tmp_path):
    from opesci.output import read_vts, write_raw, write_vts
    rng = np.random.default_rng(5)
    f = rng.standard_normal((5, 6, 7)).astype(np.float32)
    path = write_vts(str(tmp_path / 'U_3'), f, (0.1, 0.2, 0.3), 2)
    assert path.endswith('U_3.vts')
    text = open(path).read()
    assert 'WholeExtent="0 6 0 5 0 4"' in text and 'type="StructuredGrid"' in text
    g, pts = read_vts(path)
    assert np.array_equal(g, f)
    assert np.allclose(pts[4, 0, 6], [(4 - 2) * 0.1, (0 - 2) * 0.2, (6 - 2) * 0.3])
    raw = write_raw(str(tmp_path / 'U_3'), f)
    assert np.array_equal(np.fromfile(raw, dtype=np.float32).reshape(f.shape), f)
