"""Point sources, receivers and their input files.

The reference generator has no source or receiver support; the hand-written
propagator tests/src/test_ref_iso_elastic.cpp does, and its helpers live in
src/opesciIO.cpp and src/opesciHandy.cpp.  This module gives the same
behaviour on the host side of the GPU propagator:

* geometry files: a header line, then one "x y z" line per source or
  receiver (opesci_read_receivers / opesci_read_souces, src/opesciIO.cpp:
  345-396); traces are raw float32 (opesci_read_simple_binary, :296-317);
* nearest grid point of a model coordinate (test_ref_iso_elastic.cpp:229-
  283 rounds coord/h; here coordinates follow the generator's convention
  x = dx*(i - m), so i = round(x/dx) + m);
* spectral resampling of a trace to the simulation time step
  (opesci_resample_timeseries, src/opesciHandy.cpp:117-195);
* CFL time step from the largest Vp (opesci_calculate_dt, :69-89) and Lame
  constants from Vp, Vs, rho (opesci_calculate_lame_costants, :53-67);
* the Ricker wavelet (1 - 2 pi^2 f0^2 tau^2) exp(-pi^2 f0^2 tau^2),
  tau = t - 1/f0.
"""
import numpy as np

__all__ = ['ricker', 'nearest_index', 'read_geometry', 'read_receivers', 'read_sources',
           'read_simple_binary', 'resample_timeseries', 'calculate_dt', 'calculate_lame_constants']


def ricker(f0, dt, nt, t0=None):
    """Ricker wavelet of peak frequency f0 sampled at t = n*dt, n < nt, with
    its peak at t0 (default 1/f0)."""
    t = np.arange(nt, dtype=np.float64) * dt
    tau = t - (1.0 / f0 if t0 is None else t0)
    a = (np.pi * f0 * tau) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def nearest_index(coord, spacing, margin):
    """Grid index (i, j, k) of the unstaggered point nearest to a model
    coordinate; coordinate 0 is index `margin`.  Halves round away from zero,
    like C's round()."""
    out = []
    for x, h in zip(coord, spacing):
        r = float(x) / float(h)
        out.append(int(np.sign(r) * np.floor(abs(r) + 0.5)) + int(margin))
    return tuple(out)


def read_geometry(filename):
    """(n, 3) float32 coordinates from a text file: one header line, then
    whitespace-separated 'x y z' lines (blank lines skipped)."""
    rows = []
    with open(filename) as f:
        lines = f.read().splitlines()
    if not lines:
        raise IOError('geometry file %s is empty' % filename)
    for line in lines[1:]:
        if not line.strip():
            continue
        v = line.split()
        if len(v) < 3:
            raise IOError('bad geometry line in %s: %r' % (filename, line))
        rows.append([float(v[0]), float(v[1]), float(v[2])])
    return np.asarray(rows, dtype=np.float32).reshape(-1, 3)


read_receivers = read_geometry


def read_simple_binary(filename, size=None):
    """Raw float32 values of a file; with size, exactly `size` values (a
    shorter file raises instead of reading past its end, a longer one
    raises like the reference's size check)."""
    a = np.fromfile(filename, dtype=np.float32)
    if size is not None:
        if a.size < size:
            raise IOError('%s holds %d values, %d needed' % (filename, a.size, size))
        if a.size > size:
            raise IOError('%s holds %d values, more than the %d of the array' % (filename, a.size, size))
    return a


def read_sources(xyz_filename, xsrc_filename, ysrc_filename, zsrc_filename):
    """Source coordinates (n, 3) and the x, y, z traces (3, nt)."""
    xyz = read_geometry(xyz_filename)
    tr = [read_simple_binary(fn) for fn in (xsrc_filename, ysrc_filename, zsrc_filename)]
    n = min(t.size for t in tr)
    return xyz, np.stack([t[:n] for t in tr])


def _dft(re, im):
    """Forward DFT as the reference defines it: X_k = sum_t x_t e^{-2 pi i t k/n}."""
    x = np.asarray(re, dtype=np.float64) + 1j * np.asarray(im, dtype=np.float64)
    return np.fft.fft(x)


def resample_timeseries(src, dt, sdt):
    """Resample a trace sampled every sdt to the time step dt by padding
    (dt < sdt) or cutting (dt > sdt) the middle of its spectrum, keeping the
    frequency content; unitary normalisation on both transforms.  Returns the
    input when dt and sdt agree to float precision.

    Same algorithm and index arithmetic as opesci_resample_timeseries
    (src/opesciHandy.cpp:125-195), in double precision with an FFT instead of
    the O(n^2) float DFT, so results agree with it to float round-off.
    """
    src = np.asarray(src, dtype=np.float32)
    eps = np.finfo(np.float32).eps
    if abs(float(np.float32(dt)) - sdt) < eps * (float(np.float32(dt)) + sdt):
        return src.copy()
    dtf = float(np.float32(dt))
    snt = src.size
    snt2 = int(np.floor(snt * sdt / dtf + 0.5))
    ft = _dft(src, np.zeros(snt)) / np.sqrt(snt)
    ft2 = np.zeros(snt2, dtype=np.complex128)
    if dtf < sdt:
        mid = snt // 2
        ft2[:mid] = ft[:mid]
        off = snt2 - snt
        ft2[off + mid:off + snt] = ft[mid:snt]
    else:
        mid = snt2 // 2
        ft2[:mid] = ft[:mid]
        off = snt - snt2
        ft2[mid:snt2] = ft[off + mid:off + snt2]
    # inverse as the reference does it: conjugate, forward DFT, keep the real part
    out = _dft(ft2.real, -ft2.imag).real / np.sqrt(snt2)
    return out.astype(np.float32)


def calculate_dt(vp, h):
    """(6/7) (1/sqrt(3)) h / max(Vp), with h / max(Vp) in float as the
    reference evaluates it."""
    maxv = np.max(np.asarray(vp, dtype=np.float32))
    q = float(np.float32(h) / maxv)
    return float(np.float32((6.0 / 7.0) * (1.0 / np.sqrt(3.0)) * q))


def calculate_lame_constants(vp, vs, rho):
    """(mu, lambda) float32: mu = rho vs vs in float; lambda = rho (vp vp -
    2 vs vs) with vp vp rounded to float and the rest in double, the
    reference's mixed float/double arithmetic."""
    vp, vs, rho = (np.asarray(a, dtype=np.float32) for a in (vp, vs, rho))
    if not (vp.shape == vs.shape == rho.shape):
        raise ValueError('vp, vs and rho must have the same shape')
    mu = (rho * vs * vs).astype(np.float32)
    vp2 = (vp * vp).astype(np.float64)
    vs64 = vs.astype(np.float64)
    lam = (rho.astype(np.float64) * (vp2 - 2.0 * vs64 * vs64)).astype(np.float32)
    return mu, lam
