"""ORACLE loader -- test infrastructure only.

ctypes front end for ``oracle/elastic_oracle.cpp``, the CPU/OpenMP
restatement of the program opesci-fd's StaggeredGrid generator emits
(opesci/staggeredgrid.py, opesci/templates/staggered3d_tmpl.py).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module; the product package never does.

Field order everywhere is the reference ABI order ``StaggeredGrid.fields``
= vfields + sfields = U, V, W, Txx, Tyy, Tzz, Txy, Tyz, Txz
(opesci/staggeredgrid.py:69-71).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("U", "V", "W", "Txx", "Tyy", "Tzz", "Txy", "Tyz", "Txz")


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int32),
        ("n", ctypes.c_int32 * 3),
        ("ntsteps", ctypes.c_int32),
        ("real_bytes", ctypes.c_int32),
        ("media_arrays", ctypes.c_int32),
        ("snapshot", ctypes.c_int32),
        ("nthreads", ctypes.c_int32),
        ("bc_algo", ctypes.c_int32),
        ("ic_mode", ctypes.c_int32),
        ("src_on", ctypes.c_int32),
        ("src_idx", ctypes.c_int32 * 3),
        ("nrec", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("L", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("rho", ctypes.c_double),
        ("lam", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("rho_arr", ctypes.POINTER(ctypes.c_float)),
        ("vp_arr", ctypes.POINTER(ctypes.c_float)),
        ("vs_arr", ctypes.POINTER(ctypes.c_float)),
        ("src_wavelet", ctypes.POINTER(ctypes.c_float)),
        ("rec_idx", ctypes.POINTER(ctypes.c_int32)),
        ("rec_out", ctypes.POINTER(ctypes.c_float)),
    ]


_LIBS = {}


def build():
    """Compile the oracle (and the reference snapshot when /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(native=False):
    name = "liboracle_native.so" if native else "liboracle.so"
    if name not in _LIBS:
        path = os.path.join(HERE, "lib", name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.oracle_run.restype = ctypes.c_int
        lib.oracle_run.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_void_p * 9,
                                   ctypes.c_double * 9, ctypes.c_double * 9,
                                   ctypes.c_double * 2]
        lib.oracle_norms.restype = ctypes.c_int
        lib.oracle_norms.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_void_p * 9,
                                     ctypes.c_double * 9, ctypes.c_double * 9]
        lib.oracle_run_from.restype = ctypes.c_int
        lib.oracle_run_from.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_void_p * 9, ctypes.c_void_p * 9]
        lib.oracle_media.restype = ctypes.c_int
        lib.oracle_media.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_void_p * 9]
        lib.oracle_fd_weights.restype = ctypes.c_int
        lib.oracle_fd_weights.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        _LIBS[name] = lib
    return _LIBS[name]


def fd_weights(half_order):
    w = (ctypes.c_double * half_order)()
    assert load().oracle_fd_weights(half_order, w) == 0
    return list(w)


def resample(src, dt, sdt):
    """opesci_resample_timeseries restated in C (float DFT), see
    oracle_resample."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    cap = int(round(src.size * sdt / float(np.float32(dt)))) + src.size + 8
    out = np.zeros(cap, dtype=np.float32)
    lib = load()
    lib.oracle_resample.restype = ctypes.c_int
    lib.oracle_resample.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_float, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    n = lib.oracle_resample(src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), src.size, float(dt), float(sdt),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), cap)
    assert n >= 0
    return out[:n].copy()


def media_from_velocities(rho, vp, vs):
    """StaggeredGrid.set_media_params(read=False) (opesci/staggeredgrid.py:279-282)."""
    return rho, rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2


def _params(order, n, ntsteps, dt, L, double, rho, lam, mu, vp, vs, snapshot, nthreads, bc_algo,
            ic_mode):
    if lam is None or mu is None:
        _, lam, mu = media_from_velocities(rho, vp, vs)
    p = OracleParams()
    p.order = order
    p.n[:] = n
    p.ntsteps = ntsteps
    p.real_bytes = 8 if double else 4
    p.snapshot = 1 if snapshot else 0
    p.nthreads = nthreads
    p.bc_algo = bc_algo
    p.ic_mode = ic_mode
    p.L[:] = L
    p.dt = dt
    p.rho, p.lam, p.mu = rho, lam, mu
    return p


def norms(fields, order=4, n=(16, 16, 16), ntsteps=10, dt=None, L=(1.0, 1.0, 1.0), double=False,
          rho=1.0, lam=None, mu=None, vp=1.0, vs=0.5, snapshot=False, media=None):
    """L2 norms (f64-accumulated and reference-style) of externally computed
    final fields with the oracle's conventions; returns (l2_f64, l2_ref)."""
    p = _params(order, tuple(n), ntsteps, dt, L, double, rho, lam, mu, vp, vs, snapshot, 0, -1, 1)
    keep = []
    if media is not None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in media]
        keep += arrs
        p.media_arrays = 1
        p.rho_arr, p.vp_arr, p.vs_arr = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in arrs]
    dtype = np.float64 if double else np.float32
    fs = [np.ascontiguousarray(f, dtype=dtype) for f in fields]
    ptrs = (ctypes.c_void_p * 9)(*[f.ctypes.data for f in fs])
    l2 = (ctypes.c_double * 9)()
    l2r = (ctypes.c_double * 9)()
    rc = load().oracle_norms(ctypes.byref(p), ptrs, l2, l2r)
    if rc != 0:
        raise RuntimeError("oracle_norms failed with code %d" % rc)
    return list(l2), list(l2r)


def run(order=4, n=(16, 16, 16), ntsteps=10, dt=None, L=(1.0, 1.0, 1.0), double=False,
        rho=1.0, lam=None, mu=None, vp=1.0, vs=0.5, media=None, snapshot=False,
        nthreads=0, bc_algo=-1, ic_mode=0, src=None, src_idx=None, rec_idx=None,
        want_fields=True, native=False):
    """Run the oracle; returns dict(fields=[9 arrays (dim1,dim2,dim3)], l2=.., l2_ref=.., seconds=..).

    ``media`` = (rho, vp, vs) float32 arrays of shape (dim1, dim2, dim3) selects
    the reference read path; otherwise constant media from (rho, lam, mu) or
    (rho, vp, vs).
    """
    n = tuple(int(v) for v in n)
    M = order // 2
    dims = tuple(v + 1 + 2 * M for v in n)
    if lam is None or mu is None:
        _, lam, mu = media_from_velocities(rho, vp, vs)
    if dt is None:
        raise ValueError("dt required")
    p = OracleParams()
    p.order = order
    p.n[:] = n
    p.ntsteps = ntsteps
    p.real_bytes = 8 if double else 4
    p.snapshot = 1 if snapshot else 0
    p.nthreads = nthreads
    p.bc_algo = bc_algo
    p.ic_mode = ic_mode
    p.L[:] = L
    p.dt = dt
    p.rho, p.lam, p.mu = rho, lam, mu
    keep = []
    if media is not None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in media]
        for a in arrs:
            assert a.size == dims[0] * dims[1] * dims[2]
        keep += arrs
        p.media_arrays = 1
        p.rho_arr, p.vp_arr, p.vs_arr = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in arrs]
    rec_out = None
    if src is not None:
        # one trace (Txx, Tyy, Tzz alike) or three (x, y, z), each >= ntsteps long
        s = np.atleast_2d(np.asarray(src, dtype=np.float32))
        assert s.shape[0] in (1, 3) and s.shape[1] >= ntsteps
        s = np.ascontiguousarray(s[:, :max(ntsteps, 1)])
        keep.append(s)
        p.src_on = s.shape[0]
        p.src_wavelet = s.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        p.src_idx[:] = src_idx
    if rec_idx is not None:
        r = np.ascontiguousarray(rec_idx, dtype=np.int32).reshape(-1, 3)
        keep.append(r)
        p.nrec = r.shape[0]
        p.rec_idx = r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        rec_out = np.zeros((ntsteps, r.shape[0], 4), dtype=np.float32)
        p.rec_out = rec_out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    dtype = np.float64 if double else np.float32
    outs = [np.zeros(dims, dtype=dtype) for _ in range(9)] if want_fields else None
    ptrs = (ctypes.c_void_p * 9)(*([o.ctypes.data for o in outs] if outs else [None] * 9))
    l2 = (ctypes.c_double * 9)()
    l2r = (ctypes.c_double * 9)()
    tm = (ctypes.c_double * 2)()
    rc = load(native).oracle_run(ctypes.byref(p), ptrs, l2, l2r, tm)
    if rc != 0:
        raise RuntimeError("oracle_run failed with code %d" % rc)
    return dict(fields=outs, l2=list(l2), l2_ref=list(l2r), seconds=tm[0], seconds_total=tm[1],
                dims=dims, rec=rec_out)


def run_from(state, order=4, n=(16, 16, 16), ntsteps=1, dt=None, L=(1.0, 1.0, 1.0), double=True,
             rho=1.0, lam=None, mu=None, vp=1.0, vs=0.5, media=None, bc_algo=-1, nthreads=0):
    """Steps from a given state: ``state`` = 9 arrays (dim1, dim2, dim3) that
    become level 0 as they are (ghost cells included, no analytic initial
    state and no initial boundary pass); returns the 9 arrays of the level
    after ``ntsteps`` steps."""
    n = tuple(int(v) for v in n)
    M = order // 2
    dims = tuple(v + 1 + 2 * M for v in n)
    if dt is None:
        raise ValueError("dt required")
    p = _params(order, n, ntsteps, dt, L, double, rho, lam, mu, vp, vs, False, nthreads, bc_algo, 1)
    keep = []
    if media is not None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in media]
        keep += arrs
        p.media_arrays = 1
        p.rho_arr, p.vp_arr, p.vs_arr = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in arrs]
    dtype = np.float64 if double else np.float32
    ins = [np.ascontiguousarray(a, dtype=dtype) for a in state]
    for a in ins:
        assert a.shape == dims, (a.shape, dims)
    outs = [np.zeros(dims, dtype=dtype) for _ in range(9)]
    rc = load().oracle_run_from(ctypes.byref(p), (ctypes.c_void_p * 9)(*[a.ctypes.data for a in ins]),
                                (ctypes.c_void_p * 9)(*[o.ctypes.data for o in outs]))
    if rc != 0:
        raise RuntimeError("oracle_run_from failed with code %d" % rc)
    return outs


MEDIA_NAMES = ("beta1", "beta2", "beta3", "lambda", "mu", "mu12", "mu13", "mu23", "beta")


def media(rho, vp, vs, order=4, double=True):
    """The read path's effective media as the oracle builds them from float32
    rho/Vp/Vs arrays of the full grid shape: dict name -> float64 array."""
    rho, vp, vs = (np.ascontiguousarray(a, dtype=np.float32) for a in (rho, vp, vs))
    M = order // 2
    dims = rho.shape
    n = tuple(v - 1 - 2 * M for v in dims)
    p = _params(order, n, 0, 1.0, (1.0, 1.0, 1.0), double, 1.0, 1.0, 1.0, 1.0, 0.5, False, 0, -1, 1)
    p.media_arrays = 1
    flat = [a.reshape(-1) for a in (rho, vp, vs)]
    p.rho_arr, p.vp_arr, p.vs_arr = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in flat]
    outs = [np.zeros(dims, dtype=np.float64) for _ in range(9)]
    rc = load().oracle_media(ctypes.byref(p), (ctypes.c_void_p * 9)(*[o.ctypes.data for o in outs]))
    if rc != 0:
        raise RuntimeError("oracle_media failed with code %d" % rc)
    return dict(zip(MEDIA_NAMES, outs))


def snapshot_test3d(nthreads=0):
    """The configuration of benchmarks/test3d.cpp: 4th order, 100 cells, dt=0.002,
    1000 steps, float, mu = lambda = 0.5, rho = beta = 1 (test3d.cpp:52-68)."""
    return run(order=4, n=(100, 100, 100), ntsteps=1000, dt=0.002, rho=1.0, lam=0.5, mu=0.5,
               snapshot=True, nthreads=nthreads, want_fields=False)
