"""ctypes binding of libopesci_hip.so (include/opesci_hip.h).

This is the Python side of the drop-in boundary: the reference loads its
generated library with ctypes.cdll and calls opesci_execute /
opesci_convergence through ctypes structs (opesci/grid.py:35-156); the same
happens here against the one prebuilt gfx950 library, after the model
description is handed over with opesci_model_load / opesci_ctx_create.
No compute happens in Python and there is no fallback: a missing library or
GPU raises.
"""
import ctypes
import os

import numpy as np

from .compilation import library_path

ABI_VERSION = 2
FLAG_FUSED_STEP = 1
FLAG_POINTWISE_SOLUTION = 2
FIELDS = ('U', 'V', 'W', 'Txx', 'Tyy', 'Tzz', 'Txy', 'Tyz', 'Txz')
MEDIA_IDS = {'rho': 0, 'vp': 1, 'vs': 2, 'beta': 3, 'lambda': 4, 'mu': 5, 'beta1': 6, 'beta2': 7, 'beta3': 8,
             'mu12': 9, 'mu13': 10, 'mu23': 11}

_i32 = ctypes.c_int32
_f64 = ctypes.c_double
_vp = ctypes.c_void_p


class OpesciModelDesc(ctypes.Structure):
    _fields_ = [
        ('abi_version', _i32), ('dims', _i32 * 3), ('space_order', _i32), ('ntsteps', _i32),
        ('real_bytes', _i32), ('storage', _i32), ('media_arrays', _i32), ('flags', _i32),
        ('dt', _f64), ('spacing', _f64 * 3), ('media_const', _f64 * 12),
        ('cn', _f64 * 36), ('cs', _f64 * 24), ('cv', _f64 * 36),
        ('rho', _vp), ('vp', _vp), ('vs', _vp),
        ('nstmts', _i32), ('nterms', _i32), ('ncode', _i32), ('pad0', _i32),
        ('stmts', _vp), ('terms', _vp), ('bc_code', _vp),
        ('sol_off', _i32 * 9), ('sol_len', _i32 * 9), ('nsol', _i32), ('pad1', _i32),
        ('sol_code', _vp),
        ('src_on', _i32), ('src_idx', _i32 * 3), ('src_wavelet', _vp),
        ('nrec', _i32), ('pad2', _i32), ('rec_idx', _vp),
        ('tile_y', _i32), ('x_chunk', _i32), ('kernel_variant', _i32), ('use_graph', _i32),
        ('ndterms', _i32), ('pad3', _i32), ('dterms', _vp),
    ]


class OpesciGrid(ctypes.Structure):
    _fields_ = [(n, _vp) for n in FIELDS]


class OpesciProfilingC(ctypes.Structure):
    _fields_ = [('rtime', ctypes.c_float), ('ptime', ctypes.c_float), ('mflops', ctypes.c_float)]


_LIB = {}
EXPORTS = ('opesci_model_load', 'opesci_execute', 'opesci_convergence', 'opesci_free',
           'opesci_ctx_create', 'opesci_ctx_destroy', 'opesci_ctx_init', 'opesci_ctx_run',
           'opesci_ctx_l2', 'opesci_ctx_download', 'opesci_ctx_upload', 'opesci_ctx_receivers', 'opesci_ctx_receivers_f64',
           'opesci_ctx_steps', 'opesci_ctx_kernel_times', 'opesci_ctx_stats', 'opesci_ctx_configure',
           'opesci_batch_create', 'opesci_batch_run', 'opesci_batch_destroy', 'opesci_ctx_gradient',
           'opesci_ctx_misfit_forward', 'opesci_ctx_gradient_reverse', 'opesci_ctx_download_gradient', 'opesci_ctx_upload_media', 'opesci_ctx_download_media',
           'opesci_ctx_l2_noalloc', 'opesci_last_error',
           'opesci_abi_version', 'opesci_device_info')


def load(path=None):
    path = path or library_path()
    if path not in _LIB:
        if not os.path.exists(path):
            raise OSError('%s is not built (run __graft_entry__.build())' % path)
        lib = ctypes.CDLL(path)
        lib.opesci_last_error.restype = ctypes.c_char_p
        lib.opesci_abi_version.restype = ctypes.c_int
        for n in EXPORTS:
            f = getattr(lib, n)
            if n not in ('opesci_last_error',):
                f.restype = ctypes.c_int
        lib.opesci_ctx_create.argtypes = [ctypes.POINTER(OpesciModelDesc), ctypes.POINTER(_vp)]
        lib.opesci_ctx_destroy.argtypes = [_vp]
        lib.opesci_ctx_init.argtypes = [_vp]
        lib.opesci_ctx_run.argtypes = [_vp, _i32, ctypes.POINTER(_f64)]
        lib.opesci_ctx_l2.argtypes = [_vp, ctypes.POINTER(_f64)]
        lib.opesci_ctx_l2_noalloc.argtypes = [_vp, _vp, _vp]
        lib.opesci_ctx_download.argtypes = [_vp, _i32, _i32, _vp]
        lib.opesci_ctx_upload.argtypes = [_vp, _i32, _i32, _vp]
        lib.opesci_ctx_receivers.argtypes = [_vp, _vp, ctypes.c_int64]
        lib.opesci_ctx_receivers_f64.argtypes = [_vp, _vp, ctypes.c_int64]
        lib.opesci_ctx_steps.argtypes = [_vp, ctypes.POINTER(_i32)]
        lib.opesci_ctx_kernel_times.argtypes = [_vp, _i32, ctypes.POINTER(_f64)]
        lib.opesci_ctx_stats.argtypes = [_vp, ctypes.POINTER(_i32)]
        lib.opesci_ctx_configure.argtypes = [_vp, _i32, _i32, _i32]
        lib.opesci_batch_create.argtypes = [ctypes.POINTER(_vp), _i32, ctypes.POINTER(_vp)]
        lib.opesci_batch_run.argtypes = [_vp, _i32, ctypes.POINTER(_f64)]
        lib.opesci_batch_destroy.argtypes = [_vp]
        lib.opesci_ctx_gradient.argtypes = [_vp, _vp, ctypes.c_int64, _i32, ctypes.POINTER(_f64)]
        lib.opesci_ctx_misfit_forward.argtypes = [_vp, _vp, ctypes.c_int64, _i32, ctypes.POINTER(_f64)]
        lib.opesci_ctx_gradient_reverse.argtypes = [_vp]
        lib.opesci_ctx_download_gradient.argtypes = [_vp, _i32, _vp]
        lib.opesci_ctx_upload_media.argtypes = [_vp, _i32, _vp]
        lib.opesci_ctx_download_media.argtypes = [_vp, _i32, _vp]
        lib.opesci_model_load.argtypes = [ctypes.POINTER(OpesciModelDesc)]
        lib.opesci_execute.argtypes = [_vp, _vp]
        lib.opesci_convergence.argtypes = [_vp, _vp]
        lib.opesci_free.argtypes = [_vp]
        lib.opesci_device_info.argtypes = [ctypes.c_char_p, _i32, ctypes.c_char_p, _i32]
        if lib.opesci_abi_version() != ABI_VERSION:
            raise OSError('libopesci_hip ABI %d, expected %d' % (lib.opesci_abi_version(), ABI_VERSION))
        _LIB[path] = lib
    return _LIB[path]


def _hip_runtimes():
    """Files of the HIP runtimes mapped into this process."""
    try:
        with open('/proc/self/maps') as f:
            return sorted({ln.split()[-1] for ln in f if 'libamdhip64' in ln})
    except OSError:
        return []


def _is_device_tensor(a):
    """A torch device tensor, usable here only when torch and this library
    share one HIP runtime.  torch's wheel carries its own runtime under the
    SONAME this library links against: imported first, torch's runtime is
    the one this library binds to; imported after the library, a second
    runtime (and HSA runtime) loads beside the first, and their device
    pointers and streams are not each other's."""
    if not (hasattr(a, 'data_ptr') and getattr(a, 'is_cuda', False)):
        return False
    rts = _hip_runtimes()
    if len(rts) > 1:
        raise RuntimeError('two HIP runtimes in this process (%s): import torch before loading the opesci '
                           'library so that both use one' % ', '.join(rts))
    return True


def _check(rc, what):
    if rc != 0:
        msg = load().opesci_last_error()
        raise RuntimeError('%s failed (code %d): %s' % (what, rc, msg.decode() if msg else ''))


def device_string():
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(64)
    load().opesci_device_info(name, 256, arch, 64)
    return '%s (%s)' % (name.value.decode(), arch.value.decode())


def make_desc(plan):
    """Pack a ModelPlan into OpesciModelDesc; returns (desc, keepalive)."""
    d = OpesciModelDesc()
    keep = []

    def ptr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data

    d.abi_version = ABI_VERSION
    d.dims[:] = plan.dims
    d.space_order = plan.order
    d.ntsteps = plan.ntsteps
    d.real_bytes = plan.real_bytes
    d.storage = plan.storage_code
    d.media_arrays = 1 if plan.media_arrays else 0
    d.flags = ((FLAG_FUSED_STEP if getattr(plan, 'fuse_step', False) else 0) |
               (FLAG_POINTWISE_SOLUTION if getattr(plan, 'pointwise_solution', False) else 0))
    d.dt = plan.dt
    d.spacing[:] = plan.spacing
    d.media_const[:] = [float(v) for v in plan.media_const]
    d.cn[:] = [float(v) for v in plan.cn.reshape(-1)]
    d.cs[:] = [float(v) for v in plan.cs.reshape(-1)]
    d.cv[:] = [float(v) for v in plan.cv.reshape(-1)]
    if plan.media_arrays:
        d.rho, d.vp, d.vs = [ptr(np.ascontiguousarray(a, dtype=np.float32)) for a in plan.media]
    d.nstmts = len(plan.stmts)
    d.nterms = len(plan.terms)
    d.ncode = len(plan.bc_code)
    d.stmts = ptr(plan.stmts)
    d.terms = ptr(plan.terms)
    d.bc_code = ptr(plan.bc_code)
    d.sol_off[:] = plan.sol_off
    d.sol_len[:] = plan.sol_len
    d.nsol = len(plan.sol_code)
    d.sol_code = ptr(plan.sol_code)
    if plan.src_wavelet is not None:
        w = np.ascontiguousarray(plan.src_wavelet, dtype=np.float32)
        if w.ndim != 2 or w.shape[0] not in (1, 3) or w.shape[1] != max(plan.ntsteps, 1):
            raise ValueError('source traces must be (1 or 3, ntsteps)')
        d.src_on = w.shape[0]
        d.src_idx[:] = [int(v) for v in plan.src_idx]
        d.src_wavelet = ptr(w)
    if plan.rec_idx is not None:
        r = np.ascontiguousarray(plan.rec_idx, dtype=np.int32).reshape(-1, 3)
        d.nrec = r.shape[0]
        d.rec_idx = ptr(r)
    d.tile_y = int(plan.tile_y or 0)
    d.x_chunk = int(plan.x_chunk or 0)
    d.kernel_variant = int(plan.kernel_variant or 0)
    d.use_graph = int(plan.use_graph)
    dterms = getattr(plan, 'dterms', None)
    if dterms is not None and len(dterms):
        d.ndterms = len(dterms)
        d.dterms = ptr(dterms)
    return d, keep


# ---------------------------------------------------------------- reference-named ABI
def execute(plan, arg_grid, arg_prof):
    lib = load()
    desc, keep = make_desc(plan)
    _check(lib.opesci_model_load(ctypes.byref(desc)), 'opesci_model_load')
    _check(lib.opesci_execute(ctypes.byref(arg_grid), ctypes.byref(arg_prof)), 'opesci_execute')


def convergence(plan, arg_grid):
    real = ctypes.c_double if plan.real_bytes == 8 else ctypes.c_float
    out = (real * 9)()
    _check(load().opesci_convergence(ctypes.byref(arg_grid), out), 'opesci_convergence')
    return [float(v) for v in out]


def free(arg_grid):
    _check(load().opesci_free(ctypes.byref(arg_grid)), 'opesci_free')


# ---------------------------------------------------------------- context API
class Context(object):
    """One device-resident model (the handle API of include/opesci_hip.h)."""

    def __init__(self, plan):
        self.lib = load()
        self.plan = plan
        self.desc, self._keep = make_desc(plan)
        h = _vp()
        _check(self.lib.opesci_ctx_create(ctypes.byref(self.desc), ctypes.byref(h)), 'opesci_ctx_create')
        self.h = h
        self.dtype = np.float64 if plan.real_bytes == 8 else np.float32

    def close(self):
        if self.h:
            _check(self.lib.opesci_ctx_destroy(self.h), 'opesci_ctx_destroy')
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        _check(self.lib.opesci_ctx_init(self.h), 'opesci_ctx_init')

    def run(self, nsteps=None):
        s = _f64()
        n = self.plan.ntsteps if nsteps is None else nsteps
        _check(self.lib.opesci_ctx_run(self.h, n, ctypes.byref(s)), 'opesci_ctx_run')
        return s.value

    def l2(self):
        out = (_f64 * 9)()
        _check(self.lib.opesci_ctx_l2(self.h, out), 'opesci_ctx_l2')
        return list(out)

    def download(self, field, level=-1):
        f = FIELDS.index(field) if isinstance(field, str) else field
        a = np.zeros(self.plan.dims, dtype=self.dtype)
        _check(self.lib.opesci_ctx_download(self.h, f, level, a.ctypes.data), 'opesci_ctx_download')
        return a

    def upload(self, field, level, arr):
        f = FIELDS.index(field) if isinstance(field, str) else field
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.shape == tuple(self.plan.dims)
        _check(self.lib.opesci_ctx_upload(self.h, f, level, a.ctypes.data), 'opesci_ctx_upload')

    def fields(self, level=-1):
        return [self.download(i, level) for i in range(9)]

    def steps(self):
        s = _i32()
        _check(self.lib.opesci_ctx_steps(self.h, ctypes.byref(s)), 'opesci_ctx_steps')
        return s.value

    def receivers(self, dtype=np.float32):
        """Traces [step][receiver][U, V, W, pressure]; dtype float64 returns
        the unrounded traces of an FP64 context."""
        n = self.steps() * max(int(self.desc.nrec), 0) * 4
        f64 = np.dtype(dtype) == np.float64
        out = np.zeros(max(n, 1), dtype=np.float64 if f64 else np.float32)
        fn = 'opesci_ctx_receivers_f64' if f64 else 'opesci_ctx_receivers'
        _check(getattr(self.lib, fn)(self.h, out.ctypes.data, n), fn)
        return out[:n].reshape(-1, int(self.desc.nrec), 4) if n else out[:0]

    def kernel_times(self, nsteps):
        out = (_f64 * 3)()
        _check(self.lib.opesci_ctx_kernel_times(self.h, nsteps, out), 'opesci_ctx_kernel_times')
        return list(out)

    def stats(self, full=False):
        """(boundary launches per step, deep z-face statements of the stress
        phase, launches they take[, 1 if the stress sweep applies them in
        place])."""
        out = (_i32 * 5)()
        _check(self.lib.opesci_ctx_stats(self.h, out), 'opesci_ctx_stats')
        return tuple(out[:4] if full else out[:3])

    def fused_margin(self):
        """F of the fused step's box [F, dim - F) in x and y, 0 when the
        stress and velocity updates run as separate launches."""
        out = (_i32 * 5)()
        _check(self.lib.opesci_ctx_stats(self.h, out), 'opesci_ctx_stats')
        return int(out[4])

    def configure(self, kernel_variant=0, tile_y=0, x_chunk=0):
        """Interior launch configuration: kernel_variant 0 sweep (two z points
        per lane for FP32 arithmetic), 1 gather, 2 sweep with one point per
        lane; results are the same for every one."""
        _check(self.lib.opesci_ctx_configure(self.h, kernel_variant, tile_y, x_chunk), 'opesci_ctx_configure')

    def _observed(self, observed):
        """(pointer, size, keep-alive) of the observed traces: a float32
        device tensor as it is (torch's stream synchronised first), anything
        else as a contiguous float32 host array."""
        if _is_device_tensor(observed):
            import torch
            t = observed.detach().to(torch.float32).contiguous()
            torch.cuda.current_stream(t.device).synchronize()
            return t.data_ptr(), t.numel(), t
        obs = np.ascontiguousarray(observed, dtype=np.float32)
        return obs.ctypes.data, obs.size, obs

    def gradient(self, observed, checkpoint_every=32):
        """Misfit J = 1/2 sum (U,V,W receiver traces - observed)^2 and, left on
        the device, dJ/d(media arrays) by the adjoint (see download_gradient).
        observed: (ntsteps, nrec, 3) float32, host array or device tensor."""
        ptr, n, keep = self._observed(observed)
        j = _f64()
        _check(self.lib.opesci_ctx_gradient(self.h, ptr, n, int(checkpoint_every), ctypes.byref(j)),
               'opesci_ctx_gradient')
        del keep
        return j.value

    def misfit_forward(self, observed, checkpoint_every=32):
        """First half of gradient(): the forward pass with checkpoints and
        the misfit J; gradient_reverse() then runs the adjoint."""
        ptr, n, keep = self._observed(observed)
        j = _f64()
        _check(self.lib.opesci_ctx_misfit_forward(self.h, ptr, n, int(checkpoint_every), ctypes.byref(j)),
               'opesci_ctx_misfit_forward')
        del keep
        return j.value

    def gradient_reverse(self):
        """Second half of gradient(): the reverse sweep after misfit_forward()."""
        _check(self.lib.opesci_ctx_gradient_reverse(self.h), 'opesci_ctx_gradient_reverse')

    def download_gradient(self, media, out=None):
        """dJ/d(media array) of the last reverse sweep, dense (dim1, dim2, dim3);
        into `out` (a contiguous device tensor of the compute type) if given."""
        mid = MEDIA_IDS[media] if isinstance(media, str) else int(media)
        if out is not None:
            import torch
            self._check_tensor(out)
            torch.cuda.current_stream(out.device).synchronize()
            _check(self.lib.opesci_ctx_download_gradient(self.h, mid, out.data_ptr()), 'opesci_ctx_download_gradient')
            return out
        a = np.zeros(self.plan.dims, dtype=self.dtype)
        _check(self.lib.opesci_ctx_download_gradient(self.h, mid, a.ctypes.data), 'opesci_ctx_download_gradient')
        return a

    def _check_tensor(self, t):
        import torch
        want = torch.float64 if self.dtype == np.float64 else torch.float32
        if not _is_device_tensor(t) or t.dtype != want or not t.is_contiguous() \
                or tuple(t.shape) != tuple(self.plan.dims):
            raise ValueError('expected a contiguous %s device tensor of shape %s' % (want, tuple(self.plan.dims)))

    def download_media(self, media):
        """One media array as the device holds it (read path), dense (dim1, dim2, dim3)."""
        mid = MEDIA_IDS[media] if isinstance(media, str) else int(media)
        a = np.zeros(self.plan.dims, dtype=self.dtype)
        _check(self.lib.opesci_ctx_download_media(self.h, mid, a.ctypes.data), 'opesci_ctx_download_media')
        return a

    def upload_media(self, media, arr):
        """Replace one media array (dense, compute type) and refresh the
        media-dependent boundary coefficients."""
        mid = MEDIA_IDS[media] if isinstance(media, str) else int(media)
        if _is_device_tensor(arr):
            import torch
            t = arr.detach().to(torch.float64 if self.dtype == np.float64 else torch.float32).contiguous()
            self._check_tensor(t)
            torch.cuda.current_stream(t.device).synchronize()
            _check(self.lib.opesci_ctx_upload_media(self.h, mid, t.data_ptr()), 'opesci_ctx_upload_media')
            return
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.shape != tuple(self.plan.dims):
            raise ValueError('media array shape %s, grid %s' % (a.shape, self.plan.dims))
        _check(self.lib.opesci_ctx_upload_media(self.h, mid, a.ctypes.data), 'opesci_ctx_upload_media')
