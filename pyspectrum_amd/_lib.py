"""ctypes binding of libpsx.so (the gfx950 HIP library) and its status codes.

The library is built in-tree by ``python -m pyspectrum_amd.build``.  There is
no fallback: if the shared library is missing or no GPU is visible, every
call raises.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsx.so")

PSX_OK = 0
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}
SURVEY_SUMS = 13  # PSX_SURVEY_SUMS

_c = ctypes
_p = _c.c_void_p
_i = _c.c_int
_i64 = _c.c_int64
_d = _c.c_double
_f = _c.c_float

class PsxBox(ctypes.Structure):
    """psx_box of include/psx.h: one box of psx_bk_periodic_batch."""
    _fields_ = [("xyz", ctypes.c_void_p), ("w", ctypes.c_void_p), ("np", ctypes.c_int64),
                ("lbox", ctypes.c_double), ("sumw", ctypes.c_double),
                ("wsum_abs", ctypes.c_double), ("nbar", ctypes.c_double),
                ("b3", ctypes.c_double), ("b6", ctypes.c_double)]


# name -> argtypes (all return int status)
SIGNATURES = {
    "psx_abi_version": [],
    "psx_ctx_create": [_c.POINTER(_p), _i],
    "psx_ctx_destroy": [_p],
    "psx_ctx_set_stream": [_p, _p],
    "psx_ctx_synchronize": [_p],
    "psx_ctx_workspace_bytes": [_p, _c.POINTER(_i64)],
    "psx_ctx_release_workspace": [_p],
    "psx_ctx_kernel_timing": [_p, _i],
    "psx_ctx_kernel_time": [_p, _c.c_char_p, _c.POINTER(_i64), _c.POINTER(_d)],
    "psx_assign_quad": [_p, _p, _p, _i64, _i, _f, _f, _f, _f, _d, _p, _i],
    "psx_fft3d": [_p, _p, _i, _i, _i],
    "psx_fcomb_periodic": [_p, _p, _p, _i, _d, _i],
    "psx_fft_periodic": [_p, _p, _p, _i64, _i, _d, _d, _d, _p, _i],
    "psx_fft_survey_mono": [_p, _p, _p, _i64, _i, _d, _d, _p, _i],
    "psx_survey_catalogue": [_p, _p, _p, _p, _i64, _p, _p, _i, _d, _d, _d, _p, _p, _p],
    "psx_reflect_delta": [_p, _p, _p, _i, _i],
    "psx_quad_weights": [_p, _p, _p, _i64, _i, _i, _i, _i, _p],
    "psx_five_delta2g_1": [_p, _p, _p, _p, _i, _i, _i],
    "psx_five_delta2g_2": [_p, _p, _p, _p, _p, _p, _i, _i, _i],
    "psx_build_quad": [_p, _p, _p, _i, _i, _i],
    "psx_hermitian_half": [_p, _p, _i, _i64, _i64, _i64, _p, _i, _i],
    "psx_bin_power": [_p, _p, _i, _i, _i, _d, _p, _i],
    "psx_bin_power_noalloc": [_p, _p, _i, _i, _i, _d, _p, _i],
    "psx_pk_pbox_rsd": [_p, _p, _i, _i, _i, _i, _i, _p, _i],
    "psx_mock_gaussian": [_p, _i, _d, _c.c_uint64, _d, _d, _p, _c.POINTER(_d)],
    "psx_mock_sample": [_p, _p, _i, _d, _i64, _c.c_uint64, _d, _d, _i, _p],
    "psx_mock_velocities": [_p, _i64, _c.c_uint64, _d, _p],
    "psx_apply_rsd": [_p, _p, _p, _i64, _d, _d, _i, _p],
    "psx_bk_num_triangles": [_i, _i, _i, _c.POINTER(_i64)],
    "psx_bk_triangles": [_i, _i, _i, _p],
    "psx_bk_shell_fields": [_p, _p, _i, _i, _i, _i, _p, _i],
    "psx_bk_contract": [_p, _p, _i, _i, _i, _i, _p, _i],
    "psx_bk_raw": [_p, _p, _i, _i, _i, _i, _p, _p, _i],
    "psx_bk_counts": [_p, _i, _i, _i, _i, _p],
    "psx_bk_workspace_bytes": [_p, _i, _i, _i, _i, _i64, _i, _c.POINTER(_i64)],
    "psx_bk_finalize": [_p, _p, _p, _p, _p, _i, _i, _d, _d, _d, _d, _d, _p],
    "psx_bk_periodic": [_p, _p, _p, _i64, _i, _d, _d, _d, _i, _i, _i, _p, _p, _d, _d, _d, _d, _d,
                        _p, _p, _i],
    "psx_bk_periodic_batch": [_p, _i, _c.POINTER(PsxBox), _i, _i, _i, _i, _p, _p, _d, _d, _p, _p,
                              _p, _i],
}


class PsxError(RuntimeError):
    """A libpsx call returned a non-zero status."""


_lib = None
_lock = threading.Lock()


def lib():
    """Load libpsx.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.isfile(LIB_PATH):
                    raise PsxError("libpsx.so is not built (run `python -m pyspectrum_amd.build`)")
                L = ctypes.CDLL(LIB_PATH)
                for name, args in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = _i
                L.psx_last_error.argtypes = []
                L.psx_last_error.restype = _c.c_char_p
                _lib = L
    return _lib


def call(name, *args):
    """Call ``name`` and raise PsxError with the library's message on failure."""
    st = getattr(lib(), name)(*args)
    if st != PSX_OK:
        msg = lib().psx_last_error().decode(errors="replace")
        raise PsxError("%s failed (status %d): %s" % (name, st, msg))
    return st


class Context:
    """One library context on one device; work goes to torch's current stream.

    ``call(name, *args)`` binds the calling thread's current stream and runs
    the entry point under the context's lock, so threads may share a context:
    their calls are serialised, and the library orders each stream switch
    after the work already queued on the previous stream (the workspace is
    never used by two streams at once)."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise PsxError("no GPU visible: the HIP path has no CPU fallback")
        self.device = device
        self._h = _p()
        self._lock = threading.RLock()
        with torch.cuda.device(device):
            call("psx_ctx_create", ctypes.byref(self._h), device)

    @property
    def handle(self):
        """Bind the current stream and return the raw handle (single-threaded use)."""
        import torch
        stream = torch.cuda.current_stream(self.device).cuda_stream
        call("psx_ctx_set_stream", self._h, _p(stream))
        return self._h

    def call(self, name, *args):
        """``name(ctx, *args)`` on the calling thread's current stream."""
        with self._lock:
            return call(name, self.handle, *args)

    def bk_workspace_bytes(self, Ngrid, step, Ncut, Nmax, npart, precision):
        """Upper bound of the workspace a B(k) call on npart particles allocates."""
        n = _i64()
        with self._lock:
            call("psx_bk_workspace_bytes", self._h, Ngrid, step, Ncut, Nmax, int(npart),
                 PREC[precision], ctypes.byref(n))
        return n.value

    def workspace_bytes(self):
        n = _i64()
        call("psx_ctx_workspace_bytes", self._h, ctypes.byref(n))
        return n.value

    def release_workspace(self):
        """Free the workspace (plans stay) once the context's stream is idle."""
        with self._lock:
            call("psx_ctx_release_workspace", self._h)

    def kernel_timing(self, enable):
        """Start (clearing earlier records) or stop the event timing of the
        main kernels' launches on this context's stream."""
        with self._lock:
            call("psx_ctx_kernel_timing", self._h, 1 if enable else 0)

    def kernel_time(self, kernel):
        """(launches, summed device ms) of ``kernel`` since timing was enabled."""
        n, ms = _i64(), _d()
        with self._lock:
            call("psx_ctx_kernel_time", self._h, kernel.encode(), ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value

    def close(self):
        if self._h:
            lib().psx_ctx_destroy(self._h)
            self._h = _p()


_contexts = {}
_contexts_lock = threading.Lock()


def context(device=0):
    """The process-wide default context of a device (created once)."""
    with _contexts_lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = _contexts[device] = Context(device)
        return ctx


def ptr(t):
    """Raw device pointer of a torch tensor (or None -> NULL)."""
    return None if t is None else _p(t.data_ptr())


def triangles(step, Ncut, Nmax):
    """int32 [ntri, 3] triangle list in the reference's loop order (host)."""
    import numpy as np
    n = _i64()
    call("psx_bk_num_triangles", step, Ncut, Nmax, ctypes.byref(n))
    out = np.zeros((max(n.value, 1), 3), dtype=np.int32)
    call("psx_bk_triangles", step, Ncut, Nmax, out.ctypes.data_as(_p))
    return out[: n.value]
