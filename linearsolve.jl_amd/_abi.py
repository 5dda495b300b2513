"""ctypes binding of liblsx.so, the gfx950 C-ABI library declared in include/lsx.h.

This is the Python form of the `ccall` layer a Julia package would use: the same entry points
with the same argument types.  Loading fails loudly when the library has not been built: there
is no CPU fallback behind this module.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblsx.so")

LSX_OK, LSX_EINVAL, LSX_EHIP, LSX_ENOMEM, LSX_ESTATE, LSX_ETIMEOUT, LSX_EAGAIN = 0, -1, -2, -3, -4, -5, -6
LSX_F64, LSX_F32, LSX_BF16 = 0, 1, 2

_STATUS = {
    LSX_EINVAL: "invalid argument",
    LSX_EHIP: "HIP runtime error",
    LSX_ENOMEM: "out of device memory / over budget",
    LSX_ESTATE: "call out of order",
    LSX_ETIMEOUT: "device-side wait exceeded its bound",
    LSX_EAGAIN: "not ready",
}


class LSXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liblsx error {code} ({_STATUS.get(code, '?')}): {msg}")
        self.code = code


class RefineStats(ctypes.Structure):
    _fields_ = [("mode_used", ctypes.c_int32), ("iterations", ctypes.c_int32), ("outer", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("backward_error", ctypes.c_double), ("escalated", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


PROFILE_CLASSES = ["panel", "laswp_trsm", "gemm", "perm", "trsv", "refine", "convert", "other", "gemm_small"]
_NPC = len(PROFILE_CLASSES)


class Profile(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int64 * _NPC), ("ms", ctypes.c_double * _NPC), ("flops", ctypes.c_double * _NPC),
                ("bytes", ctypes.c_double * _NPC)]


class GmresStats(ctypes.Structure):
    _fields_ = [("retcode", ctypes.c_int32), ("pad", ctypes.c_int32), ("iters", ctypes.c_int64),
                ("resid", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("dtype", ctypes.c_int32), ("n", ctypes.c_int64),
                ("A", ctypes.c_void_p), ("b", ctypes.c_void_p), ("nnz", ctypes.c_int64),
                ("rowptr", ctypes.c_void_p), ("colind", ctypes.c_void_p), ("vals", ctypes.c_void_p),
                ("pl_diag", ctypes.c_void_p), ("memory", ctypes.c_int32), ("restart", ctypes.c_int32),
                ("abstol", ctypes.c_double), ("reltol", ctypes.c_double), ("maxiters", ctypes.c_int64),
                ("synth", ctypes.c_int32), ("grid_m", ctypes.c_int32), ("seed", ctypes.c_uint64), ("pe", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("retcode", ctypes.c_int32), ("info", ctypes.c_int64),
                ("iters", ctypes.c_int64), ("resid", ctypes.c_double), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double)]


# (name, restype, argtypes) for every symbol of include/lsx.h
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double
SIGNATURES = [
    ("lsx_version", ctypes.c_int, []),
    ("lsx_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    ("lsx_ctx_destroy", ctypes.c_int, [_P]),
    ("lsx_last_error", ctypes.c_char_p, [_P]),
    ("lsx_ctx_set_wait_bound", ctypes.c_int, [_P, _D]),
    ("lsx_malloc", ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    ("lsx_free", ctypes.c_int, [_P, _P]),
    ("lsx_memcpy", ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    ("lsx_synchronize", ctypes.c_int, [_P]),
    ("lsx_lu_create", ctypes.c_int, [_P, _I64, ctypes.c_int, ctypes.POINTER(_P)]),
    ("lsx_lu_destroy", ctypes.c_int, [_P]),
    ("lsx_lu_factor", ctypes.c_int, [_P, _P, _I64]),
    ("lsx_lu_get_factors", ctypes.c_int, [_P, _P, _I64, _P]),
    ("lsx_lu_solve", ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64]),
    ("lsx_lu_solve_adjoint", ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64]),
    ("lsx_lu_residual_norm", ctypes.c_int, [_P, _P, _I64, _I64, _P, _I64, _P, _I64, _P]),
    ("lsx_residual_norm_noalloc", ctypes.c_int, [_P, _I64, _P, _I64, _I64, _P, _I64, _P, _I64, _P]),
    ("lsx_lu_profile", ctypes.c_int, [_P, _I32, ctypes.POINTER(Profile)]),
    ("lsx_lu_refine", ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _D, _I32, _I32, ctypes.POINTER(RefineStats)]),
    ("lsx_batch_create", ctypes.c_int, [_P, _I64, _P, ctypes.c_int, ctypes.POINTER(_P)]),
    ("lsx_batch_destroy", ctypes.c_int, [_P]),
    ("lsx_batch_factor_solve", ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    ("lsx_batch_set_tiles", ctypes.c_int, [_P, ctypes.c_int32, _P, _P]),
    ("lsx_batch_set_refinement", ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int32]),
    ("lsx_batch_stats", ctypes.c_int, [_P, _P, _P]),
    ("lsx_csr_create", ctypes.c_int, [_P, _I64, _I64, _P, _P, _P, ctypes.POINTER(_P)]),
    ("lsx_csr_destroy", ctypes.c_int, [_P]),
    ("lsx_csr_spmv", ctypes.c_int, [_P, _P, _P]),
    ("lsx_gmres_solve", ctypes.c_int, [_P, _P, _P, _I32, _I32, _D, _D, _I64, _P, _P, ctypes.POINTER(GmresStats)]),
    ("lsx_gmres_max_memory", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    ("lsx_gmres_solve_batched", ctypes.c_int, [_P, _I64, _P, _P, _I32, _I32, _D, _D, _I64, _P, _P,
                                               ctypes.POINTER(GmresStats)]),
    ("lsx_serve_create", ctypes.c_int, [_P, ctypes.c_size_t, _I32, ctypes.POINTER(_P)]),
    ("lsx_serve_destroy", ctypes.c_int, [_P]),
    ("lsx_serve_submit", ctypes.c_int, [_P, ctypes.POINTER(Request), _P, ctypes.POINTER(ctypes.c_uint64)]),
    ("lsx_serve_poll", ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(Result)]),
    ("lsx_serve_drain", ctypes.c_int, [_P]),
    ("lsx_serve_clock", ctypes.c_double, [_P]),
    ("lsx_serve_stats", ctypes.c_int, [_P, _P]),
]

_lib = None
_lock = threading.Lock()


def lib():
    """Loads liblsx.so once; raises if it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(f"{LIB_PATH} is not built: the gfx950 HIP library is required")
                # One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64.so.7 (same
                # soname as the system one liblsx.so is linked against); whichever is loaded first
                # serves both, and torch fails to initialise on a runtime other than its own.  So
                # when torch is installed it is loaded before liblsx.so.
                try:
                    import torch  # noqa: F401
                except ImportError:
                    pass
                L = ctypes.CDLL(LIB_PATH)
                for name, res, args in SIGNATURES:
                    f = getattr(L, name)
                    f.restype = res
                    f.argtypes = args
                _lib = L
    return _lib


class Context:
    """One GPU (lsx_ctx).  `Context.get(device)` returns a process-wide instance per device."""

    _instances = {}

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.lsx_ctx_create(device, ctypes.byref(h))
        if rc != LSX_OK:
            raise LSXError(rc, f"lsx_ctx_create({device}) failed")
        self.handle = h
        self.device = device

    @classmethod
    def get(cls, device=0):
        with _lock:
            if device not in cls._instances:
                cls._instances[device] = None
        if cls._instances[device] is None:
            cls._instances[device] = Context(device)
        return cls._instances[device]

    def check(self, rc, what):
        if rc < 0:
            msg = lib().lsx_last_error(self.handle)
            raise LSXError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def set_wait_bound(self, seconds):
        """Bound (s) of the device-side waits of later calls; 0 restores the defaults."""
        self.check(lib().lsx_ctx_set_wait_bound(self.handle, float(seconds)), "lsx_ctx_set_wait_bound")
