"""ctypes binding of libsvdh.so (include/svdh.h).

The library is loaded from the package directory (single-svdlib_amd/libsvdh.so).
There is no fallback: if the HIP library is missing or no GPU is present,
the first call that needs it raises.
"""
import ctypes as C
import os
import threading

import numpy as np

from . import error, srchash

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libsvdh.so")

SVDH_F64, SVDH_F32, SVDH_BF16 = 0, 1, 2
PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2
NORM_QR, NORM_LU, NORM_NONE = 0, 1, 2
KEEP_F32, KEEP_BF16 = 0x1, 0x2
KC_SPMV, KC_SPMM, KC_DENSE, KC_BLAS1, KC_GEMV, KC_SPMM_T, KC_SDDMM, KC_INGEST = 0, 1, 2, 3, 4, 5, 6, 7


class Diagnostics_(C.Structure):
    _fields_ = [
        ("non_zero", C.c_int64),
        ("dimensions", C.c_int64),
        ("iterations", C.c_int64),
        ("transposed", C.c_int32),
        ("lanczos_steps", C.c_int64),
        ("ritz_values_stabilized", C.c_int64),
        ("significant_values", C.c_int64),
        ("singular_values", C.c_int64),
        ("end_interval", C.c_double * 2),
        ("kappa", C.c_double),
        ("random_seed", C.c_uint32),
    ]


class Result_(C.Structure):
    _fields_ = [
        ("d", C.c_int64),
        ("u_rows", C.c_int64),
        ("u_cols", C.c_int64),
        ("vt_rows", C.c_int64),
        ("vt_cols", C.c_int64),
        ("u", C.POINTER(C.c_double)),
        ("s", C.POINTER(C.c_double)),
        ("vt", C.POINTER(C.c_double)),
        ("diagnostics", Diagnostics_),
        ("power_iters_run", C.c_int64),
        ("residual_estimate", C.c_double),
        ("qr_fallbacks", C.c_int64),
        ("u32", C.POINTER(C.c_float)),
        ("s32", C.POINTER(C.c_float)),
        ("vt32", C.POINTER(C.c_float)),
    ]


class BackwardInfo_(C.Structure):
    _fields_ = [
        ("loss", C.c_double),
        ("cg_iterations", C.c_int64),
        ("cg_relative_residual", C.c_double),
    ]


LOSS_SIGMA_SUM, LOSS_PROJECTION = 0, 1

_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)

# (name, restype, argtypes) for every entry point declared in include/svdh.h
SIGNATURES = [
    ("svdh_ctx_create", _i32, [_i32, C.POINTER(_vp)]),
    ("svdh_ctx_destroy", _i32, [_vp]),
    ("svdh_ctx_synchronize", _i32, [_vp]),
    ("svdh_last_error", C.c_char_p, []),
    ("svdh_build_id", C.c_char_p, []),
    ("svdh_ctx_profile", _i32, [_vp, _i32]),
    ("svdh_ctx_kernel_time", _i32, [_vp, _i32, _f64p, _i64p, _f64p]),
    ("svdh_csr_create", _i32, [_vp, _i64, _i64, _i64, _i64p, _i64p, _vp, _i32, C.c_uint32, C.POINTER(_vp)]),
    ("svdh_mask_create", _i32, [_vp, _vp, C.POINTER(C.c_uint8), C.POINTER(_vp)]),
    ("svdh_coo_create", _i32, [_vp, _i64, _i64, _i64, _i64p, _i64p, _vp, _i32, C.c_uint32, C.POINTER(_vp)]),
    ("svdh_csc_create", _i32, [_vp, _i64, _i64, _i64, _i64p, _i64p, _vp, _i32, C.c_uint32, C.POINTER(_vp)]),
    ("svdh_op_destroy", _i32, [_vp]),
    ("svdh_op_shape", _i32, [_vp, _i64p, _i64p, _i64p]),
    ("svdh_spmv", _i32, [_vp, _vp, _i32, _vp, _vp, _i32, _i32]),
    ("svdh_spmv_noalloc", _i32, [_vp, _vp, _i32, _vp, _vp, _i32]),
    ("svdh_spmm", _i32, [_vp, _vp, _i32, _vp, _i64, _i64, _vp, _i64, _i32]),
    ("svdh_spmm_qta", _i32, [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _i32]),
    ("svdh_spmm_centered", _i32, [_vp, _vp, _i32, _vp, _i64, _i64, _vp, _i64, _vp, _i32]),
    ("svdh_spmm_qta_centered", _i32, [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i32]),
    ("svdh_column_means", _i32, [_vp, _vp, _vp, _i32]),
    ("svdh_las2", _i32, [_vp, _vp, _i64, _i64, C.c_double, C.c_double, C.c_double, C.c_uint32, _f64p, _i32,
                         C.POINTER(Result_)]),
    ("svdh_las2_legacy", _i32, [_vp, _vp, _i64, _i64, C.c_double, C.c_double, C.c_double, C.c_uint32, _f64p,
                                C.POINTER(Result_)]),
    ("svdh_las2_device", _i32, [_vp, _vp, _i64, _i64, C.c_double, C.c_double, C.c_double, C.c_uint32, _f64p, _i32,
                                _vp, _i64, _f64p, _vp, _i64, C.POINTER(Result_)]),
    ("svdh_las2_start_vector", _i32, [C.c_uint32, _i64, _f64p]),
    ("svdh_rsvd", _i32, [_vp, _vp, _i64, _i64, _i64, _i32, _i32, C.c_uint64, _i32, _f64p, _i32, C.c_double, _i64,
                         C.POINTER(Result_)]),
    ("svdh_rsvd_device", _i32, [_vp, _vp, _i64, _i64, _i64, _i32, _i32, C.c_uint64, _i32, _f64p, _i32, C.c_double,
                                _i64, _vp, _i64, _f64p, _vp, _i64, C.POINTER(Result_)]),
    ("svdh_rsvd_batched", _i32, [_vp, _i64, C.POINTER(_vp), _i64, _i64, _i64, _i32, C.c_uint64, _i32,
                                 C.POINTER(_f64p), _i32, C.POINTER(Result_)]),
    ("svdh_rsvd_batched_device", _i32, [_vp, _i64, C.POINTER(_vp), _i64, _i64, _i64, _i32, C.c_uint64, _i32,
                                        C.POINTER(_f64p), _i32, C.POINTER(_vp), _i64p, C.POINTER(_vp), _i64p,
                                        C.POINTER(Result_)]),
    ("svdh_backward", _i32, [_vp, _vp, C.POINTER(Result_), _i32, _i64, _f64p, _i32, C.c_double, _i64, _vp, _i32,
                             C.POINTER(BackwardInfo_)]),
    ("svdh_result_free", None, [C.POINTER(Result_)]),
]

_lib = None
_lock = threading.Lock()
_contexts = {}


def lib():
    """The loaded libsvdh.so (loaded on first use; raises if it is missing)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(f"{LIB_PATH} is missing: build it with `python single-svdlib_amd/build.py`")
                h = C.CDLL(LIB_PATH)
                for name, res, args in SIGNATURES:
                    f = getattr(h, name)
                    f.restype = res
                    f.argtypes = args
                built = h.svdh_build_id().decode()
                want = srchash.source_hash()
                if want is not None and built != want:
                    raise ImportError(f"{LIB_PATH} was built from other sources (build id {built[:16]}, sources "
                                      f"{want[:16]}): rebuild with `python single-svdlib_amd/build.py`")
                _lib = h
    return _lib


def check(status):
    if status != 0:
        msg = lib().svdh_last_error().decode(errors="replace")
        raise error.from_status(status, msg)


def context(device=0):
    """The process-wide context of a device (one HIP stream)."""
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        h = _vp()
        check(lib().svdh_ctx_create(device, C.byref(h)))
        with _lock:
            _contexts.setdefault(device, h)
            ctx = _contexts[device]
    return ctx


def ptr(a, t=_f64p):
    return a.ctypes.data_as(t)


_hip = None


def hip():
    """The HIP runtime libsvdh.so is linked against (libamdhip64.so.7 of the
    ROCm install). A second HIP runtime in the process (PyTorch ships its own)
    does not see the GPU once this one holds it, so device buffers for the
    device-output entry points come from here."""
    global _hip
    if _hip is None:
        lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(_vp), C.c_size_t]
        h.hipFree.argtypes = [_vp]
        h.hipMemcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [_vp, C.c_int, C.c_size_t]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipMemset):
            f.restype = C.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A C-contiguous FP64/FP32 array in device memory (hipMalloc)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(d) for d in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = max(int(np.prod(self.shape)) * self.dtype.itemsize, 16)
        p = _vp()
        st = hip().hipMalloc(C.byref(p), self.nbytes)
        if st != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed with status {st}")
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(self.shape)
        if hip().hipMemcpy(self.ptr, a.ctypes.data_as(_vp), a.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy host->device failed")
        return self

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if hip().hipMemcpy(out.ctypes.data_as(_vp), self.ptr, out.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy device->host failed")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
