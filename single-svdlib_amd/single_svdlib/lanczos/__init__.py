"""lanczos::{svd, svd_dim, svd_dim_seed, svd_las2} (src/lanczos/mod.rs:27-189)
on the GPU LAS2 driver of libsvdh.so.

T is the operator's value dtype: float64 runs the FP64 path, float32 the
reference's svd_las2::<f32> control flow (f32 epsilons) with FP32 sparse
values and FP64 vectors and recurrence.
"""
import ctypes as C

import numpy as np

from .. import _ffi
from ..utils import as_operator, result_to_svdrec
from . import masked  # noqa: F401


def svd(a):
    """mod.rs:27-36: svd_las2(A, 0, 0, [-1e-30, 1e-30], 1e-6, 0)"""
    return svd_las2(a, 0, 0, [-1.0e-30, 1.0e-30], 1.0e-6, 0)


def svd_dim(a, dimensions):
    """mod.rs:45-55"""
    return svd_las2(a, dimensions, 0, [-1.0e-30, 1.0e-30], 1.0e-6, 0)


def svd_dim_seed(a, dimensions, random_seed):
    """mod.rs:65-86"""
    return svd_las2(a, dimensions, 0, [-1.0e-30, 1.0e-30], 1.0e-6, random_seed)


def svd_las2(a, dimensions, iterations, end_interval, kappa, random_seed, start=None):
    """mod.rs:104-189. start (extension): a start vector of the working
    dimension replacing the random one (the reference's own behaviour for a
    non-zero w0, mod.rs:595-606)."""
    op = as_operator(a)
    prec = _ffi.PREC_F64 if op.dtype == np.float64 else _ffi.PREC_F32
    res = _ffi.Result_()
    sv = None
    if start is not None:
        sv = np.ascontiguousarray(start, dtype=np.float64)
    _ffi.check(_ffi.lib().svdh_las2(op._ctx, op._handle, int(dimensions), int(iterations), float(end_interval[0]),
                                    float(end_interval[1]), float(kappa), int(random_seed) & 0xFFFFFFFF,
                                    _ffi.ptr(sv) if sv is not None else None, prec, C.byref(res)))
    return result_to_svdrec(res, op.dtype)
