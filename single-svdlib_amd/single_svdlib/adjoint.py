"""Adjoint of the truncated SVD with respect to the CSR values (backward()).

No counterpart in the reference; the C entry point is svdh_backward
(include/svdh.h). Losses:
  "sigma_sum":  L = sum_{i<k} s_i            -> grad[e] = sum_i u_i[r_e] v_i[c_e]
  "projection": L = |U_k^T y|^2              -> complement correction through k
                shifted systems (s_i^2 I - A A^T) w_i = (I - U_k U_k^T) y
                solved by conjugate gradients on the GPU, then a two-pair SDDMM.
"""
import ctypes as C

import numpy as np

from . import _ffi
from .utils import as_operator

_LOSSES = {"sigma_sum": _ffi.LOSS_SIGMA_SUM, "projection": _ffi.LOSS_PROJECTION}


def backward(matrix, rec, loss="sigma_sum", k=None, y=None, precision=None, tol=0.0, max_iters=0,
             grad_dtype=np.float64):
    """Returns (grad, info): grad has nnz entries in the CSR order of the
    underlying matrix; info = {"loss", "cg_iterations", "cg_relative_residual"}."""
    op = as_operator(matrix)
    kind = _LOSSES[loss] if isinstance(loss, str) else int(loss)
    k = int(rec.d if k is None else k)
    if precision is None:
        prec = _ffi.PREC_F64 if op.dtype == np.float64 else _ffi.PREC_F32
    else:
        prec = {"f64": _ffi.PREC_F64, "f32": _ffi.PREC_F32, "bf16": _ffi.PREC_BF16}[precision]
    # single-precision factors go over as they are (the library reads either)
    single = all(np.asarray(a).dtype == np.float32 for a in (rec.u, rec.s, rec.vt))
    fdt = np.float32 if single else np.float64
    u = np.ascontiguousarray(rec.u, dtype=fdt)
    s = np.ascontiguousarray(rec.s, dtype=fdt)
    vt = np.ascontiguousarray(rec.vt, dtype=fdt)
    res = _ffi.Result_()
    res.d = int(rec.d)
    res.u_rows, res.u_cols = u.shape
    res.vt_rows, res.vt_cols = vt.shape
    if single:
        res.u32 = u.ctypes.data_as(C.POINTER(C.c_float))
        res.s32 = s.ctypes.data_as(C.POINTER(C.c_float))
        res.vt32 = vt.ctypes.data_as(C.POINTER(C.c_float))
    else:
        res.u = _ffi.ptr(u)
        res.s = _ffi.ptr(s)
        res.vt = _ffi.ptr(vt)
    yv = None
    if y is not None:
        yv = np.ascontiguousarray(y, dtype=np.float64)
    gdt = np.dtype(grad_dtype)
    base = getattr(op, "matrix", None)  # a masked view's gradient covers the whole underlying pattern
    nnz = (base if base is not None else op)._shape()[2]
    grad = np.zeros(nnz, dtype=gdt)
    info = _ffi.BackwardInfo_()
    _ffi.check(_ffi.lib().svdh_backward(op._ctx, op._handle, C.byref(res), kind, k,
                                        _ffi.ptr(yv) if yv is not None else None, prec, float(tol),
                                        int(max_iters), grad.ctypes.data_as(C.c_void_p),
                                        _ffi.SVDH_F64 if gdt == np.float64 else _ffi.SVDH_F32, C.byref(info)))
    return grad, {"loss": info.loss, "cg_iterations": info.cg_iterations,
                  "cg_relative_residual": info.cg_relative_residual}
