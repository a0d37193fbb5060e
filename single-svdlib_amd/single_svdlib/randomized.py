"""randomized::{randomized_svd, PowerIterationNormalizer, svd_flip}
(src/randomized/mod.rs:14-438) on the GPU randomized driver of libsvdh.so.
"""
import ctypes as C
import enum

import numpy as np

from . import _ffi
from .error import Las2Error
from .utils import as_operator, result_to_svdrec


class PowerIterationNormalizer(enum.Enum):
    """randomized/mod.rs:14-19 (None is spelled NONE in Python)."""
    QR = _ffi.NORM_QR
    LU = _ffi.NORM_LU
    NONE = _ffi.NORM_NONE


def randomized_svd(m, target_rank, n_oversamples, n_power_iters, power_iteration_normalizer=PowerIterationNormalizer.QR,
                   mean_center=False, seed=None, verbose=False, *, precision=None, tol=0.0, max_power_iters=None,
                   omega=None):
    """randomized/mod.rs:24-199. Returns SvdRec with u m x d, s, vt d x n.

    Extensions (keyword-only): precision "f64" | "f32" | "bf16" (default: the
    operator's dtype); tol > 0 continues power iterations past n_power_iters
    (up to max_power_iters) until the Ritz residual estimate reaches tol;
    omega injects the n x (rank + oversample) Gaussian test matrix.
    """
    op = as_operator(m)
    if precision is None:
        prec = _ffi.PREC_F64 if op.dtype == np.float64 else _ffi.PREC_F32
    else:
        prec = {"f64": _ffi.PREC_F64, "f32": _ffi.PREC_F32, "bf16": _ffi.PREC_BF16}[precision]
    norm = power_iteration_normalizer
    if isinstance(norm, str):
        norm = PowerIterationNormalizer[norm.upper()]
    om = None
    if omega is not None:
        om = np.asfortranarray(omega, dtype=np.float64)
    res = _ffi.Result_()
    mx = int(max_power_iters) if max_power_iters is not None else max(int(n_power_iters), 1000 if tol > 0 else 0)
    _ffi.check(_ffi.lib().svdh_rsvd(op._ctx, op._handle, int(target_rank), int(n_oversamples), int(n_power_iters),
                                    norm.value, int(bool(mean_center)), int(seed or 0) & 0xFFFFFFFFFFFFFFFF,
                                    int(seed is not None), _ffi.ptr(om) if om is not None else None, prec,
                                    float(tol), mx, C.byref(res)))
    if verbose:
        print(f"SVD | Randomized | power iterations run: {res.power_iters_run}")
    return result_to_svdrec(res, op.dtype)


def randomized_svd_batched(matrices, target_rank, n_oversamples, n_power_iters,
                           power_iteration_normalizer=PowerIterationNormalizer.QR, seed=None, *, precision=None,
                           omegas=None):
    """Many randomized_svd solves over column masks of one matrix (a list of
    MaskedCSRMatrix views of the same CsrMatrix), every SpMM round of all of
    them in one launch. Returns one SvdRec per matrix, equal to what
    randomized_svd(matrix, ..., mean_center=False, seed, omega=omegas[b])
    returns for it. omegas (optional) holds one ncols_b x (rank + oversample)
    test matrix per matrix (None entries draw from seed)."""
    ops = [as_operator(m) for m in matrices]
    if not ops:
        return []
    norm = power_iteration_normalizer
    if isinstance(norm, str):
        norm = PowerIterationNormalizer[norm.upper()]
    dtype = ops[0].dtype
    if precision is None:
        prec = _ffi.PREC_F64 if dtype == np.float64 else _ffi.PREC_F32
    else:
        prec = {"f64": _ffi.PREC_F64, "f32": _ffi.PREC_F32}[precision]
    handles = (C.c_void_p * len(ops))(*[o._handle for o in ops])
    outs = (_ffi.Result_ * len(ops))()
    om_ptrs = None
    keep = []
    if omegas is not None:
        if len(omegas) != len(ops):
            raise ValueError("one omega per matrix")
        om_ptrs = (_ffi._f64p * len(ops))()
        for i, om in enumerate(omegas):
            if om is not None:
                a = np.asfortranarray(om, dtype=np.float64)
                keep.append(a)
                om_ptrs[i] = _ffi.ptr(a)
    _ffi.check(_ffi.lib().svdh_rsvd_batched(ops[0]._ctx, len(ops), handles, int(target_rank), int(n_oversamples),
                                            int(n_power_iters), norm.value, int(seed or 0) & 0xFFFFFFFFFFFFFFFF,
                                            int(seed is not None), om_ptrs, prec, outs))
    return [result_to_svdrec(outs[i], dtype) for i in range(len(ops))]


def svd_flip(u=None, v=None, u_based_decision=True):
    """randomized/mod.rs:332-438, in place on numpy arrays (u m x k columns,
    v k x n rows)."""
    if u is None and v is None:
        raise Las2Error("svdlibrs/svdLas2: Both u and v cannot be None")
    if u_based_decision:
        if u is None:
            raise Las2Error("svdlibrs/svdLas2: u cannot be None when u_based_decision is true")
        idx = np.argmax(np.abs(u), axis=0)
        signs = np.where(u[idx, np.arange(u.shape[1])] < 0, -1.0, 1.0).astype(u.dtype)
        u *= signs[None, :]
        if v is not None:
            k = min(v.shape[0], signs.size)
            v[:k] *= signs[:k, None]
    else:
        if v is None:
            raise Las2Error("svdlibrs/svdLas2: v cannot be None when u_based_decision is false")
        idx = np.argmax(np.abs(v), axis=1)
        signs = np.where(v[np.arange(v.shape[0]), idx] < 0, -1.0, 1.0).astype(v.dtype)
        v *= signs[:, None]
        if u is not None:
            k = min(u.shape[1], signs.size)
            u[:, :k] *= signs[None, :k]
