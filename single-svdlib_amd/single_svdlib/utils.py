"""Boundary types: SMat, SvdRec, Diagnostics, SvdFloat, determine_chunk_size
(src/utils.rs:10-113), and CsrMatrix, the device-resident operator that
replaces nalgebra_sparse::CsrMatrix as an SMat (src/lanczos/mod.rs:1506-1655).
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Protocol

import numpy as np

from . import _ffi


def determine_chunk_size(nrows, num_threads=None):
    """utils.rs:10-20: max(64, ceil(nrows / (4 * threads)))."""
    t = num_threads or os.cpu_count() or 1
    return max(64, -(-nrows // (4 * t)))


class SvdFloat:
    """utils.rs:81-113 constants for f32 / f64."""

    @staticmethod
    def eps(dtype):
        return float(np.finfo(dtype).eps)

    @staticmethod
    def eps34(dtype):
        return float(np.finfo(dtype).eps) ** 0.75

    @staticmethod
    def compare(a, b, dtype=np.float64):
        return abs(b - a) < float(np.finfo(dtype).eps)


@dataclass
class Diagnostics:
    """utils.rs:66-79"""
    non_zero: int
    dimensions: int
    iterations: int
    transposed: bool
    lanczos_steps: int
    ritz_values_stabilized: int
    significant_values: int
    singular_values: int
    end_interval: list
    kappa: float
    random_seed: int


@dataclass
class SvdRec:
    """utils.rs:43-50. LAS2 returns u as d x m (rows are left vectors);
    randomized_svd returns u as m x d (columns are left vectors), as the
    reference does (src/lanczos/mod.rs:172, src/randomized/mod.rs:175-185)."""
    d: int
    u: np.ndarray
    s: np.ndarray
    vt: np.ndarray
    diagnostics: Diagnostics
    extra: dict = field(default_factory=dict)

    @property
    def ut(self):
        """README/north-star name of the left-vector field (utils.rs:46 calls it u)."""
        return self.u

    def recompose(self):
        """src/lanczos/mod.rs:1408-1413: u . diag(s) . vt (valid for the m x d layout)."""
        return self.u @ np.diag(self.s) @ self.vt

    def backward(self, matrix, loss="sigma_sum", k=None, y=None, **kw):
        """Gradient of a loss of these triplets with respect to the CSR values
        of `matrix` (the operator they were computed from); see
        single_svdlib.adjoint.backward."""
        from .adjoint import backward
        return backward(matrix, self, loss=loss, k=k, y=y, **kw)


class SMat(Protocol):
    """The reference's plugin trait (src/utils.rs:22-33)."""

    def nrows(self) -> int: ...
    def ncols(self) -> int: ...
    def nnz(self) -> int: ...
    def svd_opa(self, x, y, transposed: bool) -> None: ...
    def compute_column_means(self): ...
    def multiply_with_dense(self, dense, result, transpose_self: bool) -> None: ...
    def multiply_with_dense_centered(self, dense, result, transpose_self: bool, means) -> None: ...
    def multiply_transposed_by_dense(self, q, result) -> None: ...
    def multiply_transposed_by_dense_centered(self, q, result, means) -> None: ...


def result_to_svdrec(res, dtype, extra=None):
    """svdh_result -> SvdRec with arrays of dtype; the host arrays are FP64
    (u, s, vt) or, for FP32 / BF16 solves, single precision (u32, s32, vt32)."""
    d = res.d
    single = bool(res.u32) or bool(res.s32)
    pu, ps, pvt = (res.u32, res.s32, res.vt32) if single else (res.u, res.s, res.vt)
    nu, nvt = res.u_rows * res.u_cols, res.vt_rows * res.vt_cols
    u = np.ctypeslib.as_array(pu, (max(nu, 1),))[:nu] if pu else np.zeros(nu)
    vt = np.ctypeslib.as_array(pvt, (max(nvt, 1),))[:nvt] if pvt else np.zeros(nvt)
    s = np.ctypeslib.as_array(ps, (max(d, 1),))[:d] if ps else np.zeros(d)
    dg = res.diagnostics
    diag = Diagnostics(
        non_zero=dg.non_zero, dimensions=dg.dimensions, iterations=dg.iterations, transposed=bool(dg.transposed),
        lanczos_steps=dg.lanczos_steps, ritz_values_stabilized=dg.ritz_values_stabilized,
        significant_values=dg.significant_values, singular_values=dg.singular_values,
        end_interval=[dtype(dg.end_interval[0]), dtype(dg.end_interval[1])], kappa=dtype(dg.kappa),
        random_seed=int(dg.random_seed))
    rec = SvdRec(d=d, u=u.reshape(res.u_rows, res.u_cols).astype(dtype),
                 s=s.astype(dtype), vt=vt.reshape(res.vt_rows, res.vt_cols).astype(dtype), diagnostics=diag,
                 extra=dict(extra or {}))
    rec.extra.update(power_iters_run=res.power_iters_run, residual_estimate=res.residual_estimate,
                     qr_fallbacks=res.qr_fallbacks)
    _ffi.lib().svdh_result_free(C.byref(res))
    return rec


class _DeviceOp:
    """SMat methods over a device operator handle."""

    _handle = None
    dtype = np.float64

    def nrows(self):
        return self._shape()[0]

    def ncols(self):
        return self._shape()[1]

    def nnz(self):
        return self._shape()[2]

    def _shape(self):
        r, c, z = C.c_int64(), C.c_int64(), C.c_int64()
        _ffi.check(_ffi.lib().svdh_op_shape(self._handle, C.byref(r), C.byref(c), C.byref(z)))
        return r.value, c.value, z.value

    def _vec_dtype(self):
        return (_ffi.SVDH_F64, np.float64) if self.dtype == np.float64 else (_ffi.SVDH_F32, np.float32)

    def svd_opa(self, x, y, transposed):
        """y = A x (transposed False) or A^T x; y is overwritten (mod.rs:1520-1603)."""
        nr, nc, _ = self._shape()
        nin, nout = (nr, nc) if transposed else (nc, nr)
        if len(x) != nin:
            raise ValueError(f"svd_opa: x must be A.ncols() in length, x = {len(x)}, A.ncols = {nin}")
        if len(y) != nout:
            raise ValueError(f"svd_opa: y must be A.nrows() in length, y = {len(y)}, A.nrows = {nout}")
        code, dt = self._vec_dtype()
        xv = np.ascontiguousarray(x, dtype=dt)
        yv = np.empty(nout, dtype=dt)
        _ffi.check(_ffi.lib().svdh_spmv(self._ctx, self._handle, int(bool(transposed)), xv.ctypes.data,
                                        yv.ctypes.data, code, 0))
        y[:] = yv

    def multiply_with_dense(self, dense, result, transpose_self):
        """result = A dense, or A^T dense (masked.rs:265-471); 2-D arrays."""
        nr, nc, _ = self._shape()
        rows_out, rows_in = (nc, nr) if transpose_self else (nr, nc)
        if dense.shape[0] != rows_in:
            raise ValueError("Dense matrix has incompatible row count")
        if result.shape[0] != rows_out:
            raise ValueError("Result matrix has incompatible row count")
        if result.shape[1] != dense.shape[1]:
            raise ValueError("Result matrix has incompatible column count")
        code, dt = self._vec_dtype()
        x = np.asfortranarray(dense, dtype=dt)
        l = x.shape[1]
        y = np.zeros((rows_out, l), dtype=dt, order="F")
        _ffi.check(_ffi.lib().svdh_spmm(self._ctx, self._handle, int(bool(transpose_self)), x.ctypes.data,
                                        max(rows_in, 1), l, y.ctypes.data, max(rows_out, 1), code))
        result[...] = y

    def multiply_with_dense_centered(self, dense, result, transpose_self, means):
        """result = (A - 1 means^T) dense, or its transpose times dense
        (utils.rs:29, masked.rs:473-614; the forward product in its exact
        form A X - 1 (means^T X))."""
        nr, nc, _ = self._shape()
        rows_out, rows_in = (nc, nr) if transpose_self else (nr, nc)
        if dense.shape[0] != rows_in:
            raise ValueError("Dense matrix has incompatible row count")
        if result.shape[0] != rows_out:
            raise ValueError("Result matrix has incompatible row count")
        if result.shape[1] != dense.shape[1]:
            raise ValueError("Result matrix has incompatible column count")
        if len(means) != nc:
            raise ValueError(f"means must have A.ncols() = {nc} entries, got {len(means)}")
        code, dt = self._vec_dtype()
        x = np.asfortranarray(dense, dtype=dt)
        mu = np.ascontiguousarray(means, dtype=dt)
        l = x.shape[1]
        y = np.zeros((rows_out, l), dtype=dt, order="F")
        _ffi.check(_ffi.lib().svdh_spmm_centered(self._ctx, self._handle, int(bool(transpose_self)), x.ctypes.data,
                                                 max(rows_in, 1), l, y.ctypes.data, max(rows_out, 1), mu.ctypes.data,
                                                 code))
        result[...] = y

    def multiply_transposed_by_dense_centered(self, q, result, means):
        """result (l x ncols) = q^T (A - 1 means^T) (utils.rs:32, masked.rs:735-874)."""
        nr, nc, _ = self._shape()
        if q.shape[0] != nr:
            raise ValueError(f"Q matrix has incompatible row count: expected {nr}, got {q.shape[0]}")
        if result.shape != (q.shape[1], nc):
            raise ValueError("Result matrix has incompatible shape")
        if len(means) != nc:
            raise ValueError(f"means must have A.ncols() = {nc} entries, got {len(means)}")
        code, dt = self._vec_dtype()
        x = np.asfortranarray(q, dtype=dt)
        mu = np.ascontiguousarray(means, dtype=dt)
        l = x.shape[1]
        b = np.zeros((l, nc), dtype=dt, order="F")
        _ffi.check(_ffi.lib().svdh_spmm_qta_centered(self._ctx, self._handle, x.ctypes.data, max(nr, 1), l,
                                                     b.ctypes.data, max(l, 1), mu.ctypes.data, code))
        result[...] = b

    def multiply_transposed_by_dense(self, q, result):
        """result (l x ncols) = q^T A (masked.rs:616-733)."""
        nr, nc, _ = self._shape()
        if q.shape[0] != nr:
            raise ValueError(f"Q matrix has incompatible row count: expected {nr}, got {q.shape[0]}")
        if result.shape != (q.shape[1], nc):
            raise ValueError("Result matrix has incompatible shape")
        code, dt = self._vec_dtype()
        x = np.asfortranarray(q, dtype=dt)
        l = x.shape[1]
        b = np.zeros((l, nc), dtype=dt, order="F")
        _ffi.check(_ffi.lib().svdh_spmm_qta(self._ctx, self._handle, x.ctypes.data, max(nr, 1), l, b.ctypes.data,
                                            max(l, 1), code))
        result[...] = b

    def compute_column_means(self):
        code, dt = self._vec_dtype()
        out = np.empty(self.ncols(), dtype=dt)
        _ffi.check(_ffi.lib().svdh_column_means(self._ctx, self._handle, out.ctypes.data, code))
        return out


class CsrMatrix(_DeviceOp):
    """A CSR matrix uploaded to one GPU (the SMat operand of every solver).

    CsrMatrix(indptr, indices, data, shape, device=0, keep_bf16=False)
    CsrMatrix.from_scipy(matrix) accepts any scipy.sparse matrix (COO and CSC
    are converted to CSR on the host; duplicate entries are summed).
    data may be float64, float32 (the reference's T) or, with
    storage="bf16", kept as BF16 on the device for the mixed-precision solver.
    """

    def __init__(self, indptr, indices, data, shape, device=0, keep_f32=False, keep_bf16=False):
        self.device = device
        self._ctx = _ffi.context(device)
        data = np.asarray(data)
        if data.dtype == np.float64:
            code = _ffi.SVDH_F64
        elif data.dtype == np.float32:
            code = _ffi.SVDH_F32
        else:
            raise ValueError(f"values must be float64 or float32, got {data.dtype}")
        self.dtype = data.dtype.type
        rp = np.ascontiguousarray(indptr, dtype=np.int64)
        ci = np.ascontiguousarray(indices, dtype=np.int64)
        dv = np.ascontiguousarray(data)
        nrows, ncols = int(shape[0]), int(shape[1])
        if rp.size != nrows + 1:
            raise ValueError("indptr must have nrows + 1 entries")
        flags = (_ffi.KEEP_F32 if keep_f32 else 0) | (_ffi.KEEP_BF16 if keep_bf16 else 0)
        h = C.c_void_p()
        _ffi.check(_ffi.lib().svdh_csr_create(self._ctx, nrows, ncols, int(dv.size), rp.ctypes.data_as(_ffi._i64p),
                                              ci.ctypes.data_as(_ffi._i64p), dv.ctypes.data if dv.size else None,
                                              code, flags, C.byref(h)))
        self._handle = h

    @classmethod
    def from_scipy(cls, matrix, dtype=None, device=0, **kw):
        import scipy.sparse as sp
        a = sp.csr_matrix(matrix)
        if dtype is not None:
            a = a.astype(dtype)
        elif a.dtype not in (np.float64, np.float32):
            a = a.astype(np.float64)
        return cls(a.indptr, a.indices, a.data, a.shape, device=device, **kw)

    def __del__(self):
        h, self._handle = self._handle, None
        if h and _ffi._lib is not None:
            _ffi.lib().svdh_op_destroy(h)


def _value_code(data):
    data = np.asarray(data)
    if data.dtype == np.float64:
        return _ffi.SVDH_F64, np.ascontiguousarray(data)
    if data.dtype == np.float32:
        return _ffi.SVDH_F32, np.ascontiguousarray(data)
    raise ValueError(f"values must be float64 or float32, got {data.dtype}")


class CooMatrix(CsrMatrix):
    """A COO matrix (nalgebra_sparse::CooMatrix, the operand of SMat for
    CooMatrix, src/lanczos/mod.rs:1657-1740) converted to the device CSR
    operator on the GPU. Repeated (row, column) pairs are summed in input
    order (the reference's products add every triplet).

    CooMatrix(rows, cols, data, shape, device=0, keep_f32=False, keep_bf16=False)
    """

    def __init__(self, rows, cols, data, shape, device=0, keep_f32=False, keep_bf16=False):
        self.device = device
        self._ctx = _ffi.context(device)
        code, dv = _value_code(data)
        self.dtype = dv.dtype.type
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        if r.size != dv.size or c.size != dv.size:
            raise ValueError("rows, cols and data must have one entry per triplet")
        flags = (_ffi.KEEP_F32 if keep_f32 else 0) | (_ffi.KEEP_BF16 if keep_bf16 else 0)
        h = C.c_void_p()
        _ffi.check(_ffi.lib().svdh_coo_create(self._ctx, int(shape[0]), int(shape[1]), int(dv.size),
                                              r.ctypes.data_as(_ffi._i64p), c.ctypes.data_as(_ffi._i64p),
                                              dv.ctypes.data if dv.size else None, code, flags, C.byref(h)))
        self._handle = h

    @classmethod
    def from_scipy(cls, matrix, dtype=None, device=0, **kw):
        """From any scipy.sparse matrix, its COO triplets as stored (a coo_matrix
        keeps repeated pairs)."""
        import scipy.sparse as sp
        a = matrix if getattr(matrix, "format", None) == "coo" else sp.coo_matrix(matrix)
        data = a.data.astype(dtype) if dtype is not None else a.data
        if data.dtype not in (np.float64, np.float32):
            data = data.astype(np.float64)
        return cls(a.row, a.col, data, a.shape, device=device, **kw)


class CscMatrix(CsrMatrix):
    """A CSC matrix (nalgebra_sparse::CscMatrix csc_data(), the operand of SMat
    for CscMatrix, src/lanczos/mod.rs:1415-1504) converted to the device CSR
    operator on the GPU.

    CscMatrix(indptr, indices, data, shape, device=0, keep_f32=False, keep_bf16=False)
    """

    def __init__(self, indptr, indices, data, shape, device=0, keep_f32=False, keep_bf16=False):
        self.device = device
        self._ctx = _ffi.context(device)
        code, dv = _value_code(data)
        self.dtype = dv.dtype.type
        cp = np.ascontiguousarray(indptr, dtype=np.int64)
        ri = np.ascontiguousarray(indices, dtype=np.int64)
        nrows, ncols = int(shape[0]), int(shape[1])
        if cp.size != ncols + 1:
            raise ValueError("indptr must have ncols + 1 entries")
        flags = (_ffi.KEEP_F32 if keep_f32 else 0) | (_ffi.KEEP_BF16 if keep_bf16 else 0)
        h = C.c_void_p()
        _ffi.check(_ffi.lib().svdh_csc_create(self._ctx, nrows, ncols, int(dv.size), cp.ctypes.data_as(_ffi._i64p),
                                              ri.ctypes.data_as(_ffi._i64p), dv.ctypes.data if dv.size else None,
                                              code, flags, C.byref(h)))
        self._handle = h

    @classmethod
    def from_scipy(cls, matrix, dtype=None, device=0, **kw):
        import scipy.sparse as sp
        a = sp.csc_matrix(matrix)
        data = a.data.astype(dtype) if dtype is not None else a.data
        if data.dtype not in (np.float64, np.float32):
            data = data.astype(np.float64)
        return cls(a.indptr, a.indices, data, a.shape, device=device, **kw)


def as_operator(a):
    """Accept a device operator or any scipy.sparse matrix."""
    if isinstance(a, _DeviceOp):
        return a
    import scipy.sparse as sp
    if sp.issparse(a):
        return CsrMatrix.from_scipy(a)
    raise TypeError(f"expected a single_svdlib operator or scipy.sparse matrix, got {type(a).__name__}")
