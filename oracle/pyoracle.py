"""ctypes binding of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F64_EPS = float(np.finfo(np.float64).eps)
F32_EPS = float(np.finfo(np.float32).eps)


def _load():
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `python oracle/build.py`")
    return C.CDLL(path)


_lib = _load()
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class _Op(C.Structure):
    _fields_ = [
        ("nrows", C.c_int64),
        ("ncols", C.c_int64),
        ("rowptr", _i64p),
        ("colidx", _i64p),
        ("vals", _f64p),
        ("n_masked", C.c_int64),
        ("m2o", _i64p),
        ("o2m", _i64p),
        ("nthreads", C.c_int),
    ]


class _Diag(C.Structure):
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


class _Rec(C.Structure):
    _fields_ = [
        ("d", C.c_int64),
        ("m", C.c_int64),
        ("n", C.c_int64),
        ("u", _f64p),
        ("s", _f64p),
        ("vt", _f64p),
        ("diag", _Diag),
    ]


_lib.orc_opa.argtypes = [C.POINTER(_Op), _f64p, _f64p, C.c_int]
_lib.orc_multiply_with_dense.argtypes = [C.POINTER(_Op), _f64p, C.c_int64, _f64p, C.c_int]
_lib.orc_multiply_transposed_by_dense.argtypes = [C.POINTER(_Op), _f64p, C.c_int64, _f64p]
_f32p = C.POINTER(C.c_float)
_lib.orc_multiply_with_dense_f32.argtypes = [C.POINTER(_Op), _f32p, _f32p, C.c_int64, _f32p, C.c_int]
_lib.orc_multiply_transposed_by_dense_f32.argtypes = [C.POINTER(_Op), _f32p, _f32p, C.c_int64, _f32p]
_lib.orc_compute_column_means.argtypes = [C.POINTER(_Op), _f64p]
_lib.orc_nnz.argtypes = [C.POINTER(_Op)]
_lib.orc_nnz.restype = C.c_int64
_lib.orc_pythag.argtypes = [C.c_double, C.c_double, C.c_double]
_lib.orc_pythag.restype = C.c_double
_lib.orc_imtqlb.argtypes = [C.c_int64, _f64p, _f64p, _f64p, C.c_int64, C.c_double]
_lib.orc_imtql2.argtypes = [C.c_int64, C.c_int64, _f64p, _f64p, _f64p, C.c_int64, C.c_double]
_lib.orc_start_vector.argtypes = [C.c_uint32, C.c_int64, _f64p]
_lib.orc_las2.argtypes = [C.POINTER(_Op), C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                          C.c_uint32, _f64p, C.c_double, C.POINTER(_Rec)]
_lib.orc_las2_legacy.argtypes = [C.POINTER(_Op), C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                                 C.c_uint32, _f64p, C.POINTER(_Rec)]
_lib.orc_imtqlb_legacy.argtypes = [C.c_int64, _f64p, _f64p, _f64p, C.c_double]
_lib.orc_imtql2_legacy.argtypes = [C.c_int64, C.c_int64, _f64p, _f64p, _f64p, C.c_double]
_lib.orc_svdrec_free.argtypes = [C.POINTER(_Rec)]
_lib.orc_last_error.restype = C.c_char_p
_lib.orc_determine_chunk_size.argtypes = [C.c_int64, C.c_int]
_lib.orc_determine_chunk_size.restype = C.c_int
_i64p = C.POINTER(C.c_int64)
_f32p = C.POINTER(C.c_float)
_lib.orc_coo_to_csr.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p]
_lib.orc_coo_to_csr.restype = C.c_int64
_lib.orc_coo_to_csr_f32.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64p, _i64p, _f32p, _i64p, _i64p, _f32p]
_lib.orc_coo_to_csr_f32.restype = C.c_int64


def _p(a, t=_f64p):
    return a.ctypes.data_as(t)


def determine_chunk_size(nrows, nthreads):
    return _lib.orc_determine_chunk_size(nrows, nthreads)


class OracleOp:
    """FP64 CSR (optionally column-masked) operator restating SMat.

    csr: scipy.sparse matrix (converted to CSR, values to float64).
    columns: optional kept columns (MaskedCSRMatrix::with_columns semantics).
    """

    def __init__(self, csr, columns=None, mask=None, nthreads=1):
        import scipy.sparse as sp
        a = sp.csr_matrix(csr)
        self.rowptr = np.ascontiguousarray(a.indptr, dtype=np.int64)
        self.colidx = np.ascontiguousarray(a.indices, dtype=np.int64)
        self.vals = np.ascontiguousarray(a.data, dtype=np.float64)
        self.shape0 = a.shape
        self._st = _Op()
        self._st.nrows, self._st.ncols = a.shape
        self._st.rowptr = _p(self.rowptr, _i64p)
        self._st.colidx = _p(self.colidx, _i64p) if self.colidx.size else None
        self._st.vals = _p(self.vals) if self.vals.size else None
        self._st.nthreads = nthreads
        if columns is not None and mask is None:
            mask = np.zeros(a.shape[1], dtype=bool)
            mask[np.asarray(columns, dtype=np.int64)] = True
        if mask is not None:
            mask = np.asarray(mask, dtype=bool)
            self.m2o = np.nonzero(mask)[0].astype(np.int64)
            self.o2m = np.full(a.shape[1], -1, dtype=np.int64)
            self.o2m[self.m2o] = np.arange(self.m2o.size, dtype=np.int64)
            self._st.n_masked = self.m2o.size
            self._st.m2o = _p(self.m2o, _i64p) if self.m2o.size else _p(np.zeros(1, np.int64), _i64p)
            self._st.o2m = _p(self.o2m, _i64p)
        else:
            self.m2o = None

    @property
    def nrows(self):
        return self._st.nrows

    @property
    def ncols(self):
        return self._st.n_masked if self.m2o is not None else self._st.ncols

    def nnz(self):
        return _lib.orc_nnz(C.byref(self._st))

    def opa(self, x, transposed=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n_out = self.ncols if transposed else self.nrows
        y = np.empty(n_out, dtype=np.float64)
        _lib.orc_opa(C.byref(self._st), _p(x), _p(y), int(bool(transposed)))
        return y

    def multiply_with_dense(self, dense, transpose_self=False):
        """dense, result: 2-D arrays (any order); computed column-major."""
        d = np.asfortranarray(dense, dtype=np.float64)
        l = d.shape[1]
        rows = self.ncols if transpose_self else self.nrows
        out = np.zeros((rows, l), dtype=np.float64, order="F")
        _lib.orc_multiply_with_dense(C.byref(self._st), _p(d), l, _p(out), int(bool(transpose_self)))
        return out

    def multiply_transposed_by_dense(self, q):
        q = np.asfortranarray(q, dtype=np.float64)
        l = q.shape[1]
        out = np.zeros((l, self.ncols), dtype=np.float64, order="F")
        _lib.orc_multiply_transposed_by_dense(C.byref(self._st), _p(q), l, _p(out))
        return out

    def _vals32(self):
        if getattr(self, "_v32", None) is None:
            self._v32 = np.ascontiguousarray(self.vals, dtype=np.float32)
        return self._v32

    def multiply_with_dense_f32(self, dense, transpose_self=False):
        """multiply_with_dense with T = f32: values rounded to f32, f32
        products and accumulators (the reference instantiated with f32)."""
        d = np.asfortranarray(dense, dtype=np.float32)
        l = d.shape[1]
        rows = self.ncols if transpose_self else self.nrows
        out = np.zeros((rows, l), dtype=np.float32, order="F")
        v = self._vals32()
        _lib.orc_multiply_with_dense_f32(C.byref(self._st), _p(v, _f32p), _p(d, _f32p), l, _p(out, _f32p),
                                         int(bool(transpose_self)))
        return out

    def multiply_transposed_by_dense_f32(self, q):
        q = np.asfortranarray(q, dtype=np.float32)
        l = q.shape[1]
        out = np.zeros((l, self.ncols), dtype=np.float32, order="F")
        v = self._vals32()
        _lib.orc_multiply_transposed_by_dense_f32(C.byref(self._st), _p(v, _f32p), _p(q, _f32p), l, _p(out, _f32p))
        return out

    def multiply_with_dense_centered(self, dense, means, transpose_self=False):
        """SMat::multiply_with_dense_centered (masked.rs:473-614) for the
        implicit A - 1 mu^T, mu = means (ncols). Forward: A X - 1 (mu^T X),
        the exact form; masked.rs:496-509 subtracts (sum_j mu_j)(sum_i X_ic)
        (defect D7). Transposed: A^T Y - mu (1^T Y) (masked.rs:576-588)."""
        d = np.asarray(dense, dtype=np.float64)
        mu = np.asarray(means, dtype=np.float64)
        out = self.multiply_with_dense(d, transpose_self)
        if not transpose_self:
            return out - np.ones((out.shape[0], 1)) @ (mu[None, :] @ d)
        return out - mu[:, None] @ d.sum(axis=0)[None, :]

    def multiply_transposed_by_dense_centered(self, q, means):
        """SMat::multiply_transposed_by_dense_centered (masked.rs:735-874):
        Q^T A - (Q^T 1) mu^T; the reference spreads the correction over its
        row chunks by their row fractions (masked.rs:839-848), which sums to
        the same term."""
        q = np.asarray(q, dtype=np.float64)
        mu = np.asarray(means, dtype=np.float64)
        return self.multiply_transposed_by_dense(q) - q.sum(axis=0)[:, None] @ mu[None, :]

    def column_means(self):
        out = np.empty(self.ncols, dtype=np.float64)
        _lib.orc_compute_column_means(C.byref(self._st), _p(out))
        return out


def pythag(a, b, eps=F64_EPS):
    return _lib.orc_pythag(a, b, eps)


def imtqlb(d, e, max_iter=100, eps=F64_EPS):
    """Eigenvalues (ascending) and bottom-row bounds of the tridiagonal with
    diagonal d and sub-diagonal e[1:] (e[0] ignored)."""
    d = np.array(d, dtype=np.float64)
    e = np.array(e, dtype=np.float64)
    bnd = np.zeros_like(d)
    st = _lib.orc_imtqlb(d.size, _p(d), _p(e), _p(bnd), max_iter, eps)
    if st:
        raise RuntimeError(_lib.orc_last_error().decode())
    return d, bnd


def imtql2(d, e, max_iter=100, eps=F64_EPS):
    """Eigenvalues (ascending) and eigenvectors (columns of z)."""
    d = np.array(d, dtype=np.float64)
    e = np.array(e, dtype=np.float64)
    n = d.size
    z = np.eye(n, dtype=np.float64)
    st = _lib.orc_imtql2(n, n, _p(d), _p(e), _p(z), max_iter, eps)
    if st:
        raise RuntimeError(_lib.orc_last_error().decode())
    return d, z


def imtqlb_legacy(d, e):
    """legacy.rs:961-1067: imtqlb with the absolute convergence test and at
    most 30 iterations per eigenvalue (RuntimeError after)."""
    d = np.array(d, dtype=np.float64)
    e = np.array(e, dtype=np.float64)
    bnd = np.zeros_like(d)
    st = _lib.orc_imtqlb_legacy(d.size, _p(d), _p(e), _p(bnd), F64_EPS)
    if st:
        raise RuntimeError(_lib.orc_last_error().decode())
    return d, bnd


def imtql2_legacy(d, e):
    """legacy.rs:1575-1692: imtql2 with the absolute convergence test and at
    most 30 iterations per eigenvalue (RuntimeError after)."""
    d = np.array(d, dtype=np.float64)
    e = np.array(e, dtype=np.float64)
    n = d.size
    z = np.eye(n, dtype=np.float64)
    st = _lib.orc_imtql2_legacy(n, n, _p(d), _p(e), _p(z), F64_EPS)
    if st:
        raise RuntimeError(_lib.orc_last_error().decode())
    return d, z


def start_vector(seed, n):
    out = np.empty(n, dtype=np.float64)
    _lib.orc_start_vector(seed, n, _p(out))
    return out


def las2(op, dimensions=0, iterations=0, end_interval=(-1e-30, 1e-30), kappa=1e-6, random_seed=0,
         start=None, eps=F64_EPS):
    """LAS2 restatement. Returns (status, dict) with d, u (d x m), s, vt (d x n), diagnostics."""
    rec = _Rec()
    sv = None
    if start is not None:
        sv = np.ascontiguousarray(start, dtype=np.float64)
    st = _lib.orc_las2(C.byref(op._st), dimensions, iterations, end_interval[0], end_interval[1], kappa,
                       random_seed, _p(sv) if sv is not None else None, eps, C.byref(rec))
    return _unpack(st, rec)


def las2_legacy(op, dimensions=0, iterations=0, end_interval=(-1e-30, 1e-30), kappa=1e-6, random_seed=0,
                start=None):
    """legacy::svdLAS2 restatement (legacy.rs:569-663): FP64 only, the legacy
    control flow. Returns (status, dict) as las2 does."""
    rec = _Rec()
    sv = None
    if start is not None:
        sv = np.ascontiguousarray(start, dtype=np.float64)
    st = _lib.orc_las2_legacy(C.byref(op._st), dimensions, iterations, end_interval[0], end_interval[1], kappa,
                              random_seed, _p(sv) if sv is not None else None, C.byref(rec))
    return _unpack(st, rec)


def _unpack(st, rec):
    if st:
        return st, {"error": _lib.orc_last_error().decode()}
    d, m, n = rec.d, rec.m, rec.n
    out = {
        "d": d,
        "s": np.ctypeslib.as_array(rec.s, (max(d, 1),))[:d].copy() if d else np.zeros(0),
        "u": np.ctypeslib.as_array(rec.u, (max(d * m, 1),))[: d * m].reshape(d, m).copy() if d else np.zeros((0, m)),
        "vt": np.ctypeslib.as_array(rec.vt, (max(d * n, 1),))[: d * n].reshape(d, n).copy() if d else np.zeros((0, n)),
        "diagnostics": {f: getattr(rec.diag, f) for f, _ in _Diag._fields_ if f != "end_interval"},
    }
    out["diagnostics"]["end_interval"] = list(rec.diag.end_interval)
    _lib.orc_svdrec_free(C.byref(rec))
    return 0, out


def coo_to_csr(m, n, rows, cols, vals):
    """COO triplets -> (rowptr, cols, vals) of the CSR with entries in (row,
    column) order and duplicates summed in input order, in the dtype of vals
    (float64 or float32). Raises ValueError on an index out of range."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    vals = np.ascontiguousarray(vals)
    nnz = rows.size
    rp = np.zeros(m + 1, dtype=np.int64)
    oc = np.zeros(max(nnz, 1), dtype=np.int64)
    if vals.dtype == np.float32:
        ov = np.zeros(max(nnz, 1), dtype=np.float32)
        u = _lib.orc_coo_to_csr_f32(m, n, nnz, _p(rows, _i64p), _p(cols, _i64p), _p(vals, _f32p), _p(rp, _i64p),
                                    _p(oc, _i64p), _p(ov, _f32p))
    else:
        vals = vals.astype(np.float64, copy=False)
        ov = np.zeros(max(nnz, 1), dtype=np.float64)
        u = _lib.orc_coo_to_csr(m, n, nnz, _p(rows, _i64p), _p(cols, _i64p), _p(vals), _p(rp, _i64p),
                                _p(oc, _i64p), _p(ov))
    if u < 0:
        raise ValueError("COO index out of range")
    return rp, oc[:u].copy(), ov[:u].copy()
