"""legacy::{svd, svd_dim, svd_dim_seed, svdLAS2} (src/legacy.rs:438-663): the
f64 port of SVDLIBC's LAS2 that the reference keeps beside lanczos, on the
GPU LAS2 driver of libsvdh.so in its legacy mode.

Operators are the FP64 device operators (CsrMatrix, CscMatrix, CooMatrix of
float64, or a scipy.sparse matrix, converted to float64); the legacy SMat
trait is f64 only (legacy.rs:438-443). Results are SvdRec with `ut` (d x m,
rows are the left vectors), `s`, `vt` and Diagnostics (legacy.rs:454-489);
errors are the SvdLibError variants (legacy/error.rs).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from .utils import CsrMatrix, Diagnostics, _DeviceOp, result_to_svdrec


@dataclass
class SvdRec:
    """legacy.rs:454-460"""
    d: int
    ut: np.ndarray
    s: np.ndarray
    vt: np.ndarray
    diagnostics: Diagnostics

    def recompose(self):
        """legacy.rs:2022-2027: ut^T . diag(s) . vt"""
        return self.ut.T @ np.diag(self.s) @ self.vt


def svd(a):
    """legacy.rs:498-500: svdLAS2(A, 0, 0, [-1e-30, 1e-30], 1e-6, 0)"""
    return svdLAS2(a, 0, 0, [-1.0e-30, 1.0e-30], 1.0e-6, 0)


def svd_dim(a, dimensions):
    """legacy.rs:510-512"""
    return svdLAS2(a, dimensions, 0, [-1.0e-30, 1.0e-30], 1.0e-6, 0)


def svd_dim_seed(a, dimensions, random_seed):
    """legacy.rs:523-525"""
    return svdLAS2(a, dimensions, 0, [-1.0e-30, 1.0e-30], 1.0e-6, random_seed)


def _as_f64_operator(a):
    if isinstance(a, _DeviceOp):
        if a.dtype != np.float64:
            raise TypeError("legacy::SMat is f64 only: build the operator from float64 values")
        return a
    import scipy.sparse as sp
    if sp.issparse(a):
        return CsrMatrix.from_scipy(a.astype(np.float64))
    raise TypeError(f"expected a single_svdlib operator or scipy.sparse matrix, got {type(a).__name__}")


def svdLAS2(a, dimensions, iterations, end_interval, kappa, random_seed, start=None):  # noqa: N802
    """legacy.rs:569-663. random_seed 0 draws one from the OS. start
    (extension): a start vector of the working dimension replacing the
    random one, as in lanczos.svd_las2."""
    op = _as_f64_operator(a)
    res = _ffi.Result_()
    sv = None
    if start is not None:
        sv = np.ascontiguousarray(start, dtype=np.float64)
    _ffi.check(_ffi.lib().svdh_las2_legacy(op._ctx, op._handle, int(dimensions), int(iterations),
                                           float(end_interval[0]), float(end_interval[1]), float(kappa),
                                           int(random_seed) & 0xFFFFFFFF, _ffi.ptr(sv) if sv is not None else None,
                                           C.byref(res)))
    rec = result_to_svdrec(res, np.float64)
    return SvdRec(d=rec.d, ut=rec.u, s=rec.s, vt=rec.vt, diagnostics=rec.diagnostics)
