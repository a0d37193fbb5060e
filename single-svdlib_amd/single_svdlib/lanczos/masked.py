"""MaskedCSRMatrix (src/lanczos/masked.rs:13-875): a column-subset view of a
device CsrMatrix. The view borrows the matrix (it must outlive the view);
the remap tables are built once on the device and the masked nnz is counted
once (masked.rs:89-102 rescans on every call).
"""
import ctypes as C

import numpy as np

from .. import _ffi
from ..utils import CsrMatrix, _DeviceOp


class MaskedCSRMatrix(_DeviceOp):
    def __init__(self, matrix, column_mask):
        if not isinstance(matrix, CsrMatrix):
            raise TypeError("MaskedCSRMatrix needs a single_svdlib CsrMatrix")
        mask = np.asarray(column_mask, dtype=bool)
        if mask.size != matrix.ncols():
            raise ValueError("Column mask must have the same length as the number of columns in the matrix")
        self.matrix = matrix
        self.column_mask = mask
        self.masked_to_original = np.nonzero(mask)[0]
        self.dtype = matrix.dtype
        self.device = matrix.device
        self._ctx = matrix._ctx
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        h = C.c_void_p()
        _ffi.check(_ffi.lib().svdh_mask_create(self._ctx, matrix._handle,
                                               m.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(h)))
        self._handle = h

    @classmethod
    def new(cls, matrix, column_mask):
        """masked.rs:21-46"""
        return cls(matrix, column_mask)

    @classmethod
    def with_columns(cls, matrix, columns):
        """masked.rs:48-55"""
        mask = np.zeros(matrix.ncols(), dtype=bool)
        for c in columns:
            if not 0 <= c < matrix.ncols():
                raise ValueError("Column index out of bounds")
            mask[c] = True
        return cls(matrix, mask)

    def uses_all_columns(self):
        """masked.rs:57-59"""
        return self.masked_to_original.size == self.matrix.ncols() and bool(self.column_mask.all())

    def ensure_identical_results_mode(self):
        """masked.rs:61-65"""
        small = self.matrix.nrows() <= 5 and self.matrix.ncols() <= 5
        return small and self.uses_all_columns()

    def __del__(self):
        h, self._handle = self._handle, None
        if h and _ffi._lib is not None:
            _ffi.lib().svdh_op_destroy(h)
