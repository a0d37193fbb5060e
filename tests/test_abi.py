"""The C-ABI library loads and exports every entry point declared in
include/svdh.h; host-only entry points agree with the oracle. No GPU needed."""
import ctypes as C
import os
import re

import numpy as np

import oracle
from single_svdlib import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "svdh.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|void|const char \*)\s*\*?\s*(svdh_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    names = declared_symbols()
    assert len(names) >= 18
    assert {"svdh_las2", "svdh_rsvd", "svdh_spmv", "svdh_spmm", "svdh_csr_create"} <= set(names)


def test_library_exports_every_declared_symbol():
    lib = _ffi.lib()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _ffi.SIGNATURES}
    assert set(declared_symbols()) == bound


def test_start_vector_matches_oracle():
    n = 1001
    out = np.empty(n)
    assert _ffi.lib().svdh_las2_start_vector(3141, n, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert np.array_equal(out, oracle.start_vector(3141, n))


def test_null_arguments_are_errors_not_crashes():
    lib = _ffi.lib()
    assert lib.svdh_ctx_create(0, None) == 10
    assert lib.svdh_las2_start_vector(1, 4, None) == 10
    assert lib.svdh_op_shape(None, None, None, None) == 10
    assert b"null" in lib.svdh_last_error()
