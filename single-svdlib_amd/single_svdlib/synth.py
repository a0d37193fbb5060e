"""Benchmark workloads as deterministic synthetic CSR matrices (synth/synth.c).

config1(): 2,000 x 1,500, 15 distinct uniform columns per row, U(0,1), seed 7
config2(): 100,000 x 20,000, nnz ~ Poisson(100) per row, Zipf(1.1) column
           popularity, log1p(Poisson(2)+1) values, seed 20261203
config3(): 500,000 x 30,000, nnz = min(floor(Pareto(2, 20)), 5000), uniform
           columns, N(0,1) values, seed 20261204
config4(): 200,000 x 25,000, nnz ~ Poisson(60), uniform columns,
           log1p(Poisson(2)+1) values, seed 20261205, plus 64 column masks of
           sizes uniform in [500, 8000] (numpy default_rng(seed + 1))
Seeds not fixed by the workload definitions are 20261201 + config number.
"""
import ctypes as C
import hashlib
import os

import numpy as np
import scipy.sparse as sp

from ._ffi import PKG_DIR

LIB_PATH = os.path.join(PKG_DIR, "libsynth.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python single-svdlib_amd/build.py`")
        h = C.CDLL(LIB_PATH)
        i64p = C.POINTER(C.c_int64)
        h.synth_row_lengths.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_uint64, C.c_double, C.c_double,
                                        C.c_double, i64p]
        h.synth_fill.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                 i64p, i64p, C.POINTER(C.c_double)]
        _lib = h
    return _lib


def generate(kind, m, n, seed, p1=0.0, p2=0.0, p3=0.0, dtype=np.float64):
    lib = _load()
    i64p = C.POINTER(C.c_int64)
    lens = np.empty(m, dtype=np.int64)
    lib.synth_row_lengths(kind, m, n, seed, p1, p2, p3, lens.ctypes.data_as(i64p))
    rowptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lens, out=rowptr[1:])
    nnz = int(rowptr[-1])
    colidx = np.empty(max(nnz, 1), dtype=np.int64)
    vals = np.empty(max(nnz, 1), dtype=np.float64)
    lib.synth_fill(kind, m, n, seed, p1, p2, p3, rowptr.ctypes.data_as(i64p), colidx.ctypes.data_as(i64p),
                   vals.ctypes.data_as(C.POINTER(C.c_double)))
    a = sp.csr_matrix((vals[:nnz].astype(dtype), colidx[:nnz], rowptr), shape=(m, n))
    return a


def config1():
    return generate(0, 2000, 1500, 7, 15)


def config2(dtype=np.float32):
    return generate(1, 100_000, 20_000, 20261203, 100.0, 1.1, dtype=dtype)


def config3(dtype=np.float64):
    return generate(2, 500_000, 30_000, 20261204, 2.0, 20.0, 5000.0, dtype=dtype)


def config4(dtype=np.float32):
    a = generate(3, 200_000, 25_000, 20261205, 60.0, dtype=dtype)
    rng = np.random.default_rng(20261206)
    masks = []
    for _ in range(64):
        size = int(rng.integers(500, 8001))
        masks.append(np.sort(rng.choice(25_000, size, replace=False)))
    return a, masks


def checksum(a):
    """SHA-256 over rowptr (int64), colidx (int64) and values (as stored)."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(a.indptr, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(a.indices, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(a.data).tobytes())
    return h.hexdigest()
