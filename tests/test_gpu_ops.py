"""GPU parity of the SMat operators (libsvdh.so through the C ABI) against
the FP64 oracle on identical inputs.

Tolerance: the GPU sums each row in a different order than the oracle, so
each output element must lie within  (nnz_row + 2) * eps * (|A| |x|)_i  of
the oracle (the standard floating-point bound for a dot product), with eps of
the accumulation type; FP32 outputs add one rounding of |y|.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import random_csr, skewed_csr

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


@pytest.fixture(scope="module")
def svdlib():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)  # fails loudly without a GPU or without libsvdh.so
    return single_svdlib


def bound(a, x, eps):
    a = sp.csr_matrix(a)
    n = np.diff(a.indptr).astype(np.float64) + 2.0
    if x.ndim == 2:
        n = n[:, None]
    return n * eps * (abs(a) @ np.abs(x))


def check_close(got, want, b):
    err = np.abs(np.asarray(got, dtype=np.float64) - want)
    assert np.all(err <= b + 1e-300), float(np.max(err - b))


MATRICES = {
    "uniform": lambda: random_csr(700, 500, 0.02, seed=1),
    "wide": lambda: random_csr(90, 2000, 0.05, seed=2),
    "skewed": lambda: skewed_csr(6000, 2000, seed=3, mean_row=20, max_row=3000),
    "empty_rows": lambda: random_csr(400, 300, 0.002, seed=4),
    "single": lambda: sp.csr_matrix(np.array([[2.5]])),
}


@pytest.mark.parametrize("name", list(MATRICES))
def test_spmv_f64(svdlib, name):
    a = MATRICES[name]()
    print(f"case {name} shape={a.shape} nnz={a.nnz}", flush=True)
    op = svdlib.CsrMatrix.from_scipy(a)
    ref = oracle.OracleOp(a)
    rng = np.random.default_rng(0)
    x = rng.normal(size=a.shape[1])
    y = np.empty(a.shape[0])
    op.svd_opa(x, y, False)
    check_close(y, ref.opa(x), bound(a, x, EPS64))
    xt = rng.normal(size=a.shape[0])
    yt = np.empty(a.shape[1])
    op.svd_opa(xt, yt, True)
    check_close(yt, ref.opa(xt, True), bound(a.T, xt, EPS64))
    # deterministic: a second call gives identical bits
    y2 = np.empty_like(yt)
    op.svd_opa(xt, y2, True)
    assert np.array_equal(yt, y2)


@pytest.mark.parametrize("name", ["uniform", "skewed"])
def test_spmv_f32(svdlib, name):
    a = MATRICES[name]().astype(np.float32)
    op = svdlib.CsrMatrix.from_scipy(a)
    a64 = a.astype(np.float64)
    ref = oracle.OracleOp(a64)
    x = np.random.default_rng(1).normal(size=a.shape[1]).astype(np.float32)
    y = np.empty(a.shape[0], dtype=np.float32)
    op.svd_opa(x, y, False)
    want = ref.opa(x.astype(np.float64))
    check_close(y, want, bound(a64, x.astype(np.float64), EPS64) + EPS32 * np.abs(want))


def pareto_rows(m, n, seed, dtype, mean_row=6, long_rows=3, long_len=25_000):
    """Pareto row lengths with uniform columns, a few rows far longer than the
    rest, about 10 % empty rows and 3 empty columns: more than 2^20 nonzeros,
    so the SpMV takes the source-tiled path."""
    rng = np.random.default_rng(seed)
    lens = np.minimum((mean_row / 2 / rng.uniform(size=m) ** 0.5).astype(np.int64), 4000)
    lens[rng.random(m) < 0.1] = 0
    lens[rng.choice(m, long_rows, replace=False)] = long_len
    rows = np.repeat(np.arange(m), lens)
    cols = rng.integers(0, n, size=rows.size)
    a = sp.csr_matrix((rng.normal(size=rows.size), (rows, cols)), shape=(m, n + 3))
    a.sum_duplicates()
    return a.astype(dtype)


@pytest.mark.parametrize("shape,mean_row", [((300_000, 40_000), 6), ((150_000, 9_000), 14)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_source_tiled(svdlib, shape, mean_row, dtype):
    # the source vector staged in LDS one block at a time (three blocks for
    # A x and one block, written directly, for 9,000 columns; 8 to 19 blocks
    # for A^T x), block partials added in order: within the dot-product bound
    # of the FP64 oracle, and the same bits on a second call
    a = pareto_rows(*shape, seed=11, dtype=dtype, mean_row=mean_row)
    assert a.nnz > 1 << 20
    print(f"case tiled {a.shape} nnz={a.nnz} dtype={np.dtype(dtype).name}", flush=True)
    op = svdlib.CsrMatrix.from_scipy(a)
    a64 = a.astype(np.float64)
    ref = oracle.OracleOp(a64)
    rng = np.random.default_rng(2)
    for transposed in (False, True):
        # vectors travel in the operator's dtype; FP32 outputs add one rounding
        nin = a.shape[0] if transposed else a.shape[1]
        x = rng.normal(size=nin).astype(dtype)
        y = np.empty(a.shape[1] if transposed else a.shape[0], dtype=dtype)
        op.svd_opa(x, y, transposed)
        want = ref.opa(x.astype(np.float64), transposed)
        b = bound(a64.T if transposed else a64, x.astype(np.float64), EPS64)
        if dtype == np.float32:
            b = b + EPS32 * np.abs(want)
        check_close(y, want, b)
        y2 = np.empty_like(y)
        op.svd_opa(x, y2, transposed)
        assert np.array_equal(y, y2)
        if dtype == np.float32:
            # FP32 values with FP64 vectors (the LAS2 FP32-SpMV mode): FP64 bound
            from single_svdlib import _ffi
            x64 = rng.normal(size=nin)
            y64 = np.empty(y.size)
            _ffi.check(_ffi.lib().svdh_spmv(op._ctx, op._handle, int(transposed), x64.ctypes.data,
                                            y64.ctypes.data, _ffi.SVDH_F64, 0))
            check_close(y64, ref.opa(x64, transposed), bound(a64.T if transposed else a64, x64, EPS64))


@pytest.mark.parametrize("name", ["uniform", "wide", "skewed", "empty_rows"])
@pytest.mark.parametrize("l", [1, 7, 60])
def test_spmm_f64(svdlib, name, l):
    a = MATRICES[name]()
    print(f"case {name} l={l}", flush=True)
    op = svdlib.CsrMatrix.from_scipy(a)
    ref = oracle.OracleOp(a)
    rng = np.random.default_rng(l)
    X = rng.normal(size=(a.shape[1], l))
    Y = np.zeros((a.shape[0], l))
    op.multiply_with_dense(X, Y, False)
    check_close(Y, ref.multiply_with_dense(X), bound(a, X, EPS64))
    Yin = rng.normal(size=(a.shape[0], l))
    Z = np.zeros((a.shape[1], l))
    op.multiply_with_dense(Yin, Z, True)
    check_close(Z, ref.multiply_with_dense(Yin, True), bound(a.T, Yin, EPS64))
    B = np.zeros((l, a.shape[1]))
    op.multiply_transposed_by_dense(Yin, B)
    check_close(B, ref.multiply_transposed_by_dense(Yin), bound(a.T, Yin, EPS64).T)


@pytest.mark.parametrize("name", ["uniform", "skewed"])
def test_spmm_f32(svdlib, name):
    a = MATRICES[name]().astype(np.float32)
    a64 = a.astype(np.float64)
    op = svdlib.CsrMatrix.from_scipy(a)
    ref = oracle.OracleOp(a64)
    X = np.random.default_rng(2).normal(size=(a.shape[1], 60)).astype(np.float32)
    Y = np.zeros((a.shape[0], 60), dtype=np.float32)
    op.multiply_with_dense(X, Y, False)
    want = ref.multiply_with_dense(X.astype(np.float64))
    check_close(Y, want, bound(a64, X.astype(np.float64), EPS32) + EPS32 * np.abs(want))


@pytest.mark.parametrize("keep", [400, 1450])
def test_masked_ops_equal_physical_subset(svdlib, keep):
    # 400 of 1500 columns: the forward products read a compacted copy of the
    # kept entries; 1450: they skip dropped entries through the column remap
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = skewed_csr(5000, 1500, seed=8, mean_row=16, max_row=2000)
    cols = np.sort(np.random.default_rng(3).choice(1500, keep, replace=False))
    base = svdlib.CsrMatrix.from_scipy(a)
    m = MaskedCSRMatrix.with_columns(base, cols)
    sub = a[:, cols]
    ref = oracle.OracleOp(a, columns=cols)
    assert m.nrows() == 5000 and m.ncols() == keep and m.nnz() == sub.nnz == ref.nnz()
    rng = np.random.default_rng(4)
    x = rng.normal(size=keep)
    y = np.empty(5000)
    m.svd_opa(x, y, False)
    check_close(y, ref.opa(x), bound(sub, x, EPS64))
    xt = rng.normal(size=5000)
    yt = np.empty(keep)
    m.svd_opa(xt, yt, True)
    check_close(yt, ref.opa(xt, True), bound(sub.T, xt, EPS64))
    X = rng.normal(size=(keep, 30))
    Y = np.zeros((5000, 30))
    m.multiply_with_dense(X, Y, False)
    check_close(Y, ref.multiply_with_dense(X), bound(sub, X, EPS64))
    Yin = rng.normal(size=(5000, 30))
    Z = np.zeros((keep, 30))
    m.multiply_with_dense(Yin, Z, True)
    check_close(Z, ref.multiply_with_dense(Yin, True), bound(sub.T, Yin, EPS64))
    means = m.compute_column_means()
    assert np.allclose(means, ref.column_means(), rtol=1e-12, atol=1e-14)
    assert m.uses_all_columns() is False


def test_empty_matrix_and_shape_errors(svdlib):
    a = sp.csr_matrix((5, 3))
    op = svdlib.CsrMatrix.from_scipy(a)
    y = np.ones(5)
    op.svd_opa(np.ones(3), y, False)
    assert np.all(y == 0)
    with pytest.raises(ValueError):
        op.svd_opa(np.ones(4), y, False)
    with pytest.raises(ValueError):
        svdlib.CsrMatrix(np.array([0, 1]), np.array([7]), np.array([1.0]), (1, 3))


def tall_zipf(m, n, seed, dtype=np.float64):
    """A tall matrix (m >= 65,536 rows, so A^T products cut rows at source
    blocks): Zipf-popular columns give long CSC rows (chunks), short ones
    (segments) and empty columns."""
    rng = np.random.default_rng(seed)
    nnz = 6 * m
    pop = 1.0 / np.arange(1, n + 1) ** 1.2
    pop /= pop.sum()
    cols = rng.choice(n, size=nnz, p=pop)
    rows = rng.integers(0, m, size=nnz)
    a = sp.csr_matrix((rng.normal(size=nnz), (rows, cols)), shape=(m, n + 5))  # 5 empty columns
    a.sum_duplicates()
    return a.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmm_transposed_source_blocks(svdlib, dtype):
    a = tall_zipf(80_000, 700, seed=5, dtype=dtype)
    print(f"case tall {a.shape} nnz={a.nnz} dtype={np.dtype(dtype).name}", flush=True)
    op = svdlib.CsrMatrix.from_scipy(a)
    a64 = a.astype(np.float64)
    rng = np.random.default_rng(1)
    eps = EPS64 if dtype == np.float64 else EPS32
    for l in (8, 30, 60):
        Yin = rng.normal(size=(a.shape[0], l)).astype(dtype)
        Z = np.zeros((a.shape[1], l), dtype=dtype)
        op.multiply_with_dense(Yin, Z, True)
        want = a64.T @ Yin.astype(np.float64)
        check_close(Z, want, bound(a64.T, Yin.astype(np.float64), eps) + eps * np.abs(want))
        assert np.all(Z[-5:] == 0)
    y = rng.normal(size=a.shape[0])
    z = np.empty(a.shape[1])
    op.svd_opa(y.astype(np.float64), z, True)
    check_close(z, a64.T @ y, bound(a64.T, y, EPS64 if dtype == np.float64 else EPS32))


def zipf_rows(m, n, per_row, seed, dtype):
    """Short rows (about per_row entries, Zipf(1.1)-popular columns) plus 5
    empty columns: A X stages the popular panel rows, and A^T Y takes the 256
    most popular columns on the matrix cores."""
    rng = np.random.default_rng(seed)
    nnz = per_row * m
    pop = 1.0 / np.arange(1, n + 1) ** 1.1
    pop /= pop.sum()
    cols = rng.permutation(n)[rng.choice(n, size=nnz, p=pop)]
    rows = rng.integers(0, m, size=nnz)
    a = sp.csr_matrix((rng.normal(size=nnz), (rows, cols)), shape=(m, n + 5))
    a.sum_duplicates()
    return a.astype(dtype)


@pytest.mark.parametrize("h", [128, 256])
def test_spmm_transposed_hot_columns_on_matrix_cores(svdlib, h, monkeypatch):
    # FP32 A^T Y with the h most popular columns as a dense tile product on
    # the matrix cores (one or two 16-slot tiles per wave) and the rest by the
    # gather kernel: equals the exact product within the FP error bound, as
    # does the path without the split (whose rounding differs, which shows
    # that the split ran); rows past the last full 32-row block and empty
    # columns included. The operator takes h on its first A^T Y.
    import os
    monkeypatch.setenv("SVDH_SPMM_THOT_H", str(h))
    a = zipf_rows(70_001, 3000, 20, seed=9, dtype=np.float32)
    print(f"case zipf rows {a.shape} nnz={a.nnz} h={h}", flush=True)
    op = svdlib.CsrMatrix.from_scipy(a)
    a64 = a.astype(np.float64)
    rng = np.random.default_rng(4)
    for l in (8, 30, 60, 64):
        Yin = rng.normal(size=(a.shape[0], l)).astype(np.float32)
        Z = np.zeros((a.shape[1], l), dtype=np.float32)
        op.multiply_with_dense(Yin, Z, True)
        want = a64.T @ Yin.astype(np.float64)
        tol = bound(a64.T, Yin.astype(np.float64), EPS32) + EPS32 * np.abs(want)
        check_close(Z, want, tol)
        assert np.all(Z[-5:] == 0)
        os.environ["SVDH_SPMM_THOT"] = "0"
        try:
            Z2 = np.zeros_like(Z)
            op.multiply_with_dense(Yin, Z2, True)
        finally:
            del os.environ["SVDH_SPMM_THOT"]
        check_close(Z2, want, tol)
        assert not np.array_equal(Z, Z2)
    X = rng.normal(size=(a.shape[1], 60)).astype(np.float32)
    Y = np.zeros((a.shape[0], 60), dtype=np.float32)
    op.multiply_with_dense(X, Y, False)
    want = a64 @ X.astype(np.float64)
    check_close(Y, want, bound(a64, X.astype(np.float64), EPS32) + EPS32 * np.abs(want))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmm_hot_columns_staged_in_lds(svdlib, dtype):
    # Zipf columns: A X stages the most popular panel rows in LDS and sums
    # each row's staged entries first; both it and the path without staging
    # equal the exact product within the FP error bound
    import os
    a = skewed_csr(8000, 3000, seed=12, mean_row=30, max_row=200, zipf=1.2).astype(dtype)
    op = svdlib.CsrMatrix.from_scipy(a)
    a64 = a.astype(np.float64)
    rng = np.random.default_rng(2)
    eps = EPS64 if dtype == np.float64 else EPS32
    for l in (8, 30, 60):
        X = rng.normal(size=(3000, l)).astype(dtype)
        Y = np.zeros((8000, l), dtype=dtype)
        op.multiply_with_dense(X, Y, False)
        want = a64 @ X.astype(np.float64)
        check_close(Y, want, bound(a64, X.astype(np.float64), eps) + eps * np.abs(want))
        os.environ["SVDH_SPMM_HOT"] = "0"
        try:
            Y2 = np.zeros_like(Y)
            op.multiply_with_dense(X, Y2, False)
        finally:
            del os.environ["SVDH_SPMM_HOT"]
        check_close(Y2, want, bound(a64, X.astype(np.float64), eps) + eps * np.abs(want))
        assert not np.array_equal(Y, np.zeros_like(Y))


@pytest.mark.parametrize("name", ["uniform", "wide", "skewed"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("masked", [False, True])
def test_spmm_centered(svdlib, name, dtype, masked):
    # SMat::multiply_with_dense_centered / multiply_transposed_by_dense_centered
    # (utils.rs:29, 32; masked.rs:473-614, 735-874) against the oracle's
    # centred products (the exact forward form, defect D7 corrected)
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = MATRICES[name]().astype(dtype)
    print(f"case centred {name} {np.dtype(dtype).name} masked={masked}", flush=True)
    n = a.shape[1]
    cols = np.sort(np.random.default_rng(7).choice(n, max(1, n * 2 // 3), replace=False)) if masked else None
    base = svdlib.CsrMatrix.from_scipy(a)
    op = MaskedCSRMatrix.with_columns(base, cols) if masked else base
    a64 = a.astype(np.float64)
    ref = oracle.OracleOp(a64, columns=cols) if masked else oracle.OracleOp(a64)
    sub = a64[:, cols] if masked else a64
    mu = ref.column_means().astype(dtype)
    mu64 = mu.astype(np.float64)
    eps = EPS64 if dtype == np.float64 else EPS32
    rng = np.random.default_rng(3)
    nc = sub.shape[1]
    X = rng.normal(size=(nc, 9)).astype(dtype)
    Yin = rng.normal(size=(a.shape[0], 9)).astype(dtype)
    X64, Y64 = X.astype(np.float64), Yin.astype(np.float64)

    got = np.zeros((a.shape[0], 9), dtype=dtype)
    op.multiply_with_dense_centered(X, got, False, mu)
    want = ref.multiply_with_dense_centered(X64, mu64)
    tol = bound(sub, X64, eps) + (nc + 2) * eps * (np.abs(mu64) @ np.abs(X64))[None, :] + 2 * eps * np.abs(want)
    check_close(got, want, tol)

    got = np.zeros((nc, 9), dtype=dtype)
    op.multiply_with_dense_centered(Yin, got, True, mu)
    want = ref.multiply_with_dense_centered(Y64, mu64, True)
    tol = (bound(sub.T, Y64, eps) + (a.shape[0] + 2) * eps * np.abs(mu64)[:, None] * np.abs(Y64).sum(0)[None, :]
           + 2 * eps * np.abs(want))
    check_close(got, want, tol)

    got = np.zeros((9, nc), dtype=dtype)
    op.multiply_transposed_by_dense_centered(Yin, got, mu)
    want = ref.multiply_transposed_by_dense_centered(Y64, mu64)
    check_close(got, want, tol.T)
