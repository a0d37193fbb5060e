"""GPU LAS2 (lanczos::svd*) against the reference's known answers and the
FP64 oracle on identical inputs and identical start vectors.

Bars: FP64 singular values within 1e-10 relative of the oracle; singular
vectors within a max principal angle of 1e-8 rad; diagnostics equal. FP32
(svd_las2::<f32> control flow): singular values within 1e-5 relative of the
FP64 truth.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import golden, max_principal_angle, random_csr, triplet_residual

pytestmark = pytest.mark.gpu

DIAG_KEYS = ("non_zero", "dimensions", "iterations", "lanczos_steps", "ritz_values_stabilized",
             "significant_values", "singular_values", "random_seed")


@pytest.fixture(scope="module")
def lib():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)
    return single_svdlib


def cfg1_matrix(seed=7):
    """2,000 x 1,500, 15 distinct uniform columns per row, U(0,1) values."""
    rng = np.random.default_rng(seed)
    m, n, k = 2000, 1500, 15
    cols = np.stack([rng.choice(n, k, replace=False) for _ in range(m)])
    rows = np.repeat(np.arange(m), k)
    return sp.csr_matrix((rng.uniform(size=m * k), (rows, cols.ravel())), shape=(m, n))


def compare_with_oracle(a, rec, r, sv_tol=1e-10, angle_tol=1e-8, diag=True):
    assert rec.d == r["d"]
    assert np.allclose(rec.s, r["s"], rtol=sv_tol, atol=0), np.max(np.abs(rec.s - r["s"]) / r["s"])
    if rec.d:
        assert max_principal_angle(rec.vt.T, r["vt"].T) < angle_tol
        assert max_principal_angle(rec.u.T, r["u"].T) < angle_tol
    if diag:
        for k in DIAG_KEYS:
            assert getattr(rec.diagnostics, k) == r["diagnostics"][k], k
        assert rec.diagnostics.transposed == bool(r["diagnostics"]["transposed"])


def test_legacy_3x3_known_answers(lib):
    g = golden("legacy_kats.json")
    a = lib.CsrMatrix.from_scipy(sp.csc_matrix(np.array(g["matrix_3x3"])))
    rec = lib.lanczos.svd_dim_seed(a, 0, 3141)
    dump = g["svdlibrs_3x3_seed3141"]
    assert rec.d == 3
    assert np.allclose(rec.s, dump["s"], rtol=0, atol=1e-12)
    for i in range(3):
        for got, want in ((rec.u[i], np.array(dump["ut"][i])), (rec.vt[i], np.array(dump["vt"][i]))):
            assert np.allclose(np.sign(got @ want) * got, want, atol=1e-11)
    assert np.allclose(rec.u.T @ np.diag(rec.s) @ rec.vt, g["matrix_3x3"], atol=1e-12)
    for k in DIAG_KEYS:
        assert getattr(rec.diagnostics, k) == dump["diagnostics"][k], k


def test_basic_2x2_and_identity(lib):
    g = golden("legacy_kats.json")
    rec = lib.lanczos.svd(lib.CsrMatrix.from_scipy(sp.coo_matrix(np.array(g["basic_2x2"]["matrix"]))))
    assert rec.d == 2 and np.allclose(rec.s, g["basic_2x2"]["s"], atol=1e-12)
    assert np.allclose(rec.u.T @ np.diag(rec.s) @ rec.vt, g["basic_2x2"]["matrix"], atol=1e-12)
    rec = lib.lanczos.svd(lib.CsrMatrix.from_scipy(sp.identity(3, format="csc")))
    assert rec.d == 1 and abs(rec.s[0] - 1.0) < 1e-12


@pytest.mark.parametrize("shape,density,k,seed", [((60, 40), 0.2, 8, 42), ((40, 90), 0.15, 6, 7),
                                                   ((300, 200), 0.05, 12, 1), ((800, 500), 0.01, 20, 3)])
def test_matches_oracle(lib, shape, density, k, seed):
    a = random_csr(*shape, density, seed=seed)
    print(f"case shape={shape} k={k}", flush=True)
    rec = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a), k, seed)
    st, r = oracle.las2(oracle.OracleOp(a), k, 0, random_seed=seed)
    assert st == 0
    compare_with_oracle(a, rec, r)
    s_true = np.linalg.svd(a.toarray(), compute_uv=False)
    assert np.allclose(rec.s, s_true[: rec.d], rtol=1e-10)


def test_source_tiled_operator_matches_oracle(lib):
    # 1.2 M nonzeros, 150,000 x 20,000: both SpMV directions of A^T (A x) take
    # the source-tiled layout (2 and 10 source blocks); Pareto row lengths with
    # a few long rows
    rng = np.random.default_rng(23)
    m, n = 150_000, 20_000
    lens = np.minimum((4 / rng.uniform(size=m) ** 0.5).astype(np.int64), 3000)
    lens[rng.choice(m, 4, replace=False)] = 12_000
    rows = np.repeat(np.arange(m), lens)
    a = sp.csr_matrix((rng.normal(size=rows.size), (rows, rng.integers(0, n, size=rows.size))), shape=(m, n))
    a.sum_duplicates()
    assert a.nnz > 1 << 20
    print(f"case tiled LAS2 shape={a.shape} nnz={a.nnz}", flush=True)
    rec = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a), 6, 5)
    st, r = oracle.las2(oracle.OracleOp(a, nthreads=4), 6, 0, random_seed=5)
    assert st == 0
    compare_with_oracle(a, rec, r, diag=False)
    assert triplet_residual(a, rec.u.T, rec.s, rec.vt) < 1e-8
    # FP32 values with the FP64 recurrence through the same fused operator
    rec32 = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a.astype(np.float32)), 6, 5)
    assert rec32.d >= 4
    assert np.allclose(rec32.s[:4], r["s"][:4], rtol=1e-5)


def test_cfg1_las2_f64(lib):
    a = cfg1_matrix()
    rec = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a), 10, 42)
    st, r = oracle.las2(oracle.OracleOp(a, nthreads=4), 10, 0, random_seed=42)
    assert st == 0
    compare_with_oracle(a, rec, r)
    assert triplet_residual(a, rec.u.T, rec.s, rec.vt) < 1e-8


def test_injected_start_vector(lib):
    a = random_csr(200, 150, 0.05, seed=11)
    start = np.random.default_rng(0).uniform(-1, 1, 150)
    rec = lib.lanczos.svd_las2(lib.CsrMatrix.from_scipy(a), 10, 0, [-1e-30, 1e-30], 1e-6, 99, start=start)
    st, r = oracle.las2(oracle.OracleOp(a), 10, 0, random_seed=99, start=start)
    compare_with_oracle(a, rec, r)


def test_masked_equals_physical_subset(lib):
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(1400, 1000, 0.01, seed=42, low=0.1, high=10.0)
    cols = np.sort(np.random.default_rng(1).choice(1000, 300, replace=False))
    base = lib.CsrMatrix.from_scipy(a)
    r1 = lib.lanczos.svd_dim_seed(MaskedCSRMatrix.with_columns(base, cols), 10, 5)
    r2 = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a[:, cols]), 10, 5)
    assert r1.d == r2.d
    assert np.allclose(r1.s, r2.s, rtol=1e-10)


def test_f32_mode(lib):
    a = random_csr(500, 300, 0.03, seed=3).astype(np.float32)
    rec = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a), 8, 9)
    assert rec.s.dtype == np.float32
    s_true = np.linalg.svd(a.astype(np.float64).toarray(), compute_uv=False)
    assert np.allclose(rec.s, s_true[: rec.d], rtol=1e-5)
    st, r = oracle.las2(oracle.OracleOp(a.astype(np.float64)), 8, 0, random_seed=9, eps=oracle.F32_EPS)
    assert st == 0 and rec.d == r["d"]


def test_error_paths(lib):
    from single_svdlib import error
    with pytest.raises(error.Las2Error):
        lib.lanczos.svd(lib.CsrMatrix.from_scipy(sp.csr_matrix(np.ones((1, 5)))))
    with pytest.raises(error.StartvError):
        lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(sp.csr_matrix((6, 4))), 0, 1)


@pytest.mark.parametrize("shape", [(900, 400), (300, 800)])  # the second takes the transposed path
def test_device_output_equals_host_output(lib, shape):
    # svdh_las2_device writes U (d x m) and V^T (d x n) into device buffers;
    # same solve, so the values and vectors equal the host hand-over bit for bit
    import ctypes as C
    from single_svdlib import _ffi
    a = random_csr(shape[0], shape[1], 0.03, seed=41)
    op = lib.CsrMatrix.from_scipy(a)
    rec = lib.lanczos.svd_dim_seed(op, 6, 123)
    m, n = shape
    k = 6
    u_dev = _ffi.DeviceArray((k, m))
    vt_dev = _ffi.DeviceArray((k, n))
    s = np.zeros(k)
    meta = _ffi.Result_()
    _ffi.check(_ffi.lib().svdh_las2_device(op._ctx, op._handle, 6, 0, -1e-30, 1e-30, 1e-6, 123, None,
                                           _ffi.PREC_F64, u_dev.ptr, m, _ffi.ptr(s), vt_dev.ptr, n,
                                           C.byref(meta)))
    d = meta.d
    assert d == rec.d and meta.diagnostics.lanczos_steps == rec.diagnostics.lanczos_steps
    assert np.array_equal(s[:d], rec.s)
    assert np.array_equal(u_dev.numpy()[:d], rec.u) and np.array_equal(vt_dev.numpy()[:d], rec.vt)
    assert bool(meta.diagnostics.transposed) == (n >= 1.2 * m)


@pytest.mark.parametrize("case", ["cfg1", "wide", "long_vectors", "rank_deficient", "masked"])
def test_pipelined_steps_equal_synchronous_steps(lib, case, monkeypatch):
    # With the pipeline, step j + 1 is enqueued before the host reads step j's
    # scalars and dropped when the host side changes the state
    # (reorthogonalisation, rnm <= tol, restarts). Every step must run the same
    # kernels on the same values as the synchronous driver: equal bits.
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    rng = np.random.default_rng(3)
    if case == "cfg1":
        a, k = cfg1_matrix(), 10
    elif case == "wide":
        a, k = random_csr(300, 900, 0.03, seed=5), 40
    elif case == "long_vectors":
        a, k = random_csr(30_000, 12_000, 0.0005, seed=9), 12   # vectors above the one-workgroup chain
    elif case == "rank_deficient":
        b = sp.csr_matrix(rng.normal(size=(400, 12))) @ sp.csr_matrix(rng.normal(size=(12, 300)))
        a, k = sp.csr_matrix(b), 30                             # invariant subspace: restarts
    else:
        a, k = random_csr(2000, 900, 0.02, seed=13), 15
    op = lib.CsrMatrix.from_scipy(a)
    if case == "masked":
        op = MaskedCSRMatrix.with_columns(op, np.sort(rng.choice(900, 500, replace=False)))
    print(f"case pipeline {case} shape={a.shape} k={k}", flush=True)
    monkeypatch.setenv("SVDH_LAS2_PIPELINE", "1")
    r1 = lib.lanczos.svd_dim_seed(op, k, 21)
    monkeypatch.setenv("SVDH_LAS2_PIPELINE", "0")
    r0 = lib.lanczos.svd_dim_seed(op, k, 21)
    assert r1.d == r0.d and r1.d >= 2
    for key in DIAG_KEYS:
        assert getattr(r1.diagnostics, key) == getattr(r0.diagnostics, key), key
    np.testing.assert_array_equal(r1.s, r0.s)
    np.testing.assert_array_equal(r1.u, r0.u)
    np.testing.assert_array_equal(r1.vt, r0.vt)


def test_f32_solve_hands_over_single_precision_arrays(lib):
    # svd_las2::<f32> returns SvdRec<f32>: u32 / s32 / vt32, the FP64 device
    # results of the same solve rounded to nearest, bit for bit
    import ctypes as C
    from single_svdlib import _ffi
    a = random_csr(700, 400, 0.03, seed=43).astype(np.float32)
    op = lib.CsrMatrix.from_scipy(a)
    res = _ffi.Result_()
    _ffi.check(_ffi.lib().svdh_las2(op._ctx, op._handle, 8, 0, -1e-30, 1e-30, 1e-6, 77, None, _ffi.PREC_F32,
                                    C.byref(res)))
    d = res.d
    assert d >= 2 and not res.u and not res.s and res.u32 and res.s32 and res.vt32
    u32 = np.ctypeslib.as_array(res.u32, (d * 700,)).reshape(d, 700).copy()
    s32 = np.ctypeslib.as_array(res.s32, (d,)).copy()
    vt32 = np.ctypeslib.as_array(res.vt32, (d * 400,)).reshape(d, 400).copy()
    _ffi.lib().svdh_result_free(C.byref(res))
    u_dev, vt_dev = _ffi.DeviceArray((8, 700)), _ffi.DeviceArray((8, 400))
    s = np.zeros(8)
    meta = _ffi.Result_()
    _ffi.check(_ffi.lib().svdh_las2_device(op._ctx, op._handle, 8, 0, -1e-30, 1e-30, 1e-6, 77, None, _ffi.PREC_F32,
                                           u_dev.ptr, 700, _ffi.ptr(s), vt_dev.ptr, 400, C.byref(meta)))
    assert meta.d == d
    np.testing.assert_array_equal(s32, s[:d].astype(np.float32))
    np.testing.assert_array_equal(u32, u_dev.numpy()[:d].astype(np.float32))
    np.testing.assert_array_equal(vt32, vt_dev.numpy()[:d].astype(np.float32))
