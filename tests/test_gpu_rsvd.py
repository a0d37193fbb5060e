"""GPU randomized SVD (randomized::randomized_svd) against the FP64 oracle
restatement with the same injected test matrix Omega, plus the reference's
invariants (shapes, positive descending values, power iterations improve the
fit, normalisers span the same space, masked == physical subset).

Bars: FP64 singular values within 1e-10 relative of the oracle and a max
principal angle of 1e-8 rad for U and V; FP32 within 1e-5 relative and
1e-4 rad; BF16 ladder reaches the FP32 answer within 1e-5 in the adaptive mode.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import max_principal_angle, random_csr, skewed_csr, triplet_residual

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)
    return single_svdlib


def lowrank_plus_noise(m, n, r, seed, noise=1e-3, density=0.2):
    rng = np.random.default_rng(seed)
    mask = sp.random(m, n, density=density, random_state=seed, format="csr")
    dense = (rng.normal(size=(m, r)) * np.linspace(10, 1, r)) @ rng.normal(size=(r, n))
    vals = dense[mask.nonzero()] + noise * rng.normal(size=mask.nnz)
    return sp.csr_matrix((vals, mask.nonzero()), shape=(m, n))


@pytest.mark.parametrize("dtype,sv_tol,ang_tol", [(np.float64, 1e-10, 1e-8), (np.float32, 1e-5, 1e-3)])
@pytest.mark.parametrize("q", [0, 2])
def test_matches_oracle_same_omega(lib, dtype, sv_tol, ang_tol, q):
    a = random_csr(3000, 800, 0.01, seed=5)
    rng = np.random.default_rng(1)
    om = rng.normal(size=(800, 30))
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a.astype(dtype)), 20, 10, q,
                                        lib.randomized.PowerIterationNormalizer.QR, False, 42, False, omega=om)
    a64 = a.astype(dtype).astype(np.float64)
    u, s, vt = oracle.randomized_svd(oracle.OracleOp(a64), 20, 10, q, "QR", omega=om)
    assert rec.d == 20 and rec.u.shape == (3000, 20) and rec.vt.shape == (20, 800)
    assert np.allclose(rec.s, s, rtol=sv_tol, atol=0), np.max(np.abs(rec.s - s) / s)
    # well-separated leading triplets span the same subspaces
    assert max_principal_angle(rec.u[:, :10], u[:, :10]) < ang_tol
    assert max_principal_angle(rec.vt[:10].T, vt[:10].T) < ang_tol


def test_shapes_order_positive(lib):
    # randomized/mod.rs:651-690
    a = random_csr(5000, 1000, 0.01, seed=42)
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 20, 10, 2,
                                        lib.randomized.PowerIterationNormalizer.QR, False, 42, False)
    assert rec.d == 20
    assert rec.u.shape == (5000, 20) and rec.vt.shape == (20, 1000)
    assert np.all(rec.s > 0) and np.all(np.diff(rec.s) <= 0)
    dg = rec.diagnostics
    assert (dg.lanczos_steps, dg.iterations, dg.dimensions, dg.random_seed) == (0, 2, 20, 42)


def test_power_iterations_improve_fit(lib):
    # randomized/mod.rs:692-777
    rng = np.random.default_rng(987)
    dense = rng.uniform(-1, 1, (100, 10)) @ rng.uniform(-1, 1, (10, 50)) + rng.uniform(-0.01, 0.01, (100, 50))
    a = lib.CsrMatrix.from_scipy(sp.csr_matrix(dense))
    errs = []
    for q in (0, 1, 3, 5):
        rec = lib.randomized.randomized_svd(a, 10, 5, q, lib.randomized.PowerIterationNormalizer.QR, False, 42, False)
        errs.append(np.linalg.norm(dense - rec.recompose()) / np.linalg.norm(dense))
    assert min(errs[1:]) < 0.9 * errs[0] or errs[0] < 1e-3


@pytest.mark.parametrize("norm", ["LU", "NONE"])
def test_normalizers(lib, norm):
    a = random_csr(2000, 600, 0.02, seed=11)
    om = np.random.default_rng(5).normal(size=(600, 15))
    ref = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 10, 5, 2, "QR", False, 1, False, omega=om)
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 10, 5, 2, norm, False, 1, False, omega=om)
    assert np.allclose(rec.s, ref.s, rtol=1e-8)


def test_adaptive_reaches_tolerance(lib):
    a = skewed_csr(20000, 3000, seed=21, mean_row=30, max_row=2000).astype(np.float32)
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 20, 10, 3, "QR", False, 7, False, tol=1e-5,
                                        max_power_iters=300)
    assert rec.extra["residual_estimate"] <= 1e-5
    assert triplet_residual(a, rec.u, rec.s, rec.vt) < 3e-5


@pytest.mark.parametrize("rank,over,dtype", [(20, 10, np.float32), (50, 14, np.float32), (60, 20, np.float32),
                                             (20, 10, np.float64)])
def test_adaptive_residual_estimate_is_the_fp64_residual(lib, rank, over, dtype):
    # the on-device residual check (matrix-core path for FP32 panels up to 64
    # wide, vector path otherwise) must report the residual of the triplets it
    # returns, as recomputed from them in FP64 on the CPU
    a = skewed_csr(20000, 3000, seed=23, mean_row=30, max_row=2000).astype(dtype)
    tol = 1e-5 if dtype == np.float32 else 1e-8
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), rank, over, 0, "QR", False, 3, False, tol=tol,
                                        max_power_iters=400)
    est = rec.extra["residual_estimate"]
    true = triplet_residual(a, rec.u, rec.s, rec.vt)
    assert est <= tol
    assert abs(est - true) <= 0.05 * true, (est, true)


def test_bf16_ladder_matches_f32(lib):
    a = skewed_csr(20000, 3000, seed=22, mean_row=30, max_row=2000).astype(np.float32)
    op = lib.CsrMatrix.from_scipy(a, keep_bf16=True)
    r32 = lib.randomized.randomized_svd(op, 20, 10, 3, "QR", False, 7, False, tol=1e-5, max_power_iters=300)
    r16 = lib.randomized.randomized_svd(op, 20, 10, 3, "QR", False, 7, False, tol=1e-5, max_power_iters=300,
                                        precision="bf16")
    assert r16.extra["residual_estimate"] <= 1e-5
    assert np.allclose(r16.s, r32.s, rtol=1e-5)


def test_masked_equals_physical_subset(lib):
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(3000, 1200, 0.01, seed=3)
    cols = np.sort(np.random.default_rng(2).choice(1200, 500, replace=False))
    om = np.random.default_rng(3).normal(size=(500, 25))
    base = lib.CsrMatrix.from_scipy(a)
    r1 = lib.randomized.randomized_svd(MaskedCSRMatrix.with_columns(base, cols), 15, 10, 2, "QR", False, 4, False,
                                       omega=om)
    r2 = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a[:, cols]), 15, 10, 2, "QR", False, 4, False,
                                       omega=om)
    assert np.allclose(r1.s, r2.s, rtol=1e-10)
    assert max_principal_angle(r1.u[:, :8], r2.u[:, :8]) < 1e-8


def test_rank_deficient_and_tiny(lib):
    # l > rank: the panel Gram is singular; the eigen fallback keeps it usable
    rng = np.random.default_rng(0)
    dense = rng.normal(size=(30, 3)) @ rng.normal(size=(3, 10))
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(sp.csr_matrix(dense)), 3, 3, 2, "QR", False, 42,
                                        False)
    s_true = np.linalg.svd(dense, compute_uv=False)[:3]
    assert np.allclose(rec.s, s_true, rtol=1e-8)


def test_device_output_matches_host_output(lib):
    # svdh_rsvd_device leaves U and V^T in HBM at caller-chosen leading
    # dimensions; the solve is deterministic, so it equals svdh_rsvd bit for bit.
    import ctypes as C
    from single_svdlib import _ffi
    a = random_csr(4000, 900, 0.01, seed=12)
    op = lib.CsrMatrix.from_scipy(a)
    k, ldu, ldvt = 20, 23, 905
    ref = lib.randomized.randomized_svd(op, k, 10, 2, "QR", False, 9, True)
    u = _ffi.DeviceArray((4000, ldu)).upload(np.full((4000, ldu), 7.0))
    vt = _ffi.DeviceArray((k, ldvt)).upload(np.full((k, ldvt), 7.0))
    s = np.zeros(k)
    meta = _ffi.Result_()

    def call(ld_u):
        return _ffi.lib().svdh_rsvd_device(_ffi.context(0), op._handle, k, 10, 2, _ffi.NORM_QR, 0, 9, 1, None,
                                           _ffi.PREC_F64, 0.0, 0, u.ptr, ld_u, _ffi.ptr(s), vt.ptr, ldvt,
                                           C.byref(meta))
    _ffi.check(call(ldu))
    assert meta.d == k and not meta.u and not meta.vt
    uh, vth = u.numpy(), vt.numpy()
    assert np.array_equal(s, ref.s)
    assert np.array_equal(uh[:, :k], ref.u) and np.all(uh[:, k:] == 7.0)
    assert np.array_equal(vth[:, :900], ref.vt) and np.all(vth[:, 900:] == 7.0)
    with pytest.raises(ValueError, match="leading dimension"):
        _ffi.check(call(k - 1))


@pytest.mark.parametrize("q", [0, 2])
def test_mean_center_matches_oracle(lib, q):
    # randomized/mod.rs:46-53 + 440-469: products with A - 1 mu^T, the column
    # means folded in as rank-one corrections on the device.
    a = random_csr(2500, 700, 0.02, seed=21)
    a.data += 1.0  # non-zero means
    om = np.random.default_rng(4).normal(size=(700, 25))
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 15, 10, q, "QR", True, 3, False, omega=om)
    u, s, vt = oracle.randomized_svd(oracle.OracleOp(a), 15, 10, q, "QR", mean_center=True, omega=om)
    assert np.allclose(rec.s, s, rtol=1e-10, atol=0), np.max(np.abs(rec.s - s) / s)
    assert max_principal_angle(rec.u[:, :8], u[:, :8]) < 1e-8
    assert max_principal_angle(rec.vt[:8].T, vt[:8].T) < 1e-8


def test_mean_center_adaptive_reaches_centered_svd(lib):
    # the adaptive solver on the centered operator converges to the dense
    # SVD of A - 1 mu^T (numpy LAPACK), in FP32 and FP64
    a = lowrank_plus_noise(1500, 400, 12, seed=8, noise=1e-2)
    a.data += 0.5
    dense = a.toarray()
    s_true = np.linalg.svd(dense - dense.mean(axis=0), compute_uv=False)[:10]
    for dt, tol in ((np.float64, 1e-9), (np.float32, 1e-5)):
        rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a.astype(dt)), 10, 10, 0, "QR", True, 1, False,
                                            tol=1e-7 if dt == np.float64 else 1e-5)
        assert np.allclose(rec.s, s_true, rtol=tol * 10), np.max(np.abs(rec.s - s_true) / s_true)
        assert triplet_residual(dense - dense.mean(axis=0), rec.u, rec.s, rec.vt) < 1e-4


@pytest.mark.parametrize("rank,over", [(60, 20), (100, 28)])
def test_wide_panels_match_oracle(lib, rank, over):
    # l = rank + oversample above 64 exercises the multi-block Gram (the
    # off-diagonal 64 x 64 column groups) and the wide apply/residual paths
    a = random_csr(4000, 1500, 0.02, seed=31)
    om = np.random.default_rng(6).normal(size=(1500, rank + over))
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a.astype(np.float32)), rank, over, 2, "QR", False, 5,
                                        False, omega=om)
    a64 = a.astype(np.float32).astype(np.float64)
    u, s, vt = oracle.randomized_svd(oracle.OracleOp(a64), rank, over, 2, "QR", omega=om)
    assert np.allclose(rec.s, s, rtol=2e-5, atol=0), np.max(np.abs(rec.s - s) / s)
    assert max_principal_angle(rec.u[:, :5], u[:, :5]) < 1e-3


def test_device_test_matrix_is_the_documented_generator(lib):
    # the Omega drawn on the device for a seed (k_gauss) is the counter-based
    # Box-Muller stream restated by oracle.gaussian_matrix: the seeded solve
    # equals the solve with that matrix injected (FP64: the device's log and
    # cospi may differ from numpy's in the last bit)
    a = random_csr(2000, 700, 0.02, seed=44)
    op = lib.CsrMatrix.from_scipy(a)
    for seed in (0, 12345):
        om = oracle.gaussian_matrix(700, 25, seed)
        r1 = lib.randomized.randomized_svd(op, 15, 10, 1, "QR", False, seed, False)
        r2 = lib.randomized.randomized_svd(op, 15, 10, 1, "QR", False, seed, False, omega=om)
        assert np.allclose(r1.s, r2.s, rtol=1e-12, atol=0), np.max(np.abs(r1.s - r2.s) / r2.s)
        assert max_principal_angle(r1.u[:, :10], r2.u[:, :10]) < 1e-9


@pytest.mark.parametrize("norm,sv_tol", [("LU", 1e-10), ("NONE", 1e-8)])
def test_normalizers_match_oracle(lib, norm, sv_tol):
    # randomized/mod.rs:102-105, 126-127, 239-293 against the oracle's LU
    # (column L2 normalisation) and None (no normalisation) loops with the
    # same Omega; None's panel grows as s_i^(2q+1), so its final QR works on
    # a panel of condition ~1e6 here and the bar is looser
    a = random_csr(2000, 600, 0.02, seed=11)
    om = np.random.default_rng(5).normal(size=(600, 15))
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 10, 5, 2, norm, False, 1, False, omega=om)
    u, s, vt = oracle.randomized_svd(oracle.OracleOp(a), 10, 5, 2, norm, omega=om)
    err = float(np.max(np.abs(rec.s - s) / s))
    assert err < sv_tol, err
    assert max_principal_angle(rec.u[:, :5], u[:, :5]) < 1e-6


def graded_matrix(m, n, r, span, seed, floor=1e-9):
    """A (m x n, stored as CSR) = U diag(s) V^T + noise with random dense
    orthonormal U, V and s log-spaced from 1 down to 1/span over r values."""
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.normal(size=(m, r)))
    v, _ = np.linalg.qr(rng.normal(size=(n, r)))
    s = np.logspace(0, -np.log10(span), r)
    return sp.csr_matrix((u * s) @ v.T + floor * rng.normal(size=(m, n)))


@pytest.mark.parametrize("span", [1e4, 1e6])
def test_graded_spectrum_matches_oracle(lib, span):
    # the top-l singular values span 1e4 / 1e6: the randomized solve (implicit
    # CholeskyQR, SVD of B from the Gram of Z = B^T) against the oracle's
    # Householder QR and direct SVD of B (randomized/mod.rs:96-173) with the
    # same Omega, at the FP64 bar for every returned value
    a = graded_matrix(2000, 300, 30, span, seed=3)
    om = np.random.default_rng(1).normal(size=(300, 30))
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 20, 10, 2, "QR", False, 1, False, omega=om)
    u, s, vt = oracle.randomized_svd(oracle.OracleOp(a), 20, 10, 2, "QR", omega=om)
    err = float(np.max(np.abs(rec.s - s) / s))
    assert s[-1] / s[0] < 10.0 / span ** (2 / 3)  # the returned values do span the range
    assert err < 1e-10, err
    assert max_principal_angle(rec.u, u) < 1e-8
    assert max_principal_angle(rec.vt.T, vt.T) < 1e-8
    assert rec.extra["qr_fallbacks"] == 0


def test_single_precision_host_arrays(lib):
    # FP32 solves hand over u32 / s32 / vt32 (u, s, vt NULL): the FP64 device
    # results of the same solve rounded to nearest, bit for bit
    import ctypes as C
    from single_svdlib import _ffi
    a = random_csr(3000, 700, 0.01, seed=14).astype(np.float32)
    op = lib.CsrMatrix.from_scipy(a)
    k = 12
    res = _ffi.Result_()
    _ffi.check(_ffi.lib().svdh_rsvd(_ffi.context(0), op._handle, k, 8, 2, _ffi.NORM_QR, 0, 5, 1, None,
                                    _ffi.PREC_F32, 0.0, 0, C.byref(res)))
    assert not res.u and not res.s and not res.vt and res.u32 and res.s32 and res.vt32
    u32 = np.ctypeslib.as_array(res.u32, (3000 * k,)).reshape(3000, k).copy()
    s32 = np.ctypeslib.as_array(res.s32, (k,)).copy()
    vt32 = np.ctypeslib.as_array(res.vt32, (k * 700,)).reshape(k, 700).copy()
    _ffi.lib().svdh_result_free(C.byref(res))
    assert not res.u32 and not res.s32 and not res.vt32
    u = _ffi.DeviceArray((3000, k))
    vt = _ffi.DeviceArray((k, 700))
    s = np.zeros(k)
    meta = _ffi.Result_()
    _ffi.check(_ffi.lib().svdh_rsvd_device(_ffi.context(0), op._handle, k, 8, 2, _ffi.NORM_QR, 0, 5, 1, None,
                                           _ffi.PREC_F32, 0.0, 0, u.ptr, k, _ffi.ptr(s), vt.ptr, 700, C.byref(meta)))
    assert u32.dtype == np.float32
    np.testing.assert_array_equal(u32, u.numpy().astype(np.float32))
    np.testing.assert_array_equal(vt32, vt.numpy().astype(np.float32))
    np.testing.assert_array_equal(s32, s.astype(np.float32))
    # the Python mirror reads them as the reference's SvdRec<f32>
    rec = lib.randomized.randomized_svd(op, k, 8, 2, "QR", False, 5, True)
    assert rec.u.dtype == np.float32
    np.testing.assert_array_equal(rec.u, u32)
