"""The benchmark workloads at their full sizes, checked through properties
that hold at any size (the oracle's CPU solves take minutes there):

- randomized SVD, 100,000 x 20,000 Zipf CSR (nnz 9,998,163), rank 50 +
  oversample 10 to a residual of 1e-5: the residual the solver reports is
  the FP64 residual of the triplets it returns, recomputed on the CPU; the
  singular vectors are orthonormal; the values are positive and descending;
- LAS2, 500,000 x 30,000 Pareto-row CSR (nnz 19,698,610), k = 100: the FP64
  solve's triplets satisfy A v = s u and A^T u = s v to 1e-8, its vectors are
  orthonormal; the FP32-value solve of the same matrix agrees with it;
- batched randomized SVD over the 64 column masks of the 200,000 x 25,000
  Poisson(60) CSR: two of the batch's results equal separate solves of the
  same masked views bit for bit (at 200,000 rows the A^T Y rows are cut at
  source blocks, so this covers the order of the piece sums), and, with
  injected test matrices, agree with the FP64 oracle on the physical column
  subsets to the FP32 bar.

The matrices come from the deterministic generator (single_svdlib.synth);
their SHA-256 checksums pin them.
"""
import numpy as np
import pytest

from helpers import max_principal_angle, triplet_residual

pytestmark = pytest.mark.gpu

# SHA-256 of (rowptr int64, colidx int64, FP32 values) as single_svdlib.synth.checksum computes it
SHA = {
    "100k_x_20k": "d144cf5c9fad6b4d5bdf934df9f043f4bd2c87680bd9ee5ad02dccf59501f966",
    "500k_x_30k": "d236152c926adbe5031f16057669d7b2c85f40e14d44b5fbbcfef1c826e95cf1",
    "200k_x_25k": "24a469a690b004d84333f005637f3a9ff47b982cfd817f9dc9a093d37195f7d1",
}


@pytest.fixture(scope="module")
def lib():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)
    return single_svdlib


def orthonormality_error(rows):
    """max |Q Q^T - I| for the rows of Q (FP64)."""
    q = np.asarray(rows, dtype=np.float64)
    return float(np.max(np.abs(q @ q.T - np.eye(q.shape[0]))))


def test_rsvd_100k_x_20k_reaches_tolerance(lib):
    from single_svdlib import synth
    a = synth.config2(np.float32)
    print(f"case rsvd 100,000 x 20,000 nnz={a.nnz}", flush=True)
    assert a.shape == (100_000, 20_000) and a.nnz == 9_998_163 and synth.checksum(a) == SHA["100k_x_20k"]
    rec = lib.randomized.randomized_svd(lib.CsrMatrix.from_scipy(a), 50, 10, 3, "QR", False, 42, False, tol=1e-5,
                                        max_power_iters=200)
    est = rec.extra["residual_estimate"]
    true = triplet_residual(a, rec.u, rec.s, rec.vt)
    assert est <= 1e-5
    assert abs(est - true) <= 0.05 * true, (est, true)
    assert rec.u.shape == (100_000, 50) and rec.vt.shape == (50, 20_000)
    assert orthonormality_error(rec.u.T) < 1e-5
    assert orthonormality_error(rec.vt) < 1e-5
    assert np.all(rec.s > 0) and np.all(np.diff(rec.s) <= 0)


def test_las2_500k_x_30k_triplets(lib):
    from single_svdlib import synth
    a32 = synth.config3(np.float32)
    a64 = a32.astype(np.float64)  # the same matrix in both solves
    print(f"case las2 500,000 x 30,000 nnz={a32.nnz}", flush=True)
    assert a32.shape == (500_000, 30_000) and a32.nnz == 19_698_610 and synth.checksum(a32) == SHA["500k_x_30k"]
    r64 = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a64), 100, 42)
    assert r64.d == 100 and r64.diagnostics.significant_values >= 100
    assert triplet_residual(a64, r64.u.T, r64.s, r64.vt) < 1e-8
    assert orthonormality_error(r64.u) < 1e-8
    assert orthonormality_error(r64.vt) < 1e-8
    assert np.all(r64.s > 0) and np.all(np.diff(r64.s) <= 0)
    r32 = lib.lanczos.svd_dim_seed(lib.CsrMatrix.from_scipy(a32), 100, 42)
    assert r32.s.dtype == np.float32 and r32.d == 100
    assert triplet_residual(a64, r32.u.T, r32.s.astype(np.float64), r32.vt) < 5e-4
    assert np.allclose(r32.s, r64.s, rtol=1e-4, atol=0)


def test_batched_64_masks_equal_separate_solves(lib):
    import oracle
    from single_svdlib import synth
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a, masks = synth.config4(np.float32)
    print(f"case batched 64 masks of 200,000 x 25,000 nnz={a.nnz}", flush=True)
    assert synth.checksum(a) == SHA["200k_x_25k"] and sum(len(c) for c in masks) == 272_846
    base = lib.CsrMatrix.from_scipy(a)
    views = [MaskedCSRMatrix.with_columns(base, c) for c in masks]
    recs = lib.randomized.randomized_svd_batched(views, 20, 10, 3, "QR", seed=7)
    assert len(recs) == 64
    for i in (0, 63):
        one = lib.randomized.randomized_svd(views[i], 20, 10, 3, "QR", False, 7, False)
        du = float(np.max(np.abs(recs[i].u - one.u)))
        dv = float(np.max(np.abs(recs[i].vt - one.vt)))
        print(f"mask {i}: max |dU| {du:.3e}, max |dV^T| {dv:.3e}", flush=True)
        np.testing.assert_array_equal(recs[i].s, one.s)
        np.testing.assert_array_equal(recs[i].u, one.u)
        np.testing.assert_array_equal(recs[i].vt, one.vt)
    for r, c in zip(recs, masks):
        assert r.vt.shape == (20, len(c)) and np.all(np.diff(r.s) <= 0)
    # accuracy: with injected test matrices, masks 0 and 63 against the FP64
    # oracle restatement on their physical column subsets (FP32 bar 1e-5)
    rng = np.random.default_rng(11)
    oms = [rng.normal(size=(len(c), 30)) if i in (0, 63) else None for i, c in enumerate(masks)]
    recs = lib.randomized.randomized_svd_batched(views, 20, 10, 3, "QR", seed=7, omegas=oms)
    a64 = a.astype(np.float64)
    for i in (0, 63):
        u, s, vt = oracle.randomized_svd(oracle.OracleOp(a64[:, masks[i]].tocsr(), nthreads=8), 20, 10, 3, "QR",
                                         omega=oms[i])
        err = float(np.max(np.abs(recs[i].s - s) / s))
        ang = max_principal_angle(recs[i].u[:, :5], u[:, :5])
        print(f"mask {i} vs oracle: sigma rel err {err:.2e}, top-5 U angle {ang:.2e}", flush=True)
        assert err < 1e-5
        assert ang < 1e-3


def test_rsvd_100k_x_20k_bf16_ladder_and_oracle(lib):
    # the mixed-precision leg of the same workload: BF16 values and BF16
    # input panels with FP32 accumulation until the residual estimate is
    # 1e-2, FP32 after. Checked against the FP64 truth of the committed cfg5
    # fixture (ARPACK + Rayleigh-Ritz on the CPU: the top 40 values), against
    # the FP32 solve, and through its FP64 residual; plus the FP32 fixed-count
    # solve against the FP64 oracle with the same injected Omega.
    import oracle
    from helpers import golden
    from single_svdlib import synth
    a = synth.config2(np.float32)
    print(f"case rsvd bf16 100,000 x 20,000 nnz={a.nnz}", flush=True)
    assert synth.checksum(a) == SHA["100k_x_20k"]
    truth = np.asarray(golden("cfg5_grad.json")["sigma"])
    op = lib.CsrMatrix.from_scipy(a, keep_bf16=True)
    r16 = lib.randomized.randomized_svd(op, 50, 10, 3, "QR", False, 42, False, tol=1e-5, max_power_iters=200,
                                        precision="bf16")
    r32 = lib.randomized.randomized_svd(op, 50, 10, 3, "QR", False, 42, False, tol=1e-5, max_power_iters=200)
    res16 = triplet_residual(a, r16.u, r16.s, r16.vt)
    e16 = float(np.max(np.abs(r16.s[:40] - truth) / truth))
    e32 = float(np.max(np.abs(r32.s[:40] - truth) / truth))
    print(f"bf16: {r16.extra['power_iters_run']} iterations, residual {res16:.2e}, sigma vs FP64 truth {e16:.2e}; "
          f"f32: {r32.extra['power_iters_run']} iterations, sigma vs truth {e32:.2e}", flush=True)
    assert r16.extra["residual_estimate"] <= 1e-5 and res16 <= 1.05e-5
    assert e16 < 1e-6 and e32 < 1e-6
    assert np.allclose(r16.s, r32.s, rtol=1e-6, atol=0)
    # fixed count (q = 3, the workload as specified) with an injected Omega
    om = np.random.default_rng(5).normal(size=(20_000, 60))
    q3 = lib.randomized.randomized_svd(op, 50, 10, 3, "QR", False, 42, False, omega=om)
    u, s, vt = oracle.randomized_svd(oracle.OracleOp(a.astype(np.float64), nthreads=16), 50, 10, 3, "QR", omega=om)
    e = float(np.max(np.abs(q3.s - s) / s))
    print(f"q=3 f32 vs FP64 oracle (same Omega): sigma rel err {e:.2e}", flush=True)
    assert e < 1e-5
