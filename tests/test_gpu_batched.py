"""Batched masked randomized SVD (svdh_rsvd_batched): every SpMM round of all
masks in one launch. Each mask's result must equal its own randomized_svd
(same kernels, same reduction orders: bit for bit), and masked == physical
subset (masked.rs:914-1000) within 1e-10."""
import numpy as np
import pytest

from helpers import max_principal_angle, random_csr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)
    return single_svdlib


def masks_of(n, sizes, seed):
    rng = np.random.default_rng(seed)
    return [np.sort(rng.choice(n, s, replace=False)) for s in sizes]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_equals_separate_solves(lib, dtype):
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(3000, 1200, 0.02, seed=17).astype(dtype)
    base = lib.CsrMatrix.from_scipy(a)
    cols = masks_of(1200, [40, 700, 120, 1200, 333, 90], seed=5)
    views = [MaskedCSRMatrix.with_columns(base, c) for c in cols]
    recs = lib.randomized.randomized_svd_batched(views, 10, 10, 2, "QR", seed=9)
    assert len(recs) == len(views)
    for v, r in zip(views, recs):
        one = lib.randomized.randomized_svd(v, 10, 10, 2, "QR", False, 9, False)
        assert r.u.shape == one.u.shape and r.vt.shape == one.vt.shape
        np.testing.assert_array_equal(r.s, one.s)
        np.testing.assert_array_equal(r.u, one.u)
        np.testing.assert_array_equal(r.vt, one.vt)
        assert r.diagnostics.iterations == 2 and r.extra["power_iters_run"] == 2


def test_batched_calls_of_changing_widths_and_sizes_equal_separate_solves(lib):
    # the batch panels are reused from call to call: their padding columns
    # past the panel width must read as zero whatever an earlier call with
    # another width, leading dimension or mask size left there
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(2500, 1000, 0.02, seed=23).astype(np.float32)
    base = lib.CsrMatrix.from_scipy(a)
    # (rank, oversample) -> panel width l and leading dimension: 30/32,
    # 26/32 (narrower, same ld), 30/32 again, 21/24, 30/32 with larger masks
    plan = [((20, 10), [60, 200]), ((16, 10), [60, 200]), ((20, 10), [60, 200]), ((11, 10), [300, 90]),
            ((20, 10), [900, 700, 400])]
    for i, ((rank, over), sizes) in enumerate(plan):
        cols = masks_of(1000, sizes, seed=40 + i)
        views = [MaskedCSRMatrix.with_columns(base, c) for c in cols]
        recs = lib.randomized.randomized_svd_batched(views, rank, over, 2, "QR", seed=11)
        for v, r in zip(views, recs):
            one = lib.randomized.randomized_svd(v, rank, over, 2, "QR", False, 11, False)
            np.testing.assert_array_equal(r.s, one.s)
            np.testing.assert_array_equal(r.u, one.u)
            np.testing.assert_array_equal(r.vt, one.vt)


@pytest.mark.parametrize("dtype,sv_tol,ang_tol", [(np.float64, 1e-10, 1e-8), (np.float32, 1e-5, 1e-3)])
def test_batched_matches_oracle_on_physical_subsets(lib, dtype, sv_tol, ang_tol):
    # every mask against the FP64 oracle restatement of randomized_svd
    # (randomized/mod.rs:24-199) on the physical column subset a[:, cols]
    # (the masked == physical invariant, masked.rs:914-1000), with the same
    # injected test matrix per mask
    import oracle
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(3000, 1000, 0.02, seed=31).astype(dtype)
    base = lib.CsrMatrix.from_scipy(a)
    cols = masks_of(1000, [300, 800, 150, 1000, 520], seed=12)
    rng = np.random.default_rng(3)
    oms = [rng.normal(size=(len(c), 20)) for c in cols]
    recs = lib.randomized.randomized_svd_batched([MaskedCSRMatrix.with_columns(base, c) for c in cols], 10, 10, 2,
                                                 "QR", seed=1, omegas=oms)
    a64 = a.astype(np.float64)
    for c, om, r in zip(cols, oms, recs):
        u, s, vt = oracle.randomized_svd(oracle.OracleOp(a64[:, c].tocsr()), 10, 10, 2, "QR", omega=om)
        assert r.d == 10 and r.u.shape == (3000, 10) and r.vt.shape == (10, len(c))
        err = float(np.max(np.abs(r.s - s) / s))
        assert err < sv_tol, err
        assert max_principal_angle(r.u[:, :5], u[:, :5]) < ang_tol
        assert max_principal_angle(r.vt[:5].T, vt[:5].T) < ang_tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_injected_omegas_equal_separate_solves(lib, dtype):
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(2500, 900, 0.02, seed=37).astype(dtype)
    base = lib.CsrMatrix.from_scipy(a)
    cols = masks_of(900, [200, 900, 450], seed=4)
    rng = np.random.default_rng(8)
    oms = [rng.normal(size=(len(c), 16)) for c in cols]
    oms[1] = None  # drawn from the seed, as the separate solve with that seed
    views = [MaskedCSRMatrix.with_columns(base, c) for c in cols]
    recs = lib.randomized.randomized_svd_batched(views, 8, 8, 2, "QR", seed=6, omegas=oms)
    for v, om, r in zip(views, oms, recs):
        one = lib.randomized.randomized_svd(v, 8, 8, 2, "QR", False, 6, False, omega=om)
        np.testing.assert_array_equal(r.s, one.s)
        np.testing.assert_array_equal(r.u, one.u)
        np.testing.assert_array_equal(r.vt, one.vt)


def test_batched_masked_equals_physical(lib):
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(2500, 900, 0.03, seed=23)
    base = lib.CsrMatrix.from_scipy(a)
    cols = masks_of(900, [300, 650, 120], seed=8)
    recs = lib.randomized.randomized_svd_batched([MaskedCSRMatrix.with_columns(base, c) for c in cols], 8, 8, 3,
                                                 "QR", seed=4)
    phys = lib.randomized.randomized_svd_batched([lib.CsrMatrix.from_scipy(a[:, c].tocsr()) for c in cols], 8, 8,
                                                 3, "QR", seed=4)
    for r, p in zip(recs, phys):
        assert np.allclose(r.s, p.s, rtol=1e-10)
        assert max_principal_angle(r.u[:, :4], p.u[:, :4]) < 1e-8


def test_batched_rejects_mixed_widths_and_normalizers(lib):
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(500, 300, 0.05, seed=2)
    base = lib.CsrMatrix.from_scipy(a)
    small = MaskedCSRMatrix.with_columns(base, np.arange(5))     # rank clamps to 5 -> other width
    big = MaskedCSRMatrix.with_columns(base, np.arange(100))
    with pytest.raises(ValueError, match="panel width"):
        lib.randomized.randomized_svd_batched([small, big], 10, 5, 1, "QR", seed=1)
    with pytest.raises(ValueError, match="QR"):
        lib.randomized.randomized_svd_batched([big], 10, 5, 1, "LU", seed=1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_device_outputs_equal_host_outputs(lib, dtype):
    # device outputs are FP64; an FP32 solve's host arrays are those values
    # rounded to single precision
    import ctypes as C
    from single_svdlib import _ffi
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(2000, 800, 0.02, seed=29, dtype=dtype)
    base = lib.CsrMatrix.from_scipy(a)
    cols = masks_of(800, [150, 500, 60], seed=2)
    views = [MaskedCSRMatrix.with_columns(base, c) for c in cols]
    recs = lib.randomized.randomized_svd_batched(views, 6, 6, 2, "QR", seed=3)
    prec = _ffi.PREC_F64 if dtype == np.float64 else _ffi.PREC_F32
    k = 6
    us = [_ffi.DeviceArray((2000, k + 2)) for _ in cols]
    vts = [_ffi.DeviceArray((k, len(c) + 3)) for c in cols]
    nb = len(cols)
    metas = (_ffi.Result_ * nb)()
    handles = (C.c_void_p * nb)(*[v._handle for v in views])
    up = (C.c_void_p * nb)(*[u.ptr for u in us])
    vp = (C.c_void_p * nb)(*[v.ptr for v in vts])
    ldu = (C.c_int64 * nb)(*[k + 2] * nb)
    ldvt = (C.c_int64 * nb)(*[len(c) + 3 for c in cols])
    _ffi.check(_ffi.lib().svdh_rsvd_batched_device(_ffi.context(0), nb, handles, 6, 6, 2, _ffi.NORM_QR, 3, 1,
                                                   None, prec, up, ldu, vp, ldvt, metas))
    for b in range(nb):
        assert metas[b].d == k and not metas[b].u and not metas[b].vt
        s = np.ctypeslib.as_array(metas[b].s, (k,)).copy()
        _ffi.lib().svdh_result_free(C.byref(metas[b]))
        assert recs[b].u.dtype == dtype and recs[b].vt.dtype == dtype
        np.testing.assert_array_equal(s.astype(dtype), recs[b].s)
        np.testing.assert_array_equal(us[b].numpy()[:, :k].astype(dtype), recs[b].u)
        np.testing.assert_array_equal(vts[b].numpy()[:, :len(cols[b])].astype(dtype), recs[b].vt)


def test_batched_persistent_autotuned_and_static_launches_agree(lib, tmp_path, monkeypatch):
    # the persistent work-queue launch, its autotuner (every candidate timed
    # on the first call of a workload, the choice stored in the tuning file)
    # and the static one-workgroup-per-block launch compute every row the same
    # way: all equal the separate solves bit for bit
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a = random_csr(4000, 1500, 0.02, seed=41).astype(np.float32)
    base = lib.CsrMatrix.from_scipy(a)
    cols = masks_of(1500, [300, 1400, 800, 50, 1100], seed=13)
    views = [MaskedCSRMatrix.with_columns(base, c) for c in cols]
    tune = tmp_path / "tune.txt"
    monkeypatch.setenv("SVDH_TUNE_FILE", str(tune))
    first = lib.randomized.randomized_svd_batched(views, 10, 10, 2, "QR", seed=5)
    text = tune.read_text()
    assert "spmm_batched dir=0" in text and "spmm_batched dir=1" in text, text
    again = lib.randomized.randomized_svd_batched(views, 10, 10, 2, "QR", seed=5)
    monkeypatch.setenv("SVDH_BATCH_TUNE", "0")
    default = lib.randomized.randomized_svd_batched(views, 10, 10, 2, "QR", seed=5)
    monkeypatch.setenv("SVDH_BATCH_PERSISTENT", "0")
    static = lib.randomized.randomized_svd_batched(views, 10, 10, 2, "QR", seed=5)
    for i, v in enumerate(views):
        one = lib.randomized.randomized_svd(v, 10, 10, 2, "QR", False, 5, False)
        for r in (first[i], again[i], default[i], static[i]):
            np.testing.assert_array_equal(r.s, one.s)
            np.testing.assert_array_equal(r.u, one.u)
            np.testing.assert_array_equal(r.vt, one.vt)
