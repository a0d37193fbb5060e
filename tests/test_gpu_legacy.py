"""GPU legacy::svdLAS2 (single_svdlib.legacy, svdh_las2_legacy) against the
reference's known answers, which its own tests run on legacy::svd /
svd_dim_seed (legacy.rs:2159-2235), and against the FP64 oracle's legacy
mode on identical inputs (the same seeded start vector).

Bars as for lanczos: singular values within 1e-10 relative of the oracle,
singular vectors within a max principal angle of 1e-8 rad, diagnostics equal.
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


def compare_with_oracle(rec, r):
    assert rec.d == r["d"]
    assert np.allclose(rec.s, r["s"], rtol=1e-10, atol=0), np.max(np.abs(rec.s - r["s"]) / r["s"])
    assert max_principal_angle(rec.vt.T, r["vt"].T) < 1e-8
    assert max_principal_angle(rec.ut.T, r["u"].T) < 1e-8
    for k in DIAG_KEYS:
        assert getattr(rec.diagnostics, k) == r["diagnostics"][k], k
    assert rec.diagnostics.transposed == bool(r["diagnostics"]["transposed"])


def test_legacy_known_answers(lib):
    g = golden("legacy_kats.json")
    rec = lib.legacy.svd_dim_seed(lib.CsrMatrix.from_scipy(sp.csc_matrix(np.array(g["matrix_3x3"]))), 0, 3141)
    dump = g["svdlibrs_3x3_seed3141"]
    assert rec.d == 3 and np.allclose(rec.s, dump["s"], rtol=0, atol=1e-12)
    for i in range(3):
        for got, want in ((rec.ut[i], np.array(dump["ut"][i])), (rec.vt[i], np.array(dump["vt"][i]))):
            assert np.allclose(np.sign(got @ want) * got, want, atol=1e-11)
    for k in DIAG_KEYS:
        assert getattr(rec.diagnostics, k) == dump["diagnostics"][k], k
    assert np.allclose(rec.recompose(), g["matrix_3x3"], atol=1e-12)  # legacy.rs:2169-2181
    rec = lib.legacy.svd(lib.CooMatrix.from_scipy(sp.coo_matrix(np.array(g["basic_2x2"]["matrix"]))))
    assert rec.d == 2 and np.allclose(rec.s, g["basic_2x2"]["s"], atol=1e-12)
    rec = lib.legacy.svd(lib.CscMatrix.from_scipy(sp.identity(3, format="csc")))
    assert rec.d == g["identity_3x3"]["d"] and abs(rec.s[0] - 1.0) < 1e-12


def test_legacy_format_independence(lib):
    # legacy.rs:2158-2165: COO, CSC and CSR give equal results
    coo = sp.coo_matrix(([3.0, 4.0], ([1, 2], [0, 1])), shape=(4, 4))
    recs = [lib.legacy.svd_dim_seed(f, 3, 12345) for f in (lib.CooMatrix.from_scipy(coo),
                                                            lib.CscMatrix.from_scipy(coo.tocsc()),
                                                            lib.CsrMatrix.from_scipy(coo.tocsr()))]
    for r in recs[1:]:
        assert r.d == recs[0].d
        np.testing.assert_allclose(r.s, recs[0].s, rtol=1e-14)
        np.testing.assert_allclose(np.abs(r.vt), np.abs(recs[0].vt), atol=1e-14)


@pytest.mark.parametrize("shape,density,k,seed", [((60, 40), 0.2, 8, 42), ((40, 90), 0.15, 6, 7),
                                                   ((300, 200), 0.05, 12, 1), ((100, 100), 0.0098, 50, 42),
                                                   ((800, 500), 0.01, 20, 3)])
def test_legacy_matches_oracle(lib, shape, density, k, seed):
    a = random_csr(*shape, density, seed=seed)
    print(f"case legacy shape={shape} k={k}", flush=True)
    rec = lib.legacy.svd_dim_seed(lib.CsrMatrix.from_scipy(a), k, seed)
    st, r = oracle.las2_legacy(oracle.OracleOp(a), k, 0, random_seed=seed)
    assert st == 0
    compare_with_oracle(rec, r)
    assert triplet_residual(a, rec.ut.T, rec.s, rec.vt) < 1e-8
    s_true = np.linalg.svd(a.toarray(), compute_uv=False)
    assert np.allclose(rec.s, s_true[: rec.d], rtol=1e-10)


def test_legacy_and_current_agree(lib):
    # lib.rs:60-110: the same matrix through legacy and lanczos: equal rank,
    # singular values to 1e-12 absolute, recompositions to 1e-8 of max |a_ij|
    # (the oracle's two modes differ by 2.6e-13 and 7.9e-10 here)
    a = random_csr(400, 300, 0.03, seed=19)
    op = lib.CsrMatrix.from_scipy(a)
    old = lib.legacy.svd_dim_seed(op, 0, 17)
    new = lib.lanczos.svd_dim_seed(op, 0, 17)
    assert old.d == new.d
    np.testing.assert_allclose(old.s, new.s, rtol=0, atol=1e-12)
    rec_old = old.recompose()
    rec_new = new.u.T @ np.diag(new.s) @ new.vt
    assert np.max(np.abs(rec_old - rec_new)) < 1e-8 * np.abs(a).max()


def test_legacy_injected_start_vector(lib):
    a = random_csr(200, 150, 0.05, seed=11)
    start = np.random.default_rng(0).uniform(-1, 1, 150)
    rec = lib.legacy.svdLAS2(lib.CsrMatrix.from_scipy(a), 10, 0, [-1e-30, 1e-30], 1e-6, 99, start=start)
    st, r = oracle.las2_legacy(oracle.OracleOp(a), 10, 0, random_seed=99, start=start)
    assert st == 0
    compare_with_oracle(rec, r)


def test_legacy_errors(lib):
    from single_svdlib import error
    # a NaN entry reaches the tridiagonal: legacy imtqlb gives up after 30
    # iterations (ImtqlbError), as the oracle's legacy mode does
    a = random_csr(50, 40, 0.2, seed=1).tolil()
    a[3, 4] = np.nan
    a = a.tocsr()
    st, r = oracle.las2_legacy(oracle.OracleOp(a), 5, 0, random_seed=3)
    assert st == 1
    with pytest.raises(error.ImtqlbError, match="after 30 iterations"):
        lib.legacy.svd_dim_seed(lib.CsrMatrix.from_scipy(a), 5, 3)
    with pytest.raises(error.Las2Error, match="svdLAS2: insufficient dimensions"):
        lib.legacy.svd(lib.CsrMatrix.from_scipy(sp.csr_matrix(np.ones((1, 5)))))
    with pytest.raises(TypeError, match="f64"):
        lib.legacy.svd(lib.CsrMatrix.from_scipy(random_csr(30, 20, 0.2, seed=2).astype(np.float32)))
