"""COO and CSC input formats converted on the device (svdh_coo_create,
svdh_csc_create) against the oracle's COO -> CSR (entries in (row, column)
order, repeated pairs summed left to right in input order).

The device operator is read back exactly through A I: every output element
is one stored value times 1.0 plus exact zeros, so the dense matrix it gives
must equal the oracle's CSR bit for bit (values of repeated pairs included).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import max_principal_angle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sl():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)
    return single_svdlib


def dense_of(op, n, dtype):
    m = op.nrows()
    eye = np.eye(n, dtype=dtype)
    out = np.zeros((m, n), dtype=dtype)
    for c0 in range(0, n, 64):
        w = min(64, n - c0)
        part = np.zeros((m, w), dtype=dtype)
        op.multiply_with_dense(eye[:, c0:c0 + w], part, False)
        out[:, c0:c0 + w] = part
    return out


def oracle_dense(m, n, r, c, v):
    rp, oc, ov = oracle.coo_to_csr(m, n, r, c, v)
    return sp.csr_matrix((ov, oc, rp), shape=(m, n)).toarray(), ov.size


def coo_case(m, n, nnz, seed, dtype, integer=False):
    rng = np.random.default_rng(seed)
    r = rng.integers(0, m, nnz)
    c = rng.integers(0, n, nnz)
    v = (rng.integers(-9, 10, nnz) if integer else rng.normal(size=nnz)).astype(dtype)
    return r, c, v


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,nnz,seed", [(300, 120, 5000, 1), (1, 90, 400, 2), (500, 1, 700, 3), (64, 64, 64, 4),
                                          (3000, 100, 100_000, 5)])
def test_coo_matches_oracle_bit_for_bit(sl, dtype, m, n, nnz, seed):
    r, c, v = coo_case(m, n, nnz, seed, dtype)
    print(f"case coo m={m} n={n} nnz={nnz} {np.dtype(dtype).name}", flush=True)
    op = sl.CooMatrix(r, c, v, (m, n))
    want, u = oracle_dense(m, n, r, c, v)
    assert op.nnz() == u
    if nnz >= 400:
        assert u < nnz  # repeated pairs were there and were merged
    got = dense_of(op, n, dtype)
    assert got.dtype == want.dtype and np.array_equal(got, want)


def test_coo_sums_repeated_pairs_in_input_order(sl):
    # 1e16 + 1 - 1e16 depends on the order of the additions
    r = np.array([0, 1, 0, 0, 1, 2])
    c = np.array([2, 0, 2, 2, 0, 1])
    v = np.array([1e16, 3.0, 1.0, -1e16, 4.0, 5.0])
    op = sl.CooMatrix(r, c, v, (3, 3))
    got = dense_of(op, 3, np.float64)
    assert got[0, 2] == (1e16 + 1.0) - 1e16 and got[1, 0] == 7.0 and got[2, 1] == 5.0 and op.nnz() == 3


def test_coo_edge_shapes_and_errors(sl):
    op = sl.CooMatrix(np.zeros(0), np.zeros(0), np.zeros(0), (4, 5))
    assert op.nnz() == 0 and np.array_equal(dense_of(op, 5, np.float64), np.zeros((4, 5)))
    with pytest.raises(ValueError):
        sl.CooMatrix(np.array([0, 4]), np.array([0, 0]), np.array([1.0, 2.0]), (4, 5))
    with pytest.raises(ValueError):
        sl.CooMatrix(np.array([0, 1]), np.array([0, -2]), np.array([1.0, 2.0]), (4, 5))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csc_matches_oracle_bit_for_bit(sl, dtype):
    rng = np.random.default_rng(7)
    m, n = 400, 150
    a = sp.random(m, n, density=0.05, random_state=8, format="csc", dtype=np.float64).astype(dtype)
    # a column with a repeated row index and two empty columns
    rows = [a.indices[a.indptr[j]:a.indptr[j + 1]] for j in range(n)]
    vals = [a.data[a.indptr[j]:a.indptr[j + 1]] for j in range(n)]
    rows[3] = np.array([5, 9, 5]); vals[3] = np.array([1.5, 2.0, -0.25], dtype=dtype)
    rows[10] = np.array([], dtype=np.int64); vals[10] = np.array([], dtype=dtype)
    rows[n - 1] = np.array([], dtype=np.int64); vals[n - 1] = np.array([], dtype=dtype)
    indptr = np.concatenate([[0], np.cumsum([len(x) for x in rows])])
    ri = np.concatenate(rows).astype(np.int64)
    dv = np.concatenate(vals).astype(dtype)
    op = sl.CscMatrix(indptr, ri, dv, (m, n))
    cols = np.repeat(np.arange(n), np.diff(indptr))
    want, u = oracle_dense(m, n, ri, cols, dv)
    assert op.nnz() == u
    assert np.array_equal(dense_of(op, n, dtype), want)
    with pytest.raises(ValueError):
        bad = indptr.copy(); bad[5] = bad[6] + 1
        sl.CscMatrix(bad, ri, dv, (m, n))
    assert rng is not None


def test_solvers_on_coo_and_csc_equal_csr(sl):
    # the same matrix through the three input formats: identical operators, so
    # identical LAS2 and randomized results
    rng = np.random.default_rng(11)
    a = sp.random(2000, 700, density=0.02, random_state=12, format="csr", dtype=np.float64)
    coo = a.tocoo()
    perm = rng.permutation(coo.nnz)  # triplets in a scrambled order
    ops = [sl.CsrMatrix.from_scipy(a), sl.CooMatrix(coo.row[perm], coo.col[perm], coo.data[perm], a.shape),
           sl.CscMatrix.from_scipy(a.tocsc())]
    recs = [sl.lanczos.svd_dim_seed(op, 8, 17) for op in ops]
    for rec in recs[1:]:
        assert np.array_equal(rec.s, recs[0].s)
        assert rec.diagnostics.lanczos_steps == recs[0].diagnostics.lanczos_steps
    om = np.random.default_rng(3).normal(size=(700, 18))
    rs = [sl.randomized.randomized_svd(op, 8, 10, 2, "QR", False, 5, False, omega=om) for op in ops]
    for rec in rs[1:]:
        assert np.array_equal(rec.s, rs[0].s)
    u, s, vt = np.linalg.svd(a.toarray(), full_matrices=False)
    assert np.allclose(recs[0].s, s[:8], rtol=1e-10)
    assert max_principal_angle(recs[1].vt.T, vt[:8].T) < 1e-8
