"""Pin the CPU oracle (oracle/) before anything is compared with it.

Checks: the reference's own known answers (tests/golden/legacy_kats.json,
extracted from src/legacy.rs), library routines the algorithms reduce to
(math.hypot, LAPACK eigh_tridiagonal and SVD, scipy.sparse products),
invariants the reference tests assert (masked == physical subset, format
independence, power iterations improve the fit), and planted-defect checks
showing that each test would catch the reference defect it guards against.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle
from helpers import golden, max_principal_angle, random_csr, triplet_residual

EPS = np.finfo(np.float64).eps


# ---------------------------------------------------------------- pythag (D1)

def _pythag_reference_new(a, b):
    """src/lanczos/mod.rs:361-380 as written (update r = (s/u)^2 drops r)."""
    n = max(abs(a), abs(b))
    if n <= 0:
        return 0.0
    p = n
    r = (min(abs(a), abs(b)) / p) ** 2
    t = 4.0 + r
    for _ in range(200):
        if abs(t - 4.0) < EPS:
            break
        s = r / t
        u = 1.0 + 2.0 * s
        p = p * u
        r = (s / u) ** 2
        t = 4.0 + r
    return p


def test_pythag_matches_hypot():
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.normal(size=500) * 10.0 ** rng.integers(-8, 8, 500), [0.0, 3.0, 4.0, 1e-300]])
    for a, b in zip(vals, np.roll(vals, 7)):
        h = math.hypot(a, b)
        assert abs(oracle.pythag(a, b) - h) <= 4 * EPS * h


def test_pythag_check_catches_d1():
    # the same tolerance rejects the reference's current form
    errs = [abs(_pythag_reference_new(a, b) - math.hypot(a, b)) / math.hypot(a, b)
            for a, b in [(3.0, 4.0), (1.0, 1.0), (2.0, 7.0)]]
    assert max(errs) > 1e-4


# ------------------------------------------------------- tridiagonal QL solvers

def _tridiag(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n), rng.normal(size=n - 1)


@pytest.mark.parametrize("n", [2, 3, 7, 20, 41])
def test_imtqlb_eigenvalues_and_bounds(n):
    d, off = _tridiag(n, n)
    w, z = sla.eigh_tridiagonal(d, off)
    ev, bnd = oracle.imtqlb(d, np.concatenate([[0.0], off]))
    scale = np.abs(w).max()
    assert np.max(np.abs(ev - w)) <= 1e-13 * scale
    # bnd = first components of the eigenvectors (SVDLIBC's Ritz bound factor)
    assert np.max(np.abs(np.abs(bnd) - np.abs(z[0, :]))) <= 1e-10


def _imtqlb_stale_p(n, d, e, maxit=100):
    """imtqlb with p = d[l], f = bnd[l] read once per l (mod.rs:455-457, D12)."""
    d = list(d)
    e = list(e)
    bnd = [0.0] * n
    sf = math.sqrt(n)
    bnd[0] = 1.0
    last = n - 1
    for i in range(1, n):
        e[i - 1] = e[i]
    e[last] = 0.0
    for l in range(n):
        it = 0
        p, f = d[l], bnd[l]
        while it <= maxit:
            m = l
            while m < last:
                test = abs(d[m]) + abs(d[m + 1])
                if abs(e[m]) <= EPS * 100 * max(test, 1) * sf:
                    break
                m += 1
            if m == l:
                i = l
                while i >= 1 and p < d[i - 1]:
                    d[i] = d[i - 1]
                    bnd[i] = bnd[i - 1]
                    i -= 1
                d[i] = p
                bnd[i] = f
                break
            if it == maxit:
                e[l] = 0.0
                break
            it += 1
            g = (d[l + 1] - p) / (2 * e[l])
            r = math.hypot(g, 1.0)
            g = d[m] - p + e[l] / (g + (r if g >= 0 else -r))
            s = c = 1.0
            p = 0.0
            i = m - 1
            while True:
                f = s * e[i]
                b = c * e[i]
                r = math.hypot(f, g)
                e[i + 1] = r
                s, c = f / r, g / r
                g = d[i + 1] - p
                r = (d[i] - g) * s + 2 * c * b
                p = s * r
                d[i + 1] = g + p
                g = c * r - b
                f = bnd[i + 1]
                bnd[i + 1] = s * bnd[i] + c * f
                bnd[i] = c * bnd[i] - s * f
                if i == l:
                    break
                i -= 1
            d[l] -= p
            e[l] = g
            e[m] = 0.0
    return np.array(d)


def test_imtqlb_check_catches_d12():
    d, off = _tridiag(12, 3)
    w = sla.eigh_tridiagonal(d, off, eigvals_only=True)
    bad = np.sort(_imtqlb_stale_p(12, d, np.concatenate([[0.0], off])))
    assert np.max(np.abs(bad - w)) > 1e-6 * np.abs(w).max()


@pytest.mark.parametrize("n", [2, 5, 16, 60])
def test_imtql2_eigenpairs(n):
    d, off = _tridiag(n, 100 + n)
    w, z = sla.eigh_tridiagonal(d, off)
    ev, zz = oracle.imtql2(d, np.concatenate([[0.0], off]))
    assert np.max(np.abs(ev - w)) <= 1e-13 * np.abs(w).max()
    cos = np.abs(np.sum(zz * z, axis=0))
    assert np.min(cos) >= 1 - 1e-10


# ----------------------------------------------------------- SMat operators

@pytest.mark.parametrize("shape,density", [((50, 30), 0.2), ((200, 700), 0.03), ((1, 5), 0.6), ((31, 1), 0.5)])
@pytest.mark.parametrize("nthreads", [1, 3])
def test_opa_matches_scipy(shape, density, nthreads):
    a = random_csr(*shape, density, seed=shape[0])
    op = oracle.OracleOp(a, nthreads=nthreads)
    rng = np.random.default_rng(1)
    x = rng.normal(size=shape[1])
    y = rng.normal(size=shape[0])
    assert np.allclose(op.opa(x), a @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(op.opa(y, transposed=True), a.T @ y, rtol=1e-13, atol=1e-13)
    assert op.nnz() == a.nnz


def test_masked_ops_equal_physical_subset():
    a = random_csr(120, 90, 0.1, seed=4)
    cols = np.array([1, 3, 5, 7, 40, 41, 89])
    op = oracle.OracleOp(a, columns=cols, nthreads=2)
    sub = a[:, cols]
    rng = np.random.default_rng(2)
    x = rng.normal(size=cols.size)
    y = rng.normal(size=120)
    assert op.ncols == cols.size and op.nnz() == sub.nnz
    assert np.allclose(op.opa(x), sub @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(op.opa(y, True), sub.T @ y, rtol=1e-13, atol=1e-13)
    X = rng.normal(size=(cols.size, 6))
    Y = rng.normal(size=(120, 6))
    assert np.allclose(op.multiply_with_dense(X), sub @ X, rtol=1e-12, atol=1e-12)
    assert np.allclose(op.multiply_with_dense(Y, True), sub.T @ Y, rtol=1e-12, atol=1e-12)
    assert np.allclose(op.multiply_transposed_by_dense(Y), Y.T @ sub, rtol=1e-12, atol=1e-12)
    assert np.allclose(op.column_means(), np.asarray(sub.mean(axis=0)).ravel(), rtol=1e-13, atol=1e-13)


def test_transposed_panel_reads_column_major_rows():
    # D5: the transposed product must read row i of a column-major panel as
    # dense[(i, c)]; a row-major read of the same buffer gives another answer.
    a = random_csr(40, 25, 0.3, seed=9)
    op = oracle.OracleOp(a)
    Y = np.random.default_rng(3).normal(size=(40, 5))
    good = op.multiply_with_dense(Y, True)
    assert np.allclose(good, a.T @ Y, rtol=1e-12)
    scrambled = np.asfortranarray(Y).ravel(order="F").reshape(40, 5)  # row-major view of column-major data
    assert not np.allclose(a.T @ scrambled, good)


# ------------------------------------------------------------------ LAS2 KATs

def test_las2_legacy_3x3_known_answers():
    g = golden("legacy_kats.json")
    a = sp.csc_matrix(np.array(g["matrix_3x3"]))
    st, r = oracle.las2(oracle.OracleOp(a), 0, 0, random_seed=3141)
    assert st == 0
    dump = g["svdlibrs_3x3_seed3141"]
    assert r["d"] == 3
    assert np.allclose(r["s"], g["asserted_s_3x3"]["s"], rtol=0, atol=1e-12)
    assert np.allclose(r["s"], dump["s"], rtol=0, atol=1e-12)
    # vectors equal the documented ones up to a common sign per triplet
    for i in range(3):
        for got, want in ((r["u"][i], np.array(dump["ut"][i])), (r["vt"][i], np.array(dump["vt"][i]))):
            sgn = np.sign(got @ want)
            assert np.allclose(sgn * got, want, atol=1e-11)
    rec = r["u"].T @ np.diag(r["s"]) @ r["vt"]
    assert np.allclose(rec, g["matrix_3x3"], atol=1e-12)
    dg = r["diagnostics"]
    for key in ("non_zero", "dimensions", "iterations", "lanczos_steps", "ritz_values_stabilized",
                "significant_values", "singular_values", "random_seed"):
        assert dg[key] == dump["diagnostics"][key], key
    assert bool(dg["transposed"]) == dump["diagnostics"]["transposed"]
    assert dg["kappa"] == pytest.approx(dump["diagnostics"]["kappa"])


def test_las2_basic_2x2_and_identity():
    g = golden("legacy_kats.json")
    st, r = oracle.las2(oracle.OracleOp(sp.csr_matrix(np.array(g["basic_2x2"]["matrix"]))), 0, 0, random_seed=1)
    assert st == 0 and r["d"] == 2
    assert np.allclose(r["s"], g["basic_2x2"]["s"], atol=1e-12)
    assert np.allclose(r["u"].T @ np.diag(r["s"]) @ r["vt"], g["basic_2x2"]["matrix"], atol=1e-12)
    st, r = oracle.las2(oracle.OracleOp(sp.identity(3, format="csc")), 0, 0, random_seed=7)
    assert st == 0 and r["d"] == g["identity_3x3"]["d"]
    assert abs(r["s"][0] - 1.0) < 1e-12


def test_las2_format_independence():
    # legacy.rs:2158-2165: COO, CSC and CSR inputs give identical results
    coo = sp.coo_matrix(([3.0, 4.0], ([1, 2], [0, 1])), shape=(4, 4))
    outs = [oracle.las2(oracle.OracleOp(f), 3, 0, random_seed=12345) for f in (coo, coo.tocsc(), coo.tocsr())]
    for st, r in outs:
        assert st == 0
    for st, r in outs[1:]:
        assert np.array_equal(r["s"], outs[0][1]["s"])
        assert np.array_equal(r["vt"], outs[0][1]["vt"])


@pytest.mark.parametrize("shape,density,k", [((60, 40), 0.2, 8), ((40, 90), 0.15, 6), ((300, 200), 0.05, 12)])
def test_las2_matches_lapack(shape, density, k):
    a = random_csr(*shape, density, seed=shape[1])
    st, r = oracle.las2(oracle.OracleOp(a), k, 0, random_seed=42)
    assert st == 0 and r["d"] >= 2
    s_true = np.linalg.svd(a.toarray(), compute_uv=False)
    d = r["d"]
    assert np.all(np.diff(r["s"]) <= 0)
    assert np.allclose(r["s"], s_true[:d], rtol=1e-10, atol=0)
    assert triplet_residual(a, r["u"].T, r["s"], r["vt"]) < 1e-8
    assert r["diagnostics"]["transposed"] == int(shape[1] >= 1.2 * shape[0])


def test_las2_masked_equals_physical_subset():
    # masked.rs:914-1000
    a = random_csr(14, 10, 0.3, seed=42, low=0.1, high=10.0)
    cols = [1, 3, 5, 7]
    st1, r1 = oracle.las2(oracle.OracleOp(a, columns=cols), 0, 0, random_seed=5)
    st2, r2 = oracle.las2(oracle.OracleOp(a[:, cols]), 0, 0, random_seed=5)
    assert st1 == 0 and st2 == 0 and r1["d"] == r2["d"]
    assert np.allclose(np.sort(r1["s"]), np.sort(r2["s"]), atol=1e-10)


def test_las2_all_true_mask_equals_csr():
    # lib.rs:111-154
    a = random_csr(50, 30, 0.2, seed=12345)
    st1, r1 = oracle.las2(oracle.OracleOp(a), 0, 0, random_seed=12345)
    st2, r2 = oracle.las2(oracle.OracleOp(a, mask=np.ones(30, bool)), 0, 0, random_seed=12345)
    assert st1 == 0 and st2 == 0 and r1["d"] == r2["d"]
    assert np.allclose(r1["s"], r2["s"], rtol=1e-3)


def test_las2_very_sparse_runs():
    # lib.rs:156-164: 100 x 100 at 0.98% density, dims 50
    a = random_csr(100, 100, 0.0098, seed=42)
    st, r = oracle.las2(oracle.OracleOp(a), 50, 0, random_seed=42)
    assert st == 0
    s_true = np.linalg.svd(a.toarray(), compute_uv=False)
    assert np.allclose(r["s"], s_true[: r["d"]], rtol=1e-8)


def test_las2_error_paths():
    st, r = oracle.las2(oracle.OracleOp(sp.csr_matrix(np.ones((1, 5)))), 0, 0, random_seed=1)
    assert st == 5  # Las2Error: insufficient dimensions
    st, r = oracle.las2(oracle.OracleOp(sp.csr_matrix((6, 4))), 0, 0, random_seed=1)
    assert st == 2  # StartvError: A = 0


def test_las2_f32_epsilon_mode():
    a = random_csr(80, 60, 0.1, seed=3).astype(np.float32).astype(np.float64)
    st, r = oracle.las2(oracle.OracleOp(a), 6, 0, random_seed=9, eps=oracle.F32_EPS)
    assert st == 0
    s_true = np.linalg.svd(a.toarray(), compute_uv=False)
    assert np.allclose(r["s"], s_true[: r["d"]], rtol=1e-5)


# ------------------------------------------------- legacy::svdLAS2 (f64)

@pytest.mark.parametrize("n", [2, 3, 7, 20, 41])
def test_imtqlb_legacy_eigenvalues_and_bounds(n):
    d, off = _tridiag(n, 200 + n)
    w, z = sla.eigh_tridiagonal(d, off)
    ev, bnd = oracle.imtqlb_legacy(d, np.concatenate([[0.0], off]))
    assert np.max(np.abs(ev - w)) <= 1e-13 * np.abs(w).max()
    assert np.max(np.abs(np.abs(bnd) - np.abs(z[0, :]))) <= 1e-10


@pytest.mark.parametrize("n", [2, 5, 16, 61])
def test_imtql2_legacy_eigenpairs(n):
    d, off = _tridiag(n, 300 + n)
    w, z = sla.eigh_tridiagonal(d, off)
    ev, zz = oracle.imtql2_legacy(d, np.concatenate([[0.0], off]))
    assert np.max(np.abs(ev - w)) <= 1e-13 * np.abs(w).max()
    t = np.diag(d) + np.diag(off, 1) + np.diag(off, -1)
    assert np.max(np.abs(t @ zz - zz * ev)) <= 1e-12 * np.abs(w).max()
    assert np.max(np.abs(zz.T @ zz - np.eye(n))) <= 1e-12


@pytest.mark.parametrize("legacy", [False, True])
def test_imtql2_keeps_the_absolute_underflow_test(legacy):
    # legacy.rs:1634 and mod.rs:970 end a QL sweep early when compare(r, 0),
    # i.e. |r| < 2.2e-16 absolute, where EISPACK tests r == 0. On this
    # tridiagonal one rotation radius is 5.1e-17: the eigenvalues stay exact
    # but some eigenvectors lose accuracy (residual 5.5e-2). The restatement
    # keeps the reference's test; an exact-zero test would leave 6e-15.
    n = 60
    d, off = _tridiag(n, 360)
    t = np.diag(d) + np.diag(off, 1) + np.diag(off, -1)
    e = np.concatenate([[0.0], off])
    ev, zz = oracle.imtql2_legacy(d, e) if legacy else oracle.imtql2(d, e)
    assert np.max(np.abs(ev - sla.eigh_tridiagonal(d, off, eigvals_only=True))) <= 1e-13 * np.abs(ev).max()
    res = np.abs(t @ zz - zz * ev).max(axis=0)
    assert 0.05 < res.max() < 0.06
    assert np.sum(res > 1e-10) == 6


def test_imtql_legacy_stop_after_30_iterations():
    # legacy.rs:1015-1019, 1610-1614: a tridiagonal whose convergence test can
    # never pass (a NaN entry) ends in the error after 30 QL iterations
    with pytest.raises(RuntimeError, match="after 30 iterations"):
        oracle.imtqlb_legacy([np.nan, 1.0, 2.0], [0.0, 1.0, 1.0])
    with pytest.raises(RuntimeError, match="after 30 iterations"):
        oracle.imtql2_legacy([np.nan, 1.0, 2.0], [0.0, 1.0, 1.0])


def test_las2_legacy_mode_known_answers():
    # the reference's tests run these on legacy::svd / svd_dim_seed themselves
    # (legacy.rs:2169-2235; the 3x3 dump is svdlibrs' output for seed 3141)
    g = golden("legacy_kats.json")
    st, r = oracle.las2_legacy(oracle.OracleOp(sp.csc_matrix(np.array(g["matrix_3x3"]))), 0, 0, random_seed=3141)
    assert st == 0 and r["d"] == 3
    dump = g["svdlibrs_3x3_seed3141"]
    assert np.allclose(r["s"], dump["s"], rtol=0, atol=1e-12)
    for i in range(3):
        for got, want in ((r["u"][i], np.array(dump["ut"][i])), (r["vt"][i], np.array(dump["vt"][i]))):
            assert np.allclose(np.sign(got @ want) * got, want, atol=1e-11)
    for key in ("non_zero", "dimensions", "iterations", "lanczos_steps", "ritz_values_stabilized",
                "significant_values", "singular_values", "random_seed"):
        assert r["diagnostics"][key] == dump["diagnostics"][key], key
    st, r = oracle.las2_legacy(oracle.OracleOp(sp.csr_matrix(np.array(g["basic_2x2"]["matrix"]))), 0, 0,
                               random_seed=1)
    assert st == 0 and r["d"] == 2 and np.allclose(r["s"], g["basic_2x2"]["s"], atol=1e-12)
    st, r = oracle.las2_legacy(oracle.OracleOp(sp.identity(3, format="csc")), 0, 0, random_seed=7)
    assert st == 0 and r["d"] == g["identity_3x3"]["d"] and abs(r["s"][0] - 1.0) < 1e-12


@pytest.mark.parametrize("shape,density,k", [((60, 40), 0.2, 8), ((40, 90), 0.15, 6), ((300, 200), 0.05, 12),
                                             ((100, 100), 0.0098, 50)])
def test_las2_legacy_mode_matches_lapack_and_current(shape, density, k):
    # LAPACK truth, and lib.rs:60-110: legacy and lanczos agree on d, the
    # singular values and the recomposition
    a = random_csr(*shape, density, seed=shape[1])
    st, r = oracle.las2_legacy(oracle.OracleOp(a), k, 0, random_seed=42)
    st2, r2 = oracle.las2(oracle.OracleOp(a), k, 0, random_seed=42)
    assert st == 0 and st2 == 0 and r["d"] >= 2 and r["d"] == r2["d"]
    s_true = np.linalg.svd(a.toarray(), compute_uv=False)
    assert np.allclose(r["s"], s_true[: r["d"]], rtol=1e-10, atol=0)
    assert np.allclose(r["s"], r2["s"], rtol=1e-10, atol=0)
    assert triplet_residual(a, r["u"].T, r["s"], r["vt"]) < 1e-8
    rec = r["u"].T @ np.diag(r["s"]) @ r["vt"]
    rec2 = r2["u"].T @ np.diag(r2["s"]) @ r2["vt"]
    assert np.max(np.abs(rec - rec2)) < 1e-8 * max(1.0, np.abs(rec).max())


def test_las2_legacy_mode_errors():
    # a NaN in A reaches the tridiagonal: legacy imtqlb stops after 30
    # iterations (ImtqlbError, status 1) where the current imtqlb forces
    # convergence and imtql2 fails later (status 4)
    a = random_csr(50, 40, 0.2, seed=1).tolil()
    a[3, 4] = np.nan
    a = a.tocsr()
    st, r = oracle.las2_legacy(oracle.OracleOp(a), 5, 0, random_seed=3)
    assert st == 1 and "30 iterations" in r["error"]
    st, r = oracle.las2_legacy(oracle.OracleOp(sp.csr_matrix(np.ones((1, 5)))), 0, 0, random_seed=1)
    assert st == 5 and "svdLAS2: insufficient dimensions" in r["error"]


def test_start_vector_closed_form():
    v = oracle.start_vector(42, 5)
    phi = 0x9E3779B97F4A7C15
    mask = (1 << 64) - 1
    key = ((42 * phi) & mask) ^ 0x5356444C49423031
    for i in range(5):
        z = (key + (i + 1) * phi) & mask
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        assert v[i] == 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0


# ------------------------------------------------------------------ randomized

def test_rsvd_oracle_exact_on_low_rank():
    rng = np.random.default_rng(0)
    a = sp.csr_matrix(rng.normal(size=(80, 6)) @ rng.normal(size=(6, 50)))
    op = oracle.OracleOp(a)
    om = rng.normal(size=(50, 10))
    u, s, vt = oracle.randomized_svd(op, 6, 4, 0, "QR", omega=om)
    s_true = np.linalg.svd(a.toarray(), compute_uv=False)[:6]
    assert np.allclose(s, s_true, rtol=1e-10)
    assert triplet_residual(a, u, s, vt) < 1e-9


def test_rsvd_oracle_power_iterations_improve_fit():
    # randomized/mod.rs:692-777
    rng = np.random.default_rng(987)
    dense = rng.uniform(-1, 1, (100, 10)) @ rng.uniform(-1, 1, (10, 50)) + rng.uniform(-0.01, 0.01, (100, 50))
    op = oracle.OracleOp(sp.csr_matrix(dense))
    om = rng.normal(size=(50, 15))
    errs = []
    for q in (0, 1, 3, 5):
        u, s, vt = oracle.randomized_svd(op, 10, 5, q, "QR", omega=om)
        errs.append(np.linalg.norm(dense - u @ np.diag(s) @ vt) / np.linalg.norm(dense))
    assert min(errs[1:]) < 0.9 * errs[0] or errs[0] < 1e-3


@pytest.mark.parametrize("norm", ["QR", "LU", "NONE"])
def test_rsvd_oracle_normalizers_span_same_space(norm):
    a = random_csr(300, 120, 0.05, seed=11)
    op = oracle.OracleOp(a)
    om = np.random.default_rng(5).normal(size=(120, 15))
    u0, s0, vt0 = oracle.randomized_svd(op, 10, 5, 2, "QR", omega=om)
    u1, s1, vt1 = oracle.randomized_svd(op, 10, 5, 2, norm, omega=om)
    assert np.allclose(s0, s1, rtol=1e-8)


def test_rsvd_oracle_mean_center_matches_dense():
    # randomized/mod.rs:585-649 shape: rank-3 factors plus column means, so the
    # centred matrix has rank <= 3 and l = 6 recovers it exactly
    rng = np.random.default_rng(123)
    dense = rng.uniform(-1, 1, (30, 3)) @ rng.uniform(-1, 1, (3, 10)) + np.arange(10) * 2.0
    op = oracle.OracleOp(sp.csr_matrix(dense))
    om = rng.normal(size=(10, 6))
    u, s, vt = oracle.randomized_svd(op, 3, 3, 2, "QR", mean_center=True, omega=om)
    centered = dense - dense.mean(axis=0)
    s_true = np.linalg.svd(centered, compute_uv=False)[:3]
    assert np.allclose(s, s_true, rtol=1e-6)


def _adjoint_case(seed=3):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    a = sp.random(40, 25, density=0.3, random_state=seed, format="csr")
    a.data = rng.normal(size=a.nnz)
    y = rng.normal(size=40)
    return a, y


@pytest.mark.parametrize("loss", ["sigma_sum", "projection"])
def test_adjoint_oracle_matches_finite_differences(loss):
    # the oracle gradient is pinned by central differences of the loss itself
    a, y = _adjoint_case()
    k = 5
    g = oracle.adjoint.gradient_on_pattern(a, loss, k, y)
    rng = np.random.default_rng(1)
    picks = rng.choice(a.nnz, 20, replace=False)
    h = 1e-6
    for e in picks:
        ap, am = a.copy(), a.copy()
        ap.data[e] += h
        am.data[e] -= h
        fd = (oracle.adjoint.svd_loss(ap.toarray(), loss, k, y) - oracle.adjoint.svd_loss(am.toarray(), loss, k, y)) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * max(1.0, np.max(np.abs(g))), (e, fd, g[e])


def test_adjoint_oracle_planted_mistake_fails_fd():
    # dropping the u_i (A^T w_i)^T half of the complement term is caught by FD
    a, y = _adjoint_case()
    k = 5
    dense = a.toarray()
    u, s, vt = np.linalg.svd(dense, full_matrices=True)
    c = u[:, :k].T @ y
    coef = u.T @ y
    lam = np.zeros(40)
    lam[:25] = s ** 2
    g = np.zeros_like(dense)
    for i in range(k):
        w = u[:, k:] @ (coef[k:] / (s[i] ** 2 - lam[k:]))
        g += 2.0 * c[i] * s[i] * np.outer(w, vt[i])  # planted: second term missing
    rows = np.repeat(np.arange(40), np.diff(a.indptr))
    gp = g[rows, a.indices]
    h = 1e-6
    worst = 0.0
    for e in range(0, a.nnz, 7):
        ap, am = a.copy(), a.copy()
        ap.data[e] += h
        am.data[e] -= h
        fd = (oracle.adjoint.svd_loss(ap.toarray(), "projection", k, y)
              - oracle.adjoint.svd_loss(am.toarray(), "projection", k, y)) / (2 * h)
        worst = max(worst, abs(fd - gp[e]))
    assert worst > 1e-3


def _coo_case(m, n, nnz, seed, integer=True, dtype=np.float64):
    rng = np.random.default_rng(seed)
    r = rng.integers(0, m, nnz)
    c = rng.integers(0, n, nnz)
    v = (rng.integers(-7, 8, nnz) if integer else rng.normal(size=nnz)).astype(dtype)
    return r, c, v


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_coo_to_csr_oracle_matches_scipy(dtype):
    # integer-valued entries: every order of summation gives the same sums, so
    # structure and values must equal scipy's COO -> CSR exactly
    for m, n, nnz, seed in ((50, 40, 600, 0), (1, 7, 30, 1), (300, 2, 900, 2), (20, 30, 0, 3)):
        r, c, v = _coo_case(m, n, nnz, seed, dtype=dtype)
        rp, oc, ov = oracle.coo_to_csr(m, n, r, c, v)
        a = sp.coo_matrix((v, (r, c)), shape=(m, n)).tocsr()
        a.sum_duplicates()
        a.sort_indices()
        assert ov.dtype == dtype
        assert np.array_equal(rp, a.indptr) and np.array_equal(oc, a.indices) and np.array_equal(ov, a.data)
        if nnz > 40:
            assert oc.size < nnz  # the case holds duplicates, so dropping or keeping them both would fail


def test_coo_to_csr_oracle_sums_duplicates_in_input_order():
    # FP64 sums are not associative: 1e16 + 1 - 1e16 depends on the order. The
    # oracle adds each run left to right in input order.
    r = np.array([0, 1, 0, 0, 1])
    c = np.array([2, 0, 2, 2, 0])
    v = np.array([1e16, 3.0, 1.0, -1e16, 4.0])
    rp, oc, ov = oracle.coo_to_csr(2, 3, r, c, v)
    assert list(rp) == [0, 1, 2] and list(oc) == [2, 0]
    assert ov[0] == (1e16 + 1.0) - 1e16 and ov[1] == 7.0


def test_coo_to_csr_oracle_product_equals_triplet_sum():
    # SMat for CooMatrix (mod.rs:1657-1740) adds every triplet into y
    r, c, v = _coo_case(200, 150, 3000, 5, integer=False)
    rp, oc, ov = oracle.coo_to_csr(200, 150, r, c, v)
    x = np.random.default_rng(1).normal(size=150)
    y = np.zeros(200)
    np.add.at(y, r, v * x[c])
    a = sp.csr_matrix((ov, oc, rp), shape=(200, 150))
    assert np.allclose(a @ x, y, rtol=1e-12, atol=1e-12)


def test_coo_to_csr_oracle_rejects_out_of_range():
    with pytest.raises(ValueError):
        oracle.coo_to_csr(3, 3, np.array([0, 3]), np.array([0, 0]), np.array([1.0, 2.0]))
    with pytest.raises(ValueError):
        oracle.coo_to_csr(3, 3, np.array([0, 1]), np.array([0, -1]), np.array([1.0, 2.0]))


# ------------------------------------------------------- centred products (f1)

@pytest.mark.parametrize("masked", [False, True])
def test_centered_products_match_dense(masked):
    a = random_csr(60, 40, 0.2, seed=9)
    cols = np.sort(np.random.default_rng(2).choice(40, 25, replace=False)) if masked else np.arange(40)
    op = oracle.OracleOp(a, columns=cols) if masked else oracle.OracleOp(a)
    dense = a.toarray()[:, cols]
    mu = op.column_means()
    assert np.allclose(mu, dense.mean(axis=0), rtol=1e-13, atol=1e-14)
    c = dense - mu[None, :]
    rng = np.random.default_rng(4)
    x = rng.normal(size=(len(cols), 7))
    y = rng.normal(size=(60, 7))
    assert np.allclose(op.multiply_with_dense_centered(x, mu), c @ x, rtol=1e-12, atol=1e-12)
    assert np.allclose(op.multiply_with_dense_centered(y, mu, True), c.T @ y, rtol=1e-12, atol=1e-12)
    assert np.allclose(op.multiply_transposed_by_dense_centered(y, mu), y.T @ c, rtol=1e-12, atol=1e-12)
    # the reference's forward form (masked.rs:496-509, defect D7) subtracts
    # (sum_j mu_j)(sum_i X_ic) from every row; the dense check rejects it
    d7 = op.multiply_with_dense(x) - mu.sum() * x.sum(axis=0)[None, :]
    assert not np.allclose(d7, c @ x, rtol=1e-6, atol=1e-6)


# ------------------------------------------------------------ svd_flip (f3)

def test_svd_flip_oracle_invariants():
    rng = np.random.default_rng(6)
    u = rng.normal(size=(30, 5))
    v = rng.normal(size=(5, 12))
    u[3, 2] = -50.0  # a negative dominant entry
    u[:, 4] = 0.0    # an all-zero column keeps its sign
    fu, fv, err = oracle.svd_flip(u, v, True)
    assert err is None
    for j in range(5):
        col = fu[:, j]
        assert col[np.argmax(np.abs(col))] >= 0
        assert np.array_equal(np.abs(col), np.abs(u[:, j]))
    assert np.array_equal(fu[:, 2], -u[:, 2]) and np.array_equal(fv[2], -v[2])
    assert np.array_equal(fu[:, 4], u[:, 4])
    # flipping twice changes nothing more; the product u v is unchanged
    gu, gv, _ = oracle.svd_flip(fu, fv, True)
    assert np.array_equal(gu, fu) and np.array_equal(gv, fv)
    assert np.allclose(fu @ fv, u @ v, rtol=0, atol=1e-12)
    # v-based decision
    hu, hv, _ = oracle.svd_flip(u, v, False)
    for i in range(5):
        assert hv[i, np.argmax(np.abs(hv[i]))] >= 0
    assert np.allclose(hu @ hv, u @ v, rtol=0, atol=1e-12)
    # ties: the first entry of largest magnitude decides (strict >)
    t = np.array([[2.0], [-2.0]])
    assert oracle.svd_flip(t, None, True)[0][0, 0] == 2.0
    t = np.array([[-2.0], [2.0]])
    assert oracle.svd_flip(t, None, True)[0][0, 0] == 2.0
    assert oracle.svd_flip(None, None, True)[2] == "Both u and v cannot be None"
    assert oracle.svd_flip(None, v, True)[2] == "u cannot be None when u_based_decision is true"
    assert oracle.svd_flip(u, None, False)[2] == "v cannot be None when u_based_decision is false"


@pytest.mark.parametrize("decide_u", [True, False])
def test_host_svd_flip_matches_oracle(decide_u):
    from single_svdlib.error import Las2Error
    from single_svdlib.randomized import svd_flip
    rng = np.random.default_rng(8)
    for k_u, k_v in ((6, 6), (6, 4), (4, 6)):
        u = rng.normal(size=(40, k_u))
        v = rng.normal(size=(k_v, 25))
        u[5, 1] = -9.0
        v[2, 3] = -9.0
        ru, rv, err = oracle.svd_flip(u, v, decide_u)
        hu, hv = u.copy(), v.copy()
        svd_flip(hu, hv, decide_u)
        assert err is None
        assert np.array_equal(hu, ru) and np.array_equal(hv, rv)
    with pytest.raises(Las2Error, match="Both u and v cannot be None"):
        svd_flip(None, None)
    with pytest.raises(Las2Error, match="u cannot be None"):
        svd_flip(None, np.ones((2, 2)), True)
    with pytest.raises(Las2Error, match="v cannot be None"):
        svd_flip(np.ones((2, 2)), None, False)


# ------------------------------------------------- Gaussian test matrix (a19)

def test_gaussian_matrix_distribution_and_layout():
    from scipy import stats
    g = oracle.gaussian_matrix(20000, 6, 42)
    assert g.shape == (20000, 6)
    x = g.ravel()
    assert abs(x.mean()) < 4 / np.sqrt(x.size)
    assert abs(x.var() - 1.0) < 4 * np.sqrt(2.0 / x.size)
    assert stats.kstest(x, "norm").pvalue > 1e-3
    # columns are independent streams: no correlation above the noise level
    c = np.corrcoef(g.T)
    assert np.max(np.abs(c - np.eye(6))) < 5 / np.sqrt(20000)
    # column-major counter order: the first rows x 1 block is the
    # rows x 1 matrix of the same seed, and another seed differs
    assert np.array_equal(oracle.gaussian_matrix(20000, 1, 42)[:, 0], g[:, 0])
    assert not np.allclose(oracle.gaussian_matrix(20000, 1, 43)[:, 0], g[:, 0])


# ------------------------------------------------ sparse adjoint path (a26)

@pytest.mark.parametrize("loss", ["sigma_sum", "projection"])
@pytest.mark.parametrize("extra", [0, 4])
def test_adjoint_sparse_path_matches_dense(loss, extra):
    a, y = _adjoint_case(seed=11)
    k = 5
    gd = oracle.adjoint.gradient_on_pattern(a, loss, k, y)
    tr = oracle.adjoint.top_triplets(a, k + extra)
    gs, info = oracle.adjoint.gradient_sparse(a, loss, k, y, triplets=tr, tol=1e-13)
    assert np.max(np.abs(gs - gd)) < 1e-10 * np.max(np.abs(gd))
    if loss == "projection":
        assert info["cg_relative_residual"] <= 1e-13


def test_adjoint_sparse_path_planted_mistake_fails():
    # dropping the explicit part of the complement (the u_k .. u_{kk-1}
    # directions) must show against the dense gradient
    a, y = _adjoint_case(seed=11)
    k = 5
    gd = oracle.adjoint.gradient_on_pattern(a, "projection", k, y)
    u, s, vt = oracle.adjoint.top_triplets(a, k + 4)
    w, _, _ = oracle.adjoint.complement_solutions(a, y, u, s, k, tol=1e-13)
    w_bad, _, _ = oracle.adjoint.complement_solutions(a, y, u[:, :k], s[:k], k, tol=1e-13)
    assert np.max(np.abs(w - w_bad)) < 1e-9 * np.max(np.abs(w))  # both solve the same systems
    # but a complement projected onto range(I - U_kk U_kk^T) only misses terms
    w_miss = w - u[:, k:] @ (u[:, k:].T @ w)
    rows = np.repeat(np.arange(a.shape[0]), np.diff(a.indptr))
    c = u[:, :k].T @ y
    atw = a.T @ w_miss
    g_bad = (np.einsum("ei,ei->e", (w_miss * (2 * c * s[:k]))[rows], vt[:k, a.indices].T)
             + np.einsum("ei,ei->e", (u[:, :k] * 2 * c)[rows], atw[a.indices]))
    assert np.max(np.abs(g_bad - gd)) > 1e-3 * np.max(np.abs(gd))


def test_cfg5_fixture_is_consistent():
    # the committed FP64 gradients agree with their own central differences
    # (each from fresh FP64 solves of the perturbed matrix), and the sampled
    # nonzeros are where the fixture says they are in the regenerated matrix
    from single_svdlib import synth
    fix = golden("cfg5_grad.json")
    a = synth.config2(np.float32)
    assert synth.checksum(a) == fix["sha256"]
    idx = np.asarray(fix["idx"])
    rows = np.repeat(np.arange(a.shape[0]), np.diff(a.indptr))
    assert np.array_equal(rows[idx], fix["rows"]) and np.array_equal(a.indices[idx], fix["cols"])
    pos = np.asarray(fix["fd_positions"])
    for loss in ("sigma_sum", "projection"):
        g = np.asarray(fix["grad_" + loss])
        fd = np.asarray(fix["fd_" + loss])
        assert np.max(np.abs(g[pos] - fd)) < 1e-5 * np.max(np.abs(g))
    # the sigma-sum loss is the sum of the top 32 values listed
    assert abs(sum(fix["sigma"][:32]) - fix["loss_sigma_sum"]) < 1e-9 * fix["loss_sigma_sum"]
