"""FP64 restatement of randomized::randomized_svd (TEST INFRASTRUCTURE ONLY).

Follows src/randomized/mod.rs:24-199: rank = min(r, m, n), l = rank + p,
Y = A Omega, q power iterations (Z = A^T Y, normalise Z, Y = A Z, normalise
Y), Q = qr(Y), B = Q^T A, SVD(B), U = Q U_B, truncation to
min(target_rank, #singular values).  The products use the oracle's C
restatement of the SMat methods (masked.rs:265-471, 616-733); the dense QR
and SVD use LAPACK through numpy in place of nalgebra 0.34.1 (third-party,
not in the container: sign conventions of Q and of singular vectors are
parity unpinned, so callers compare singular values and subspaces).

Centering (mean_center) uses the exact rank-one corrections
  (A - 1 mu^T) X   = A X   - 1 (mu^T X)
  (A - 1 mu^T)^T Y = A^T Y - mu (1^T Y)
  Q^T (A - 1 mu^T) = Q^T A - (Q^T 1) mu^T
which masked.rs:576-588 and 843-848 implement for the transposed products;
the forward product at masked.rs:496-509 subtracts (sum_j mu_j)(sum_i X_ic)
instead of sum_j mu_j X_jc (defect D7), corrected here.
"""
import numpy as np


def _mult(op, dense, transpose, means):
    out = op.multiply_with_dense(dense, transpose_self=transpose)
    if means is not None:
        if not transpose:
            out = out - np.ones((out.shape[0], 1)) @ (means[None, :] @ dense)
        else:
            out = out - means[:, None] @ dense.sum(axis=0)[None, :]
    return out


def _qr_q(a):
    # nalgebra qr().q(): thin Q, rows x min(rows, cols)
    return np.linalg.qr(a, mode="reduced")[0]


def _normalize_columns(a):
    # randomized/mod.rs:239-293: columns with norm <= 1e-10 left unscaled
    a = a.copy()
    norms = np.sqrt((a * a).sum(axis=0))
    keep = norms > 1e-10
    a[:, keep] /= norms[keep]
    return a


def randomized_svd(op, target_rank, n_oversamples, n_power_iters, normalizer="QR", mean_center=False,
                   omega=None):
    """Returns (u m x d, s d, vt d x n). omega (n x l) must be supplied."""
    m, n = op.nrows, op.ncols
    rank = min(target_rank, m, n)
    l = rank + n_oversamples
    means = op.column_means() if mean_center else None
    if omega is None:
        raise ValueError("the oracle needs the test matrix Omega (n x l)")
    omega = np.asarray(omega, dtype=np.float64)
    assert omega.shape == (n, l)
    y = _mult(op, omega, False, means)
    for _ in range(n_power_iters):
        z = _mult(op, y, True, means)
        if normalizer == "QR":
            z = _qr_q(z)
        elif normalizer == "LU":
            z = _normalize_columns(z)
        y = _mult(op, z, False, means)
        if normalizer == "QR":
            y = _qr_q(y)
        elif normalizer == "LU":
            y = _normalize_columns(y)
    q = _qr_q(y)
    b = op.multiply_transposed_by_dense(q)
    if means is not None:
        b = b - q.sum(axis=0)[:, None] @ means[None, :]
    u_b, s, vt = np.linalg.svd(b, full_matrices=False)
    u = q @ u_b
    k = min(target_rank, s.size)
    return u[:, :k], s[:k], vt[:k]
