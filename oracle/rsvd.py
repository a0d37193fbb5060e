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


def _mult(op, dense, transpose, means, f32=False):
    if f32:
        out = op.multiply_with_dense_f32(dense, transpose_self=transpose)
    else:
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
                   omega=None, dtype=np.float64):
    """Returns (u m x d, s d, vt d x n). omega (n x l) must be supplied.
    dtype np.float32 runs the reference's T = f32 instantiation: f32 sparse
    products (oracle C, f32 accumulators), f32 QR and SVD (LAPACK single
    precision for nalgebra's), no mean centring."""
    m, n = op.nrows, op.ncols
    rank = min(target_rank, m, n)
    l = rank + n_oversamples
    f32 = np.dtype(dtype) == np.float32
    if f32 and mean_center:
        raise ValueError("the f32 oracle path has no mean centring")
    means = op.column_means() if mean_center else None
    if omega is None:
        raise ValueError("the oracle needs the test matrix Omega (n x l)")
    omega = np.asarray(omega, dtype=dtype)
    assert omega.shape == (n, l)
    y = _mult(op, omega, False, means, f32)
    for _ in range(n_power_iters):
        z = _mult(op, y, True, means, f32)
        if normalizer == "QR":
            z = _qr_q(z)
        elif normalizer == "LU":
            z = _normalize_columns(z)
        y = _mult(op, z, False, means, f32)
        if normalizer == "QR":
            y = _qr_q(y)
        elif normalizer == "LU":
            y = _normalize_columns(y)
    q = _qr_q(y)
    b = op.multiply_transposed_by_dense_f32(q) if f32 else op.multiply_transposed_by_dense(q)
    if means is not None:
        b = b - q.sum(axis=0)[:, None] @ means[None, :]
    u_b, s, vt = np.linalg.svd(b, full_matrices=False)
    u = q @ u_b
    k = min(target_rank, s.size)
    return u[:, :k], s[:k], vt[:k]


def svd_flip(u=None, v=None, u_based_decision=True):
    """randomized::svd_flip (src/randomized/mod.rs:332-438), loop for loop
    on numpy copies (u m x k, v k x n). u_based_decision: the sign of each
    column j of u is that of its first entry of largest magnitude (strict >
    from 0, so an all-zero column keeps +1); the same signs multiply the first
    min(k, rows of v) rows of v. Otherwise the rows of v decide and scale the
    first columns of u. Returns (u, v, error message or None)."""
    if u is None and v is None:
        return None, None, "Both u and v cannot be None"
    u = None if u is None else np.array(u, dtype=np.float64)
    v = None if v is None else np.array(v, dtype=np.float64)
    if u_based_decision:
        if u is None:
            return None, v, "u cannot be None when u_based_decision is true"
        nrows, ncols = u.shape
        signs = [1.0] * ncols
        for j in range(ncols):
            max_abs, max_idx = 0.0, 0
            for i in range(nrows):
                if abs(u[i, j]) > max_abs:
                    max_abs, max_idx = abs(u[i, j]), i
            if u[max_idx, j] < 0.0:
                signs[j] = -1.0
        for j in range(ncols):
            for i in range(nrows):
                u[i, j] *= signs[j]
        if v is not None:
            for i in range(min(v.shape[0], len(signs))):
                for j in range(v.shape[1]):
                    v[i, j] *= signs[i]
    else:
        if v is None:
            return u, None, "v cannot be None when u_based_decision is false"
        nrows, ncols = v.shape
        signs = [1.0] * nrows
        for i in range(nrows):
            max_abs, max_idx = 0.0, 0
            for j in range(ncols):
                if abs(v[i, j]) > max_abs:
                    max_abs, max_idx = abs(v[i, j]), j
            if v[i, max_idx] < 0.0:
                signs[i] = -1.0
        for i in range(nrows):
            for j in range(ncols):
                v[i, j] *= signs[i]
        if u is not None:
            for j in range(min(u.shape[1], len(signs))):
                for i in range(u.shape[0]):
                    u[i, j] *= signs[j]
    return u, v, None


_PHI = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def gaussian_matrix(rows, cols, seed):
    """The build's counter-based replacement of generate_random_matrix
    (src/randomized/mod.rs:299-313; rand_distr's ziggurat stream is not
    reproducible without the crate): element (i, j) takes counter
    e = j * rows + i (the column-major fill order of DMatrix::from_fn) and
      a = mix64(key + (2e + 1) phi), b = mix64(key + (2e + 2) phi),
      u1 = ((a >> 11) + 0.5) 2^-53, u2 = (b >> 11) 2^-53,
      z  = sqrt(-2 ln u1) cos(2 pi u2)                    (Box-Muller)
    with phi = 0x9E3779B97F4A7C15, key = seed phi XOR 0x4F4D454741303031 and
    mix64 the splitmix64 finaliser. Returns the rows x cols FP64 matrix."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * _PHI) ^ np.uint64(0x4F4D454741303031)
        e = np.arange(rows * cols, dtype=np.uint64)
        a = _mix64(key + (np.uint64(2) * e + np.uint64(1)) * _PHI)
        b = _mix64(key + (np.uint64(2) * e + np.uint64(2)) * _PHI)
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return z.reshape(cols, rows).T.copy()
