"""FP64 dense restatement of the SVD adjoint used by svdh_backward.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). The reference has no
backward pass; this module is the checker for the build's own adjoint and is
pinned by central finite differences of the loss itself (tests/test_oracle.py).

For A = U S V^T (numpy's full SVD, small dense inputs):
  sigma_sum:  L = sum_{i<k} s_i,     dL/dA = U_k V_k^T
  projection: L = |U_k^T y|^2,       dL/dA = sum_i 2 c_i (s_i w_i v_i^T + u_i (A^T w_i)^T)
              with c_i = u_i^T y and w_i = sum_{j >= k} u_j (u_j^T y) / (s_i^2 - lambda_j),
              lambda_j = s_j^2 for j < min(m, n), 0 for the left null space.
"""
import numpy as np


def svd_loss(a, loss, k, y=None):
    u, s, _ = np.linalg.svd(np.asarray(a, dtype=np.float64), full_matrices=False)
    if loss == "sigma_sum":
        return float(s[:k].sum())
    c = u[:, :k].T @ y
    return float(c @ c)


def gradient_dense(a, loss, k, y=None):
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    u, s, vt = np.linalg.svd(a, full_matrices=True)
    v = vt.T
    if loss == "sigma_sum":
        return u[:, :k] @ v[:, :k].T
    r = min(m, n)
    lam = np.zeros(m)
    lam[:r] = s[:r] ** 2
    c = u[:, :k].T @ y
    coef = u.T @ y
    g = np.zeros_like(a)
    for i in range(k):
        w = u[:, k:] @ (coef[k:] / (s[i] ** 2 - lam[k:]))
        g += 2.0 * c[i] * (s[i] * np.outer(w, v[:, i]) + np.outer(u[:, i], a.T @ w))
    return g


def gradient_on_pattern(csr, loss, k, y=None):
    """grad[e] = dL/dA[r_e, c_e] in CSR order."""
    g = gradient_dense(csr.toarray(), loss, k, y)
    rows = np.repeat(np.arange(csr.shape[0]), np.diff(csr.indptr))
    return g[rows, csr.indices]


# ---------------------------------------------------------------------------
# Sparse FP64 path (matrices too large for a dense SVD): the same gradients
# from the top triplets of the sparse matrix and a CPU conjugate-gradient
# solve of the complement systems. Pinned against the dense path above on
# small inputs (tests/test_oracle.py), which is itself pinned by central
# finite differences.
# ---------------------------------------------------------------------------

def top_triplets(csr, k, tol=1e-14, polish=3, seed=1):
    """Top-k singular triplets (u m x k, s k, vt k x n) of a sparse matrix in
    FP64, descending: ARPACK (scipy.sparse.linalg.svds) on A^T A, then
    `polish` rounds of block subspace iteration with Rayleigh-Ritz on the
    k-column block (each round: X <- A^T (A X), orthonormalise, SVD of A X)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    a = sp.csr_matrix(csr, dtype=np.float64)
    u, s, vt = sla.svds(a, k=k, tol=tol, random_state=seed)
    order = np.argsort(-s)
    v = vt[order].T
    for _ in range(polish):
        v, _ = np.linalg.qr(a.T @ (a @ v))
    # Rayleigh-Ritz: A V = U_r S W^T, so A (V W) = U_r S
    av = a @ v
    u_r, s, wt = np.linalg.svd(av, full_matrices=False)
    v = v @ wt.T
    return u_r, s, np.ascontiguousarray(v.T)


def _pattern_rows(csr):
    return np.repeat(np.arange(csr.shape[0]), np.diff(csr.indptr))


def complement_solutions(csr, y, u, s, k, tol=1e-12, maxiter=20000):
    """W (m x k): w_i solves (s_i^2 I - A A^T) w_i = P y on range(P),
    P = I - U_kk U_kk^T for the kk >= k given triplets. The part in
    span(u_k .. u_{kk-1}) is explicit, sum_j u_j (u_j^T y) / (s_i^2 - s_j^2);
    the part in range(I - U_kk U_kk^T) comes from k conjugate-gradient
    solves run together on the CPU (the operator is positive definite there
    when s_{k-1} > s_kk), to the relative residual tol.
    Returns (W, iterations, max relative residual)."""
    import scipy.sparse as sp
    a = sp.csr_matrix(csr, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    kk = u.shape[1]
    assert kk >= k and s.size >= kk
    y = np.asarray(y, dtype=np.float64)

    def proj(x):  # x - U_kk (U_kk^T x), column-wise
        return x - u @ (u.T @ x)

    shifts = s[:k] ** 2
    coef = u[:, k:kk].T @ y  # u_j^T y for the explicit part
    w_exp = np.zeros((a.shape[0], k))
    for i in range(k):
        if kk > k:
            w_exp[:, i] = u[:, k:kk] @ (coef / (shifts[i] - s[k:kk] ** 2))
    b = proj(y[:, None])[:, 0]
    bn = float(np.linalg.norm(b))
    w = np.zeros((a.shape[0], k))
    if bn == 0.0:
        return w_exp, 0, 0.0
    r = np.repeat(b[:, None], k, axis=1)
    p = r.copy()
    rho = (r * r).sum(axis=0)
    done = np.zeros(k, dtype=bool)
    it = 0
    rel = 1.0
    while it < maxiter:
        q = proj(p * shifts[None, :] - a @ (a.T @ p))
        gam = (p * q).sum(axis=0)
        alpha = np.where(done | (gam <= 0), 0.0, rho / np.where(gam == 0, 1.0, gam))
        w += p * alpha[None, :]
        r -= q * alpha[None, :]
        rr = (r * r).sum(axis=0)
        it += 1
        relc = np.sqrt(rr) / bn
        rel = float(relc.max())
        done |= relc <= tol
        if done.all():
            break
        beta = np.where(done, 0.0, rr / rho)
        rho = np.where(done, rho, rr)
        p = r + p * beta[None, :]
    return proj(w) + w_exp, it, rel


def gradient_sparse(csr, loss, k, y=None, triplets=None, tol=1e-12):
    """grad[e] = dL/dA[r_e, c_e] in CSR order, FP64, from the top triplets
    (u m x kk, s, vt kk x n; kk >= k, computed by top_triplets when not given)
    of the sparse matrix:
      sigma_sum:  grad[e] = sum_{i<k} u_i[r_e] v_i[c_e]
      projection: grad[e] = sum_{i<k} 2 c_i (s_i w_i[r_e] v_i[c_e] + u_i[r_e] (A^T w_i)[c_e])
    with c_i = u_i^T y and w_i from complement_solutions.
    Returns (grad, info) with info = {"cg_iterations", "cg_relative_residual"}."""
    import scipy.sparse as sp
    a = sp.csr_matrix(csr, dtype=np.float64)
    if triplets is None:
        triplets = top_triplets(a, k + 8)
    u, s, vt = triplets
    rows = _pattern_rows(a)
    cols = a.indices
    info = {"cg_iterations": 0, "cg_relative_residual": 0.0}
    if loss == "sigma_sum":
        return np.einsum("ei,ei->e", u[rows, :k], vt[:k, cols].T), info
    c = u[:, :k].T @ y
    w, it, rel = complement_solutions(a, y, u, s, k, tol=tol)
    atw = a.T @ w
    left1 = w * (2.0 * c * s[:k])[None, :]
    left2 = u[:, :k] * (2.0 * c)[None, :]
    g = np.einsum("ei,ei->e", left1[rows], vt[:k, cols].T) + np.einsum("ei,ei->e", left2[rows], atw[cols])
    info.update(cg_iterations=it, cg_relative_residual=rel)
    return g, info


def loss_sparse(csr, loss, k, y=None, kk=None, tol=1e-14):
    """The loss of the sparse matrix from its top triplets (top_triplets with
    kk >= k columns)."""
    u, s, _ = top_triplets(csr, kk or k + 8, tol=tol)
    if loss == "sigma_sum":
        return float(s[:k].sum())
    c = u[:, :k].T @ y
    return float(c @ c)
