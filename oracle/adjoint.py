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
