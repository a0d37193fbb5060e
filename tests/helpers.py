"""Shared helpers for the tests: random sparse matrices and comparisons."""
import json
import os

import numpy as np
import scipy.sparse as sp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def random_csr(m, n, density, seed, dtype=np.float64, low=-10.0, high=10.0):
    rng = np.random.default_rng(seed)
    nnz = max(1, int(round(m * n * density)))
    flat = rng.choice(m * n, size=min(nnz, m * n), replace=False)
    rows, cols = np.divmod(flat, n)
    vals = rng.uniform(low, high, size=rows.size)
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, n)).astype(dtype)


def skewed_csr(m, n, seed, mean_row=20, max_row=3000, zipf=1.1, dtype=np.float64):
    """Rows with Pareto lengths and Zipf-popular columns: long rows and long columns."""
    rng = np.random.default_rng(seed)
    lens = np.minimum((mean_row / 2 / rng.uniform(size=m) ** 0.5).astype(np.int64), min(max_row, n))
    pop = 1.0 / np.arange(1, n + 1) ** zipf
    pop /= pop.sum()
    perm = rng.permutation(n)
    rows, cols = [], []
    for r in range(m):
        k = lens[r]
        if k == 0:
            continue
        c = rng.choice(n, size=k, replace=False, p=pop)
        rows.append(np.full(k, r))
        cols.append(perm[c])
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = rng.normal(size=rows.size)
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, n)).astype(dtype)


def max_principal_angle(a, b):
    """Largest principal angle (radians) between the column spaces of a and b,
    from its sine |Q_b - Q_a Q_a^T Q_b|_2 (accurate for small angles, where
    arccos of a cosine near 1 is not)."""
    qa, _ = np.linalg.qr(np.asarray(a, dtype=np.float64))
    qb, _ = np.linalg.qr(np.asarray(b, dtype=np.float64))
    r = qb - qa @ (qa.T @ qb)
    s = np.linalg.norm(r, 2)
    return float(np.arcsin(min(s, 1.0)))


def triplet_residual(a, u_cols, s, vt_rows):
    """max_i sqrt(|A v_i - s_i u_i|^2 + |A^T u_i - s_i v_i|^2) / s_i in FP64."""
    a = sp.csr_matrix(a, dtype=np.float64)
    u = np.asarray(u_cols, dtype=np.float64)
    v = np.asarray(vt_rows, dtype=np.float64).T
    s = np.asarray(s, dtype=np.float64)
    r1 = a @ v - u * s
    r2 = a.T @ u - v * s
    return float(np.max(np.sqrt((r1 * r1).sum(0) + (r2 * r2).sum(0)) / s))
