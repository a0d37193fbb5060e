"""Write tests/golden/cfg5_grad.json: the FP64 CPU gradients of the adjoint
workload on the 100,000 x 20,000 Zipf CSR (the randomized-SVD benchmark
matrix, nnz 9,998,163), at 1,000 seeded nonzeros, plus central finite
differences of both losses at 8 of them.

Losses (k = 32):  L1 = sum_{i<k} s_i   and   L2 = |U_k^T y|^2, y ~ N(0, 1).

Everything here is the oracle (oracle/adjoint.py) in FP64 on the CPU:
  * the top 48 triplets from ARPACK (scipy svds, tol 1e-14) polished by block
    subspace iteration with Rayleigh-Ritz;
  * the sigma-sum gradient u_i[r] v_i[c] summed over i < 32;
  * the projection gradient with the complement systems
    (s_i^2 I - A A^T) w_i = P y solved explicitly over u_32 .. u_47 and by
    conjugate gradients (relative residual 1e-12) on the rest;
  * central differences (L(A + h E_e) - L(A - h E_e)) / 2h with h = 0.25,
    each loss from a fresh FP64 solve of the perturbed matrix (the
    perturbation moves the spectrum by about h |u_i[r] v_i[c]| ~ 1e-5, far
    below the 0.8 gap between s_32 and s_33, so the O(h^2) terms vanish at
    this scale and the difference is limited by the solves' 1e-14 accuracy).
The script asserts that the differences agree with the gradients, which pins
the sparse oracle at full size. Run time: about ten minutes on 8 cores.

Usage: python tests/golden/make_cfg5.py [--cache DIR]
(--cache keeps each finite-difference point in DIR as it completes, so an
interrupted run resumes where it stopped.)
"""
import json
import os
import sys
import time
from multiprocessing import get_context

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for _p in (ROOT, os.path.join(ROOT, "single-svdlib_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

OUT = os.path.join(HERE, "cfg5_grad.json")
SHA = "d144cf5c9fad6b4d5bdf934df9f043f4bd2c87680bd9ee5ad02dccf59501f966"
K = 32
KK = 48
Y_SEED = 20261205
IDX_SEED = 20261206
FD_SEED = 20261207
N_IDX = 1000
N_FD = 8
H = 0.25


def matrix():
    from single_svdlib import synth
    a32 = synth.config2(np.float32)
    assert synth.checksum(a32) == SHA, "generator output changed"
    return a32.astype(np.float64)


CACHE = None


def _set_cache(path):
    global CACHE
    CACHE = path


def fd_point(e):
    import oracle
    if CACHE:
        path = os.path.join(CACHE, f"fd_{e}.json")
        if os.path.exists(path):
            with open(path) as f:
                return tuple(json.load(f))
    a = matrix()
    y = np.random.default_rng(Y_SEED).standard_normal(a.shape[0])
    out = {}
    for sign in (1.0, -1.0):
        ap = a.copy()
        ap.data[e] += sign * H
        u, s, _ = oracle.adjoint.top_triplets(ap, KK, tol=1e-14, polish=3)
        c = u[:, :K].T @ y
        out[sign] = (float(s[:K].sum()), float(c @ c))
    res = ((out[1.0][0] - out[-1.0][0]) / (2 * H), (out[1.0][1] - out[-1.0][1]) / (2 * H))
    if CACHE:
        with open(os.path.join(CACHE, f"fd_{e}.json"), "w") as f:
            json.dump(list(res), f)
    return res


def main():
    import oracle
    t0 = time.time()
    a = matrix()
    m, n = a.shape
    y = np.random.default_rng(Y_SEED).standard_normal(m)
    idx = np.sort(np.random.default_rng(IDX_SEED).choice(a.nnz, N_IDX, replace=False))
    fd_pos = np.sort(np.random.default_rng(FD_SEED).choice(N_IDX, N_FD, replace=False))
    print(f"matrix {a.shape} nnz {a.nnz}", flush=True)
    tr = oracle.adjoint.top_triplets(a, KK, tol=1e-14, polish=3)
    u, s, vt = tr
    print(f"triplets in {time.time() - t0:.1f} s; s_32 {s[K - 1]:.6f} s_33 {s[K]:.6f}", flush=True)
    g1, _ = oracle.adjoint.gradient_sparse(a, "sigma_sum", K, y, triplets=tr)
    g2, info = oracle.adjoint.gradient_sparse(a, "projection", K, y, triplets=tr, tol=1e-12)
    print(f"gradients in {time.time() - t0:.1f} s; CG {info}", flush=True)
    c = u[:, :K].T @ y
    # spawned workers: forking after the threaded BLAS/OpenMP calls above
    # leaves the children waiting on locks the parent's threads held
    with get_context("spawn").Pool(4, initializer=_set_cache, initargs=(CACHE,)) as pool:
        fds = pool.map(fd_point, [int(idx[p]) for p in fd_pos])
    fd1 = np.array([f[0] for f in fds])
    fd2 = np.array([f[1] for f in fds])
    e1 = np.abs(fd1 - g1[idx[fd_pos]]) / np.abs(g1[idx]).max()
    e2 = np.abs(fd2 - g2[idx[fd_pos]]) / np.abs(g2[idx]).max()
    print(f"FD vs gradient (of max |g|): sigma_sum {e1.max():.2e}, projection {e2.max():.2e}", flush=True)
    assert e1.max() < 1e-4 and e2.max() < 1e-4, (e1, e2)
    rows = np.repeat(np.arange(m), np.diff(a.indptr))
    doc = {
        "about": "FP64 CPU gradients of sum_{i<32} s_i and |U_32^T y|^2 w.r.t. the CSR values of the "
                 "100,000 x 20,000 Zipf CSR (synth config 2, FP32 values), written by make_cfg5.py",
        "sha256": SHA,
        "k": K,
        "triplets_used": KK,
        "y_seed": Y_SEED,
        "idx_seed": IDX_SEED,
        "fd_seed": FD_SEED,
        "fd_h": H,
        "cg_iterations": int(info["cg_iterations"]),
        "cg_relative_residual": info["cg_relative_residual"],
        "sigma": [float(x) for x in s[:40]],
        "loss_sigma_sum": float(s[:K].sum()),
        "loss_projection": float(c @ c),
        "idx": [int(x) for x in idx],
        "rows": [int(x) for x in rows[idx]],
        "cols": [int(x) for x in a.indices[idx]],
        "grad_sigma_sum": [float(x) for x in g1[idx]],
        "grad_projection": [float(x) for x in g2[idx]],
        "fd_positions": [int(x) for x in fd_pos],
        "fd_sigma_sum": [float(x) for x in fd1],
        "fd_projection": [float(x) for x in fd2],
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)
    print(f"wrote {OUT} in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    if "--cache" in sys.argv:
        CACHE = sys.argv[sys.argv.index("--cache") + 1]
        os.makedirs(CACHE, exist_ok=True)
    main()
