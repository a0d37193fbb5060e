"""GPU adjoint (svdh_backward) against the FP64 oracle gradient
(oracle/adjoint.py, itself pinned by finite differences).

Bars: with exact FP64 factors, the sigma-sum gradient to 1e-12 and the
projection-loss gradient to 1e-8 of max |grad| (CG tolerance 1e-12); FP32
factors and arithmetic to 1e-4; end to end from the GPU's own randomized SVD
to 1e-6 (FP64).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import random_csr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)
    return single_svdlib


def exact_rec(lib, a, k):
    """An SvdRec holding numpy's FP64 top-k triplets (randomized layout)."""
    from single_svdlib.utils import Diagnostics, SvdRec
    u, s, vt = np.linalg.svd(a.toarray(), full_matrices=False)
    diag = Diagnostics(0, k, 0, False, 0, k, k, k, [0.0, 0.0], 0.0, 0)
    return SvdRec(d=k, u=np.ascontiguousarray(u[:, :k]), s=s[:k].copy(), vt=np.ascontiguousarray(vt[:k]),
                  diagnostics=diag)


def case(seed=5, m=300, n=120):
    rng = np.random.default_rng(seed)
    a = random_csr(m, n, 0.08, seed=seed)
    y = rng.normal(size=m)
    return a, y


@pytest.mark.parametrize("loss", ["sigma_sum", "projection"])
def test_backward_f64_matches_oracle(lib, loss):
    a, y = case()
    k = 8
    rec = exact_rec(lib, a, k)
    g, info = lib.adjoint.backward(lib.CsrMatrix.from_scipy(a), rec, loss=loss, k=k, y=y, tol=1e-12)
    ref = oracle.adjoint.gradient_on_pattern(a, loss, k, y)
    scale = np.max(np.abs(ref))
    tol = 1e-12 if loss == "sigma_sum" else 1e-8
    assert np.max(np.abs(g - ref)) <= tol * scale, np.max(np.abs(g - ref)) / scale
    assert np.isclose(info["loss"], oracle.adjoint.svd_loss(a.toarray(), loss, k, y), rtol=1e-12)
    if loss == "projection":
        assert info["cg_relative_residual"] <= 1e-12 and info["cg_iterations"] > 0


@pytest.mark.parametrize("loss", ["sigma_sum", "projection"])
def test_backward_f32(lib, loss):
    a, y = case(seed=9)
    k = 6
    rec = exact_rec(lib, a, k)
    g, _ = lib.adjoint.backward(lib.CsrMatrix.from_scipy(a.astype(np.float32)), rec, loss=loss, k=k, y=y,
                                precision="f32", grad_dtype=np.float32)
    ref = oracle.adjoint.gradient_on_pattern(a.astype(np.float32).astype(np.float64), loss, k, y)
    assert g.dtype == np.float32
    assert np.max(np.abs(g - ref)) <= 1e-4 * np.max(np.abs(ref))


def test_backward_masked_equals_physical_subset(lib):
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a, y = case(seed=11, m=250, n=150)
    cols = np.sort(np.random.default_rng(3).choice(150, 90, replace=False))
    sub = a[:, cols].tocsr()
    k = 5
    rec = exact_rec(lib, sub, k)
    base = lib.CsrMatrix.from_scipy(a)
    gm, _ = lib.adjoint.backward(MaskedCSRMatrix.with_columns(base, cols), rec, loss="projection", k=k, y=y,
                                 tol=1e-12)
    gp, _ = lib.adjoint.backward(lib.CsrMatrix.from_scipy(sub), rec, loss="projection", k=k, y=y, tol=1e-12)
    keep = np.isin(a.indices, cols)
    assert gm.shape == (a.nnz,)
    assert np.all(gm[~keep] == 0.0)
    # the kept entries of the full pattern, in CSR order, are the subset's entries in CSR order
    assert np.max(np.abs(gm[keep] - gp)) <= 1e-10 * np.max(np.abs(gp))


def test_backward_end_to_end_from_gpu_solves(lib):
    # factors from the GPU's own randomized SVD (adaptive to 1e-12) and LAS2
    # (u in the d x m layout) give the oracle gradient
    a, y = case(seed=13, m=400, n=150)
    op = lib.CsrMatrix.from_scipy(a)
    k = 6
    rec = lib.randomized.randomized_svd(op, k, 14, 2, "QR", False, 1, False, tol=1e-12, max_power_iters=400)
    ref = oracle.adjoint.gradient_on_pattern(a, "projection", k, y)
    g, _ = rec.backward(op, loss="projection", k=k, y=y, tol=1e-12)
    assert np.max(np.abs(g - ref)) <= 1e-6 * np.max(np.abs(ref))
    las = lib.lanczos.svd_dim_seed(op, k, 7)
    assert las.u.shape == (k, 400)
    g2, _ = las.backward(op, loss="sigma_sum", k=k)
    ref2 = oracle.adjoint.gradient_on_pattern(a, "sigma_sum", k)
    assert np.max(np.abs(g2 - ref2)) <= 1e-8 * np.max(np.abs(ref2))


@pytest.mark.parametrize("layout", ["columns", "rows"])
@pytest.mark.parametrize("fdt,precision", [(np.float64, "f64"), (np.float32, "f32"), (np.float64, "f32"),
                                           (np.float32, "f64")])
def test_backward_factor_layouts_and_hand_over(lib, layout, fdt, precision):
    # The sigma-sum gradient is the SDDMM sum_i u_i[r] v_i[c] of whatever
    # factors the record holds, so any factors pin the factor upload (vectors
    # as columns, m x d, or as rows, d x m as LAS2 returns them; FP64 or FP32
    # arrays into FP64 or FP32 panels) and the gradient's hand-over. 140,000
    # rows of 10 FP64 factors (11 MB) and 1.1 M nonzeros (8.8 MB of FP64
    # gradient) take the staged paths; the direct paths are covered by the
    # small cases above.
    from single_svdlib.utils import Diagnostics, SvdRec
    m, n, d, k = 140_000, 3_000, 10, 7
    print(f"case layouts {layout} {np.dtype(fdt).name} {precision} {m}x{n} d={d} k={k}", flush=True)
    rng = np.random.default_rng(17)
    a = random_csr(m, n, 0.0026, seed=17)
    u = rng.normal(size=(m, d)).astype(fdt)
    vt = rng.normal(size=(d, n)).astype(fdt)
    s = np.sort(rng.uniform(1.0, 2.0, size=d))[::-1].astype(fdt)
    diag = Diagnostics(0, d, 0, False, 0, d, d, d, [0.0, 0.0], 0.0, 0)
    uu = np.ascontiguousarray(u if layout == "columns" else u.T)
    rec = SvdRec(d=d, u=uu, s=s, vt=np.ascontiguousarray(vt), diagnostics=diag)
    mat = a if precision == "f64" else a.astype(np.float32)
    g, info = lib.adjoint.backward(lib.CsrMatrix.from_scipy(mat), rec, loss="sigma_sum", k=k, precision=precision)
    coo = a.tocoo()
    u64, vt64 = u.astype(np.float64), vt.astype(np.float64)
    ref = np.einsum("ij,ij->i", u64[coo.row, :k], vt64[:k, coo.col].T)
    # the gradient in CSR order of the matrix (tocoo keeps it)
    scale = np.max(np.abs(ref))
    tol = 1e-12 if (precision == "f64" and fdt == np.float64) else 1e-5
    assert g.shape == (a.nnz,)
    assert np.max(np.abs(g - ref)) <= tol * scale, np.max(np.abs(g - ref)) / scale
    assert np.isclose(info["loss"], float(np.sum(s[:k].astype(np.float64))), rtol=1e-6)
