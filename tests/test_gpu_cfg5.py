"""The adjoint workload at full size: gradients of sum_{i<32} s_i and of
|U_32^T y|^2 with respect to the CSR values of the 100,000 x 20,000 Zipf
matrix (nnz 9,998,163), checked on the 1,000 seeded nonzeros of the FP64 CPU
fixture (tests/golden/cfg5_grad.json: the sparse oracle's gradients, pinned
there by central differences at 8 of the nonzeros).

1. Kernel and CG parity: the GPU backward applied to FP64 CPU triplets of the
   same matrix (ARPACK + Rayleigh-Ritz on the CPU) in FP64, FP32 and BF16
   (BF16 factor panels in the SDDMM, FP32 accumulation).
2. End to end: factors from the GPU randomized SVD with the BF16 sketch
   ladder to a residual of 1e-6, then the FP32 and BF16 backward, against the
   fixture and against the central differences. The factors' own error sets
   the bar here: the top-32 subspace moves by about residual * s / (s_32 -
   s_33) ~ 2.5e-4 at this spectrum (s_32 = 204.47, s_33 = 203.66).
Errors are max |g - g_ref| over the sampled nonzeros, relative to max |g_ref|.
"""
import numpy as np
import pytest

import oracle
from helpers import golden

pytestmark = pytest.mark.gpu

SHA = "d144cf5c9fad6b4d5bdf934df9f043f4bd2c87680bd9ee5ad02dccf59501f966"


@pytest.fixture(scope="module")
def setup():
    import single_svdlib as sl
    from single_svdlib import _ffi, synth
    _ffi.context(0)
    fix = golden("cfg5_grad.json")
    a = synth.config2(np.float32)
    assert synth.checksum(a) == SHA == fix["sha256"]
    idx = np.asarray(fix["idx"])
    rows = np.repeat(np.arange(a.shape[0]), np.diff(a.indptr))
    assert np.array_equal(rows[idx], fix["rows"]) and np.array_equal(a.indices[idx], fix["cols"])
    y = np.random.default_rng(fix["y_seed"]).standard_normal(a.shape[0])
    op = sl.CsrMatrix.from_scipy(a, keep_bf16=True)
    return sl, a, op, fix, idx, y


def errors(g, fix, idx, loss):
    ref = np.asarray(fix["grad_" + loss])
    scale = np.max(np.abs(ref))
    e_fix = float(np.max(np.abs(g[idx] - ref)) / scale)
    pos = np.asarray(fix["fd_positions"])
    e_fd = float(np.max(np.abs(g[idx[pos]] - np.asarray(fix["fd_" + loss]))) / scale)
    return e_fix, e_fd


# measured (MI355X): f64 3.6e-13 / 3.0e-12 (CG to 1e-12), f32 6.4e-8 / 7.1e-5
# (CG to its 1e-5 default), bf16 3.2e-3 / 3.7e-3 (the factors' BF16 rounding,
# 2^-9 relative per entry, summed over 32 terms)
@pytest.mark.parametrize("prec,bar_sigma,bar_proj", [("f64", 1e-11, 1e-10), ("f32", 1e-6, 5e-4),
                                                     ("bf16", 1e-2, 1e-2)])
def test_cfg5_backward_of_exact_factors(setup, prec, bar_sigma, bar_proj):
    from single_svdlib.utils import Diagnostics, SvdRec
    sl, a, op, fix, idx, y = setup
    print(f"case cfg5 exact factors {prec}", flush=True)
    if not hasattr(test_cfg5_backward_of_exact_factors, "trip"):
        test_cfg5_backward_of_exact_factors.trip = oracle.adjoint.top_triplets(a.astype(np.float64), 40,
                                                                              tol=1e-14, polish=2)
    u, s, vt = test_cfg5_backward_of_exact_factors.trip
    assert abs(s[31] - fix["sigma"][31]) < 1e-9 * s[0]
    k = 32
    diag = Diagnostics(0, k, 0, False, 0, k, k, k, [0.0, 0.0], 0.0, 0)
    rec = SvdRec(d=k, u=np.ascontiguousarray(u[:, :k]), s=s[:k].copy(), vt=np.ascontiguousarray(vt[:k]),
                 diagnostics=diag)
    g1, _ = rec.backward(op, loss="sigma_sum", k=k, precision=prec)
    g2, info = rec.backward(op, loss="projection", k=k, y=y, precision=prec,
                            tol=1e-12 if prec == "f64" else 0.0)
    e1, f1 = errors(g1, fix, idx, "sigma_sum")
    e2, f2 = errors(g2, fix, idx, "projection")
    print(f"{prec}: sigma_sum {e1:.2e} (FD {f1:.2e}); projection {e2:.2e} (FD {f2:.2e}), "
          f"CG {info['cg_iterations']} iterations to {info['cg_relative_residual']:.1e}", flush=True)
    assert e1 < bar_sigma and f1 < bar_sigma + 1e-6
    assert e2 < bar_proj and f2 < bar_proj + 1e-6


def test_cfg5_end_to_end_from_bf16_sketch(setup):
    sl, a, op, fix, idx, y = setup
    print("case cfg5 end to end", flush=True)
    rec = sl.randomized.randomized_svd(op, 50, 10, 3, "QR", False, 42, False, tol=1e-6, max_power_iters=400,
                                       precision="bf16")
    assert rec.extra["residual_estimate"] <= 1e-6
    assert np.allclose(rec.s[:32], fix["sigma"][:32], rtol=1e-6, atol=0)
    for prec in ("f32", "bf16"):
        g1, _ = rec.backward(op, loss="sigma_sum", k=32, precision=prec)
        g2, info = rec.backward(op, loss="projection", k=32, y=y, precision=prec)
        e1, f1 = errors(g1, fix, idx, "sigma_sum")
        e2, f2 = errors(g2, fix, idx, "projection")
        print(f"end to end {prec}: sigma_sum {e1:.2e} (FD {f1:.2e}); projection {e2:.2e} (FD {f2:.2e})", flush=True)
        # measured: f32 1.3e-6 / 7.1e-5, bf16 3.2e-3 / 3.8e-3
        bar = 1e-3 if prec == "f32" else 1e-2
        assert max(e1, f1, e2, f2) < bar
