"""CPU oracle for the sparse-SVD hot path of single-svdlib v1.0.9.

TEST INFRASTRUCTURE ONLY. This package is the checker of the HIP product
path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import it, and only as the checker or the timed CPU baseline; the product
(single-svdlib_amd/) never imports, links or calls it.

Parity status of each function (pinned = checked by tests/ against a golden
vector, a closed form or a library routine chosen so a plausible mistake fails):
  opa / multiply_with_dense / multiply_transposed_by_dense / column_means
      pinned: scipy.sparse FP64 products and dense numpy on small inputs
  pythag      pinned: math.hypot
  imtqlb, imtql2  pinned: scipy.linalg.eigh_tridiagonal (LAPACK)
  las2        pinned: the reference's KATs (legacy.rs:137-283, 2184-2235),
              numpy.linalg.svd on random sparse matrices, and invariants
  las2_legacy, imtqlb_legacy, imtql2_legacy  pinned: the same KATs, which
              the reference runs on legacy::svdLAS2 itself; numpy.linalg.svd;
              scipy.linalg.eigh_tridiagonal; the 30-iteration error forced by
              a tridiagonal that needs more
  start_vector pinned by its own closed form (the reference's ChaCha12
              stream is parity unpinned: no Rust toolchain here)
  randomized_svd (oracle/rsvd.py) pinned: numpy.linalg.svd truth for
              rank-deficient/low-rank inputs and the Halko bound invariants;
              nalgebra's Householder QR / SVD sign conventions are parity
              unpinned (not in the container), so comparisons use subspace
              angles and singular values.
  svd_flip    pinned: the sign invariant (largest-magnitude entry of each
              decided vector positive), involution and error cases
  gaussian_matrix  pinned: moments and a Kolmogorov-Smirnov test against
              N(0, 1) and the column-major counter order (the reference's
              rand_distr stream is parity unpinned: crate not present)
  adjoint (oracle/adjoint.py, no reference counterpart) pinned: central
              finite differences of the loss through numpy.linalg.svd; the
              sparse FP64 path (top triplets + CPU CG) against the dense one,
              and at full size against central differences (make_cfg5.py)
  centred products (multiply_*_centered): dense (A - 1 mu^T) X in numpy
  coo_to_csr  pinned: scipy.sparse COO -> CSR (structure exact; sums of
              duplicates exact on integer-valued inputs, 1 ulp otherwise),
              the triplet-sum product of SMat for CooMatrix (mod.rs:1657-1740),
              and a sequential left-to-right sum of each duplicate run.
"""
from .pyoracle import (  # noqa: F401
    OracleOp,
    las2,
    las2_legacy,
    imtqlb,
    imtql2,
    imtqlb_legacy,
    imtql2_legacy,
    pythag,
    start_vector,
    determine_chunk_size,
    coo_to_csr,
    F64_EPS,
    F32_EPS,
)
from .rsvd import gaussian_matrix, randomized_svd, svd_flip  # noqa: F401
from . import adjoint  # noqa: F401
