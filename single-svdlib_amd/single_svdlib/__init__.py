"""single_svdlib: MI355X-native drop-in for the sparse-SVD hot path of
single-svdlib v1.0.9 (module layout of src/lib.rs:1-9; legacy is src/legacy.rs).

    from single_svdlib import lanczos, randomized, CsrMatrix
    a = CsrMatrix.from_scipy(scipy_csr)          # uploaded once to the GPU
    rec = lanczos.svd_dim_seed(a, 10, 42)        # LAS2 on the GPU
    rec = randomized.randomized_svd(a, 50, 10, 3, randomized.PowerIterationNormalizer.QR, False, 42, False)
"""
from . import adjoint, error, lanczos, legacy, randomized  # noqa: F401
from .utils import (  # noqa: F401
    CooMatrix,
    CscMatrix,
    CsrMatrix,
    Diagnostics,
    SMat,
    SvdFloat,
    SvdRec,
    determine_chunk_size,
)

laczos = lanczos  # README.md:39 spells the module `laczos`
