/*
 * svdh: MI355X-native (gfx950) sparse SVD hot path of single-svdlib v1.0.9.
 *
 * C ABI. Every entry point returns an int32_t svdh_status; on failure
 * svdh_last_error() returns a thread-local message. The ABI never aborts.
 * A context owns one HIP stream and is not thread-safe: callers that share
 * a context across threads must serialise calls (the Rust binding wraps it in
 * a Mutex, which is what lets its operator type be Sync).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   SMat<T> trait                          src/utils.rs:22-33
 *   CsrMatrix as SMat                      src/lanczos/mod.rs:1506-1655
 *   MaskedCSRMatrix                        src/lanczos/masked.rs:13-875
 *   lanczos::{svd, svd_dim, svd_dim_seed, svd_las2}   src/lanczos/mod.rs:27-189
 *   randomized::randomized_svd             src/randomized/mod.rs:24-199
 *   SvdRec, Diagnostics                    src/utils.rs:43-79
 *   SvdLibError                            src/error.rs:3-22
 */
#ifndef SVDH_H
#define SVDH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1..6 map one-to-one onto SvdLibError (src/error.rs:5-21). */
typedef enum {
  SVDH_OK = 0,
  SVDH_ERR_IMTQLB = 1,  /* SvdLibError::ImtqlbError */
  SVDH_ERR_STARTV = 2,  /* SvdLibError::StartvError */
  SVDH_ERR_STPONE = 3,  /* SvdLibError::StponeError */
  SVDH_ERR_IMTQL2 = 4,  /* SvdLibError::Imtql2Error */
  SVDH_ERR_LAS2 = 5,    /* SvdLibError::Las2Error */
  SVDH_ERR_NDARRAY = 6, /* SvdLibError::NDArrayError */
  SVDH_ERR_SHAPE = 7,   /* shape mismatch (the reference panics via assert_eq!) */
  SVDH_ERR_ALLOC = 8,   /* device or host allocation failed */
  SVDH_ERR_HIP = 9,     /* HIP runtime error */
  SVDH_ERR_ARG = 10,    /* invalid argument (null pointer, bad enum, nnz >= 2^31) */
  SVDH_ERR_NO_DEVICE = 11
} svdh_status;

/* Element types at the boundary. */
typedef enum { SVDH_F64 = 0, SVDH_F32 = 1, SVDH_BF16 = 2 } svdh_dtype;

/* Arithmetic of a solve.
 *  SVDH_PREC_F64:  T = f64. FP64 values, vectors, panels and recurrence.
 *  SVDH_PREC_F32:  T = f32. LAS2: FP32 sparse values, FP64 vectors and
 *                  recurrence, f32::EPSILON tolerances (the reference's
 *                  svd_las2::<f32> control flow). rSVD: FP32 values and
 *                  panels with FP32 accumulation, FP64 small dense algebra.
 *  SVDH_PREC_BF16: rSVD only. BF16 values and BF16 input panels with FP32
 *                  accumulation for the early power iterations, FP32 for the
 *                  last ones (precision ladder), FP64 small dense algebra. */
typedef enum { SVDH_PREC_F64 = 0, SVDH_PREC_F32 = 1, SVDH_PREC_BF16 = 2 } svdh_precision;

/* randomized::PowerIterationNormalizer (src/randomized/mod.rs:14-19). */
typedef enum { SVDH_NORM_QR = 0, SVDH_NORM_LU = 1, SVDH_NORM_NONE = 2 } svdh_normalizer;

/* svdh_csr_create flags */
#define SVDH_CSR_KEEP_F32 0x1u  /* also keep an FP32 copy of the values */
#define SVDH_CSR_KEEP_BF16 0x2u /* also keep a BF16 copy of the values */

typedef struct svdh_ctx svdh_ctx;
typedef struct svdh_op svdh_op;

/* Diagnostics, field for field (src/utils.rs:66-79). */
typedef struct {
  int64_t non_zero;
  int64_t dimensions;
  int64_t iterations;
  int32_t transposed;
  int64_t lanczos_steps;
  int64_t ritz_values_stabilized;
  int64_t significant_values;
  int64_t singular_values;
  double end_interval[2];
  double kappa;
  uint32_t random_seed;
} svdh_diagnostics;

/* SvdRec (src/utils.rs:43-50). Host arrays owned by the library, freed by
 * svdh_result_free. u is u_rows x u_cols row-major, vt is vt_rows x vt_cols
 * row-major, s has d entries. Layouts are the reference's: LAS2 returns
 * u as d x m (rows are left vectors, src/lanczos/mod.rs:172); the randomized
 * solver returns u as m x d (columns are left vectors,
 * src/randomized/mod.rs:175-185). Values are FP64 regardless of precision. */
typedef struct {
  int64_t d;
  int64_t u_rows, u_cols, vt_rows, vt_cols;
  double *u;
  double *s;
  double *vt;
  svdh_diagnostics diagnostics;
  /* Extra (not in the reference): power iterations actually run, the
   * residual estimate the adaptive randomized solver stopped on, and how
   * many QR normalisations fell back from CholeskyQR2 to the eigen path. */
  int64_t power_iters_run;
  double residual_estimate;
  int64_t qr_fallbacks;
  /* Solves in SVDH_PREC_F32 or SVDH_PREC_BF16 hand their host arrays over
   * in single precision, as SvdRec<f32> holds them (utils.rs:43-50): u32,
   * s32 and vt32 (same shapes) are set and u, s, vt stay NULL; FP64 solves
   * set u, s, vt and leave these NULL. The values are the FP64 results
   * rounded to nearest. Freed by svdh_result_free as well. */
  float *u32;
  float *s32;
  float *vt32;
} svdh_result;

/* ---- context ---------------------------------------------------------- */

int32_t svdh_ctx_create(int32_t device, svdh_ctx **out);
int32_t svdh_ctx_destroy(svdh_ctx *ctx);
int32_t svdh_ctx_synchronize(svdh_ctx *ctx);
const char *svdh_last_error(void);
/* SHA-256 (hex) of the sources the library was built from (no GPU needed):
 * lets a binding refuse a library older than its sources. No counterpart in
 * the reference. */
const char *svdh_build_id(void);
/* Kernel timing with HIP events on the context's stream. Kernel classes:
 * 0 = SpMV, 1 = SpMM Y = A X, 2 = dense (Gram / apply / GEMM), 3 = BLAS-1,
 * 4 = basis GEMV (reorthogonalisation), 5 = SpMM Z = A^T Y, 6 = SDDMM,
 * 7 = COO / CSC ingestion (the device conversion, between the uploads of the
 * input arrays and the download of the CSR arrays). total_ms,
 * launches and the launches' algorithmic bytes accumulate until
 * svdh_ctx_profile is called again with enable = 1. */
int32_t svdh_ctx_profile(svdh_ctx *ctx, int32_t enable);
int32_t svdh_ctx_kernel_time(svdh_ctx *ctx, int32_t kernel_class, double *total_ms,
                             int64_t *launches, double *algorithmic_bytes);

/* ---- operators (SMat<T>) --------------------------------------------- */

/* Upload a CSR matrix: rowptr[nrows+1] and colidx[nnz] as usize-like int64
 * (nalgebra-sparse csr_data(), src/lanczos/mod.rs:1547), values of val_dtype.
 * Indices are narrowed to int32; nnz >= 2^31 or an index out of range is
 * SVDH_ERR_ARG. The transpose (CSC) is built once here. Replaces the
 * nalgebra_sparse::CsrMatrix operand of SMat (src/lanczos/mod.rs:1506). */
int32_t svdh_csr_create(svdh_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                        const int64_t *rowptr, const int64_t *colidx, const void *vals,
                        int32_t val_dtype, uint32_t flags, svdh_op **out);

/* A COO matrix (nalgebra_sparse::CooMatrix: row_indices, col_indices,
 * values of val_dtype F64 or F32; the operand of SMat for CooMatrix,
 * src/lanczos/mod.rs:1657-1740). Repeated (row, column) pairs are allowed:
 * the reference's products add every triplet, so they are summed here, left
 * to right in input order in the value type. The conversion runs on the
 * device (stable radix sort by (row, column), run sums, row pointers), then
 * the operator is built as by svdh_csr_create. Indices out of range or
 * nnz >= 2^31 are SVDH_ERR_ARG. */
int32_t svdh_coo_create(svdh_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rows,
                        const int64_t *cols, const void *vals, int32_t val_dtype, uint32_t flags,
                        svdh_op **out);

/* A CSC matrix (nalgebra_sparse::CscMatrix csc_data(): colptr[ncols+1],
 * row indices[nnz], values of val_dtype F64 or F32; the operand of SMat for
 * CscMatrix, src/lanczos/mod.rs:1415-1504). Converted on the device as
 * svdh_coo_create does (repeated pairs summed). */
int32_t svdh_csc_create(svdh_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *colptr,
                        const int64_t *rowidx, const void *vals, int32_t val_dtype, uint32_t flags,
                        svdh_op **out);

/* MaskedCSRMatrix::new(matrix, column_mask) (src/lanczos/masked.rs:21-46).
 * mask has ncols(base) bytes, nonzero = column kept. The view borrows base:
 * base must outlive it. Masked nnz is counted once here (masked.rs:89-102
 * rescans on every call). */
int32_t svdh_mask_create(svdh_ctx *ctx, const svdh_op *base, const uint8_t *mask,
                         svdh_op **out);
int32_t svdh_op_destroy(svdh_op *op);

/* SMat::nrows / ncols / nnz (src/utils.rs:23-25). */
int32_t svdh_op_shape(const svdh_op *op, int64_t *nrows, int64_t *ncols, int64_t *nnz);

/* SMat::svd_opa (src/utils.rs:26): y = A x, or y = A^T x when transposed.
 * y is overwritten. x and y are host pointers unless on_device != 0, in which
 * case they are device pointers valid on the context's device. vec_dtype is
 * SVDH_F64 or SVDH_F32. */
int32_t svdh_spmv(svdh_ctx *ctx, const svdh_op *op, int32_t transposed, const void *x, void *y,
                  int32_t vec_dtype, int32_t on_device);

/* svdh_spmv with device vectors (on_device = 1) that allocates nothing. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The workgroup partials go to a static device
 * scratch array of 16 MiB (2^21 FP64 partials) that each GPU the library runs
 * on holds until the process ends; a product that needs more partials (the
 * source-tiled layout needs source blocks x output rows) returns
 * SVDH_ERR_ARG. The operator must have been prepared for this direction and
 * vector dtype by one earlier svdh_spmv call (which builds its schedule and
 * value layouts); otherwise SVDH_ERR_ARG. The product is enqueued on the
 * context's stream and the call returns without waiting
 * (svdh_ctx_synchronize waits). Every call on a GPU shares that GPU's scratch
 * (each call writes every partial it reads), so calls that overlap on one
 * GPU, from contexts on different streams, race. */
int32_t svdh_spmv_noalloc(svdh_ctx *ctx, const svdh_op *op, int32_t transposed, const void *x, void *y,
                          int32_t vec_dtype);

/* SMat::multiply_with_dense (src/utils.rs:28, masked.rs:265-471).
 * transpose_self == 0: Y (nrows x l) = A X (ncols x l);
 * transpose_self != 0: Y (ncols x l) = A^T X (nrows x l).
 * Host panels are column-major (nalgebra DMatrix) with leading dimensions
 * ldx, ldy. dtype is SVDH_F64 or SVDH_F32. */
int32_t svdh_spmm(svdh_ctx *ctx, const svdh_op *op, int32_t transpose_self, const void *X,
                  int64_t ldx, int64_t l, void *Y, int64_t ldy, int32_t dtype);

/* SMat::multiply_transposed_by_dense (src/utils.rs:31, masked.rs:616-733):
 * B (l x ncols, column-major, ldb) = Q^T A with Q (nrows x l, column-major). */
int32_t svdh_spmm_qta(svdh_ctx *ctx, const svdh_op *op, const void *Q, int64_t ldq, int64_t l,
                      void *B, int64_t ldb, int32_t dtype);

/* SMat::multiply_with_dense_centered (src/utils.rs:29, masked.rs:473-614):
 * svdh_spmm of the implicitly centred matrix A - 1 mu^T, means = mu (ncols
 * entries of dtype, in the operator's column numbering, as
 * compute_column_means returns them).
 * transpose_self == 0: Y (nrows x l) = A X - 1 (mu^T X);
 * transpose_self != 0: Y (ncols x l) = A^T X - mu (1^T X).
 * The forward form is the exact one; the reference's masked.rs:496-509
 * subtracts (sum_j mu_j)(sum_i X_ic) instead. The corrections run on the
 * device as rank-one updates of the sparse product. */
int32_t svdh_spmm_centered(svdh_ctx *ctx, const svdh_op *op, int32_t transpose_self, const void *X, int64_t ldx,
                           int64_t l, void *Y, int64_t ldy, const void *means, int32_t dtype);

/* SMat::multiply_transposed_by_dense_centered (src/utils.rs:32,
 * masked.rs:735-874): B (l x ncols, column-major, ldb) =
 * Q^T (A - 1 mu^T) = Q^T A - (Q^T 1) mu^T. */
int32_t svdh_spmm_qta_centered(svdh_ctx *ctx, const svdh_op *op, const void *Q, int64_t ldq, int64_t l, void *B,
                               int64_t ldb, const void *means, int32_t dtype);

/* SMat::compute_column_means (src/utils.rs:27, masked.rs:240-263). means has
 * ncols entries of dtype (SVDH_F64 or SVDH_F32). */
int32_t svdh_column_means(svdh_ctx *ctx, const svdh_op *op, void *means, int32_t dtype);

/* ---- whole solves ----------------------------------------------------- */

/* lanczos::svd_las2 (src/lanczos/mod.rs:104-189) with svd / svd_dim /
 * svd_dim_seed as the callers' defaults. dimensions / iterations 0 = max.
 * random_seed 0 draws a seed from the OS like the reference (mod.rs:116-119).
 * start (nullable) is a start vector of the working dimension
 * (min(nrows, ncols) side chosen by the transpose rule, mod.rs:143-145);
 * when given it replaces the random start vector, which is the reference's
 * own behaviour for a non-zero w0 (mod.rs:595-606). precision is
 * SVDH_PREC_F64 or SVDH_PREC_F32. */
int32_t svdh_las2(svdh_ctx *ctx, const svdh_op *op, int64_t dimensions, int64_t iterations,
                  double endl, double endr, double kappa, uint32_t random_seed,
                  const double *start, int32_t precision, svdh_result *out);

/* legacy::svdLAS2 (src/legacy.rs:569-663; svd / svd_dim / svd_dim_seed,
 * legacy.rs:498-525, are its callers' defaults): FP64 only, the legacy
 * control flow of lanso / ritvec / imtqlb / imtql2 (legacy.rs:961-2014):
 * f64::EPSILON tolerances without the sparsity adaptation, an unreduced
 * block ends where |beta| < eps, at most 30 QL iterations per eigenvalue
 * (then SVDH_ERR_IMTQLB / SVDH_ERR_IMTQL2), significant Ritz values
 * bnd <= kappa |ritz|. out->u is d x nrows (SvdRec.ut, legacy.rs:454-460).
 * Arguments as svdh_las2. */
int32_t svdh_las2_legacy(svdh_ctx *ctx, const svdh_op *op, int64_t dimensions, int64_t iterations,
                         double endl, double endr, double kappa, uint32_t random_seed,
                         const double *start, svdh_result *out);

/* svdh_las2 with the singular vectors left in device memory (the measured
 * path; the host hand-over of U, d x nrows FP64, is PCIe-bound for tall
 * matrices). u: device, room for min(dimensions, nrows, ncols) rows of
 * nrows (row-major, leading dimension ldu >= nrows: rows are the left
 * vectors, the layout of SvdRec.u for LAS2, mod.rs:172); vt: device, same
 * count of rows of ncols (ldvt >= ncols); s: host, room for as many values.
 * meta receives d (the rows written), the shapes and the diagnostics; its
 * u, s and vt stay NULL. Same algorithm and arguments as svdh_las2. */
int32_t svdh_las2_device(svdh_ctx *ctx, const svdh_op *op, int64_t dimensions, int64_t iterations,
                         double endl, double endr, double kappa, uint32_t random_seed,
                         const double *start, int32_t precision, double *u, int64_t ldu, double *s,
                         double *vt, int64_t ldvt, svdh_result *meta);

/* Start vector the LAS2 driver draws for a seed (host only, no GPU):
 * out[i] = 2 U_i - 1 with U_i = (mix64(key + (i+1) phi) >> 11) 2^-53,
 * key = seed phi XOR 0x5356444C49423031, phi = 0x9E3779B97F4A7C15 and mix64
 * the splitmix64 finaliser. */
int32_t svdh_las2_start_vector(uint32_t seed, int64_t n, double *out);

/* randomized::randomized_svd (src/randomized/mod.rs:24-199).
 * has_seed == 0 is seed None (deterministic seed 0, randomized/mod.rs:304-307).
 * omega (nullable) injects the ncols x (rank + oversample) Gaussian test
 * matrix, column-major FP64; otherwise it is drawn on the device from a
 * counter-based generator keyed by seed.
 * tol <= 0: exactly n_power_iters power iterations (the reference).
 * tol > 0: power iterations continue past n_power_iters, up to
 * max_power_iters, until the Rayleigh-Ritz residual estimate
 * max_i sqrt(|A v_i - s_i u_i|^2 + |A^T u_i - s_i v_i|^2) / s_i over the
 * returned rank drops to tol (normaliser QR only). */
int32_t svdh_rsvd(svdh_ctx *ctx, const svdh_op *op, int64_t target_rank, int64_t n_oversamples,
                  int64_t n_power_iters, int32_t normalizer, int32_t mean_center,
                  uint64_t seed, int32_t has_seed, const double *omega, int32_t precision,
                  double tol, int64_t max_power_iters, svdh_result *out);

/* svdh_rsvd with the singular vectors left in device memory, for callers
 * that keep working on the GPU (and the throughput measurement, which
 * excludes the PCIe hand-over). k = min(target_rank, nrows, ncols);
 * u: device, m x k row-major with leading dimension ldu >= k;
 * vt: device, k x n row-major with leading dimension ldvt >= n;
 * s: host, k entries. meta receives d, the shapes, the diagnostics and the
 * extra fields of svdh_result; its u, s and vt stay NULL. Same algorithm and
 * arguments as svdh_rsvd (src/randomized/mod.rs:24-199). Returns after u and
 * vt are written. */
int32_t svdh_rsvd_device(svdh_ctx *ctx, const svdh_op *op, int64_t target_rank, int64_t n_oversamples,
                         int64_t n_power_iters, int32_t normalizer, int32_t mean_center,
                         uint64_t seed, int32_t has_seed, const double *omega, int32_t precision,
                         double tol, int64_t max_power_iters, double *u, int64_t ldu, double *s,
                         double *vt, int64_t ldvt, svdh_result *meta);

/* Batched randomized SVD: nops operators with the same number of rows
 * (typically svdh_mask_create column masks of one matrix), each solved
 * as svdh_rsvd would with tol = 0 (exactly n_power_iters power iterations),
 * QR normaliser only. Every SpMM round over all operators is one kernel
 * launch. All operators must give the same panel width
 * min(target_rank, nrows, ncols_b) + n_oversamples <= 64. omegas (nullable,
 * and each omegas[b] nullable) injects operator b's ncols_b x (rank +
 * oversample) column-major FP64 test matrix, as svdh_rsvd's omega does.
 * outs[b] receives operator b's result (free each with svdh_result_free),
 * bit for bit the result of svdh_rsvd on that operator with the same seed
 * (or omega). No counterpart in the reference, whose callers loop over
 * randomized_svd (src/randomized/mod.rs:24-199) per MaskedCSRMatrix. */
int32_t svdh_rsvd_batched(svdh_ctx *ctx, int64_t nops, const svdh_op *const *ops, int64_t target_rank,
                          int64_t n_oversamples, int64_t n_power_iters, int32_t normalizer, uint64_t seed,
                          int32_t has_seed, const double *const *omegas, int32_t precision, svdh_result *outs);

/* svdh_rsvd_batched with the singular vectors left in device memory:
 * operator b writes u[b] (m x k_b row-major, leading dimension ldu[b] >= k_b)
 * and vt[b] (k_b x n_b row-major, ldvt[b] >= n_b), k_b = min(target_rank, m,
 * n_b), FP64 device buffers from the same HIP runtime. metas[b] receives d,
 * the shapes, the diagnostics and s (host, freed by svdh_result_free); its u
 * and vt stay NULL. */
int32_t svdh_rsvd_batched_device(svdh_ctx *ctx, int64_t nops, const svdh_op *const *ops, int64_t target_rank,
                                 int64_t n_oversamples, int64_t n_power_iters, int32_t normalizer, uint64_t seed,
                                 int32_t has_seed, const double *const *omegas, int32_t precision,
                                 double *const *u, const int64_t *ldu,
                                 double *const *vt, const int64_t *ldvt, svdh_result *metas);

/* ---- adjoint (backward) ------------------------------------------------ */

/* Losses of svdh_backward. No counterpart in the reference: the adjoint sits
 * behind a backward() on SvdRec (utils.rs:43-50) in the Rust binding. */
typedef enum {
  SVDH_LOSS_SIGMA_SUM = 0,  /* L = sum_{i<k} s_i                 */
  SVDH_LOSS_PROJECTION = 1  /* L = |U_k^T y|^2, y of length nrows */
} svdh_loss;

typedef struct {
  double loss;                  /* L at the given triplets            */
  int64_t cg_iterations;        /* projection loss: CG iterations run */
  double cg_relative_residual;  /* max_i |r_i| / |P y| at exit        */
} svdh_backward_info;

/* Gradient of the loss with respect to the stored values of the operator's
 * CSR (grad has nnz entries of grad_dtype, in CSR order of the underlying
 * matrix; for a masked view the dropped columns get 0). res holds the top
 * triplets of a solve of the same operator (svdh_rsvd or svdh_las2; either u
 * layout, FP64 or single-precision arrays); k <= min(res->d, 64). The projection loss solves k
 * shifted systems (s_i^2 I - A A^T) w_i = (I - U_k U_k^T) y by conjugate
 * gradients to the relative residual tol (<= 0: 1e-10 in FP64, 1e-5 in FP32)
 * within max_iters (<= 0: 1000) iterations. precision: SVDH_PREC_F64,
 * SVDH_PREC_F32 (factor panels and SpMM arithmetic) or SVDH_PREC_BF16 (FP32
 * panels and conjugate gradients; the SDDMM gathers BF16 copies of its factor
 * panels and accumulates in FP32). */
int32_t svdh_backward(svdh_ctx *ctx, const svdh_op *op, const svdh_result *res, int32_t loss, int64_t k,
                      const double *y, int32_t precision, double tol, int64_t max_iters, void *grad,
                      int32_t grad_dtype, svdh_backward_info *info);

void svdh_result_free(svdh_result *res);

#ifdef __cplusplus
}
#endif
#endif
