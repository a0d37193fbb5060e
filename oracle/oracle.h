/*
 * CPU oracle for the sparse-SVD hot path of single-svdlib v1.0.9.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it. The product library never links or calls it.
 *
 * Everything here is FP64 and restates the reference's algorithms:
 *   - SMat::svd_opa for CsrMatrix        src/lanczos/mod.rs:1520-1603
 *   - SMat::svd_opa for MaskedCSRMatrix  src/lanczos/masked.rs:104-238
 *     (without the small-matrix fast path, defect D6)
 *   - multiply_with_dense                src/lanczos/masked.rs:265-471
 *     (column-major panel read correctly, defect D5 fixed)
 *   - multiply_transposed_by_dense       src/lanczos/masked.rs:616-733
 *   - compute_column_means               src/lanczos/masked.rs:240-263, mod.rs:1605-1627
 *   - the LAS2 driver svd_las2 / lanso / lanczos_step / purge / ortbnd /
 *     error_bound / imtqlb / imtql2 / ritvec   src/lanczos/mod.rs:104-1406
 *     with the numerical defects corrected as src/legacy.rs does them:
 *       D1  svd_pythag update r *= (s/u)^2        (legacy.rs:872-889)
 *       D2  Ritz coefficient s[(js-1-i)*js + k]   (legacy.rs:1798-1813)
 *       D3  largest d values, descending           (legacy.rs:1796-1835)
 *       D12 imtqlb re-reads p=d[l], f=bnd[l] every QL iteration
 *           (mod.rs:455-457 reads them once per l; legacy.rs:990-991)
 *
 * The LAS2 path follows SVDLIBC's control flow; SVDLIBC is
 * Copyright (c) 2002, University of Tennessee Research Foundation, BSD licence
 * (see SVDLIBC-LICENSE.txt in the reference tree).
 *
 * Parallel sections use OpenMP and restate the reference's Rayon
 * decomposition (row-parallel Ax, chunked partial A^T x with
 * determine_chunk_size, src/utils.rs:10-20), so this library is also the
 * timed CPU baseline.
 */
#ifndef SVDLIB_ORACLE_H
#define SVDLIB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; values match svdh_status in include/svdh.h. */
enum {
  ORC_OK = 0,
  ORC_ERR_IMTQLB = 1,
  ORC_ERR_STARTV = 2,
  ORC_ERR_STPONE = 3,
  ORC_ERR_IMTQL2 = 4,
  ORC_ERR_LAS2 = 5,
  ORC_ERR_SHAPE = 7,
  ORC_ERR_ALLOC = 8,
};

/* A CSR matrix (FP64 values) with an optional column-mask view.
 * When m2o is NULL the operator is the plain CSR. Otherwise the operator is
 * A[:, m2o[0..n_masked)] and o2m[c] is the masked index of original column c
 * or -1 (masked.rs:13-46). */
typedef struct {
  int64_t nrows, ncols;
  const int64_t *rowptr;
  const int64_t *colidx;
  const double *vals;
  int64_t n_masked;
  const int64_t *m2o;
  const int64_t *o2m;
  int nthreads;
} orc_op;

typedef struct {
  int64_t non_zero, dimensions, iterations;
  int32_t transposed;
  int64_t lanczos_steps, ritz_values_stabilized, significant_values, singular_values;
  double end_interval[2];
  double kappa;
  uint32_t random_seed;
} orc_diag;

/* LAS2 result. u is d x m, vt is d x n, both row-major (mod.rs:172-174). */
typedef struct {
  int64_t d, m, n;
  double *u, *s, *vt;
  orc_diag diag;
} orc_svdrec;

int64_t orc_nrows(const orc_op *op);
int64_t orc_ncols(const orc_op *op);
int64_t orc_nnz(const orc_op *op);
int orc_determine_chunk_size(int64_t nrows, int nthreads);

/* y = A x (transposed == 0) or y = A^T x (transposed != 0). */
void orc_opa(const orc_op *op, const double *x, double *y, int transposed);
/* result = A dense (transpose_self == 0) or A^T dense; dense and result are
 * column-major with l columns. */
void orc_multiply_with_dense(const orc_op *op, const double *dense, int64_t l,
                             double *result, int transpose_self);
/* result (l x ncols, column-major) = q^T A, q is nrows x l column-major. */
void orc_multiply_transposed_by_dense(const orc_op *op, const double *q, int64_t l,
                                      double *result);
/* f32 twins (T = f32): vals32 = the values rounded to f32, panels f32. */
void orc_multiply_with_dense_f32(const orc_op *op, const float *vals32, const float *dense, int64_t l,
                                 float *result, int transpose_self);
void orc_multiply_transposed_by_dense_f32(const orc_op *op, const float *vals32, const float *q, int64_t l,
                                          float *result);
void orc_compute_column_means(const orc_op *op, double *means);

/* Scalar helpers exposed for the known-answer tests. */
double orc_pythag(double a, double b, double eps);
int orc_imtqlb(int64_t n, double *d, double *e, double *bnd, int64_t max_iter, double eps);
int orc_imtql2(int64_t nm, int64_t n, double *d, double *e, double *z, int64_t max_iter,
               double eps);

/* Start vector stream of the LAS2 driver: uniform [-1, 1) from a
 * counter-based splitmix64 keyed by the 32-bit seed. */
void orc_start_vector(uint32_t seed, int64_t n, double *out);

/* Whole LAS2 solve.  eps is the machine epsilon of the reference's T
 * (f64::EPSILON or f32::EPSILON); arithmetic is FP64 either way.
 * start (nullable, length = working n) replaces the random start vector.
 * On success the caller frees the result with orc_svdrec_free. */
int orc_las2(const orc_op *op, int64_t dimensions, int64_t iterations, double endl,
             double endr, double kappa, uint32_t random_seed, const double *start,
             double eps, orc_svdrec *out);
/* legacy::svdLAS2 (src/legacy.rs:569-663): f64, plain f64::EPSILON
 * tolerances, splits at compare(bet, 0), imtqlb / imtql2 with the
 * absolute convergence test and at most 30 iterations (then an error),
 * significant values bnd <= kappa |ritz|. */
int orc_las2_legacy(const orc_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
                    double kappa, uint32_t random_seed, const double *start, orc_svdrec *out);
int orc_imtqlb_legacy(int64_t n, double *d, double *e, double *bnd, double eps);
int orc_imtql2_legacy(int64_t nm, int64_t n, double *d, double *e, double *z, double eps);
void orc_svdrec_free(orc_svdrec *rec);
const char *orc_last_error(void);

/* COO -> CSR (the input formats of SMat for CooMatrix, mod.rs:1657-1740,
 * whose products add every triplet, duplicates included): entries ordered
 * by (row, column), duplicates summed in input order in the value type
 * (FP64 in orc_coo_to_csr, FP32 in orc_coo_to_csr_f32). Outputs sized by the
 * caller: rowptr[m + 1], cols[nnz], vals[nnz]; returns the number of
 * distinct entries, or -1 on an index out of range. */
int64_t orc_coo_to_csr(int64_t m, int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols,
                       const double *vals, int64_t *rowptr, int64_t *out_cols, double *out_vals);
int64_t orc_coo_to_csr_f32(int64_t m, int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols,
                           const float *vals, int64_t *rowptr, int64_t *out_cols, float *out_vals);

#ifdef __cplusplus
}
#endif
#endif
