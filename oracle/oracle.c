/*
 * CPU oracle for the sparse-SVD hot path of single-svdlib v1.0.9 (FP64).
 * TEST INFRASTRUCTURE ONLY: see oracle.h for what may load it and for the
 * list of reference defects corrected here.
 *
 * SVDLIBC control flow: Copyright (c) 2002, University of Tennessee Research
 * Foundation. BSD licence (SVDLIBC-LICENSE.txt in the reference tree).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

const char *orc_last_error(void) { return g_err; }

/* ------------------------------------------------------------------ */
/* Operator (SMat) restatement                                          */
/* ------------------------------------------------------------------ */

int64_t orc_nrows(const orc_op *op) { return op->nrows; }

/* masked.rs:85-87: a masked view has |mask| columns. */
int64_t orc_ncols(const orc_op *op) { return op->m2o ? op->n_masked : op->ncols; }

/* masked.rs:89-102 counts nnz in masked columns by scanning all entries. */
int64_t orc_nnz(const orc_op *op) {
  if (!op->m2o) return op->rowptr[op->nrows];
  int64_t count = 0;
  for (int64_t i = 0; i < op->nrows; ++i)
    for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j)
      if (op->o2m[op->colidx[j]] >= 0) ++count;
  return count;
}

static int nthreads_of(const orc_op *op) { return op->nthreads > 0 ? op->nthreads : 1; }

/* utils.rs:10-20 */
int orc_determine_chunk_size(int64_t nrows, int nthreads) {
  int64_t target = (int64_t)nthreads * 4;
  int64_t chunk = (nrows + target - 1) / target;
  return (int)(chunk < 64 ? 64 : chunk);
}

static int uses_all_columns(const orc_op *op) {
  /* masked.rs:57-59 */
  if (!op->m2o) return 1;
  if (op->n_masked != op->ncols) return 0;
  for (int64_t c = 0; c < op->ncols; ++c)
    if (op->o2m[c] < 0) return 0;
  return 1;
}

/* Column index of entry j in the operator's column space, or -1. */
static inline int64_t mcol(const orc_op *op, int64_t j) {
  int64_t c = op->colidx[j];
  return op->m2o ? op->o2m[c] : c;
}

/* mod.rs:1551-1570 (row-parallel Ax) and masked.rs:147-196. */
static void op_ax(const orc_op *op, const double *x, double *y) {
  const int nt = nthreads_of(op);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < op->nrows; ++i) {
    double sum = 0.0;
    for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
      int64_t c = mcol(op, j);
      if (c >= 0) sum += op->vals[j] * x[c];
    }
    y[i] = sum;
  }
}

/* mod.rs:1571-1601 and masked.rs:197-237: per-chunk dense partials of
 * length ncols, combined serially in chunk order, skipping zeros.  The
 * masked variant also skips rows whose x value is zero (masked.rs:212-215). */
static void op_atx(const orc_op *op, const double *x, double *y) {
  const int nt = nthreads_of(op);
  const int64_t nr = op->nrows, nc = orc_ncols(op);
  const int64_t chunk = orc_determine_chunk_size(nr, nt);
  const int64_t nchunks = (nr + chunk - 1) / chunk;
  const int skip_zero_rows = op->m2o != NULL && !uses_all_columns(op);
  double *partial = (double *)calloc((size_t)(nchunks * nc), sizeof(double));
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    double *loc = partial + ch * nc;
    int64_t end = (ch + 1) * chunk < nr ? (ch + 1) * chunk : nr;
    for (int64_t i = ch * chunk; i < end; ++i) {
      double xv = x[i];
      if (skip_zero_rows && xv == 0.0) continue;
      for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
        int64_t c = mcol(op, j);
        if (c >= 0) loc[c] += op->vals[j] * xv;
      }
    }
  }
  for (int64_t c = 0; c < nc; ++c) y[c] = 0.0;
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    const double *loc = partial + ch * nc;
    for (int64_t c = 0; c < nc; ++c)
      if (loc[c] != 0.0) y[c] += loc[c];
  }
  free(partial);
}

void orc_opa(const orc_op *op, const double *x, double *y, int transposed) {
  if (transposed)
    op_atx(op, x, y);
  else
    op_ax(op, x, y);
}

/* masked.rs:265-471 with the column-major panel indexed correctly (D5). */
void orc_multiply_with_dense(const orc_op *op, const double *dense, int64_t l, double *result,
                             int transpose_self) {
  const int nt = nthreads_of(op);
  const int64_t nr = op->nrows, nc = orc_ncols(op);
  if (!transpose_self) {
    /* result (nr x l) = A (nr x nc) * dense (nc x l) */
#pragma omp parallel num_threads(nt)
    {
      double *row = (double *)malloc((size_t)l * sizeof(double));
#pragma omp for schedule(static)
      for (int64_t i = 0; i < nr; ++i) {
        for (int64_t c = 0; c < l; ++c) row[c] = 0.0;
        for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
          int64_t mc = mcol(op, j);
          if (mc < 0) continue;
          double v = op->vals[j];
          for (int64_t c = 0; c < l; ++c) row[c] += v * dense[mc + c * nc];
        }
        for (int64_t c = 0; c < l; ++c) result[i + c * nr] = row[c];
      }
      free(row);
    }
  } else {
    /* result (nc x l) = A^T (nc x nr) * dense (nr x l) */
    const int64_t chunk = orc_determine_chunk_size(nr, nt);
    const int64_t nchunks = (nr + chunk - 1) / chunk;
    double *partial = (double *)calloc((size_t)(nchunks * nc * l), sizeof(double));
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
      double *loc = partial + ch * nc * l; /* [masked_col * l + c] */
      int64_t end = (ch + 1) * chunk < nr ? (ch + 1) * chunk : nr;
      for (int64_t i = ch * chunk; i < end; ++i) {
        for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
          int64_t mc = mcol(op, j);
          if (mc < 0) continue;
          double v = op->vals[j];
          for (int64_t c = 0; c < l; ++c) loc[mc * l + c] += v * dense[i + c * nr];
        }
      }
    }
    for (int64_t k = 0; k < nc * l; ++k) result[k] = 0.0;
    for (int64_t ch = 0; ch < nchunks; ++ch) {
      const double *loc = partial + ch * nc * l;
      for (int64_t r = 0; r < nc; ++r)
        for (int64_t c = 0; c < l; ++c) {
          double v = loc[r * l + c];
          if (v != 0.0) result[r + c * nc] += v;
        }
    }
    free(partial);
  }
}

/* masked.rs:616-733: result (l x nc, column-major) = q^T A. */
void orc_multiply_transposed_by_dense(const orc_op *op, const double *q, int64_t l,
                                      double *result) {
  const int nt = nthreads_of(op);
  const int64_t nr = op->nrows, nc = orc_ncols(op);
  const int64_t chunk = orc_determine_chunk_size(nr, nt);
  const int64_t nchunks = (nr + chunk - 1) / chunk;
  double *partial = (double *)calloc((size_t)(nchunks * nc * l), sizeof(double));
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    double *loc = partial + ch * nc * l; /* column-major l x nc */
    int64_t end = (ch + 1) * chunk < nr ? (ch + 1) * chunk : nr;
    for (int64_t i = ch * chunk; i < end; ++i)
      for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
        int64_t mc = mcol(op, j);
        if (mc < 0) continue;
        double v = op->vals[j];
        for (int64_t c = 0; c < l; ++c) loc[c + mc * l] += q[i + c * nr] * v;
      }
  }
  for (int64_t k = 0; k < nc * l; ++k) result[k] = 0.0;
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    const double *loc = partial + ch * nc * l;
    for (int64_t k = 0; k < nc * l; ++k)
      if (loc[k] != 0.0) result[k] += loc[k];
  }
  free(partial);
}

/* The same two products with T = f32 (the reference instantiated with f32:
 * products and accumulators in f32, masked.rs:300-360, 361-470, 616-733).
 * vals32 holds the operator's values rounded to f32 (CSR order). */
void orc_multiply_with_dense_f32(const orc_op *op, const float *vals32, const float *dense, int64_t l,
                                 float *result, int transpose_self) {
  const int nt = nthreads_of(op);
  const int64_t nr = op->nrows, nc = orc_ncols(op);
  if (!transpose_self) {
#pragma omp parallel num_threads(nt)
    {
      float *row = (float *)malloc((size_t)l * sizeof(float));
#pragma omp for schedule(static)
      for (int64_t i = 0; i < nr; ++i) {
        for (int64_t c = 0; c < l; ++c) row[c] = 0.0f;
        for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
          int64_t mc = mcol(op, j);
          if (mc < 0) continue;
          float v = vals32[j];
          for (int64_t c = 0; c < l; ++c) row[c] += v * dense[mc + c * nc];
        }
        for (int64_t c = 0; c < l; ++c) result[i + c * nr] = row[c];
      }
      free(row);
    }
  } else {
    const int64_t chunk = orc_determine_chunk_size(nr, nt);
    const int64_t nchunks = (nr + chunk - 1) / chunk;
    float *partial = (float *)calloc((size_t)(nchunks * nc * l), sizeof(float));
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
      float *loc = partial + ch * nc * l;
      int64_t end = (ch + 1) * chunk < nr ? (ch + 1) * chunk : nr;
      for (int64_t i = ch * chunk; i < end; ++i)
        for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
          int64_t mc = mcol(op, j);
          if (mc < 0) continue;
          float v = vals32[j];
          for (int64_t c = 0; c < l; ++c) loc[mc * l + c] += v * dense[i + c * nr];
        }
    }
    for (int64_t k = 0; k < nc * l; ++k) result[k] = 0.0f;
    for (int64_t ch = 0; ch < nchunks; ++ch) {
      const float *loc = partial + ch * nc * l;
      for (int64_t r = 0; r < nc; ++r)
        for (int64_t c = 0; c < l; ++c) {
          float v = loc[r * l + c];
          if (v != 0.0f) result[r + c * nc] += v;
        }
    }
    free(partial);
  }
}

void orc_multiply_transposed_by_dense_f32(const orc_op *op, const float *vals32, const float *q, int64_t l,
                                          float *result) {
  const int nt = nthreads_of(op);
  const int64_t nr = op->nrows, nc = orc_ncols(op);
  const int64_t chunk = orc_determine_chunk_size(nr, nt);
  const int64_t nchunks = (nr + chunk - 1) / chunk;
  float *partial = (float *)calloc((size_t)(nchunks * nc * l), sizeof(float));
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    float *loc = partial + ch * nc * l;
    int64_t end = (ch + 1) * chunk < nr ? (ch + 1) * chunk : nr;
    for (int64_t i = ch * chunk; i < end; ++i)
      for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
        int64_t mc = mcol(op, j);
        if (mc < 0) continue;
        float v = vals32[j];
        for (int64_t c = 0; c < l; ++c) loc[c + mc * l] += q[i + c * nr] * v;
      }
  }
  for (int64_t k = 0; k < nc * l; ++k) result[k] = 0.0f;
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    const float *loc = partial + ch * nc * l;
    for (int64_t k = 0; k < nc * l; ++k)
      if (loc[k] != 0.0f) result[k] += loc[k];
  }
  free(partial);
}

/* masked.rs:240-263, mod.rs:1605-1627 */
void orc_compute_column_means(const orc_op *op, double *means) {
  const int64_t nr = op->nrows, nc = orc_ncols(op);
  for (int64_t c = 0; c < nc; ++c) means[c] = 0.0;
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = op->rowptr[i]; j < op->rowptr[i + 1]; ++j) {
      int64_t mc = mcol(op, j);
      if (mc >= 0) means[mc] += op->vals[j];
    }
  double r = 1.0 / (double)nr;
  for (int64_t c = 0; c < nc; ++c) means[c] *= r;
}

/* ------------------------------------------------------------------ */
/* BLAS-1 (mod.rs:321-424). Dots use fixed 4096-element blocks summed in */
/* block order, so results do not depend on the thread count.           */
/* ------------------------------------------------------------------ */

#define DOT_BLOCK 4096

static int g_blas_threads = 1;

static double ddot(int64_t n, const double *x, const double *y) {
  if (n < 2 * DOT_BLOCK || g_blas_threads <= 1) {
    double s = 0.0;
    int64_t nb = (n + DOT_BLOCK - 1) / DOT_BLOCK;
    for (int64_t b = 0; b < nb; ++b) {
      double p = 0.0;
      int64_t end = (b + 1) * DOT_BLOCK < n ? (b + 1) * DOT_BLOCK : n;
      for (int64_t i = b * DOT_BLOCK; i < end; ++i) p += x[i] * y[i];
      s += p;
    }
    return s;
  }
  int64_t nb = (n + DOT_BLOCK - 1) / DOT_BLOCK;
  double *part = (double *)malloc((size_t)nb * sizeof(double));
#pragma omp parallel for num_threads(g_blas_threads) schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    double p = 0.0;
    int64_t end = (b + 1) * DOT_BLOCK < n ? (b + 1) * DOT_BLOCK : n;
    for (int64_t i = b * DOT_BLOCK; i < end; ++i) p += x[i] * y[i];
    part[b] = p;
  }
  double s = 0.0;
  for (int64_t b = 0; b < nb; ++b) s += part[b];
  free(part);
  return s;
}

static void daxpy(int64_t n, double a, const double *x, double *y) {
  if (n >= 1000 && g_blas_threads > 1) {
#pragma omp parallel for num_threads(g_blas_threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
  } else {
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
  }
}

static void datx(int64_t n, double a, const double *x, double *y) {
  for (int64_t i = 0; i < n; ++i) y[i] = a * x[i];
}

static void dscal(int64_t n, double a, double *x) {
  for (int64_t i = 0; i < n; ++i) x[i] *= a;
}

static void dcopy_rev(int64_t n, int64_t offset, const double *x, double *y) {
  /* mod.rs:417-424 */
  for (int64_t i = 0; i < n; ++i) y[offset + n - 1 - i] = x[offset + i];
}

static int64_t idamax(int64_t n, const double *x) {
  /* mod.rs:335-350 */
  int64_t imax = 0;
  for (int64_t i = 1; i < n; ++i)
    if (fabs(x[i]) > fabs(x[imax])) imax = i;
  return imax;
}

static double fsign(double a, double b) {
  /* mod.rs:353-358 */
  return ((a >= 0.0 && b >= 0.0) || (a < 0.0 && b < 0.0)) ? a : -a;
}

static int cmp_eps(double a, double b, double eps) { return fabs(b - a) < eps; }

/* legacy.rs:872-889 (the D1-correct update r *= (s/u)^2). */
double orc_pythag(double a, double b, double eps) {
  double n = fmax(fabs(a), fabs(b));
  if (!(n > 0.0)) return 0.0;
  double p = n;
  double q = fmin(fabs(a), fabs(b)) / p;
  double r = q * q;
  double t = 4.0 + r;
  while (!cmp_eps(t, 4.0, eps)) {
    double s = r / t;
    double u = 1.0 + 2.0 * s;
    p *= u;
    double su = s / u;
    r *= su * su;
    t = 4.0 + r;
  }
  return p;
}

/* ------------------------------------------------------------------ */
/* Tridiagonal QL (mod.rs:428-584 and 901-1028)                         */
/* ------------------------------------------------------------------ */

int orc_imtqlb(int64_t n, double *d, double *e, double *bnd, int64_t max_iter, double eps) {
  if (n == 1) return ORC_OK;
  const double size_factor = sqrt((double)n);
  bnd[0] = 1.0;
  const int64_t last = n - 1;
  for (int64_t i = 1; i <= last; ++i) {
    bnd[i] = 0.0;
    e[i - 1] = e[i];
  }
  e[last] = 0.0;
  int64_t i = 0;
  int issues = 0;
  for (int64_t l = 0; l <= last; ++l) {
    int64_t iteration = 0;
    while (iteration <= max_iter) {
      int64_t m = l;
      while (m < n) {
        if (m == last) break;
        double test = fabs(d[m]) + fabs(d[m + 1]);
        double tol = eps * 100.0 * fmax(test, 1.0) * size_factor; /* mod.rs:467-472 */
        if (fabs(e[m]) <= tol) break;
        ++m;
      }
      /* D12: p and f are taken afresh every iteration (legacy.rs:990-991). */
      double p = d[l];
      double f = bnd[l];
      if (m == l) {
        int exchange = 1;
        if (l > 0) {
          i = l;
          while (i >= 1 && exchange) {
            if (p < d[i - 1]) {
              d[i] = d[i - 1];
              bnd[i] = bnd[i - 1];
              --i;
            } else {
              exchange = 0;
            }
          }
        }
        if (exchange) i = 0;
        d[i] = p;
        bnd[i] = f;
        iteration = max_iter + 1;
      } else {
        if (iteration == max_iter) {
          /* mod.rs:503-517: keep going with conservative bounds. */
          issues = 1;
          for (int64_t idx = l; idx <= m; ++idx) bnd[idx] = fmax(bnd[idx], 0.1);
          e[l] = 0.0;
          break;
        }
        ++iteration;
        double g = (d[l + 1] - p) / (2.0 * e[l]);
        double r = orc_pythag(g, 1.0, eps);
        g = d[m] - p + e[l] / (g + fsign(r, g));
        double s = 1.0, c = 1.0;
        p = 0.0;
        i = m - 1;
        int underflow = 0;
        while (!underflow && i >= l) {
          f = s * e[i];
          double b = c * e[i];
          r = orc_pythag(f, g, eps);
          e[i + 1] = r;
          if (r < eps * 1000.0 * (fabs(f) + fabs(g))) { /* mod.rs:539-545 */
            underflow = 1;
            break;
          }
          if (fabs(r) < eps * 100.0) r = eps * 100.0 * fsign(1.0, r); /* mod.rs:548-552 */
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          f = bnd[i + 1];
          bnd[i + 1] = s * bnd[i] + c * f;
          bnd[i] = c * bnd[i] - s * f;
          if (i == 0) break;
          --i;
        }
        if (underflow)
          d[i + 1] -= p;
        else {
          d[l] -= p;
          e[l] = g;
        }
        e[m] = 0.0;
      }
    }
  }
  if (issues)
    fprintf(stderr, "Warning: imtqlb had some convergence issues but continued with best "
                    "estimates. Results may have reduced accuracy.\n");
  return ORC_OK;
}

/* mod.rs:901-1028. z is row-major nm x n; rotations mix columns i, i+1. */
int orc_imtql2(int64_t nm, int64_t n, double *d, double *e, double *z, int64_t max_iter,
               double eps) {
  if (n == 1) return ORC_OK;
  const int64_t last = n - 1;
  for (int64_t i = 1; i < n; ++i) e[i - 1] = e[i];
  e[last] = 0.0;
  const int64_t nnm = n * nm;
  for (int64_t l = 0; l < n; ++l) {
    int64_t iteration = 0;
    while (iteration <= max_iter) {
      int64_t m = l;
      while (m < n) {
        if (m == last) break;
        double test = fabs(d[m]) + fabs(d[m + 1]);
        if (cmp_eps(test, test + fabs(e[m]), eps)) break;
        ++m;
      }
      if (m == l) break;
      if (iteration == max_iter) {
        set_err("imtql2 no convergence to an eigenvalue after %lld iterations",
                (long long)max_iter);
        return ORC_ERR_IMTQL2;
      }
      ++iteration;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = orc_pythag(g, 1.0, eps);
      g = d[m] - d[l] + e[l] / (g + fsign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int64_t i = m - 1;
      int underflow = 0;
      while (!underflow && i >= l) {
        double f = s * e[i];
        double b = c * e[i];
        r = orc_pythag(f, g, eps);
        e[i + 1] = r;
        if (cmp_eps(r, 0.0, eps)) {
          underflow = 1;
        } else {
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int64_t k = 0; k < nnm; k += n) {
            int64_t idx = k + i;
            f = z[idx + 1];
            z[idx + 1] = s * z[idx] + c * f;
            z[idx] = c * z[idx] - s * f;
          }
          if (i == 0) break;
          --i;
        }
      }
      if (underflow)
        d[i + 1] -= p;
      else {
        d[l] -= p;
        e[l] = g;
      }
      e[m] = 0.0;
    }
  }
  for (int64_t l = 1; l < n; ++l) {
    int64_t i = l - 1, k = i;
    double p = d[i];
    for (int64_t j = l; j < n; ++j)
      if (d[j] < p) {
        k = j;
        p = d[j];
      }
    if (k != i) {
      d[k] = d[i];
      d[i] = p;
      for (int64_t j = 0; j < nnm; j += n) {
        double t = z[j + i];
        z[j + i] = z[j + k];
        z[j + k] = t;
      }
    }
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Start vector                                                          */
/* ------------------------------------------------------------------ */

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_start_vector(uint32_t seed, int64_t n, double *out) {
  const uint64_t phi = 0x9E3779B97F4A7C15ULL;
  const uint64_t key = ((uint64_t)seed * phi) ^ 0x5356444C49423031ULL;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t u = mix64(key + (uint64_t)(i + 1) * phi);
    out[i] = 2.0 * ((double)(u >> 11) * 0x1.0p-53) - 1.0;
  }
}

/* ------------------------------------------------------------------ */
/* LAS2 (mod.rs:104-1406)                                               */
/* ------------------------------------------------------------------ */

#define MAXLL 2

typedef struct {
  const orc_op *A;
  double eps;
  int legacy; /* the f64 legacy API's control flow (legacy.rs:569-663, 1924-2014) */
  int64_t nrows, ncols; /* working (possibly transposed) shape */
  int transposed;
  double *w0, *w1, *w2, *w3, *w4, *w5;
  double *alf, *eta, *oldeta, *bet, *bnd, *ritz, *temp;
  /* Store (mod.rs:193-225): slot idx+MAXLL holds q_idx, slots 0,1 hold p. */
  double **vecs;
  int64_t nvecs, capvecs;
} las2_ws;

static int store_grow(las2_ws *w, int64_t want) {
  while (w->nvecs < want) {
    if (w->nvecs == w->capvecs) {
      int64_t cap = w->capvecs ? 2 * w->capvecs : 16;
      double **nv = (double **)realloc(w->vecs, (size_t)cap * sizeof(double *));
      if (!nv) return ORC_ERR_ALLOC;
      w->vecs = nv;
      w->capvecs = cap;
    }
    w->vecs[w->nvecs] = (double *)calloc((size_t)w->ncols, sizeof(double));
    if (!w->vecs[w->nvecs]) return ORC_ERR_ALLOC;
    ++w->nvecs;
  }
  return ORC_OK;
}

static void storq(las2_ws *w, int64_t idx, const double *v) {
  store_grow(w, idx + MAXLL + 1);
  memcpy(w->vecs[idx + MAXLL], v, (size_t)w->ncols * sizeof(double));
}

static void storp(las2_ws *w, int64_t idx, const double *v) {
  store_grow(w, idx + 1);
  memcpy(w->vecs[idx], v, (size_t)w->ncols * sizeof(double));
}

#define retrq(w, idx) ((w)->vecs[(idx) + MAXLL])
#define retrp(w, idx) ((w)->vecs[(idx)])

#define SWAP(a, b)     \
  do {                 \
    double *t_ = (a);  \
    (a) = (b);         \
    (b) = t_;          \
  } while (0)

/* mod.rs:309-319 */
static void opb(las2_ws *w, const double *x, double *y) {
  orc_opa(w->A, x, w->temp, w->transposed);
  orc_opa(w->A, w->temp, y, !w->transposed);
}

/* mod.rs:587-640 */
static int startv(las2_ws *w, int64_t step, uint32_t seed, double *rnm_out) {
  const int64_t n = w->ncols;
  double rnm2 = ddot(n, w->w0, w->w0);
  for (int id = 0; id < 3; ++id) {
    if (id > 0 || step > 0 || cmp_eps(rnm2, 0.0, w->eps)) orc_start_vector(seed, n, w->w0);
    memcpy(w->w3, w->w0, (size_t)n * sizeof(double));
    opb(w, w->w3, w->w0);
    memcpy(w->w3, w->w0, (size_t)n * sizeof(double));
    rnm2 = ddot(n, w->w3, w->w3);
    if (rnm2 > 0.0) break;
  }
  if (rnm2 <= 0.0) {
    set_err("rnm2 <= 0.0, rnm2 = %g", rnm2);
    return ORC_ERR_STARTV;
  }
  if (step > 0) {
    for (int64_t i = 0; i < step; ++i) {
      double *v = retrq(w, i);
      daxpy(n, -ddot(n, w->w3, v), v, w->w0);
    }
    daxpy(n, -ddot(n, w->w4, w->w0), w->w2, w->w0);
    memcpy(w->w3, w->w0, (size_t)n * sizeof(double));
    double dot = ddot(n, w->w3, w->w3);
    rnm2 = dot <= w->eps * rnm2 ? 0.0 : dot;
  }
  *rnm_out = sqrt(rnm2);
  return ORC_OK;
}

/* mod.rs:643-670 */
static int stpone(las2_ws *w, uint32_t seed, double *rnm_out, double *tol_out) {
  const int64_t n = w->ncols;
  double rnm;
  int st = startv(w, 0, seed, &rnm);
  if (st) return st;
  if (cmp_eps(rnm, 0.0, w->eps)) {
    set_err("rnm == 0.0");
    return ORC_ERR_STPONE;
  }
  datx(n, 1.0 / rnm, w->w0, w->w1);
  dscal(n, 1.0 / rnm, w->w3);
  opb(w, w->w3, w->w0);
  w->alf[0] = ddot(n, w->w0, w->w3);
  daxpy(n, -w->alf[0], w->w1, w->w0);
  double t = ddot(n, w->w0, w->w3);
  w->alf[0] += t;
  daxpy(n, -t, w->w1, w->w0);
  memcpy(w->w4, w->w0, (size_t)n * sizeof(double));
  rnm = sqrt(ddot(n, w->w4, w->w4));
  double anorm = rnm + fabs(w->alf[0]);
  *rnm_out = rnm;
  *tol_out = sqrt(w->eps) * anorm;
  return ORC_OK;
}

/* mod.rs:815-838 */
static void ortbnd(las2_ws *w, int64_t step, double rnm, double eps1) {
  if (step < 1) return;
  if (!cmp_eps(rnm, 0.0, w->eps) && step > 1) {
    w->oldeta[0] = (w->bet[1] * w->eta[1] + (w->alf[0] - w->alf[step]) * w->eta[0] -
                    w->bet[step] * w->oldeta[0]) /
                       rnm +
                   eps1;
    for (int64_t i = 1; i + 2 <= step; ++i)
      w->oldeta[i] = (w->bet[i + 1] * w->eta[i + 1] + (w->alf[i] - w->alf[step]) * w->eta[i] +
                      w->bet[i] * w->eta[i - 1] - w->bet[step] * w->oldeta[i]) /
                         rnm +
                     eps1;
  }
  w->oldeta[step - 1] = eps1;
  SWAP(w->oldeta, w->eta);
  w->eta[step] = eps1;
}

/* mod.rs:760-813 (classical Gram-Schmidt, up to two passes) */
static void purge(las2_ws *w, int64_t ll, int64_t step, double *rnm, double tol) {
  const int64_t n = w->ncols;
  if (step < ll + 2) return;
  const double reps = sqrt(w->eps);
  const double eps1 = w->eps * sqrt((double)n);
  /* mod.rs:777 searches eta from index 0 and adds ll. */
  int64_t k = idamax(step - (ll + 1), w->eta) + ll;
  if (fabs(w->eta[k]) > reps) {
    const double reps1 = eps1 / reps;
    int iteration = 0, flag = 1;
    while (iteration < 2 && flag) {
      if (*rnm > tol) {
        double tq = 0.0, tr = 0.0;
        for (int64_t i = ll; i < step; ++i) {
          double *v = retrq(w, i);
          double t = ddot(n, v, w->w3);
          tq += fabs(t);
          daxpy(n, -t, v, w->w1);
          t = ddot(n, v, w->w4);
          tr += fabs(t);
          daxpy(n, -t, v, w->w0);
        }
        memcpy(w->w3, w->w1, (size_t)n * sizeof(double));
        double t = ddot(n, w->w0, w->w3);
        tr += fabs(t);
        daxpy(n, -t, w->w1, w->w0);
        memcpy(w->w4, w->w0, (size_t)n * sizeof(double));
        *rnm = sqrt(ddot(n, w->w4, w->w4));
        if (tq <= reps1 && tr <= *rnm * reps1) flag = 0;
      }
      ++iteration;
    }
    for (int64_t i = ll; i <= step; ++i) {
      w->eta[i] = eps1;
      w->oldeta[i] = eps1;
    }
  }
}

/* mod.rs:674-758 */
static int lanczos_step(las2_ws *w, int64_t first, int64_t last, int64_t *ll, int *enough,
                        double *rnm, double *tol, int64_t *j_out) {
  const int64_t n = w->ncols;
  const double eps1 = w->eps * sqrt((double)n);
  int64_t j = first;
  while (j < last) {
    SWAP(w->w1, w->w2);
    SWAP(w->w3, w->w4);
    storq(w, j - 1, w->w2);
    if (j - 1 < MAXLL) storp(w, j - 1, w->w4);
    w->bet[j] = *rnm;
    if (cmp_eps(*rnm, 0.0, w->eps)) {
      int st = startv(w, j, 0u, rnm);
      if (st) return st;
      if (cmp_eps(*rnm, 0.0, w->eps)) *enough = 1;
    }
    if (*enough) {
      SWAP(w->w1, w->w2);
      break;
    }
    datx(n, 1.0 / *rnm, w->w0, w->w1);
    dscal(n, 1.0 / *rnm, w->w3);
    opb(w, w->w3, w->w0);
    daxpy(n, -*rnm, w->w2, w->w0);
    w->alf[j] = ddot(n, w->w0, w->w3);
    daxpy(n, -w->alf[j], w->w1, w->w0);
    if (j <= MAXLL && fabs(w->alf[j - 1]) > 4.0 * fabs(w->alf[j])) *ll = j;
    int64_t lim = (j - 1) < *ll ? (j - 1) : *ll;
    for (int64_t i = 0; i < lim; ++i) {
      double t = ddot(n, retrp(w, i), w->w0);
      daxpy(n, -t, retrq(w, i), w->w0);
      w->eta[i] = eps1;
      w->oldeta[i] = eps1;
    }
    double t = ddot(n, w->w0, w->w4);
    daxpy(n, -t, w->w2, w->w0);
    if (w->bet[j] > 0.0) w->bet[j] += t;
    t = ddot(n, w->w0, w->w3);
    daxpy(n, -t, w->w1, w->w0);
    w->alf[j] += t;
    memcpy(w->w4, w->w0, (size_t)n * sizeof(double));
    *rnm = sqrt(ddot(n, w->w4, w->w4));
    double anorm = w->bet[j] + fabs(w->alf[j]) + *rnm;
    *tol = sqrt(w->eps) * anorm;
    ortbnd(w, j, *rnm, eps1);
    purge(w, *ll, j, rnm, *tol);
    if (*rnm <= *tol) *rnm = 0.0;
    ++j;
  }
  *j_out = j;
  return ORC_OK;
}

/* legacy.rs:961-1067: EISPACK imtqlb with the absolute convergence test
 * compare(test, test + |e[m]|) (|b - a| < eps, legacy.rs:809-812), at most
 * 30 QL iterations per eigenvalue (then ImtqlbError), underflow when
 * compare(r, 0). */
int orc_imtqlb_legacy(int64_t n, double *d, double *e, double *bnd, double eps) {
  if (n == 1) return ORC_OK;
  bnd[0] = 1.0;
  const int64_t last = n - 1;
  for (int64_t i = 1; i <= last; ++i) {
    bnd[i] = 0.0;
    e[i - 1] = e[i];
  }
  e[last] = 0.0;
  for (int64_t l = 0; l <= last; ++l) {
    for (int iteration = 0; iteration <= 30;) {
      int64_t m = l;
      while (m < last) {
        double test = fabs(d[m]) + fabs(d[m + 1]);
        if (cmp_eps(test, test + fabs(e[m]), eps)) break;
        ++m;
      }
      double p = d[l], f = bnd[l];
      if (m == l) {
        int64_t i = l;
        while (i >= 1 && p < d[i - 1]) {
          d[i] = d[i - 1];
          bnd[i] = bnd[i - 1];
          --i;
        }
        d[i] = p;
        bnd[i] = f;
        break;
      }
      if (iteration == 30) {
        set_err("imtqlb no convergence to an eigenvalue after 30 iterations");
        return ORC_ERR_IMTQLB;
      }
      ++iteration;
      double g = (d[l + 1] - p) / (2.0 * e[l]);
      double r = orc_pythag(g, 1.0, eps);
      g = d[m] - p + e[l] / (g + fsign(r, g));
      double sn = 1.0, cs = 1.0;
      p = 0.0;
      int64_t i = m - 1;
      int underflow = 0;
      for (;;) {
        f = sn * e[i];
        double b = cs * e[i];
        r = orc_pythag(f, g, eps);
        e[i + 1] = r;
        if (cmp_eps(r, 0.0, eps)) {
          underflow = 1;
          break;
        }
        sn = f / r;
        cs = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * sn + 2.0 * cs * b;
        p = sn * r;
        d[i + 1] = g + p;
        g = cs * r - b;
        f = bnd[i + 1];
        bnd[i + 1] = sn * bnd[i] + cs * f;
        bnd[i] = cs * bnd[i] - sn * f;
        if (i == l) break;
        --i;
      }
      if (underflow) {
        d[i + 1] -= p;
      } else {
        d[l] -= p;
        e[l] = g;
      }
      e[m] = 0.0;
    }
  }
  return ORC_OK;
}

/* legacy.rs:1575-1692: imtql2 with the same convergence and underflow tests,
 * 30 iterations per eigenvalue (then Imtql2Error), eigenvectors in the
 * columns of z (row-major nm x n: rotations mix columns i and i + 1 of every
 * row), then a selection sort by ascending eigenvalue with column swaps. */
int orc_imtql2_legacy(int64_t nm, int64_t n, double *d, double *e, double *z, double eps) {
  if (n == 1) return ORC_OK;
  const int64_t last = n - 1, nnm = n * nm;
  for (int64_t i = 1; i < n; ++i) e[i - 1] = e[i];
  e[last] = 0.0;
  for (int64_t l = 0; l < n; ++l) {
    for (int iteration = 0; iteration <= 30;) {
      int64_t m = l;
      while (m < last) {
        double test = fabs(d[m]) + fabs(d[m + 1]);
        if (cmp_eps(test, test + fabs(e[m]), eps)) break;
        ++m;
      }
      if (m == l) break;
      if (iteration == 30) {
        set_err("imtql2 no convergence to an eigenvalue after 30 iterations");
        return ORC_ERR_IMTQL2;
      }
      ++iteration;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = orc_pythag(g, 1.0, eps);
      g = d[m] - d[l] + e[l] / (g + fsign(r, g));
      double sn = 1.0, cs = 1.0, p = 0.0;
      int64_t i = m - 1;
      int underflow = 0;
      for (;;) {
        double f = sn * e[i];
        double b = cs * e[i];
        r = orc_pythag(f, g, eps);
        e[i + 1] = r;
        if (cmp_eps(r, 0.0, eps)) {
          underflow = 1;
          break;
        }
        sn = f / r;
        cs = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * sn + 2.0 * cs * b;
        p = sn * r;
        d[i + 1] = g + p;
        g = cs * r - b;
        for (int64_t k = 0; k < nnm; k += n) {
          f = z[k + i + 1];
          z[k + i + 1] = sn * z[k + i] + cs * f;
          z[k + i] = cs * z[k + i] - sn * f;
        }
        if (i == l) break;
        --i;
      }
      if (underflow) {
        d[i + 1] -= p;
      } else {
        d[l] -= p;
        e[l] = g;
      }
      e[m] = 0.0;
    }
  }
  for (int64_t l = 1; l < n; ++l) {
    int64_t i = l - 1, k = i;
    double p = d[i];
    for (int64_t j = l; j < n; ++j)
      if (d[j] < p) {
        k = j;
        p = d[j];
      }
    if (k != i) {
      d[k] = d[i];
      d[i] = p;
      for (int64_t j = 0; j < nnm; j += n) {
        double t = z[j + i];
        z[j + i] = z[j + k];
        z[j + k] = t;
      }
    }
  }
  return ORC_OK;
}

/* mod.rs:297-307 */
static void insert_sort(int64_t n, double *a1, double *a2) {
  for (int64_t i = 1; i < n; ++i)
    for (int64_t j = i; j >= 1; --j) {
      if (a1[j - 1] <= a1[j]) break;
      double t = a1[j - 1];
      a1[j - 1] = a1[j];
      a1[j] = t;
      t = a2[j - 1];
      a2[j - 1] = a2[j];
      a2[j] = t;
    }
}

/* mod.rs:840-899 */
static int64_t error_bound(int *enough, double endl, double endr, double *ritz, double *bnd,
                           int64_t step, double tol, double eps) {
  const double eps34 = pow(eps, 0.75);
  int64_t mid = idamax(step + 1, bnd);
  int64_t i = ((step + 1) + (step - 1)) / 2;
  while (i > mid + 1) {
    if (fabs(ritz[i - 1] - ritz[i]) < eps34 * fabs(ritz[i]) && bnd[i] > tol &&
        bnd[i - 1] > tol) {
      bnd[i - 1] = sqrt(bnd[i] * bnd[i] + bnd[i - 1] * bnd[i - 1]);
      bnd[i] = 0.0;
    }
    --i;
  }
  i = ((step + 1) - (step - 1)) / 2;
  while (i + 1 < mid) {
    if (fabs(ritz[i + 1] - ritz[i]) < eps34 * fabs(ritz[i]) && bnd[i] > tol &&
        bnd[i + 1] > tol) {
      bnd[i + 1] = sqrt(bnd[i] * bnd[i] + bnd[i + 1] * bnd[i + 1]);
      bnd[i] = 0.0;
    }
    ++i;
  }
  int64_t neig = 0;
  double gapl = ritz[step] - ritz[0];
  for (i = 0; i <= step; ++i) {
    double gap = gapl;
    if (i < step) gapl = ritz[i + 1] - ritz[i];
    gap = fmin(gap, gapl);
    if (gap > bnd[i]) bnd[i] *= bnd[i] / gap;
    if (bnd[i] <= 16.0 * eps * fabs(ritz[i])) {
      ++neig;
      if (!*enough) *enough = endl < ritz[i] && ritz[i] < endr;
    }
  }
  return neig;
}

static double sparsity_of(const orc_op *A) {
  return 1.0 - (double)orc_nnz(A) / ((double)orc_nrows(A) * (double)orc_ncols(A));
}

/* mod.rs:1263-1406 */
static int lanso(las2_ws *w, int64_t dim, int64_t iterations, double endl, double endr,
                 uint32_t seed, int64_t *neig, int64_t *j_out) {
  const double sparsity = sparsity_of(w->A);
  const int64_t max_imtqlb =
      sparsity > 0.999 ? 500 : sparsity > 0.99 ? 300 : sparsity > 0.9 ? 100 : 50;
  const double eps = w->eps;
  /* legacy.rs:1924-2014: plain eps, splits at compare(bet, 0), no extra steps */
  const double adaptive_eps =
      w->legacy ? eps : sparsity > 0.99 ? eps * 100.0 : sparsity > 0.9 ? eps * 10.0 : eps;
  double rnm, tol;
  int st = stpone(w, seed, &rnm, &tol);
  if (st) return st;
  const double eps1 = adaptive_eps * sqrt((double)w->ncols);
  w->eta[0] = eps1;
  w->oldeta[0] = eps1;
  int64_t ll = 0, first = 1, j = 0, intro = 0;
  int64_t dm = dim > 8 ? dim : 8;
  int64_t last = iterations < dm + dim ? iterations : dm + dim;
  int enough = 0;
  while (!enough) {
    if (rnm <= tol) rnm = 0.0;
    int64_t steps;
    st = lanczos_step(w, first, last, &ll, &enough, &rnm, &tol, &steps);
    if (st) return st;
    j = enough ? steps - 1 : last - 1;
    first = j + 1;
    w->bet[first] = rnm;
    int64_t l = 0;
    for (int64_t it = 0; it < j; ++it) {
      if (l > j) break;
      int64_t i = l;
      while (i <= j) {
        if (w->legacy ? cmp_eps(w->bet[i + 1], 0.0, eps) : fabs(w->bet[i + 1]) <= adaptive_eps) break;
        ++i;
      }
      if (i > j) i = j;
      int64_t sz = i - l;
      dcopy_rev(sz + 1, l, w->alf, w->ritz);
      dcopy_rev(sz, l + 1, w->bet, w->w5);
      st = w->legacy ? orc_imtqlb_legacy(sz + 1, w->ritz + l, w->w5 + l, w->bnd + l, eps)
                     : orc_imtqlb(sz + 1, w->ritz + l, w->w5 + l, w->bnd + l, max_imtqlb, eps);
      if (st) return st;
      for (int64_t m = l; m <= i; ++m) w->bnd[m] = rnm * fabs(w->bnd[m]);
      l = i + 1;
    }
    insert_sort(j + 1, w->ritz, w->bnd);
    if (j < 1) {
      set_err("error_bound: expected 'step' to be non-zero");
      return ORC_ERR_LAS2;
    }
    *neig = error_bound(&enough, endl, endr, w->ritz, w->bnd, j, tol, eps);
    if (*neig < dim) {
      if (*neig == 0) {
        last = first + 9;
        intro = first;
      } else {
        int64_t extra = !w->legacy && sparsity > 0.99 ? 5 : 0;
        int64_t q = 1 + ((j - intro) * (dim - *neig)) / *neig;
        last = first + (q > 3 ? q : 3) + extra;
      }
      if (last > iterations) last = iterations;
    } else {
      enough = 1;
    }
    enough = enough || first >= iterations;
  }
  storq(w, j, w->w1);
  *j_out = j;
  return ORC_OK;
}

/* mod.rs:1056-1259 with D2 and D3 corrected as legacy.rs:1796-1835 does. */
static int ritvec(las2_ws *w, int64_t dimensions, double kappa, int64_t steps, int64_t neig,
                  orc_svdrec *out_raw, int64_t *nsig_out) {
  const int64_t js = steps + 1, jsq = js * js;
  const double eps = w->eps;
  const double sparsity = sparsity_of(w->A);
  const double adaptive_eps =
      w->legacy ? eps : sparsity > 0.99 ? eps * 100.0 : sparsity > 0.9 ? eps * 10.0 : eps;
  const int64_t max_imtql2 =
      sparsity > 0.999 ? 500 : sparsity > 0.99 ? 300 : sparsity > 0.9 ? 200 : 50;
  double *s = (double *)calloc((size_t)jsq, sizeof(double));
  double *ev = (double *)calloc((size_t)js, sizeof(double));
  if (!s || !ev) {
    free(s);
    free(ev);
    return ORC_ERR_ALLOC;
  }
  for (int64_t i = 0; i < jsq; i += js + 1) s[i] = 1.0;
  dcopy_rev(js, 0, w->alf, ev);
  dcopy_rev(steps, 1, w->bet, w->w5);
  int st = w->legacy ? orc_imtql2_legacy(js, js, ev, w->w5, s, eps)
                     : orc_imtql2(js, js, ev, w->w5, s, max_imtql2, eps);
  if (st) {
    free(s);
    free(ev);
    return st;
  }
  double max_ev = 0.0;
  for (int64_t i = 0; i < js; ++i) max_ev = fmax(max_ev, fabs(ev[i]));
  const double akappa = !w->legacy && sparsity > 0.99 ? kappa * 10.0 : kappa;
  int64_t *sig = (int64_t *)malloc((size_t)js * sizeof(int64_t));
  int64_t nsig = 0;
  for (int64_t k = 0; k < js; ++k) {
    /* legacy.rs:1798: bnd[k] <= kappa |ritz[k]| */
    double rel = w->legacy ? kappa * fabs(w->ritz[k]) : akappa * fmax(fabs(w->ritz[k]), max_ev * adaptive_eps);
    if (w->bnd[k] <= rel && k + 1 > js - neig) sig[nsig++] = k;
  }
  const int64_t d = dimensions < nsig ? dimensions : nsig;
  const int64_t nr = w->nrows, nc = w->ncols;
  double *vt = (double *)calloc((size_t)(d * nc + 1), sizeof(double));
  double *ut = (double *)calloc((size_t)(d * nr + 1), sizeof(double));
  double *sv = (double *)calloc((size_t)(d + 1), sizeof(double));
  double *tmp = (double *)calloc((size_t)nc, sizeof(double));
  /* D3: the d largest significant Ritz pairs, largest first. */
  for (int64_t r = 0; r < d; ++r) {
    int64_t k = sig[nsig - 1 - r];
    double *v = vt + r * nc;
    for (int64_t i = 0; i < js; ++i) {
      double coef = s[(js - 1 - i) * js + k]; /* D2 */
      if (w->legacy ? coef != 0.0 : fabs(coef) > adaptive_eps) daxpy(nc, coef, retrq(w, i), v);
    }
  }
  for (int64_t r = 0; r < d; ++r) {
    double *v = vt + r * nc;
    opb(w, v, tmp);
    double t = ddot(nc, v, tmp);
    double sval = w->legacy ? sqrt(t) : sqrt(fmax(t, 0.0)); /* legacy.rs:1849 */
    sv[r] = sval;
    double *u = ut + r * nr;
    orc_opa(w->A, v, u, w->transposed);
    double scale = w->legacy ? 1.0 / sval : sval > adaptive_eps ? 1.0 / sval : 1.0 / fmax(sval, adaptive_eps);
    dscal(nr, scale, u);
  }
  free(tmp);
  free(sig);
  free(s);
  free(ev);
  out_raw->d = d;
  out_raw->u = ut;
  out_raw->s = sv;
  out_raw->vt = vt;
  *nsig_out = nsig;
  return ORC_OK;
}

static void ws_free(las2_ws *w) {
  free(w->w0);
  free(w->w1);
  free(w->w2);
  free(w->w3);
  free(w->w4);
  free(w->w5);
  free(w->alf);
  free(w->eta);
  free(w->oldeta);
  free(w->bet);
  free(w->bnd);
  free(w->ritz);
  free(w->temp);
  for (int64_t i = 0; i < w->nvecs; ++i) free(w->vecs[i]);
  free(w->vecs);
}

void orc_svdrec_free(orc_svdrec *rec) {
  free(rec->u);
  free(rec->s);
  free(rec->vt);
  rec->u = rec->s = rec->vt = NULL;
}

/* mod.rs:104-189 */
static int las2_run(const orc_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
                    double kappa, uint32_t random_seed, const double *start, double eps, int legacy,
                    orc_svdrec *out);

int orc_las2(const orc_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
             double kappa, uint32_t random_seed, const double *start, double eps,
             orc_svdrec *out) {
  return las2_run(op, dimensions, iterations, endl, endr, kappa, random_seed, start, eps, 0, out);
}

/* legacy::svdLAS2 (legacy.rs:569-663): f64, the legacy control flow. */
int orc_las2_legacy(const orc_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
                    double kappa, uint32_t random_seed, const double *start, orc_svdrec *out) {
  return las2_run(op, dimensions, iterations, endl, endr, kappa, random_seed, start, DBL_EPSILON, 1, out);
}

static int las2_run(const orc_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
                    double kappa, uint32_t random_seed, const double *start, double eps, int legacy,
                    orc_svdrec *out) {
  memset(out, 0, sizeof *out);
  g_blas_threads = nthreads_of(op);
  if (random_seed == 0) random_seed = (uint32_t)(mix64((uint64_t)&out ^ 0x1234567ULL) >> 32);
  const int64_t m = orc_nrows(op), n = orc_ncols(op);
  const int64_t mn = m < n ? m : n;
  if (dimensions == 0 || dimensions > mn) dimensions = mn;
  if (iterations == 0 || iterations > mn)
    iterations = mn;
  else if (iterations < dimensions)
    iterations = dimensions;
  if (dimensions < 2) {
    set_err("%s: insufficient dimensions: %lld", legacy ? "svdLAS2" : "svd_las2", (long long)dimensions);
    return ORC_ERR_LAS2;
  }
  las2_ws w;
  memset(&w, 0, sizeof w);
  w.A = op;
  w.eps = eps;
  w.legacy = legacy;
  w.transposed = (double)n >= (double)m * 1.2;
  w.nrows = w.transposed ? n : m;
  w.ncols = w.transposed ? m : n;
  const int64_t nc = w.ncols;
  w.w0 = (double *)calloc((size_t)nc, sizeof(double));
  w.w1 = (double *)calloc((size_t)nc, sizeof(double));
  w.w2 = (double *)calloc((size_t)nc, sizeof(double));
  w.w3 = (double *)calloc((size_t)nc, sizeof(double));
  w.w4 = (double *)calloc((size_t)nc, sizeof(double));
  w.w5 = (double *)calloc((size_t)nc, sizeof(double));
  w.alf = (double *)calloc((size_t)iterations, sizeof(double));
  w.eta = (double *)calloc((size_t)iterations, sizeof(double));
  w.oldeta = (double *)calloc((size_t)iterations, sizeof(double));
  w.bet = (double *)calloc((size_t)(iterations + 1), sizeof(double));
  w.ritz = (double *)calloc((size_t)(iterations + 1), sizeof(double));
  w.bnd = (double *)calloc((size_t)(iterations + 1), sizeof(double));
  w.temp = (double *)calloc((size_t)w.nrows, sizeof(double));
  for (int64_t i = 0; i <= iterations; ++i) w.bnd[i] = DBL_MAX;
  if (start) memcpy(w.w0, start, (size_t)nc * sizeof(double));
  int64_t neig = 0, steps = 0, nsig = 0;
  int st = lanso(&w, dimensions, iterations, endl, endr, random_seed, &neig, &steps);
  if (!st) {
    double kap = fmax(fabs(kappa), pow(eps, 0.75));
    st = ritvec(&w, dimensions, kap, steps, neig, out, &nsig);
    out->diag.kappa = kap;
  }
  if (!st) {
    out->m = m;
    out->n = n;
    if (w.transposed) {
      double *t = out->u;
      out->u = out->vt;
      out->vt = t;
    }
    out->diag.non_zero = orc_nnz(op);
    out->diag.dimensions = dimensions;
    out->diag.iterations = iterations;
    out->diag.transposed = w.transposed;
    out->diag.lanczos_steps = steps + 1;
    out->diag.ritz_values_stabilized = neig;
    out->diag.significant_values = out->d;
    out->diag.singular_values = nsig;
    out->diag.end_interval[0] = endl;
    out->diag.end_interval[1] = endr;
    out->diag.random_seed = random_seed;
  }
  ws_free(&w);
  return st;
}

/* ---- COO -> CSR ---------------------------------------------------------
 * Two stable counting sorts (by column, then by row) give the (row, column)
 * order with equal pairs in input order; each run of equal pairs is then
 * summed left to right. */
static int64_t coo_order(int64_t m, int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols,
                         int64_t *ord) {
  for (int64_t k = 0; k < nnz; ++k)
    if (rows[k] < 0 || rows[k] >= m || cols[k] < 0 || cols[k] >= n) return -1;
  int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
  int64_t *cnt = (int64_t *)calloc((size_t)((m > n ? m : n) + 1), sizeof(int64_t));
  for (int64_t k = 0; k < nnz; ++k) ++cnt[cols[k] + 1];
  for (int64_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
  for (int64_t k = 0; k < nnz; ++k) tmp[cnt[cols[k]]++] = k;
  memset(cnt, 0, sizeof(int64_t) * (size_t)((m > n ? m : n) + 1));
  for (int64_t k = 0; k < nnz; ++k) ++cnt[rows[k] + 1];
  for (int64_t i = 0; i < m; ++i) cnt[i + 1] += cnt[i];
  for (int64_t t = 0; t < nnz; ++t) {
    const int64_t k = tmp[t];
    ord[cnt[rows[k]]++] = k;
  }
  free(tmp);
  free(cnt);
  return 0;
}

#define ORC_COO_TO_CSR(NAME, T)                                                                      \
  int64_t NAME(int64_t m, int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols,         \
               const T *vals, int64_t *rowptr, int64_t *out_cols, T *out_vals) {                    \
    int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));               \
    if (coo_order(m, n, nnz, rows, cols, ord)) {                                                     \
      free(ord);                                                                                     \
      return -1;                                                                                     \
    }                                                                                                \
    int64_t u = 0, r = 0;                                                                            \
    rowptr[0] = 0;                                                                                   \
    for (int64_t t = 0; t < nnz;) {                                                                  \
      const int64_t k = ord[t];                                                                      \
      T s = vals[k];                                                                                 \
      int64_t e = t + 1;                                                                             \
      while (e < nnz && rows[ord[e]] == rows[k] && cols[ord[e]] == cols[k]) s += vals[ord[e++]];     \
      while (r < rows[k]) rowptr[++r] = u;                                                           \
      out_cols[u] = cols[k];                                                                         \
      out_vals[u++] = s;                                                                             \
      t = e;                                                                                         \
    }                                                                                                \
    while (r < m) rowptr[++r] = u;                                                                   \
    free(ord);                                                                                       \
    return u;                                                                                        \
  }
ORC_COO_TO_CSR(orc_coo_to_csr, double)
ORC_COO_TO_CSR(orc_coo_to_csr_f32, float)
