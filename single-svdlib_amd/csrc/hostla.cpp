// Host FP64 dense algebra of the svdh library: the tridiagonal implicit-QL
// solvers of LAS2 and the small symmetric eigen / Cholesky routines of the
// randomized solver. These run on l x l or js x js matrices (l <= 256,
// js = Lanczos steps + 1), next to the device work.
//
// imtqlb / imtql2 take the steps of src/lanczos/mod.rs:428-584 and 901-1028
// (the reference's sparsity-dependent iteration caps, relaxed deflation and
// forced-convergence behaviour), with src/legacy.rs's corrections:
//   - pythag updates r *= (s/u)^2 (legacy.rs:882; mod.rs:373 drops the r);
//   - imtqlb takes p = d[l] and f = bnd[l] afresh on every QL iteration
//     (legacy.rs:990-991; mod.rs:455-457 reads them once per l).
// Their control flow comes from SVDLIBC (Copyright (c) 2002, University of
// Tennessee Research Foundation, BSD licence).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "common.h"

namespace svdh {

static inline bool near(double a, double b, double eps) { return std::fabs(b - a) < eps; }

static inline double sign_of(double a, double b) {
  // |a| with the sign convention of svd_fsign (mod.rs:353-358)
  const bool same = (a >= 0.0 && b >= 0.0) || (a < 0.0 && b < 0.0);
  return same ? a : -a;
}

double pythag(double a, double b, double eps) {
  const double big = std::max(std::fabs(a), std::fabs(b));
  if (!(big > 0.0)) return 0.0;
  double p = big;
  const double q = std::min(std::fabs(a), std::fabs(b)) / big;
  double r = q * q;
  for (double t = 4.0 + r; !near(t, 4.0, eps); t = 4.0 + r) {
    const double s = r / t;
    const double u = 1.0 + 2.0 * s;
    p *= u;
    const double w = s / u;
    r *= w * w;
  }
  return p;
}

int imtqlb(int64_t n, double *d, double *e, double *bnd, int64_t max_iter, double eps) {
  if (n == 1) return SVDH_OK;
  const double size_factor = std::sqrt(static_cast<double>(n));
  const int64_t last = n - 1;
  bnd[0] = 1.0;
  for (int64_t i = 1; i <= last; ++i) {
    bnd[i] = 0.0;
    e[i - 1] = e[i];
  }
  e[last] = 0.0;
  bool warned = false;
  for (int64_t l = 0; l <= last; ++l) {
    for (int64_t it = 0; it <= max_iter;) {
      // Find the first negligible sub-diagonal at or after l.
      int64_t m = l;
      for (; m < last; ++m) {
        const double test = std::fabs(d[m]) + std::fabs(d[m + 1]);
        const double tol = eps * 100.0 * std::max(test, 1.0) * size_factor;
        if (std::fabs(e[m]) <= tol) break;
      }
      double p = d[l];
      double f = bnd[l];
      if (m == l) {
        // Insert the converged value into the ascending prefix d[0..l].
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
      if (it == max_iter) {
        warned = true;
        for (int64_t k = l; k <= m; ++k) bnd[k] = std::max(bnd[k], 0.1);
        e[l] = 0.0;
        break;
      }
      ++it;
      // Implicit shift and the QL sweep from m-1 down to l.
      double g = (d[l + 1] - p) / (2.0 * e[l]);
      double r = pythag(g, 1.0, eps);
      g = d[m] - p + e[l] / (g + sign_of(r, g));
      double s = 1.0, c = 1.0;
      p = 0.0;
      int64_t i = m - 1;
      bool underflow = false;
      for (;;) {
        f = s * e[i];
        const double b = c * e[i];
        r = pythag(f, g, eps);
        e[i + 1] = r;
        if (r < eps * 1000.0 * (std::fabs(f) + std::fabs(g))) {
          underflow = true;
          break;
        }
        if (std::fabs(r) < eps * 100.0) r = eps * 100.0 * sign_of(1.0, r);
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
  if (warned)
    std::fprintf(stderr,
                 "Warning: imtqlb had some convergence issues but continued with best estimates. "
                 "Results may have reduced accuracy.\n");
  return SVDH_OK;
}

// legacy::imtqlb (src/legacy.rs:961-1067): the EISPACK form with the
// absolute tests |b - a| < eps for convergence (test vs test + |e[m]|) and
// underflow (r vs 0), at most 30 QL iterations per eigenvalue and then
// ImtqlbError.
int imtqlb_legacy(int64_t n, double *d, double *e, double *bnd, double eps) {
  if (n == 1) return SVDH_OK;
  const int64_t last = n - 1;
  bnd[0] = 1.0;
  for (int64_t i = 1; i <= last; ++i) {
    bnd[i] = 0.0;
    e[i - 1] = e[i];
  }
  e[last] = 0.0;
  for (int64_t l = 0; l <= last; ++l) {
    for (int it = 0;; ++it) {
      int64_t m = l;
      for (; m < last; ++m) {
        const double test = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (near(test, test + std::fabs(e[m]), eps)) break;
      }
      double p = d[l];
      double f = bnd[l];
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
      if (it == 30) fail(SVDH_ERR_IMTQLB, "imtqlb no convergence to an eigenvalue after 30 iterations");
      double g = (d[l + 1] - p) / (2.0 * e[l]);
      double r = pythag(g, 1.0, eps);
      g = d[m] - p + e[l] / (g + sign_of(r, g));
      double s = 1.0, c = 1.0;
      p = 0.0;
      int64_t i = m - 1;
      bool underflow = false;
      for (;;) {
        f = s * e[i];
        const double b = c * e[i];
        r = pythag(f, g, eps);
        e[i + 1] = r;
        if (near(r, 0.0, eps)) {
          underflow = true;
          break;
        }
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
  return SVDH_OK;
}

// With max_iter = 30 this is also legacy::imtql2 (src/legacy.rs:1575-1692):
// the convergence and underflow tests and the ascending selection sort are
// the same.
int imtql2(int64_t nm, int64_t n, double *d, double *e, double *z, int64_t max_iter, double eps) {
  if (n == 1) return SVDH_OK;
  const int64_t last = n - 1;
  // The rotations mix columns i, i+1 of every row of z (row-major nm x n);
  // they run on the transpose zt (n x nm), where those columns are two
  // contiguous rows, and z is written back at the end. Same arithmetic per
  // element, cache-friendly order.
  std::vector<double> zt((size_t)n * nm);
  for (int64_t row = 0; row < nm; ++row)
    for (int64_t col = 0; col < n; ++col) zt[(size_t)col * nm + row] = z[row * n + col];
  for (int64_t i = 1; i < n; ++i) e[i - 1] = e[i];
  e[last] = 0.0;
  for (int64_t l = 0; l < n; ++l) {
    for (int64_t it = 0;; ++it) {
      int64_t m = l;
      for (; m < last; ++m) {
        const double test = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (near(test, test + std::fabs(e[m]), eps)) break;
      }
      if (m == l) break;
      if (it == max_iter)
        fail(SVDH_ERR_IMTQL2, "imtql2 no convergence to an eigenvalue after %lld iterations",
             (long long)max_iter);
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = pythag(g, 1.0, eps);
      g = d[m] - d[l] + e[l] / (g + sign_of(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int64_t i = m - 1;
      bool underflow = false;
      for (;;) {
        double f = s * e[i];
        const double b = c * e[i];
        r = pythag(f, g, eps);
        e[i + 1] = r;
        if (near(r, 0.0, eps)) {
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        // Rotate columns i and i+1 of z (rows i and i+1 of zt).
        {
          double *zi = zt.data() + (size_t)i * nm;
          double *zj = zi + nm;
          for (int64_t row = 0; row < nm; ++row) {
            const double a = zi[row], b2 = zj[row];
            zj[row] = s * a + c * b2;
            zi[row] = c * a - s * b2;
          }
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
  // Selection sort into ascending order, swapping eigenvector columns.
  for (int64_t i = 0; i + 1 < n; ++i) {
    int64_t k = i;
    double p = d[i];
    for (int64_t j = i + 1; j < n; ++j)
      if (d[j] < p) {
        k = j;
        p = d[j];
      }
    if (k != i) {
      d[k] = d[i];
      d[i] = p;
      std::swap_ranges(zt.begin() + (size_t)i * nm, zt.begin() + (size_t)(i + 1) * nm, zt.begin() + (size_t)k * nm);
    }
  }
  for (int64_t row = 0; row < nm; ++row)
    for (int64_t col = 0; col < n; ++col) z[row * n + col] = zt[(size_t)col * nm + row];
  return SVDH_OK;
}

// Symmetric eigen-decomposition of a small dense matrix (the Rayleigh-Ritz
// and normaliser solves of the randomized SVD; LAS2 keeps imtql2 above).
// Householder reduction to tridiagonal form on the trailing blocks only, the
// reflections accumulated backwards (each step touches the block it changes),
// then implicit QL with a relative deflation test. The eigenvectors are kept
// transposed, so every QL rotation combines two contiguous rows. The matrix
// is scaled by its largest entry first. V[i * n + r] is component i of the
// eigenvector of w[r]; w ascending. With V null only the eigenvalues are
// computed (the same values: the vectors do not feed back into the QL
// recurrence).
// Householder reduction of A / scale (scale = largest |A_ij|, returned; 0 for
// a zero matrix, then nothing else is set) to the tridiagonal (d, e), e[i]
// coupling i and i + 1. Step k zeroes column k below the sub-diagonal with
// H_k = I - betas[k] v v^T, v stored in row k of vs (support k+1..n-1).
static double tridiagonalize(int64_t n, const double *A, std::vector<double> &d, std::vector<double> &e,
                             std::vector<double> &vs, std::vector<double> &betas) {
  double scale = 0.0;
  for (int64_t i = 0; i < n * n; ++i) scale = std::max(scale, std::fabs(A[i]));
  if (!(scale > 0.0)) return 0.0;
  std::vector<double> a(n * n);
  const double inv = 1.0 / scale;
  for (int64_t i = 0; i < n * n; ++i) a[i] = A[i] * inv;
  d.assign(n, 0.0);
  e.assign(n, 0.0);
  vs.assign((size_t)n * n, 0.0);
  betas.assign(n, 0.0);
  std::vector<double> p(n), kv(n);
  // Reduction: step k zeroes column k below the sub-diagonal with
  // H_k = I - beta v v^T, v supported on k+1..n-1 (stored in row k of vs).
  for (int64_t k = 0; k + 2 < n; ++k) {
    double *v = vs.data() + (size_t)k * n;
    double xn = 0.0;
    for (int64_t i = k + 1; i < n; ++i) xn += a[i * n + k] * a[i * n + k];
    xn = std::sqrt(xn);
    const double x0 = a[(k + 1) * n + k];
    if (xn == 0.0) continue;
    const double alpha = x0 > 0.0 ? -xn : xn;
    v[k + 1] = x0 - alpha;
    for (int64_t i = k + 2; i < n; ++i) v[i] = a[i * n + k];
    double vn = 0.0;
    for (int64_t i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double beta = 2.0 / vn;
    betas[k] = beta;
    // column/row k become (.., alpha, 0, ..)
    a[(k + 1) * n + k] = alpha;
    a[k * n + k + 1] = alpha;
    for (int64_t i = k + 2; i < n; ++i) a[i * n + k] = a[k * n + i] = 0.0;
    // trailing block T = a[k+1.., k+1..]: p = beta T v, K = p - (beta/2)(v.p) v,
    // T <- T - v K^T - K v^T
    // p = beta T v as a sum of scaled rows (T symmetric): contiguous axpys
    double *__restrict pp = p.data();
    const double *__restrict vv = v;
    for (int64_t j = k + 1; j < n; ++j) pp[j] = 0.0;
    for (int64_t i = k + 1; i < n; ++i) {
      const double *__restrict ai = a.data() + i * n;
      const double vi = beta * vv[i];
      for (int64_t j = k + 1; j < n; ++j) pp[j] += vi * ai[j];
    }
    double vp = 0.0;
    for (int64_t i = k + 1; i < n; ++i) vp += vv[i] * pp[i];
    const double h = 0.5 * beta * vp;
    double *__restrict kk = kv.data();
    for (int64_t i = k + 1; i < n; ++i) kk[i] = pp[i] - h * vv[i];
    for (int64_t i = k + 1; i < n; ++i) {
      double *__restrict ai = a.data() + i * n;
      const double vi = vv[i], ki = kk[i];
      for (int64_t j = k + 1; j < n; ++j) ai[j] -= vi * kk[j] + ki * vv[j];
    }
  }
  for (int64_t i = 0; i < n; ++i) d[i] = a[i * n + i];
  for (int64_t i = 0; i + 1 < n; ++i) e[i] = a[(i + 1) * n + i];  // e[i] couples i and i+1
  return scale;
}

void sym_eig(int64_t n, const double *A, double *w, double *V) {
  if (n == 0) return;
  const bool vectors = V != nullptr;
  std::vector<double> d, e, vs, betas;
  const double scale = tridiagonalize(n, A, d, e, vs, betas);
  if (!(scale > 0.0)) {
    for (int64_t i = 0; i < n; ++i) w[i] = 0.0;
    if (vectors) {
      for (int64_t i = 0; i < n * n; ++i) V[i] = 0.0;
      for (int64_t i = 0; i < n; ++i) V[i * n + i] = 1.0;
    }
    return;
  }
  // Q = H_0 H_1 ... H_{n-3}, accumulated from the right end: before step k
  // only the block k+1.. of Q differs from the identity.
  std::vector<double> q(vectors ? (size_t)n * n : 0, 0.0), wv(n), zt(vectors ? (size_t)n * n : 0);
  for (int64_t i = 0; vectors && i < n; ++i) q[i * n + i] = 1.0;
  for (int64_t k = n - 3; vectors && k >= 0; --k) {
    const double beta = betas[k];
    if (beta == 0.0) continue;
    const double *v = vs.data() + (size_t)k * n;
    for (int64_t j = k + 1; j < n; ++j) wv[j] = 0.0;
    for (int64_t i = k + 1; i < n; ++i) {
      const double vi = v[i];
      const double *qi = q.data() + i * n;
      for (int64_t j = k + 1; j < n; ++j) wv[j] += vi * qi[j];
    }
    for (int64_t i = k + 1; i < n; ++i) {
      const double f = beta * v[i];
      double *qi = q.data() + i * n;
      for (int64_t j = k + 1; j < n; ++j) qi[j] -= f * wv[j];
    }
  }
  // zt row r = column r of Q
  for (int64_t i = 0; vectors && i < n; ++i)
    for (int64_t r = 0; r < n; ++r) zt[(size_t)r * n + i] = q[i * n + r];
  // Implicit QL on (d, e).
  const double eps = 2.220446049250313e-16;
  auto hyp = [](double x, double y) {
    const double r = std::sqrt(x * x + y * y);
    return (r > 1e-150 && r < 1e150) ? r : std::hypot(x, y);
  };
  for (int64_t l = 0; l < n; ++l) {
    for (int64_t iter = 0;; ++iter) {
      int64_t m = l;
      for (; m + 1 < n; ++m)
        if (std::fabs(e[m]) <= eps * (std::fabs(d[m]) + std::fabs(d[m + 1]))) break;
      if (m == l) break;
      if (iter == 60) fail(SVDH_ERR_IMTQL2, "symmetric eigensolver: no convergence after %d iterations", 60);
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hyp(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, pp = 0.0;
      int64_t i = m - 1;
      bool early = false;
      for (; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = hyp(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        if (vectors) {
          double *__restrict zi = zt.data() + (size_t)i * n;
          double *__restrict zj = zi + n;
          for (int64_t t = 0; t < n; ++t) {
            const double x = zi[t], y = zj[t];
            zj[t] = s * x + c * y;
            zi[t] = c * x - s * y;
          }
        }
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // ascending order
  std::vector<int64_t> ord(n);
  for (int64_t i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return d[x] < d[y]; });
  for (int64_t r = 0; r < n; ++r) {
    w[r] = d[ord[r]] * scale;
    if (!vectors) continue;
    const double *zr = zt.data() + (size_t)ord[r] * n;
    for (int64_t i = 0; i < n; ++i) V[i * n + r] = zr[i];
  }
}

// Smallest eigenvalue of a symmetric matrix: the tridiagonal form, then
// bisection on Sturm counts (the number of negative pivots of T - x I is the
// number of eigenvalues below x) inside the Gershgorin interval, to a
// relative width of 1e-12 of the spectral scale.
double sym_eig_min(int64_t n, const double *A) {
  if (n == 0) return 0.0;
  std::vector<double> d, e, vs, betas;
  const double scale = tridiagonalize(n, A, d, e, vs, betas);
  if (!(scale > 0.0)) return 0.0;
  double lo = INFINITY, hi = -INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const double tiny = 1e-300;
  auto below = [&](double x) {  // eigenvalues of T below x
    int64_t cnt = 0;
    double q = d[0] - x;
    for (int64_t i = 0;; ++i) {
      if (q == 0.0) q = -tiny;
      cnt += q < 0.0;
      if (i + 1 == n) break;
      q = d[i + 1] - x - e[i] * e[i] / q;
    }
    return cnt;
  };
  // 1e-12 of the spectral scale: far below what the caller needs, about 40 halvings
  const double width = 1e-12 * std::max(std::fabs(lo), std::fabs(hi));
  for (int it = 0; it < 200 && hi - lo > width; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (below(mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi) * scale;
}

// Upper Cholesky factor R (A = R^T R) by rows: row j of R from the updated
// row j of A, then the upper triangle of the trailing rows updated by an
// outer product, so every inner loop runs along contiguous rows.
bool cholesky_upper(int64_t l, const double *A, double *R) {
  std::vector<double> a(A, A + l * l);
  for (int64_t i = 0; i < l * l; ++i) R[i] = 0.0;
  for (int64_t j = 0; j < l; ++j) {
    const double s = a[j * l + j];
    if (!(s > 0.0)) return false;
    const double rjj = std::sqrt(s), inv = 1.0 / rjj;
    double *rj = R + j * l;
    const double *aj = a.data() + j * l;
    rj[j] = rjj;
    for (int64_t i = j + 1; i < l; ++i) rj[i] = aj[i] * inv;
    for (int64_t i = j + 1; i < l; ++i) {
      const double f = rj[i];
      double *ai = a.data() + i * l;
      for (int64_t c = i; c < l; ++c) ai[c] -= f * rj[c];
    }
  }
  return true;
}

// Inverse of an upper-triangular R by rows from the bottom:
// X[i] = (e_i - sum_{k > i} R[i][k] X[k]) / R[i][i].
void upper_inverse(int64_t l, const double *R, double *Ri) {
  for (int64_t i = 0; i < l * l; ++i) Ri[i] = 0.0;
  for (int64_t i = l - 1; i >= 0; --i) {
    const double inv = 1.0 / R[i * l + i];
    double *xi = Ri + i * l;
    xi[i] = inv;
    for (int64_t k = i + 1; k < l; ++k) {
      const double f = -R[i * l + k] * inv;
      if (f == 0.0) continue;
      const double *xk = Ri + k * l;
      for (int64_t c = k; c < l; ++c) xi[c] += f * xk[c];
    }
  }
}

}  // namespace svdh
