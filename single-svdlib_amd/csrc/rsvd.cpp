// Randomized SVD (Halko-style range finder with power iterations) on one
// MI355X: src/randomized/mod.rs:24-237.
//
// Panels (Omega, Y, Z, Q) are row-major in HBM with a leading dimension
// padded to 8 elements, so each SpMM row gather is a contiguous 16-byte
// vector per lane. Per step:
//   Y = A Omega, then n_power_iters times  Z = A^T Y, normalise Z,
//   Y = A Z, normalise Y;  Q = qr(Y);  B^T = A^T Q;  SVD(B) from the Gram
//   B B^T = (B^T)^T B^T (l x l, FP64) -> B = W S X^T with X = B^T W S^-1;
//   U = Q W.
// The QR normaliser is CholeskyQR2: FP64 Gram of the panel, host Cholesky,
// Y R^-1 applied on the device, twice. If a Cholesky pivot is not positive
// (rank deficient or too ill-conditioned panel) the step falls back to the
// eigen-decomposition of the Gram (Y W L^-1/2), counted in qr_fallbacks.
// The thin Q of nalgebra's Householder QR and this Q span the same space and
// differ only in column signs/rotation, which the final SVD absorbs.
//
// Adaptive mode (tol > 0): Z = A^T Q is B^T for the current Q, so its Gram
// gives the Ritz values s_i and vectors w_i of the current subspace, and the
// next power step's A Z gives every Ritz residual with one small product:
//   A v_i - s_i u_i = (A Z) w_i / s_i - s_i Q w_i,
// with u_i = Q w_i, v_i = Z w_i / s_i (and A^T u_i = s_i v_i by construction).
// The residual panel is formed and normed in FP64 from the stored panels, so
// it resolves residuals down to the panel precision, unlike the cancelling
// Gram identity |r_i|^2 = w_i^T (A Z)^T (A Z) w_i / s_i^2 - s_i^2.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.h"

namespace svdh {

namespace {

struct Panel {
  DevBuf buf;
  int64_t rows = 0, w = 0, ld = 0;
  int dt = SVDH_F32;
  void init(int64_t r, int64_t width, int dtype) {
    rows = r;
    w = width;
    dt = dtype;
    ld = (width + 7) / 8 * 8;
    buf.alloc(std::max<size_t>((size_t)r * ld * dtype_size(dtype), 64));
  }
  void *p() const { return buf.p; }
};

struct Rsvd {
  Ctx *c;
  svdh_op *op;
  int vdt;     // sparse values of the FP32/FP64 passes
  int pdt;     // panel dtype
  bool bf16;   // BF16 ladder available
  int64_t fallbacks = 0;
  DevBuf bf_in;  // BF16 copy of an input panel

  // Y = op_dir X; uses BF16 values and a BF16 copy of X when low is set.
  void mult(int dir, const Panel &X, Panel &Y, bool low) {
    Y.w = X.w;
    if (low) {
      bf_in.ensure((size_t)X.rows * X.ld * 2);
      panel_convert(c, X.rows, X.w, X.p(), X.ld, X.dt, bf_in.p, X.ld, SVDH_BF16);
      spmm(op, dir, SVDH_BF16, bf_in.p, X.ld, SVDH_BF16, Y.p(), Y.ld, Y.dt, X.w);
    } else {
      spmm(op, dir, vdt, X.p(), X.ld, X.dt, Y.p(), Y.ld, Y.dt, X.w);
    }
  }

  // One Cholesky-QR pass from a given Gram; false if the Cholesky fails.
  bool chol_pass(Panel &P, const std::vector<double> &G) {
    const int64_t w = P.w;
    std::vector<double> R(w * w), Ri(w * w);
    if (!cholesky_upper(w, G.data(), R.data())) return false;
    upper_inverse(w, R.data(), Ri.data());
    panel_apply(c, P.rows, w, w, P.p(), P.ld, P.dt, Ri.data(), P.p(), P.ld, P.dt);
    return true;
  }

  // Orthonormalise the columns of P (CholeskyQR2). G0 (optional) is P's Gram.
  void orth(Panel &P, const std::vector<double> *G0 = nullptr) {
    const int64_t w = P.w;
    std::vector<double> G(w * w);
    if (G0)
      G = *G0;
    else
      gram(c, P.rows, w, P.p(), P.ld, P.dt, G.data());
    const int64_t target = std::min(P.rows, w);
    if (target == w && chol_pass(P, G)) {
      gram(c, P.rows, w, P.p(), P.ld, P.dt, G.data());
      if (chol_pass(P, G)) return;
    }
    // Eigen fallback: keep the leading `target` directions.
    ++fallbacks;
    std::vector<double> lam(w), V(w * w);
    sym_eig(w, G.data(), lam.data(), V.data());
    const double lmax = std::max(lam[w - 1], 0.0);
    const double thr = lmax * (P.dt == SVDH_F64 ? 1e-24 : 1e-12);
    std::vector<double> M(w * target, 0.0);
    for (int64_t r = 0; r < target; ++r) {
      const int64_t k = w - 1 - r;
      if (lam[k] > thr && lam[k] > 0.0)
        for (int64_t i = 0; i < w; ++i) M[i * target + r] = V[i * w + k] / std::sqrt(lam[k]);
    }
    panel_apply(c, P.rows, w, target, P.p(), P.ld, P.dt, M.data(), P.p(), P.ld, P.dt);
    P.w = target;
    gram(c, P.rows, P.w, P.p(), P.ld, P.dt, G.data());
    G.resize(P.w * P.w);
    chol_pass(P, G);
  }

  void normalize(Panel &P, int normalizer) {
    if (normalizer == SVDH_NORM_QR)
      orth(P);
    else if (normalizer == SVDH_NORM_LU)
      panel_normalize_columns(c, P.rows, P.w, P.p(), P.ld, P.dt);
  }
};

// Eigen-decomposition of a Gram sorted by descending eigenvalue.
void desc_eig(int64_t w, const std::vector<double> &G, std::vector<double> &lam, std::vector<double> &W) {
  std::vector<double> la(w), V(w * w);
  sym_eig(w, G.data(), la.data(), V.data());
  lam.assign(w, 0.0);
  W.assign(w * w, 0.0);
  for (int64_t r = 0; r < w; ++r) {
    const int64_t k = w - 1 - r;
    lam[r] = la[k];
    for (int64_t i = 0; i < w; ++i) W[i * w + r] = V[i * w + k];
  }
}

}  // namespace

int32_t rsvd_solve(svdh_ctx *ctx, svdh_op *op, int64_t target_rank, int64_t n_oversamples, int64_t n_power_iters,
                   int32_t normalizer, int32_t mean_center, uint64_t seed, int32_t has_seed, const double *omega,
                   int32_t precision, double tol, int64_t max_power_iters, svdh_result *out) {
  if (normalizer < 0 || normalizer > 2) fail(SVDH_ERR_ARG, "bad normalizer %d", normalizer);
  if (precision < 0 || precision > 2) fail(SVDH_ERR_ARG, "bad precision %d", precision);
  if (mean_center) fail(SVDH_ERR_ARG, "mean_center is not supported by this build");
  if (target_rank < 0 || n_oversamples < 0 || n_power_iters < 0) fail(SVDH_ERR_ARG, "negative argument");
  const int64_t m = op->nrows, n = op->ncols;
  if (m == 0 || n == 0) fail(SVDH_ERR_SHAPE, "empty matrix");
  const int64_t rank = std::min(target_rank, std::min(m, n));
  const int64_t l = rank + n_oversamples;
  if (l <= 0) fail(SVDH_ERR_ARG, "rank + oversamples must be positive");
  if (l > 128) fail(SVDH_ERR_ARG, "rank + oversamples = %lld exceeds 128", (long long)l);
  const bool adaptive = tol > 0.0;
  if (adaptive && normalizer != SVDH_NORM_QR) fail(SVDH_ERR_ARG, "adaptive power iterations need the QR normaliser");
  if (max_power_iters < n_power_iters) max_power_iters = n_power_iters;

  Rsvd R;
  R.c = ctx;
  R.op = op;
  R.vdt = precision == SVDH_PREC_F64 ? SVDH_F64 : SVDH_F32;
  R.pdt = precision == SVDH_PREC_F64 ? SVDH_F64 : SVDH_F32;
  R.bf16 = precision == SVDH_PREC_BF16;
  const int pdt = R.pdt;
  const uint64_t s = has_seed ? seed : 0;

  Panel Om, Y, Z;
  Om.init(n, l, pdt);
  Y.init(m, l, pdt);
  Z.init(n, l, pdt);
  HIP_CHECK(hipMemsetAsync(Om.p(), 0, Om.buf.bytes, ctx->stream));
  HIP_CHECK(hipMemsetAsync(Y.p(), 0, Y.buf.bytes, ctx->stream));
  HIP_CHECK(hipMemsetAsync(Z.p(), 0, Z.buf.bytes, ctx->stream));
  if (omega) {
    // column-major n x l FP64 -> row-major panel
    std::vector<double> h((size_t)n * Om.ld, 0.0);
    for (int64_t j = 0; j < l; ++j)
      for (int64_t i = 0; i < n; ++i) h[i * Om.ld + j] = omega[i + j * n];
    DevBuf tmp(h.size() * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(tmp.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    panel_convert(ctx, n, l, tmp.p, Om.ld, SVDH_F64, Om.p(), Om.ld, pdt);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  } else {
    panel_gaussian(ctx, n, l, s, Om.p(), Om.ld, pdt);
  }

  // BF16 ladder: early passes run on BF16 values/panels until the residual
  // estimate (adaptive) falls below this, or for the first half of a fixed count.
  const double bf16_switch = 1e-2;
  bool low = R.bf16;
  R.mult(0, Om, Y, low);
  int64_t iters_run = 0;
  double est = -1.0;
  std::vector<double> GZ, lam, W;

  if (!adaptive) {
    for (int64_t i = 0; i < n_power_iters; ++i) {
      if (R.bf16) low = i < (n_power_iters + 1) / 2;
      R.mult(1, Y, Z, low);
      R.normalize(Z, normalizer);
      R.mult(0, Z, Y, low);
      R.normalize(Y, normalizer);
      ++iters_run;
    }
    R.orth(Y);  // randomized/mod.rs:143-144
  } else {
    R.orth(Y);
    Panel Yn;
    Yn.init(m, l, pdt);
    for (;;) {
      R.mult(1, Y, Z, low);  // Z = A^T Q = B^T for the current Q
      const int64_t w = Z.w;
      GZ.assign(w * w, 0.0);
      gram(ctx, n, w, Z.p(), Z.ld, Z.dt, GZ.data());
      desc_eig(w, GZ, lam, W);
      R.mult(0, Z, Yn, low);  // A Z = A v_i s_i for the Ritz vectors of Q
      if (iters_run >= n_power_iters) {
        // r_i = A v_i - s_i u_i = (A Z) w_i / s_i - Q w_i s_i, column norms in FP64.
        const int64_t k = std::min(target_rank, std::min(w, n));
        std::vector<double> M1(w * k, 0.0), M2(w * k, 0.0), nr2(k);
        bool ok = true;
        for (int64_t r = 0; r < k; ++r) {
          const double sr = std::sqrt(std::max(lam[r], 0.0));
          if (!(sr > 0.0)) ok = false;
          for (int64_t i = 0; i < w; ++i) {
            M1[i * k + r] = sr > 0.0 ? W[i * w + r] / sr : 0.0;
            M2[i * k + r] = -W[i * w + r] * sr;
          }
        }
        panel_residual(ctx, m, w, k, Yn.p(), Y.p(), Y.ld, pdt, M1.data(), M2.data(), nr2.data());
        est = 0.0;
        for (int64_t r = 0; r < k; ++r) est = std::max(est, std::sqrt(nr2[r]) / std::sqrt(std::max(lam[r], 0.0)));
        if (!ok) est = INFINITY;
        if (est <= tol || iters_run >= max_power_iters) break;
        if (R.bf16 && low && est <= bf16_switch) low = false;
      }
      // Next subspace: Q = orth(A Z).
      std::swap(Y.buf, Yn.buf);
      std::swap(Y.ld, Yn.ld);
      Y.w = w;
      Y.rows = m;
      R.orth(Y);
      ++iters_run;
    }
  }

  // B^T = A^T Q in full precision, SVD via its Gram (skipped when the
  // adaptive loop already holds it for the current Q).
  if (!adaptive || low) {
    R.mult(1, Y, Z, false);
    const int64_t w = Z.w;
    GZ.assign(w * w, 0.0);
    gram(ctx, n, w, Z.p(), Z.ld, Z.dt, GZ.data());
    desc_eig(w, GZ, lam, W);
  }
  const int64_t w = Z.w;
  const int64_t k = std::min(target_rank, std::min(w, n));
  std::vector<double> sv(k), Wk(w * k), Ws(w * k);
  for (int64_t r = 0; r < k; ++r) {
    sv[r] = std::sqrt(std::max(lam[r], 0.0));
    const double inv = sv[r] > 0.0 ? 1.0 / sv[r] : 0.0;
    for (int64_t i = 0; i < w; ++i) {
      Wk[i * k + r] = W[i * w + r];
      Ws[i * k + r] = W[i * w + r] * inv;
    }
  }
  Panel U, V;
  U.init(m, k, SVDH_F64);
  V.init(n, k, SVDH_F64);
  panel_apply(ctx, m, w, k, Y.p(), Y.ld, Y.dt, Wk.data(), U.p(), U.ld, SVDH_F64);
  panel_apply(ctx, n, w, k, Z.p(), Z.ld, Z.dt, Ws.data(), V.p(), V.ld, SVDH_F64);
  std::vector<double> hu((size_t)m * U.ld), hv((size_t)n * V.ld);
  HIP_CHECK(hipMemcpyAsync(hu.data(), U.p(), hu.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipMemcpyAsync(hv.data(), V.p(), hv.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));

  out->d = k;
  out->u_rows = m;
  out->u_cols = k;
  out->vt_rows = k;
  out->vt_cols = n;
  out->s = static_cast<double *>(std::calloc(std::max<int64_t>(k, 1), sizeof(double)));
  out->u = static_cast<double *>(std::calloc(std::max<int64_t>(m * k, 1), sizeof(double)));
  out->vt = static_cast<double *>(std::calloc(std::max<int64_t>(k * n, 1), sizeof(double)));
  if (!out->s || !out->u || !out->vt) fail(SVDH_ERR_ALLOC, "result allocation failed");
  for (int64_t r = 0; r < k; ++r) out->s[r] = sv[r];
  for (int64_t i = 0; i < m; ++i)
    for (int64_t r = 0; r < k; ++r) out->u[i * k + r] = hu[i * U.ld + r];
  for (int64_t r = 0; r < k; ++r)
    for (int64_t i = 0; i < n; ++i) out->vt[r * n + i] = hv[i * V.ld + r];

  svdh_diagnostics &dg = out->diagnostics;  // randomized/mod.rs:214-237
  dg.non_zero = op->nnz;
  dg.dimensions = target_rank;
  dg.iterations = n_power_iters;
  dg.transposed = 0;
  dg.lanczos_steps = 0;
  dg.ritz_values_stabilized = k;
  dg.significant_values = k;
  dg.singular_values = k;
  dg.end_interval[0] = -1e-30;
  dg.end_interval[1] = 1e-30;
  dg.kappa = 1e-6;
  dg.random_seed = (uint32_t)s;
  out->power_iters_run = iters_run;
  out->residual_estimate = est;
  out->qr_fallbacks = R.fallbacks;
  return SVDH_OK;
}

}  // namespace svdh
