// Randomized SVD (Halko-style range finder with power iterations) on one
// MI355X: src/randomized/mod.rs:24-237.
//
// Panels (Omega, Y, Z) are row-major in HBM with a leading dimension padded
// to 8 elements, so each SpMM row gather is a contiguous 16-byte vector per
// lane.
//
// QR normaliser (the reference's default) as implicit CholeskyQR: Q = Y M is
// never formed. M = R^-1 comes from the Cholesky factor of the FP64 Gram
// G = Y^T Y, accumulated from the FP32 panel (products of two FP32 values are
// exact in FP64), so Q is orthonormal to FP64 rounding without an FP32 pass
// over Y. M is folded into the small n-side panel instead: Z = A^T Q =
// (A^T Y) M. One power step is therefore two SpMMs, one Gram of the m x l
// panel and one small apply on the n x l panel; the reference's extra QR of
// Z does not change span(A Z) and is not needed. If the Cholesky factor does
// not exist (rank-deficient or too ill-conditioned panel) M is taken from the
// Gram's eigen-decomposition over its leading directions (qr_fallbacks).
// The thin Q of nalgebra's Householder QR and this Q span the same space;
// the final SVD absorbs the difference.
//
// After the last step Z = A^T Q = B^T. SVD(B) comes from the Gram
// Z^T Z = B B^T = W S^2 W^T (FP64): U = Q W = Y (M W), V = Z W S^-1.
//
// Adaptive mode (tol > 0): the next power step's A Z gives the Ritz
// residuals of the current subspace with one small product,
//   A v_i - s_i u_i = (A Z) w_i / s_i - s_i Q w_i = (A Z) w_i / s_i - Y (M w_i s_i),
// formed and normed in FP64 from the stored panels (A^T u_i = s_i v_i holds
// by construction). Checks are scheduled from the observed contraction rate.
//
// LU and None normalisers (src/randomized/mod.rs:102-105, 126-127, 239-293)
// keep the reference's explicit steps: column L2 normalisation or nothing
// between the products, then an explicit CholeskyQR2 of Y before B.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "common.h"

namespace svdh {

namespace {

// Workspace slots of the context used by the solver.
enum { WS_OMEGA = 0, WS_Y = 1, WS_Z = 2, WS_YN = 3, WS_U = 4, WS_VT = 5, WS_BF16 = 6, WS_ONES = 7, WS_SUMS = 8,
       WS_YP = 9, WS_YS = 10, WS_Z2 = 11 };

// Landing offset (doubles) of the Gram of the speculative next panel, clear
// of a Gram of width <= 128 at offset 0.
constexpr size_t kGramOff2 = 128 * 128;

struct Panel {
  void *ptr = nullptr;
  int64_t rows = 0, w = 0, ld = 0;
  int dt = SVDH_F32;
  // Row-major rows x width panel in workspace slot `slot`, cleared on the
  // stream (the padding columns past width stay zero through every kernel).
  void init(Ctx *c, int slot, int64_t r, int64_t width, int dtype) {
    rows = r;
    w = width;
    dt = dtype;
    ld = (width + 7) / 8 * 8;
    const size_t bytes = std::max<size_t>((size_t)r * ld * dtype_size(dtype), 64);
    ptr = c->workspace(slot, bytes);
    HIP_CHECK(hipMemsetAsync(ptr, 0, bytes, c->stream));
  }
  void *p() const { return ptr; }
};

// C (rows x cols) = A (rows x mid) * B[:, :cols] for row-major A and B (B has
// leading dimension ldb): row axpys, contiguous inner loops.
std::vector<double> mat_mul(const double *A, int64_t rows, int64_t mid, const double *B, int64_t ldb, int64_t cols) {
  std::vector<double> C(rows * cols, 0.0);
  for (int64_t i = 0; i < rows; ++i) {
    double *ci = C.data() + i * cols;
    for (int64_t j = 0; j < mid; ++j) {
      const double a = A[i * mid + j];
      if (a == 0.0) continue;
      const double *bj = B + j * ldb;
      for (int64_t q = 0; q < cols; ++q) ci[q] += a * bj[q];
    }
  }
  return C;
}

// fn(0) .. fn(n - 1) on up to 8 host threads (the caller's among them); the
// first exception thrown is rethrown here after every thread has finished.
template <typename F>
static void parallel_for(int64_t n, F fn) {
  const int64_t threads = std::min<int64_t>({n / 4, (int64_t)std::max(1u, std::thread::hardware_concurrency()), 8});
  if (threads <= 1) {
    for (int64_t b = 0; b < n; ++b) fn(b);
    return;
  }
  std::atomic<int64_t> next{0};
  std::exception_ptr err;
  std::mutex mu;
  auto work = [&] {
    for (int64_t b; (b = next.fetch_add(1)) < n;) {
      try {
        fn(b);
      } catch (...) {
        std::lock_guard<std::mutex> g(mu);
        if (!err) err = std::current_exception();
      }
    }
  };
  std::vector<std::thread> pool;
  for (int64_t t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

// Eigen-decomposition of a symmetric matrix sorted by descending eigenvalue.
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

// M (w x wo, row-major) with P M orthonormal, from the FP64 Gram G of the
// rows x w panel P: M = R^-1 of the Cholesky factor, or, when that does not
// exist, the eigen-decomposition over the leading directions (returns false).
bool normaliser_from_gram(int64_t w, int64_t rows, int dt, const std::vector<double> &G, std::vector<double> &M,
                          int64_t &wo) {
  std::vector<double> R(w * w);
  const int64_t target = std::min(rows, w);
  if (target == w && cholesky_upper(w, G.data(), R.data())) {
    M.assign(w * w, 0.0);
    upper_inverse(w, R.data(), M.data());
    wo = w;
    return true;
  }
  std::vector<double> lam, W;
  desc_eig(w, G, lam, W);
  const double thr = std::max(lam[0], 0.0) * (dt == SVDH_F64 ? 1e-24 : 1e-12);
  M.assign(w * target, 0.0);
  for (int64_t r = 0; r < target; ++r)
    if (lam[r] > thr && lam[r] > 0.0)
      for (int64_t i = 0; i < w; ++i) M[i * target + r] = W[i * w + r] / std::sqrt(lam[r]);
  wo = target;
  return false;
}

struct Rsvd {
  Ctx *c;
  svdh_op *op;
  int vdt;     // sparse values of the FP32/FP64 passes
  bool bf16;   // BF16 ladder available
  int64_t fallbacks = 0;
  // mean_center: column sums of A on the device (mu = sums * inv_m), or null.
  const double *sums = nullptr;
  double inv_m = 0.0;

  // Y = op_dir X; uses BF16 values and a BF16 copy of X when low is set.
  void mult(int dir, const Panel &X, Panel &Y, bool low) {
    Y.w = X.w;
    if (low) {
      void *bf_in = c->workspace(WS_BF16, (size_t)X.rows * X.ld * 2);
      panel_convert(c, X.rows, X.w, X.p(), X.ld, X.dt, bf_in, X.ld, SVDH_BF16);
      spmm(op, dir, SVDH_BF16, bf_in, X.ld, SVDH_BF16, Y.p(), Y.ld, Y.dt, X.w);
    } else {
      spmm(op, dir, vdt, X.p(), X.ld, X.dt, Y.p(), Y.ld, Y.dt, X.w);
    }
    // Centered products as exact rank-one corrections of the sparse ones
    // (the reference's multiply_*_centered, randomized/mod.rs:440-469):
    //   (A - 1 mu^T) X = A X - 1 (mu^T X),  (A - 1 mu^T)^T Y = A^T Y - mu (1^T Y).
    if (sums) {
      if (dir == 0)
        panel_rank1_correct(c, X.rows, X.w, X.p(), X.ld, X.dt, sums, Y.rows, Y.p(), Y.ld, Y.dt, nullptr, inv_m);
      else
        panel_rank1_correct(c, X.rows, X.w, X.p(), X.ld, X.dt, nullptr, Y.rows, Y.p(), Y.ld, Y.dt, sums, inv_m);
    }
  }

  // Waits for the Gram enqueued by gram_async (landing offset off, event ev)
  // and copies it out.
  void gram_wait(int64_t w, std::vector<double> &G, size_t off = 0, hipEvent_t ev = nullptr) {
    HIP_CHECK(hipEventSynchronize(ev ? ev : c->small_ev));
    G.assign(c->h_small + off, c->h_small + off + w * w);
  }

  // M (w x wo, row-major) with P M orthonormal, from the FP64 Gram of P.
  void normaliser(const Panel &P, std::vector<double> &M, int64_t &wo) {
    gram_async(c, P.rows, P.w, P.p(), P.ld, P.dt, 0);
    finish_normaliser(P, M, wo);
  }

  // Second half of normaliser: the Gram of P was enqueued by gram_async, and
  // the stream may have more work queued behind it.
  void finish_normaliser(const Panel &P, std::vector<double> &M, int64_t &wo, size_t off = 0,
                         hipEvent_t ev = nullptr) {
    std::vector<double> G;
    gram_wait(P.w, G, off, ev);
    if (!normaliser_from_gram(P.w, P.rows, P.dt, G, M, wo)) ++fallbacks;
  }

  // Explicit orthonormalisation (CholeskyQR2 with the eigen fallback).
  void orth(Panel &P) {
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> M;
      int64_t wo;
      normaliser(P, M, wo);
      panel_apply(c, P.rows, P.w, wo, P.p(), P.ld, P.dt, M.data(), P.p(), P.ld, P.dt);
      P.w = wo;
    }
  }
};

}  // namespace

// The host hand-over of a result: s (k values), U (nu values) and V^T (nv
// values, FP64 on the device) into pooled pinned arrays, in FP64 for FP64
// solves and rounded to FP32 on the device (half the bytes over PCIe) for
// FP32 and BF16 solves, as SvdRec<f32> holds them.
static void host_results(Ctx *c, int32_t precision, const double *sv, int64_t k, const double *ud, int64_t nu,
                         const double *vtd, int64_t nv, svdh_result *out) {
  if (precision == SVDH_PREC_F64) {
    out->s = result_alloc(k);
    out->u = result_alloc(nu);
    out->vt = result_alloc(nv);
    for (int64_t r = 0; r < k; ++r) out->s[r] = sv[r];
    if (nu > 0) HIP_CHECK(hipMemcpyAsync(out->u, ud, (size_t)nu * 8, hipMemcpyDeviceToHost, c->stream));
    if (nv > 0) HIP_CHECK(hipMemcpyAsync(out->vt, vtd, (size_t)nv * 8, hipMemcpyDeviceToHost, c->stream));
    return;
  }
  out->s32 = result_alloc_f32(k);
  out->u32 = result_alloc_f32(nu);
  out->vt32 = result_alloc_f32(nv);
  for (int64_t r = 0; r < k; ++r) out->s32[r] = (float)sv[r];
  if (nu + nv == 0) return;
  float *f = static_cast<float *>(c->workspace(13, (size_t)(nu + nv) * sizeof(float)));
  convert_f64_f32(c, nu, ud, f);
  convert_f64_f32(c, nv, vtd, f + nu);
  if (nu > 0) HIP_CHECK(hipMemcpyAsync(out->u32, f, (size_t)nu * 4, hipMemcpyDeviceToHost, c->stream));
  if (nv > 0) HIP_CHECK(hipMemcpyAsync(out->vt32, f + nu, (size_t)nv * 4, hipMemcpyDeviceToHost, c->stream));
}

int32_t rsvd_solve(svdh_ctx *ctx, svdh_op *op, int64_t target_rank, int64_t n_oversamples, int64_t n_power_iters,
                   int32_t normalizer, int32_t mean_center, uint64_t seed, int32_t has_seed, const double *omega,
                   int32_t precision, double tol, int64_t max_power_iters, svdh_result *out,
                   const RsvdDeviceOut *dev) {
  if (normalizer < 0 || normalizer > 2) fail(SVDH_ERR_ARG, "bad normalizer %d", normalizer);
  if (precision < 0 || precision > 2) fail(SVDH_ERR_ARG, "bad precision %d", precision);
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
  const int pdt = precision == SVDH_PREC_F64 ? SVDH_F64 : SVDH_F32;
  if (pdt == SVDH_F64 && l > 64) fail(SVDH_ERR_ARG, "FP64 panels are limited to rank + oversamples <= 64");

  Rsvd R;
  R.c = ctx;
  R.op = op;
  R.vdt = pdt;
  R.bf16 = precision == SVDH_PREC_BF16;
  const uint64_t s = has_seed ? seed : 0;
  if (mean_center) {
    // column means = A^T 1 / m (randomized/mod.rs:46-53, utils.rs:27)
    double *ones = static_cast<double *>(ctx->workspace(WS_ONES, (size_t)m * sizeof(double)));
    double *sums = static_cast<double *>(ctx->workspace(WS_SUMS, (size_t)n * sizeof(double)));
    panel_fill(ctx, m, 1, 1.0, ones, 1, SVDH_F64);
    spmv(op, 1, pdt, ones, sums, SVDH_F64);
    R.sums = sums;
    R.inv_m = 1.0 / (double)m;
  }

  Panel Y, Z;
  {
    Panel Om;
    Om.init(ctx, WS_OMEGA, n, l, pdt);
    if (omega) {
      // column-major n x l FP64 -> row-major panel
      std::vector<double> h((size_t)n * Om.ld, 0.0);
      for (int64_t j = 0; j < l; ++j)
        for (int64_t i = 0; i < n; ++i) h[i * Om.ld + j] = omega[i + j * n];
      DevBuf tmp(h.size() * sizeof(double));  // freed after the synchronisation below
      HIP_CHECK(hipMemcpyAsync(tmp.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      panel_convert(ctx, n, l, tmp.p, Om.ld, SVDH_F64, Om.p(), Om.ld, pdt);
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } else {
      panel_gaussian(ctx, n, l, s, Om.p(), Om.ld, pdt);
    }
    Y.init(ctx, WS_Y, m, l, pdt);
    Z.init(ctx, WS_Z, n, l, pdt);
    // BF16 ladder: the power steps use BF16 values and BF16 copies of the
    // input panels until the residual estimate (adaptive) drops below
    // 1e-2, or for the first half of a fixed count; the final B^T = A^T Q and
    // everything after are FP32.
    R.mult(0, Om, Y, R.bf16);
  }
  const double bf16_switch = 1e-2;
  bool low = R.bf16;
  int64_t iters_run = 0;
  double est = -1.0;
  std::vector<double> GZ, lam, W, Mq;  // Mq: Q = Y Mq
  int64_t wq = Y.w;

  if (normalizer == SVDH_NORM_QR) {
    // Pipelined: the Gram of Y is enqueued, then A^T Y (which does not depend
    // on the normaliser), and only then does the host wait for the Gram and
    // factor it while the SpMM runs.
    if (!adaptive && R.bf16) low = 0 < (n_power_iters + 1) / 2;
    gram_async(ctx, Y.rows, Y.w, Y.p(), Y.ld, Y.dt, 0);
    R.mult(1, Y, Z, low && !(!adaptive && 0 >= n_power_iters));
    R.finish_normaliser(Y, Mq, wq);
    Panel Yn;
    Yn.init(ctx, WS_YN, m, l, pdt);
    int64_t next_check = n_power_iters;
    int64_t last_check_it = -1;
    double last_est = 0.0;
    // Adaptive mode shifts each power step, Y <- (A A^T - c I) Q with c half
    // the smallest Ritz value of the block (a lower bound of sigma_l^2): the
    // unwanted spectrum [0, sigma_{l+1}^2] maps into [-c, c] and the top-k
    // directions contract against it by (s_k^2 - c) / c instead of
    // s_k^2 / s_{l+1}^2. c is refreshed every step from the eigenvalues of
    // the Gram of Z = A^T Q (Q^T A A^T Q, whose eigenvalues are the Ritz
    // values), enqueued before A Z and solved on the host while A Z runs.
    // Fixed-count mode keeps the plain power iteration.
    //
    // On top of the shift, a Chebyshev three-term recurrence (momentum):
    //   X_{t+1} = (A A^T - c I) Q_t - (c^2 / 4) Q_{t-1} R_t^-1,  Q_{t+1} R_{t+1} = X_{t+1},
    // which is the unnormalised W_{t+1} = (B - c) W_t - (c^2/4) W_{t-1} with
    // every step's QR factor divided out (c^2/2 on the first momentum step:
    // 2 (c/2)^2 T_2). Each unwanted component (|lambda - c| <= c) then decays
    // at c / 2 per step and the top-k directions contract by
    // 1 / (t + sqrt(t^2 - 1)), t = (s_k^2 - c) / c: the Chebyshev rate. In the
    // implicit form Q_{t-1} R_t^-1 = Y_{t-1} (M_{t-1} M_t). The recurrence
    // carries on through the small step-to-step changes of c (restarting
    // only when c moves by more than restart_tol relative, by default never)
    // and is dropped for good if a residual check ever goes up.
    double shift = 0.0, shift_used = 0.0;
    double shift_frac = 0.5;
    if (const char *e = std::getenv("SVDH_RSVD_SHIFT")) shift_frac = std::atof(e);
    bool momentum = true;
    if (const char *e = std::getenv("SVDH_RSVD_MOMENTUM")) momentum = std::atoi(e) != 0;
    double restart_tol = INFINITY;
    if (const char *e = std::getenv("SVDH_RSVD_RESTART")) restart_tol = std::atof(e);
    const bool eig_step = adaptive && shift_frac > 0.0;
    int64_t mom_run = 0;  // consecutive momentum steps
    Panel Yp;  // Y_{t-1}
    std::vector<double> Mp;  // M_{t-1}
    bool have_prev = false;
    if (adaptive && momentum) Yp.init(ctx, WS_YP, m, l, pdt);
    // At a residual check the next iterate is built before the check is
    // decided: the shift needs only the eigenvalues of the check's Gram, so
    // Y_{t+1} goes to Ys (A Z in Yn stays for the residual) and its Gram and
    // A^T Ys (into Z2) are enqueued; the device works through them while the
    // host finds the eigenvectors and the residual. The steps are the ones of
    // the unpipelined order: a converged check drops the speculative step,
    // and if the check turns the momentum or the BF16 ladder off, the step is
    // rebuilt from A Z without them.
    Panel Ys, Z2;
    if (adaptive) {
      Ys.init(ctx, WS_YS, m, l, pdt);
      Z2.init(ctx, WS_Z2, n, l, pdt);
    }
    std::vector<double> PM;  // M_{t-1} M_t when prepared during a check
    bool pm_ready = false;
    std::vector<double> Mq_spec;  // normaliser of the iterate built at a check
    int64_t wq_spec = 0;
    for (;;) {
      const bool final_fixed = !adaptive && iters_run >= n_power_iters;
      // FP32 panels of width <= 64 use the matrix-core panel products; the
      // step's small matrices (Mq, and M_{t-1} M_t for the momentum term)
      // travel to the device in one upload.
      const bool mfma = pdt == SVDH_F32 && Y.w <= 64 && wq <= 64 && (!have_prev || Yp.w <= 64);
      if (!(adaptive && momentum && have_prev))
        PM.clear();
      else if (!pm_ready)
        PM = mat_mul(Mp.data(), Yp.w, Y.w, Mq.data(), wq, wq);  // M_{t-1} M_t
      pm_ready = false;
      const double *dMq = nullptr, *dPM = nullptr;
      const bool check = adaptive && (iters_run >= next_check || iters_run >= max_power_iters);
      const bool want_gz = final_fixed || check || eig_step;
      bool gz_done = false;
      if (mfma) {
        std::vector<double> both(Mq);
        both.insert(both.end(), PM.begin(), PM.end());
        dMq = static_cast<const double *>(ctx->upload(both.data(), both.size() * sizeof(double)));
        dPM = dMq + Mq.size();
        const float *yin[1] = {static_cast<const float *>(Z.p())};
        const int64_t ldin[1] = {Z.ld}, liin[1] = {Y.w};
        const double *min_[1] = {dMq};
        const double al[1] = {1.0};
        const bool tri = upper_triangular(Mq.data(), Y.w, wq);
        if (want_gz) {  // the Gram of Z from the same kernel
          panel_mfma_gram(ctx, n, wq, 1, yin, ldin, liin, min_, al, static_cast<float *>(Z.p()), Z.ld, 0.0, nullptr,
                          0, 0, nullptr, tri);
          gz_done = true;
        } else {
          panel_mfma(ctx, n, wq, 1, yin, ldin, liin, min_, al, static_cast<float *>(Z.p()), Z.ld, 0.0, nullptr, 0,
                     tri);
        }
      } else {
        // Z holds A^T Y; Z = A^T Q = (A^T Y) Mq
        panel_apply(ctx, n, Y.w, wq, Z.p(), Z.ld, Z.dt, Mq.data(), Z.p(), Z.ld, Z.dt);
      }
      Z.w = wq;
      if (want_gz && !gz_done) gram_async(ctx, n, wq, Z.p(), Z.ld, Z.dt, 0);
      if (final_fixed) {
        R.gram_wait(wq, GZ);
        desc_eig(wq, GZ, lam, W);
        break;
      }
      R.mult(0, Z, Yn, low);  // A Z
      if (check || eig_step) {
        // The shift for the next iterate, from the Ritz values (if this is a
        // check that converges, the shift is unused).
        R.gram_wait(wq, GZ);
        shift = shift_frac * std::max(sym_eig_min(wq, GZ.data()), 0.0);
      }
      // Next iterate: Y_{t+1} = A Z - c Q - (c^2/4) Q_{t-1} R_t^-1
      //                      = A Z - c Y Mq - (c^2/4) Y_{t-1} (M_{t-1} M_t),
      // in place in Yn, or into Ys at a check.
      Panel *next = check ? &Ys : &Yn;
      // Returns whether the Gram of the iterate was accumulated on the way
      // (matrix-core panels: landing offset goff, event gev).
      auto build_next = [&](bool mom, double coef, size_t goff, hipEvent_t gev) -> bool {
        next->w = Yn.w;
        if (mfma) {
          const float *yin[2] = {static_cast<const float *>(Y.p()), static_cast<const float *>(Yp.p())};
          const int64_t ldin[2] = {Y.ld, Yp.ld}, liin[2] = {Y.w, Yp.w};
          const double *min_[2] = {dMq, dPM};
          const double al[2] = {-shift, -coef * shift * shift};
          const bool tri = upper_triangular(Mq.data(), Y.w, wq) && (!mom || upper_triangular(PM.data(), Yp.w, wq));
          panel_mfma_gram(ctx, m, wq, mom ? 2 : 1, yin, ldin, liin, min_, al, static_cast<float *>(next->p()),
                          next->ld, 1.0, static_cast<const float *>(Yn.p()), Yn.ld, goff, gev, tri);
          return true;
        } else {
          if (next != &Yn)
            HIP_CHECK(hipMemcpyAsync(next->p(), Yn.p(), (size_t)m * Yn.ld * dtype_size(pdt), hipMemcpyDeviceToDevice,
                                     ctx->stream));
          const double *Md = static_cast<const double *>(ctx->upload(Mq.data(), Mq.size() * sizeof(double)));
          panel_gemm_dev(ctx, m, Y.w, wq, Y.p(), Y.ld, Md, -shift, 1.0, next->p(), next->ld, pdt);
          if (mom) {
            const double *PMd = static_cast<const double *>(ctx->upload(PM.data(), PM.size() * sizeof(double)));
            panel_gemm_dev(ctx, m, Yp.w, wq, Yp.p(), Yp.ld, PMd, -coef * shift * shift, 1.0, next->p(), next->ld,
                           pdt);
          }
        }
        return false;
      };
      bool mom_used = false;
      bool next_gram = false;  // the Gram of the next iterate is already enqueued
      const bool shifted = adaptive && shift > 0.0;
      if (shifted) {
        const bool restart = std::fabs(shift - shift_used) > restart_tol * shift;
        mom_used = momentum && have_prev && !restart && !PM.empty();
        next_gram = check ? build_next(mom_used, mom_run == 0 ? 0.5 : 0.25, kGramOff2, ctx->small_ev2)
                          : build_next(mom_used, mom_run == 0 ? 0.5 : 0.25, 0, nullptr);
      } else if (check) {
        next = &Yn;  // unshifted: A Z itself, read-only from here on
      }
      if (check) {
        // the speculative next step: Gram of Y_{t+1} and A^T Y_{t+1}
        if (!next_gram) gram_async(ctx, m, next->w, next->p(), next->ld, next->dt, kGramOff2, ctx->small_ev2);
        R.mult(1, *next, Z2, low);
        desc_eig(wq, GZ, lam, W);
        const int64_t k = std::min(target_rank, std::min(wq, n));
        // r_i = (A Z) w_i / s_i - Y (Mq w_i s_i)
        std::vector<double> M1(wq * k, 0.0), nr2(k);
        std::vector<double> M2 = mat_mul(Mq.data(), Y.w, wq, W.data(), wq, k);  // Mq W, then * -s
        bool ok = true;
        for (int64_t r = 0; r < k; ++r) {
          const double sr = std::sqrt(std::max(lam[r], 0.0));
          if (!(sr > 0.0)) ok = false;
          for (int64_t i = 0; i < wq; ++i) M1[i * k + r] = sr > 0.0 ? W[i * wq + r] / sr : 0.0;
          for (int64_t i = 0; i < Y.w; ++i) M2[i * k + r] *= -sr;
        }
        panel_residual(ctx, m, Y.w, wq, k, Y.p(), Yn.p(), Y.ld, pdt, M2.data(), M1.data(), nullptr);
        // While the residual runs: the normaliser of the next iterate and its
        // momentum product (dropped if the check converges).
        std::vector<double> Mq_next;
        int64_t wq_next = 0;
        const int64_t fb = R.fallbacks;
        R.finish_normaliser(*next, Mq_next, wq_next, kGramOff2, ctx->small_ev2);
        std::vector<double> PM_next;
        if (momentum && shift > 0.0) PM_next = mat_mul(Mq.data(), Y.w, wq, Mq_next.data(), wq_next, wq_next);
        panel_residual_wait(ctx, k, nr2.data());
        est = 0.0;
        for (int64_t r = 0; r < k; ++r) est = std::max(est, std::sqrt(nr2[r]) / std::sqrt(std::max(lam[r], 0.0)));
        if (!ok || !std::isfinite(est)) est = INFINITY;
        if ((est <= tol && !low) || iters_run >= max_power_iters) {
          R.fallbacks = fb;
          break;
        }
        bool redo = false;
        if (R.bf16 && low && est <= bf16_switch) {
          low = false;
          redo = true;
        }
        // Next check: half the step count predicted from the contraction rate
        // since the last check (the rate improves as the shift settles).
        int64_t step = 1;
        if (last_check_it >= 0 && est < last_est && est > tol && !low) {
          const double rate = std::pow(est / last_est, 1.0 / (double)(iters_run - last_check_it));
          if (rate > 0.0 && rate < 1.0) {
            const double need = std::log(tol / est) / std::log(rate);
            step = std::max<int64_t>(1, std::min<int64_t>(32, (int64_t)std::floor(0.5 * need)));
          }
        }
        if (last_check_it >= 0 && est > last_est) {
          momentum = false;
          if (mom_used) redo = true;
        }
        last_check_it = iters_run;
        last_est = est;
        next_check = iters_run + step;
        if (redo) {
          // the iterate itself depends on the momentum flag only; its A^T
          // product also on the BF16 ladder
          bool rebuilt = false;
          if (shifted && mom_used && !momentum) {
            rebuilt = build_next(false, 0.0, kGramOff2, ctx->small_ev2);
            mom_used = false;
          }
          if (!rebuilt) gram_async(ctx, m, next->w, next->p(), next->ld, next->dt, kGramOff2, ctx->small_ev2);
          R.mult(1, *next, Z2, low);
          R.fallbacks = fb;
          R.finish_normaliser(*next, Mq_next, wq_next, kGramOff2, ctx->small_ev2);
          PM_next.clear();
        }
        // the prepared products become the next step's (PM only if the
        // rotation below keeps Y_t as the previous iterate)
        if (momentum && shift > 0.0 && !PM_next.empty()) {
          PM = std::move(PM_next);
          pm_ready = true;
        }
        Mq_spec = std::move(Mq_next);
        wq_spec = wq_next;
      }
      if (shifted) shift_used = shift;
      mom_run = mom_used ? mom_run + 1 : 0;
      // Adopt the next iterate: Y <- Y_{t+1}, keeping Y_t and M_t as the
      // previous iterate for the momentum term.
      if (adaptive && momentum && shift > 0.0) {
        Panel old_prev = Yp;
        Yp = Y;
        Mp = Mq;
        have_prev = true;
        Y = *next;
        *next = old_prev;
      } else {
        std::swap(Y, *next);
      }
      Y.w = Z.w;
      ++iters_run;
      if (check) {  // the speculative step's A^T Y is queued and its normaliser known
        std::swap(Z, Z2);
        Mq = std::move(Mq_spec);
        wq = wq_spec;
        continue;
      }
      if (!adaptive && R.bf16) low = iters_run < (n_power_iters + 1) / 2;
      const bool next_final = !adaptive && iters_run >= n_power_iters;
      if (!next_gram) gram_async(ctx, Y.rows, Y.w, Y.p(), Y.ld, Y.dt, 0);
      R.mult(1, Y, Z, low && !next_final);
      R.finish_normaliser(Y, Mq, wq);
    }
  } else {
    for (int64_t i = 0; i < n_power_iters; ++i) {
      if (R.bf16) low = i < (n_power_iters + 1) / 2;
      R.mult(1, Y, Z, low);
      if (normalizer == SVDH_NORM_LU) panel_normalize_columns(ctx, Z.rows, Z.w, Z.p(), Z.ld, Z.dt);
      R.mult(0, Z, Y, low);
      if (normalizer == SVDH_NORM_LU) panel_normalize_columns(ctx, Y.rows, Y.w, Y.p(), Y.ld, Y.dt);
      ++iters_run;
    }
    R.orth(Y);  // randomized/mod.rs:143-144
    Mq.assign(Y.w * Y.w, 0.0);
    for (int64_t i = 0; i < Y.w; ++i) Mq[i * Y.w + i] = 1.0;
    wq = Y.w;
    R.mult(1, Y, Z, false);
    GZ.assign(wq * wq, 0.0);
    gram(ctx, n, wq, Z.p(), Z.ld, Z.dt, GZ.data());
    desc_eig(wq, GZ, lam, W);
  }

  // B^T = Z (n x wq) = A^T Q; SVD(B) = W S X^T with X = Z W S^-1; U = Q W = Y (Mq W).
  const int64_t k = std::min(target_rank, std::min(wq, n));
  std::vector<double> sv(k), Ws(wq * k);
  const std::vector<double> MW = mat_mul(Mq.data(), Y.w, wq, W.data(), wq, k);
  for (int64_t r = 0; r < k; ++r) {
    sv[r] = std::sqrt(std::max(lam[r], 0.0));
    const double inv = sv[r] > 0.0 ? 1.0 / sv[r] : 0.0;
    for (int64_t i = 0; i < wq; ++i) Ws[i * k + r] = W[i * wq + r] * inv;
  }
  // U = Y (Mq W), row-major m x k; V^T = (Z W S^-1)^T, row-major k x n, written
  // by the transposed store of the apply kernel.
  double *ud = dev ? dev->u : static_cast<double *>(ctx->workspace(WS_U, std::max<size_t>(m * k, 1) * 8));
  double *vtd = dev ? dev->vt : static_cast<double *>(ctx->workspace(WS_VT, std::max<size_t>(k * n, 1) * 8));
  const int64_t ldu = dev ? dev->ldu : k, ldvt = dev ? dev->ldvt : n;
  panel_apply(ctx, m, Y.w, k, Y.p(), Y.ld, Y.dt, MW.data(), ud, ldu, SVDH_F64);
  panel_apply(ctx, n, wq, k, Z.p(), Z.ld, Z.dt, Ws.data(), vtd, 1, SVDH_F64, ldvt);

  out->d = k;
  out->u_rows = m;
  out->u_cols = k;
  out->vt_rows = k;
  out->vt_cols = n;
  if (dev) {
    for (int64_t r = 0; r < k; ++r) dev->s[r] = sv[r];
  } else {
    host_results(ctx, precision, sv.data(), k, ud, m * k, vtd, k * n, out);
  }
  HIP_CHECK(hipStreamSynchronize(ctx->stream));

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

// Batched randomized SVD over several operators (column masks of one matrix),
// QR normaliser, fixed power-iteration count. Every SpMM round (A_b X_b for
// all b, or A_b^T Y_b for all b) is one launch; the Grams of a round are
// enqueued together, then the next SpMM round, and only then does the host
// factor them. Each operator follows exactly the steps of rsvd_solve with
// tol = 0, so its result is the one a separate solve would return.
int32_t rsvd_batched_solve(svdh_ctx *ctx, int64_t nops, svdh_op *const *ops, int64_t target_rank,
                           int64_t n_oversamples, int64_t n_power_iters, int32_t normalizer, uint64_t seed,
                           int32_t has_seed, const double *const *omegas, int32_t precision, svdh_result *outs,
                           const BatchedDeviceOut *dev) {
  if (nops <= 0) return SVDH_OK;
  if (normalizer != SVDH_NORM_QR) fail(SVDH_ERR_ARG, "the batched solver supports the QR normaliser");
  if (precision != SVDH_PREC_F64 && precision != SVDH_PREC_F32) fail(SVDH_ERR_ARG, "bad precision %d", precision);
  if (target_rank < 0 || n_oversamples < 0 || n_power_iters < 0) fail(SVDH_ERR_ARG, "negative argument");
  const int pdt = precision == SVDH_PREC_F64 ? SVDH_F64 : SVDH_F32;
  const int64_t m = ops[0]->nrows;
  int64_t l = -1;
  for (int64_t b = 0; b < nops; ++b) {
    if (ops[b]->nrows != m) fail(SVDH_ERR_ARG, "batched operators must have the same number of rows");
    if (ops[b]->ncols == 0) fail(SVDH_ERR_SHAPE, "empty mask %lld", (long long)b);
    const int64_t lb = std::min(target_rank, std::min(m, ops[b]->ncols)) + n_oversamples;
    if (l >= 0 && lb != l) fail(SVDH_ERR_ARG, "batched operators need one panel width (rank + oversample)");
    l = lb;
  }
  if (l <= 0 || l > 64) fail(SVDH_ERR_ARG, "batched panel width %lld outside 1..64", (long long)l);
  const int64_t ld = (l + 7) / 8 * 8;
  const size_t es = dtype_size(pdt);
  const uint64_t sd = has_seed ? seed : 0;
  Ctx *c = ctx;
  // panels: Y_b, Yn_b (m x ld), Z_b and Omega_b (n_b x ld), in the context's batch workspace
  std::vector<void *> Y(nops), Yn(nops), Z(nops), Om(nops);
  c->batch_ws.resize(std::max<size_t>(c->batch_ws.size(), (size_t)nops * 4));
  c->batch_pad.resize(c->batch_ws.size());
  std::vector<int64_t> g_rows;  // operators whose Omega is drawn on the device
  std::vector<void *> g_ptr;
  for (int64_t b = 0; b < nops; ++b) {
    const int64_t nb = ops[b]->ncols;
    void *p[4];
    const int64_t rows[4] = {m, m, nb, nb};
    for (int q = 0; q < 4; ++q) {
      const size_t i = (size_t)b * 4 + q;
      DevBuf &d = c->batch_ws[i];
      Ctx::PadState &ps = c->batch_pad[i];
      const size_t bytes = (size_t)rows[q] * ld * es;
      if (d.bytes < bytes) {
        c->sync();
        d.alloc(bytes);
        ps = Ctx::PadState{};
      }
      p[q] = d.p;
      // the padding columns [l, ld) must read as zero; the panel columns
      // below l are written before they are read
      if (!(ps.l >= 0 && ps.ld == ld && ps.l <= l && ps.rows >= rows[q])) {
        HIP_CHECK(hipMemsetAsync(d.p, 0, d.bytes, c->stream));
        ps.rows = (int64_t)(d.bytes / ((size_t)ld * es));
      }
      ps.l = l;
      ps.ld = ld;
    }
    Y[b] = p[0];
    Yn[b] = p[1];
    Z[b] = p[2];
    Om[b] = p[3];
    if (omegas && omegas[b]) {
      // injected n_b x l column-major FP64 Omega, as rsvd_solve takes it
      std::vector<double> h((size_t)nb * ld, 0.0);
      for (int64_t j = 0; j < l; ++j)
        for (int64_t i = 0; i < nb; ++i) h[i * ld + j] = omegas[b][i + j * nb];
      DevBuf tmp(h.size() * sizeof(double));  // freed after the synchronisation below
      HIP_CHECK(hipMemcpyAsync(tmp.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
      panel_convert(c, nb, l, tmp.p, ld, SVDH_F64, Om[b], ld, pdt);
      HIP_CHECK(hipStreamSynchronize(c->stream));
    } else {
      g_rows.push_back(nb);
      g_ptr.push_back(Om[b]);
    }
  }
  panel_gaussian_batched(c, (int64_t)g_rows.size(), g_rows.data(), l, sd, g_ptr.data(), ld, pdt);
  std::vector<svdh_op *> opv(ops, ops + nops);
  spmm_batched(c, (int)nops, opv.data(), 0, pdt, Om.data(), ld, pdt, Y.data(), ld, pdt, l);  // Y = A Omega

  std::vector<std::vector<double>> Mq(nops);
  std::vector<int64_t> wq(nops, l);
  std::vector<int64_t> fallbacks(nops, 0);
  const int64_t per_round = std::max<int64_t>(1, (int64_t)(Ctx::kSmall / (size_t)(l * l)));
  // Grams of panels P_b (rows x l) for all b; `between` is enqueued after the
  // first group of Grams and before the host waits for them.
  auto grams = [&](const std::vector<void *> &P, const std::vector<int64_t> &rows, auto between,
                   std::vector<std::vector<double>> &G) {
    G.assign(nops, {});
    bool ran = false;
    for (int64_t b0 = 0; b0 < nops; b0 += per_round) {
      const int64_t b1 = std::min(nops, b0 + per_round);
      if (l <= 64) {  // one launch for the group
        std::vector<const void *> pb(P.begin() + b0, P.begin() + b1);
        gram_batched_async(c, b1 - b0, rows.data() + b0, l, pb.data(), ld, pdt, 0);
      } else {
        for (int64_t b = b0; b < b1; ++b) gram_async(c, rows[b], l, P[b], ld, pdt, (size_t)(b - b0) * l * l);
      }
      if (!ran) {
        between();
        ran = true;
      }
      HIP_CHECK(hipEventSynchronize(c->small_ev));
      for (int64_t b = b0; b < b1; ++b)
        G[b].assign(c->h_small + (b - b0) * l * l, c->h_small + (b - b0 + 1) * l * l);
    }
  };
  std::vector<int64_t> mrows(nops, m), nrows(nops);
  for (int64_t b = 0; b < nops; ++b) nrows[b] = ops[b]->ncols;
  std::vector<std::vector<double>> G, GZ;
  auto normalise_all = [&](const std::vector<std::vector<double>> &Gs) {
    for (int64_t b = 0; b < nops; ++b)
      if (!normaliser_from_gram(l, m, pdt, Gs[b], Mq[b], wq[b])) ++fallbacks[b];
  };
  // Gram(Y), then Z = A^T Y behind it, then the normalisers
  grams(Y, mrows, [&] { spmm_batched(c, (int)nops, opv.data(), 1, pdt, Y.data(), ld, pdt, Z.data(), ld, pdt, l); },
        G);
  normalise_all(G);
  int64_t iters_run = 0;
  bool gz_done = false;  // GZ[b] (wq[b] x wq[b]) came with the last Z = (A^T Y) Mq
  for (;;) {
    const bool final_fixed = iters_run >= n_power_iters;
    bool all_mfma = pdt == SVDH_F32 && l <= 64;
    for (int64_t b = 0; b < nops; ++b) all_mfma = all_mfma && wq[b] <= 64;
    if (all_mfma) {  // Z = (A^T Y) Mq for every mask, the same kernels as rsvd_solve, one launch per group
      std::vector<float *> zf(nops);
      for (int64_t b = 0; b < nops; ++b) zf[b] = static_cast<float *>(Z[b]);
      if (final_fixed) {  // with the Gram of Z from the same kernel, as rsvd_solve's last step
        panel_mfma_gram_batched(c, nops, nrows.data(), l, wq.data(), zf.data(), ld, Mq.data(), GZ);
        gz_done = true;
      } else {
        panel_mfma_batched(c, nops, nrows.data(), l, wq.data(), zf.data(), ld, Mq.data());
      }
    }
    for (int64_t b = 0; b < nops && !all_mfma; ++b) {  // Z = (A^T Y) Mq, the same kernels as rsvd_solve
      if (pdt == SVDH_F32 && l <= 64 && wq[b] <= 64) {
        const double *dM = static_cast<const double *>(c->upload(Mq[b].data(), Mq[b].size() * sizeof(double)));
        const float *yin[1] = {static_cast<const float *>(Z[b])};
        const int64_t ldin[1] = {ld}, liin[1] = {l};
        const double *min_[1] = {dM};
        const double al[1] = {1.0};
        panel_mfma(c, nrows[b], wq[b], 1, yin, ldin, liin, min_, al, static_cast<float *>(Z[b]), ld, 0.0, nullptr,
                   0, upper_triangular(Mq[b].data(), l, wq[b]));
      } else {
        panel_apply(c, nrows[b], l, wq[b], Z[b], ld, pdt, Mq[b].data(), Z[b], ld, pdt);
      }
    }
    if (final_fixed) break;
    for (int64_t b = 0; b < nops; ++b)
      if (wq[b] != l) fail(SVDH_ERR_ARG, "rank-deficient panel in the batched solver (operator %lld)", (long long)b);
    spmm_batched(c, (int)nops, opv.data(), 0, pdt, Z.data(), ld, pdt, Yn.data(), ld, pdt, l);  // A Z
    std::swap(Y, Yn);
    ++iters_run;
    grams(Y, mrows,
          [&] { spmm_batched(c, (int)nops, opv.data(), 1, pdt, Y.data(), ld, pdt, Z.data(), ld, pdt, l); }, G);
    normalise_all(G);
  }
  // B^T = Z_b; SVD(B) from its Gram, as in rsvd_solve
  if (!gz_done) grams(Z, nrows, [] {}, GZ);
  // every operator's small factors first (independent, on host threads; the
  // GPU waits for them), then the outputs; device outputs take U and V^T of
  // all operators in one matrix-core launch each (the per-operator arithmetic)
  std::vector<std::vector<double>> MWs(nops), Wss(nops), svs(nops);
  std::vector<int64_t> ks(nops);
  parallel_for(nops, [&](int64_t b) {
    const int64_t n = ops[b]->ncols, w = wq[b];
    std::vector<double> Gb(w * w), lam, W;
    const int64_t gld = gz_done ? w : l;
    for (int64_t i = 0; i < w; ++i)
      for (int64_t j = 0; j < w; ++j) Gb[i * w + j] = GZ[b][i * gld + j];
    desc_eig(w, Gb, lam, W);
    const int64_t k = std::min(target_rank, std::min(w, n));
    std::vector<double> sv(k), Ws(w * k);
    MWs[b] = mat_mul(Mq[b].data(), l, w, W.data(), w, k);
    for (int64_t r = 0; r < k; ++r) {
      sv[r] = std::sqrt(std::max(lam[r], 0.0));
      const double inv = sv[r] > 0.0 ? 1.0 / sv[r] : 0.0;
      for (int64_t i = 0; i < w; ++i) Ws[i * k + r] = W[i * w + r] * inv;
    }
    svs[b] = std::move(sv);
    Wss[b] = std::move(Ws);
    ks[b] = k;
  });
  for (int64_t b = 0; b < nops; ++b) {
    svdh_op *op = ops[b];
    const int64_t n = op->ncols, w = wq[b], k = ks[b];
    const std::vector<double> &sv = svs[b], &MW = MWs[b], &Ws = Wss[b];
    svdh_result *out = &outs[b];
    out->d = k;
    out->u_rows = m;
    out->u_cols = k;
    out->vt_rows = k;
    out->vt_cols = n;
    if (dev) {  // U and V^T straight into the caller's device buffers; s on the host in FP64
      out->s = result_alloc(k);
      if (!out->s) fail(SVDH_ERR_ALLOC, "result allocation failed");
      for (int64_t r = 0; r < k; ++r) out->s[r] = sv[r];
      if (dev->ldu[b] < k || dev->ldvt[b] < n) fail(SVDH_ERR_ARG, "leading dimension too small for operator %lld",
                                                    (long long)b);
    } else {
      double *ud = static_cast<double *>(ctx->workspace(WS_U, std::max<size_t>(m * k, 1) * 8));
      double *vtd = static_cast<double *>(ctx->workspace(WS_VT, std::max<size_t>(k * n, 1) * 8));
      panel_apply(c, m, l, k, Y[b], ld, pdt, MW.data(), ud, k, SVDH_F64);
      panel_apply(c, n, w, k, Z[b], ld, pdt, Ws.data(), vtd, 1, SVDH_F64, n);
      host_results(c, precision, sv.data(), k, ud, m * k, vtd, k * n, out);
      HIP_CHECK(hipStreamSynchronize(c->stream));  // U/V workspaces are reused by the next operator
    }
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
    dg.random_seed = (uint32_t)sd;
    out->power_iters_run = iters_run;
    out->residual_estimate = -1.0;
    out->qr_fallbacks = fallbacks[b];
  }
  if (dev && nops > 0) {
    const int64_t k0 = ks[0];
    bool same_k = k0 <= 128 && (pdt == SVDH_F32 || pdt == SVDH_F64);
    for (int64_t b = 1; b < nops; ++b) same_k = same_k && ks[b] == k0;
    if (same_k) {
      std::vector<int64_t> rows(nops, m), lis(nops, l), zcs1(nops, 1), lz1(nops, 1), vcs(nops), cols(nops);
      std::vector<const void *> ys(nops), zs(nops);
      std::vector<double *> us(nops), vts(nops);
      for (int64_t b = 0; b < nops; ++b) {
        ys[b] = Y[b];
        zs[b] = Z[b];
        us[b] = dev->u[b];
        vts[b] = dev->vt[b];
        vcs[b] = dev->ldvt[b];
        cols[b] = ops[b]->ncols;
      }
      panel_apply_batched(c, nops, rows.data(), lis.data(), k0, ys.data(), ld, pdt, MWs.data(), us.data(),
                          dev->ldu, zcs1.data());
      panel_apply_batched(c, nops, cols.data(), wq.data(), k0, zs.data(), ld, pdt, Wss.data(), vts.data(),
                          lz1.data(), vcs.data());
    } else {
      for (int64_t b = 0; b < nops; ++b) {
        panel_apply(c, m, l, ks[b], Y[b], ld, pdt, MWs[b].data(), dev->u[b], dev->ldu[b], SVDH_F64);
        panel_apply(c, ops[b]->ncols, wq[b], ks[b], Z[b], ld, pdt, Wss[b].data(), dev->vt[b], 1, SVDH_F64,
                    dev->ldvt[b]);
      }
    }
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return SVDH_OK;
}

}  // namespace svdh
