// Adjoint of the truncated SVD with respect to the CSR values (backward()).
//
// Given A = sum_j s_j u_j v_j^T and the top-k triplets of a solve:
//
//  * L = sum_{i<k} s_i:  dL/dA = U_k V_k^T, so on the pattern
//    grad[e] = sum_i u_i[r_e] v_i[c_e]  (one SDDMM).
//
//  * L = |U_k^T y|^2:  L depends only on span(U_k), so the rotations inside
//    the top-k block cancel and only the complement term remains. With
//    c_i = u_i^T y, P = I - U_k U_k^T and w_i the solution of
//        (s_i^2 I - A A^T) w_i = P y   on range(P)
//    (positive definite there when s_k > s_{k+1}),
//        dL/dA = sum_i 2 c_i (s_i w_i v_i^T + u_i (A^T w_i)^T),
//    i.e. an SDDMM over two factor pairs (W diag(2 c s), V) and
//    (U, A^T W diag(2 c)). The k shifted systems share the right-hand side
//    and are solved together by conjugate gradients, one A (A^T P) product
//    per iteration, with the iterates re-projected onto range(P).
//
// Factors live in HBM as row-major panels (leading dimension k rounded up to
// 8); the CG scalars are per column and travel to the host once per
// half-step. BF16 precision stores the SDDMM's factor panels in BF16
// (FP32 accumulation); everything before it runs in FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "common.h"

namespace svdh {

namespace {

// Row-major padded panel of `cols` columns from a host FP64 source.
// get(i, j) returns element (row i, column j).
template <typename F>
void upload_panel(Ctx *c, DevBuf &buf, int64_t rows, int64_t cols, int64_t ld, int dt, F get) {
  const size_t es = dtype_size(dt);
  buf.alloc(std::max<size_t>((size_t)rows * ld * es, 64));
  std::vector<char> h((size_t)rows * ld * es, 0);
  for (int64_t i = 0; i < rows; ++i)
    for (int64_t j = 0; j < cols; ++j) {
      const double x = get(i, j);
      if (dt == SVDH_F64)
        reinterpret_cast<double *>(h.data())[i * ld + j] = x;
      else
        reinterpret_cast<float *>(h.data())[i * ld + j] = (float)x;
    }
  HIP_CHECK(hipMemcpyAsync(buf.p, h.data(), h.size(), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

// memcpy by up to 8 host threads (1 MB or more each)
void parallel_copy(void *dst, const void *src, size_t bytes) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::max<size_t>(1, std::min<size_t>(std::min<size_t>(8, hw), bytes >> 20));
  const size_t per = (bytes + nt - 1) / nt;
  auto part = [&](size_t t) {
    const size_t b = std::min(bytes, t * per), e = std::min(bytes, b + per);
    if (e > b) std::memcpy(static_cast<char *>(dst) + b, static_cast<const char *>(src) + b, e - b);
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; ++t) pool.emplace_back(part, t);
  part(0);
  for (auto &t : pool) t.join();
}

// bytes of device memory to a host destination that may be pageable: one DMA
// into a pinned buffer from the result pool, then up to 8 host threads copy
// it out (the runtime's own pageable copy stages through one thread; the
// first touch of a fresh destination's pages is what it waits on).
void to_host_pageable(Ctx *c, void *dst, const void *src, size_t bytes) {
  if (bytes < ((size_t)4 << 20)) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return;
  }
  char *stage = static_cast<char *>(result_alloc_bytes(bytes));
  try {
    HIP_CHECK(hipMemcpyAsync(stage, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  } catch (...) {
    result_release(stage);
    throw;
  }
  parallel_copy(dst, stage, bytes);
  result_release(stage);
}

// The reverse: a host source that may be pageable copied into a pooled
// pinned buffer by up to 8 host threads, then one DMA.
void to_device_pageable(Ctx *c, void *dst, const void *src, size_t bytes) {
  if (bytes < ((size_t)4 << 20)) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return;
  }
  char *stage = static_cast<char *>(result_alloc_bytes(bytes));
  parallel_copy(stage, src, bytes);
  try {
    HIP_CHECK(hipMemcpyAsync(dst, stage, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  } catch (...) {
    result_release(stage);
    throw;
  }
  result_release(stage);
}

// Row-major panel (rows x ld, dtype dt, columns past cols zero) of the first
// `cols` singular vectors held on the host in a solve's result layout:
// vectors as columns (src is rows x src_cols row-major) or as rows (src is
// src_rows x rows row-major, vector j at src + j * rows). The host block goes
// over in one contiguous copy (the whole rows x src_cols matrix, or the first
// cols rows; a strided copy from pageable memory moves one row at a time)
// and the device selects, converts and transposes.
void upload_factor(Ctx *c, DevBuf &buf, int64_t rows, int64_t cols, int64_t ld, int dt, const void *src, int sdt,
                   int64_t src_cols, bool vectors_as_rows) {
  const size_t es = dtype_size(dt), ses = dtype_size(sdt);
  buf.alloc(std::max<size_t>((size_t)rows * ld * es, 64));
  if (rows == 0) return;
  HIP_CHECK(hipMemsetAsync(buf.p, 0, (size_t)rows * ld * es, c->stream));
  if (cols == 0) return;
  if (vectors_as_rows) {
    DevBuf tmp(std::max<size_t>((size_t)rows * cols * ses, 64));
    to_device_pageable(c, tmp.p, src, (size_t)rows * cols * ses);
    panel_transpose_convert(c, cols, rows, tmp.p, rows, sdt, buf.p, ld, dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));  // tmp is released here
  } else {
    DevBuf tmp(std::max<size_t>((size_t)rows * src_cols * ses, 64));
    to_device_pageable(c, tmp.p, src, (size_t)rows * src_cols * ses);
    panel_convert(c, rows, cols, tmp.p, src_cols, sdt, buf.p, ld, dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
}

void to_host(Ctx *c, const double *dev, int64_t n, std::vector<double> &out) {
  out.resize(n);
  HIP_CHECK(hipMemcpyAsync(out.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace

int32_t backward_solve(svdh_ctx *ctx, svdh_op *op, const svdh_result *res, int32_t loss, int64_t k, const double *y,
                       int32_t precision, double tol, int64_t max_iters, void *grad, int32_t grad_dtype,
                       svdh_backward_info *info) {
  const int64_t m = op->nrows, n = op->ncols, d = res->d;
  if (loss != SVDH_LOSS_SIGMA_SUM && loss != SVDH_LOSS_PROJECTION) fail(SVDH_ERR_ARG, "bad loss kind %d", loss);
  if (k < 1 || k > d || k > 64) fail(SVDH_ERR_ARG, "k = %lld outside 1..min(d = %lld, 64)", (long long)k, (long long)d);
  if (precision != SVDH_PREC_F64 && precision != SVDH_PREC_F32 && precision != SVDH_PREC_BF16)
    fail(SVDH_ERR_ARG, "bad precision %d", precision);
  if (grad_dtype != SVDH_F64 && grad_dtype != SVDH_F32) fail(SVDH_ERR_ARG, "bad gradient dtype %d", grad_dtype);
  if (loss == SVDH_LOSS_PROJECTION && !y) fail(SVDH_ERR_ARG, "the projection loss needs y");
  // factors in FP64 (u, s, vt) or single precision (u32, s32, vt32)
  const bool r64 = res->u && res->s && res->vt, r32 = res->u32 && res->s32 && res->vt32;
  if (!r64 && !r32) fail(SVDH_ERR_ARG, "result without factors");
  auto s_at = [&](int64_t i) { return r64 ? res->s[i] : (double)res->s32[i]; };
  auto u_at = [&](int64_t i) { return r64 ? res->u[i] : (double)res->u32[i]; };
  auto vt_at = [&](int64_t i) { return r64 ? res->vt[i] : (double)res->vt32[i]; };
  // u is m x d (randomized layout) or d x m (LAS2 layout, rows are left vectors)
  bool u_rows_are_vectors;
  if (res->u_rows == m && res->u_cols == d)
    u_rows_are_vectors = false;
  else if (res->u_rows == d && res->u_cols == m)
    u_rows_are_vectors = true;
  else
    fail(SVDH_ERR_SHAPE, "u is %lld x %lld, expected %lld x %lld or its transpose", (long long)res->u_rows,
         (long long)res->u_cols, (long long)m, (long long)d);
  if (res->vt_rows != d || res->vt_cols != n)
    fail(SVDH_ERR_SHAPE, "vt is %lld x %lld, expected %lld x %lld", (long long)res->vt_rows, (long long)res->vt_cols,
         (long long)d, (long long)n);
  Ctx *c = ctx;
  // panels and SpMM arithmetic: FP64, or FP32 for both FP32 and BF16 (the
  // CG needs an exact operator: BF16 inputs would stall it at their rounding)
  const int dt = precision == SVDH_PREC_F64 ? SVDH_F64 : SVDH_F32;
  // BF16: the SDDMM gathers BF16 copies of its factor panels (half the
  // gathered bytes) and accumulates in FP32
  const bool bf = precision == SVDH_PREC_BF16;
  const int64_t ld = (k + 7) / 8 * 8;
  std::vector<DevBuf> bf_panels;
  auto sddmm_factor = [&](const void *P, int64_t rows) -> const void * {
    if (!bf) return P;
    bf_panels.emplace_back(std::max<size_t>((size_t)rows * ld * 2, 64));
    void *q = bf_panels.back().p;
    panel_convert(c, rows, ld, P, ld, SVDH_F32, q, ld, SVDH_BF16);
    return q;
  };
  const int sdt = bf ? SVDH_BF16 : dt;
  auto uget = [&](int64_t i, int64_t j) {
    return u_rows_are_vectors ? u_at(j * m + i) : u_at(i * d + j);
  };
  DevBuf U, V, G;
  upload_factor(c, U, m, k, ld, dt, r64 ? (const void *)res->u : (const void *)res->u32, r64 ? SVDH_F64 : SVDH_F32,
                d, u_rows_are_vectors);
  upload_factor(c, V, n, k, ld, dt, r64 ? (const void *)res->vt : (const void *)res->vt32, r64 ? SVDH_F64 : SVDH_F32,
                n, true);
  const int64_t nnz = op->base ? op->base->nnz : op->nnz;
  G.alloc(std::max<size_t>((size_t)nnz * dtype_size(grad_dtype), 16));
  std::memset(info, 0, sizeof *info);

  if (loss == SVDH_LOSS_SIGMA_SUM) {
    double L = 0.0;
    for (int64_t i = 0; i < k; ++i) L += s_at(i);
    info->loss = L;
    sddmm(op, k, sddmm_factor(U.p, m), sddmm_factor(V.p, n), nullptr, nullptr, nullptr, nullptr, ld, ld, sdt, G.p,
          grad_dtype);
  } else {
    // c = U^T y, b = P y on the host in FP64 (O(m k))
    std::vector<double> cy(k, 0.0), b(y, y + m);
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < k; ++j) cy[j] += uget(i, j) * y[i];
    double L = 0.0;
    for (int64_t j = 0; j < k; ++j) L += cy[j] * cy[j];
    info->loss = L;
    double bn = 0.0;
    for (int64_t i = 0; i < m; ++i) {
      double t = 0.0;
      for (int64_t j = 0; j < k; ++j) t += uget(i, j) * cy[j];
      b[i] -= t;
      bn += b[i] * b[i];
    }
    bn = std::sqrt(bn);
    if (tol <= 0.0) tol = dt == SVDH_F64 ? 1e-10 : 1e-5;
    if (max_iters <= 0) max_iters = 1000;
    DevBuf W, R, P, Q, T, Z, scal;
    upload_panel(c, R, m, k, ld, dt, [&](int64_t i, int64_t) { return b[i]; });
    upload_panel(c, P, m, k, ld, dt, [&](int64_t i, int64_t) { return b[i]; });
    upload_panel(c, W, m, k, ld, dt, [&](int64_t, int64_t) { return 0.0; });
    Q.alloc((size_t)m * ld * dtype_size(dt));
    HIP_CHECK(hipMemsetAsync(Q.p, 0, Q.bytes, c->stream));
    T.alloc((size_t)std::max<int64_t>(n, 1) * ld * dtype_size(dt));
    HIP_CHECK(hipMemsetAsync(T.p, 0, T.bytes, c->stream));
    // device vectors: shifts s_i^2, -1, dots, |r|^2, cross Gram (k x k)
    scal.alloc((size_t)(4 * 64 + 64 * 64) * sizeof(double));
    double *shifts = scal.as<double>(), *neg1 = shifts + 64, *dots = shifts + 128, *rr = shifts + 192,
           *Cx = shifts + 256;
    std::vector<double> h(64, 0.0);
    for (int64_t j = 0; j < k; ++j) h[j] = s_at(j) * s_at(j);
    HIP_CHECK(hipMemcpyAsync(shifts, h.data(), 64 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<double> hm1(64, -1.0);
    HIP_CHECK(hipMemcpyAsync(neg1, hm1.data(), 64 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<double> rho(k, bn * bn), gam, rrh;
    // a column stops (alpha = beta = 0) once it meets tol: iterating a
    // converged column only amplifies rounding
    std::vector<char> done(k, 0);
    int64_t it = 0;
    double relres = bn > 0.0 ? 1.0 : 0.0;
    auto project = [&](void *X) {  // X <- X - U (U^T X)
      gram_cross(c, m, k, U.p, ld, k, X, ld, dt, Cx);
      panel_gemm_dev(c, m, k, k, U.p, ld, Cx, -1.0, 1.0, X, ld, dt);
    };
    while (bn > 0.0 && it < max_iters) {
      spmm(op, 1, dt, P.p, ld, dt, T.p, ld, dt, k);          // T = A^T P
      spmm(op, 0, dt, T.p, ld, dt, Q.p, ld, dt, k);          // Q = A A^T P
      panel_colaxpby(c, m, k, shifts, P.p, neg1, Q.p, ld, dt);  // Q = s^2 P - A A^T P
      project(Q.p);
      panel_coldots(c, m, k, P.p, Q.p, ld, dt, dots);
      to_host(c, dots, k, gam);
      std::vector<double> ha(64, 0.0);
      for (int64_t j = 0; j < k; ++j) {
        if (!(gam[j] > 0.0)) done[j] = 1;  // breakdown: no further progress on this column
        ha[j] = done[j] ? 0.0 : rho[j] / gam[j];
      }
      const double *alpha_d = static_cast<const double *>(c->upload(ha.data(), 64 * sizeof(double)));
      panel_cg_update(c, m, k, W.p, R.p, P.p, Q.p, ld, dt, alpha_d, rr);
      to_host(c, rr, k, rrh);
      ++it;
      relres = 0.0;
      bool all = true;
      for (int64_t j = 0; j < k; ++j) {
        const double rj = std::sqrt(std::max(rrh[j], 0.0)) / bn;
        relres = std::max(relres, rj);
        if (rj <= tol) done[j] = 1;
        all = all && done[j];
      }
      if (all) break;
      std::vector<double> hb(64, 0.0);
      for (int64_t j = 0; j < k; ++j) {
        hb[j] = (!done[j] && rho[j] > 0.0) ? rrh[j] / rho[j] : 0.0;
        if (!done[j]) rho[j] = rrh[j];
      }
      const double *beta_d = static_cast<const double *>(c->upload(hb.data(), 64 * sizeof(double)));
      panel_colaxpby(c, m, k, nullptr, R.p, beta_d, P.p, ld, dt);  // P = R + beta P
    }
    info->cg_iterations = it;
    info->cg_relative_residual = relres;
    project(W.p);
    Z.alloc((size_t)std::max<int64_t>(n, 1) * ld * dtype_size(dt));
    HIP_CHECK(hipMemsetAsync(Z.p, 0, Z.bytes, c->stream));
    spmm(op, 1, dt, W.p, ld, dt, Z.p, ld, dt, k);  // Z = A^T W
    std::vector<double> s1(64, 0.0), s2(64, 0.0);
    for (int64_t j = 0; j < k; ++j) {
      s1[j] = 2.0 * cy[j] * s_at(j);
      s2[j] = 2.0 * cy[j];
    }
    const double *s1_d = static_cast<const double *>(c->upload(s1.data(), 64 * sizeof(double)));
    const double *s2_d = static_cast<const double *>(c->upload(s2.data(), 64 * sizeof(double)));
    sddmm(op, k, sddmm_factor(W.p, m), sddmm_factor(V.p, n), s1_d, sddmm_factor(U.p, m), sddmm_factor(Z.p, n),
          s2_d, ld, ld, sdt, G.p, grad_dtype);
  }
  to_host_pageable(c, grad, G.p, (size_t)nnz * dtype_size(grad_dtype));
  return SVDH_OK;
}

}  // namespace svdh
