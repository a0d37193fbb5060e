// Vector and dense kernels of the svdh library (gfx950).
//
//  - fused FP64 BLAS-1 for the Lanczos step (src/lanczos/mod.rs:321-424,
//    674-758): y += a x updates followed by a dot product in one pass, the
//    scalar kept on the device; reductions are two-level with a fixed block
//    order, so results are run-to-run identical;
//  - block GEMVs over the stored Lanczos basis for purge's classical
//    Gram-Schmidt (mod.rs:760-813) and the Ritz back-transform (mod.rs:1150-1167);
//  - tall-skinny Gram matrices, panel-times-small-matrix products, column
//    normalisation and Gaussian panels for the randomized solver
//    (src/randomized/mod.rs:24-313).
#include <cmath>
#include <cstring>

#include "common.h"

namespace svdh {

namespace {

typedef uint16_t bf16_t;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum in fixed order (wave sums, then waves in order); result in
// thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double w[kBlock / 64];
  v = wave_sum_d(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) t += w[i];
  }
  return t;
}

// Dot products of the fused kernel leave one partial per workgroup in the
// slot's row of Ctx::slot_parts (kSlotParts entries per slot); whoever needs
// the value (the next fused kernel's coefficient, k_slot_finish for the host)
// sums them with slot_sum, in the same order everywhere, so the device and the
// host see the same bits. No fences or atomics between workgroups.
constexpr int kSlotParts = 1024;  // >= fused_grid's cap
constexpr int64_t kChainMaxN = 8192;  // vectors a single workgroup chain takes

// Called by whole waves; every lane returns the sum of p[0 .. nb).
__device__ __forceinline__ double slot_sum(const double *p, int nb) {
  double s = 0.0;
  for (int b = threadIdx.x & 63; b < nb; b += 64) s += p[b];
  return wave_sum_d(s);
}

struct AxpyD {
  double *y;
  const double *x;
  double imm;
  int slot;
  int nb;  // partials of the slot
  double sign;
};

struct FusedArgs {
  int64_t n;
  AxpyD up[2];
  int nups;
  const double *u;
  const double *v;
  int out_slot;
  double *copy_dst;
  const double *copy_src;
  double *parts;  // Ctx::slot_parts
};

__global__ __launch_bounds__(kBlock) void k_fused_axpy_dot(FusedArgs a) {
  double c0 = 0.0, c1 = 0.0;
  if (a.nups > 0)
    c0 = a.up[0].slot >= 0 ? a.up[0].sign * slot_sum(a.parts + a.up[0].slot * kSlotParts, a.up[0].nb)
                           : a.up[0].imm;
  if (a.nups > 1)
    c1 = a.up[1].slot >= 0 ? a.up[1].sign * slot_sum(a.parts + a.up[1].slot * kSlotParts, a.up[1].nb)
                           : a.up[1].imm;
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < a.n; i += stride) {
    if (a.nups > 0) a.up[0].y[i] += c0 * a.up[0].x[i];
    if (a.nups > 1) a.up[1].y[i] += c1 * a.up[1].x[i];
    if (a.copy_dst) a.copy_dst[i] = a.copy_src[i];
    if (a.out_slot >= 0) acc += a.u[i] * a.v[i];
  }
  if (a.out_slot < 0) return;
  const double t = block_sum(acc);
  if (threadIdx.x == 0) a.parts[a.out_slot * kSlotParts + blockIdx.x] = t;
}

// A chain of fused ops on short vectors in one workgroup: op o's dot is
// reduced in the workgroup (waves in order) before op o + 1 reads it, every
// thread touches the same entries in every op, so no other wait is needed.
// Coefficients come from slots produced earlier in the chain (the workgroup's
// own values) or before it (their partials, slot_sum). The produced slots are
// published as their single partial, to scal and to the host mirror.
constexpr int kChainMax = 8;
constexpr int kChainBlock = 1024;
struct ChainOp {
  AxpyD up;
  int nups;
  int in_chain;  // up.slot was produced earlier in this chain
  const double *u;
  const double *v;
  int out_slot;
  double *copy_dst;
  const double *copy_src;
  int sqrt_slot;  // >= 0: sqrt of the dot is published to this slot as well
};
struct ChainArgs {
  int64_t n;
  int nops;
  ChainOp ops[kChainMax];
  double *parts;
  double *scal;
  double *host;
};

__global__ __launch_bounds__(kChainBlock) void k_fused_chain(ChainArgs a) {
  __shared__ double sv[S_NSLOT];
  __shared__ double red[kChainBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int o = 0; o < a.nops; ++o) {
    const ChainOp &op = a.ops[o];
    double c0 = 0.0;
    if (op.nups > 0) {
      if (op.up.slot < 0)
        c0 = op.up.imm;
      else
        c0 = op.up.sign * (op.in_chain ? sv[op.up.slot] : slot_sum(a.parts + op.up.slot * kSlotParts, op.up.nb));
    }
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < a.n; i += kChainBlock) {
      if (op.nups > 0) op.up.y[i] += c0 * op.up.x[i];
      if (op.copy_dst) op.copy_dst[i] = op.copy_src[i];
      if (op.out_slot >= 0) acc += op.u[i] * op.v[i];
    }
    if (op.out_slot >= 0) {
      acc = wave_sum_d(acc);
      if (lane == 0) red[wv] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kChainBlock / 64; ++w) t += red[w];
        sv[op.out_slot] = t;
        a.parts[op.out_slot * kSlotParts] = t;
        a.scal[op.out_slot] = t;
        a.host[op.out_slot] = t;
        if (op.sqrt_slot >= 0) {
          const double r = sqrt(t);
          sv[op.sqrt_slot] = r;
          a.parts[op.sqrt_slot * kSlotParts] = r;
          a.scal[op.sqrt_slot] = r;
          a.host[op.sqrt_slot] = r;
        }
      }
    }
    __syncthreads();  // the value (and red) before the next op
  }
}

// scal[s] = the value of slot s (one 64-thread workgroup per slot, s = first
// + block). The workgroup of slot sqrt_from also publishes sqrt(value) as the
// single partial of slot sqrt_to (a slot no other workgroup of the launch
// reads or writes).
struct SlotCounts {
  int nb[S_NSLOT];
  int first;
  int sqrt_from, sqrt_to;
};
__global__ void k_slot_finish(double *parts, SlotCounts cnt, double *scal, double *host) {
  const int s = cnt.first + blockIdx.x;
  if (cnt.nb[s] <= 0) return;
  const double v = slot_sum(parts + s * kSlotParts, cnt.nb[s]);
  if (threadIdx.x == 0) {
    scal[s] = v;
    host[s] = v;
    if (s == cnt.sqrt_from) {
      const double r = sqrt(v);
      parts[cnt.sqrt_to * kSlotParts] = r;
      scal[cnt.sqrt_to] = r;
      host[cnt.sqrt_to] = r;
    }
  }
}

__global__ void k_scale2(int64_t n, double a, const double *x, double *y, double *z, double *cdst,
                         const double *csrc) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    y[i] = a * x[i];
    z[i] *= a;
    if (cdst) cdst[i] = csrc[i];
  }
}

// slot values sum |a_i| and sum |b_i| (i < k): one workgroup, thread t adds
// the entries t, t + 256, ... and the threads' sums are added in a fixed order.
__global__ __launch_bounds__(kBlock) void k_abs_sums(int64_t k, const double *a, const double *b, double *pa,
                                                    double *pb) {
  double sa = 0.0, sb = 0.0;
  for (int64_t i = threadIdx.x; i < k; i += kBlock) {
    sa += fabs(a[i]);
    sb += fabs(b[i]);
  }
  sa = block_sum(sa);
  sb = block_sum(sb);
  if (threadIdx.x == 0) {
    *pa = sa;
    *pb = sb;
  }
}

// The Lanczos step's scaling, out of place: s = 1 / rnm with rnm the single
// partial of a slot (rpart) or s = a from the host; y = s x, zout = s zin,
// and cdst = csrc when cdst is set.
__global__ void k_scale_step(int64_t n, const double *rpart, double a, const double *x, double *y,
                             const double *zin, double *zout, double *cdst, const double *csrc) {
  const double s = rpart ? 1.0 / rpart[0] : a;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    y[i] = s * x[i];
    zout[i] = s * zin[i];
    if (cdst) cdst[i] = csrc[i];
  }
}

__global__ void k_f64_to_f32(int64_t n, const double *x, float *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

__global__ void k_copy(int64_t n, const double *x, double *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = x[i];
}

// Partial products of the reorthogonalisation's coefficients: workgroup
// (tile, group) takes kGemvRows rows of u, v and the basis and kGemvCols
// columns, and leaves part_u[col * ntiles + tile] = Q[rows, col]^T u[rows]
// (part_v likewise); k_gemv_t_sum adds a column's tile partials in tile order.
constexpr int kGemvRows = 4096;
constexpr int kGemvCols = 8;
__global__ __launch_bounds__(kBlock) void k_gemv_t(int64_t n, const double *Q, int64_t ldq, int64_t k,
                                                  const double *u, const double *v, double *part_u,
                                                  double *part_v) {
  __shared__ double red[2 * kGemvCols][kBlock / 64];
  const int64_t r0 = (int64_t)blockIdx.x * kGemvRows;
  const int64_t r1 = min<int64_t>(n, r0 + kGemvRows);
  const int64_t c0 = (int64_t)blockIdx.y * kGemvCols;
  const int nc = (int)min<int64_t>(kGemvCols, k - c0);
  double su[kGemvCols], sv[kGemvCols];
  int64_t off[kGemvCols];  // columns past k read the last one; their sums are dropped
#pragma unroll
  for (int c = 0; c < kGemvCols; ++c) {
    su[c] = sv[c] = 0.0;
    off[c] = (c0 + min(c, nc - 1)) * ldq;
  }
  // no branch around the loads: all of a row's loads are in flight together
#pragma unroll 2
  for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const double ui = u[i], vi = v[i];
    double qv[kGemvCols];
#pragma unroll
    for (int c = 0; c < kGemvCols; ++c) qv[c] = Q[off[c] + i];
#pragma unroll
    for (int c = 0; c < kGemvCols; ++c) {
      su[c] += qv[c] * ui;
      sv[c] += qv[c] * vi;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kGemvCols; ++c) {
    const double a = wave_sum_d(su[c]), b = wave_sum_d(sv[c]);
    if (lane == 0) {
      red[c][wv] = a;
      red[kGemvCols + c][wv] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * kGemvCols) {
    const int c = threadIdx.x % kGemvCols;
    if (c < nc) {
      double t = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) t += red[threadIdx.x][w];
      double *dst = threadIdx.x < kGemvCols ? part_u : part_v;
      dst[(c0 + c) * gridDim.x + blockIdx.x] = t;
    }
  }
}

__global__ void k_gemv_t_sum(int64_t k, int ntiles, const double *part_u, const double *part_v, double *cu,
                             double *cv) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= k) return;
  double a = 0.0, b = 0.0;
  for (int t = 0; t < ntiles; ++t) {
    a += part_u[c * ntiles + t];
    b += part_v[c * ntiles + t];
  }
  cu[c] = a;
  cv[c] = b;
}

// y1 -= Q c1, y2 -= Q c2 for an n x k column-major basis: a workgroup takes
// 64 rows, its wave w the columns j = w mod 4 (each load one 512-byte row
// segment), and the four waves' sums are added in wave order.
__global__ __launch_bounds__(kBlock) void k_gemv_n(int64_t n, const double *Q, int64_t ldq, int64_t k,
                                                  const double *c1, double *y1, const double *c2, double *y2) {
  __shared__ double s1[1024], s2[1024];
  __shared__ double red[2][kBlock / 64][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t i = blockIdx.x * (int64_t)64 + lane;
  const bool ok = i < n;
  double a1 = 0.0, a2 = 0.0;
  for (int64_t j0 = 0; j0 < k; j0 += 1024) {
    const int64_t kk = min<int64_t>(1024, k - j0);
    __syncthreads();
    for (int t = threadIdx.x; t < kk; t += kBlock) {
      s1[t] = c1[j0 + t];
      s2[t] = c2[j0 + t];
    }
    __syncthreads();
    if (ok) {
      const double *q = Q + j0 * ldq + (ok ? i : 0);
#pragma unroll 4
      for (int64_t j = wv; j < kk; j += kBlock / 64) {
        const double qv = q[j * ldq];
        a1 += qv * s1[j];
        a2 += qv * s2[j];
      }
    }
  }
  red[0][wv][lane] = a1;
  red[1][wv][lane] = a2;
  __syncthreads();
  if (wv == 0 && ok) {
    double t1 = 0.0, t2 = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
      t1 += red[0][w][lane];
      t2 += red[1][w][lane];
    }
    y1[i] -= t1;
    y2[i] -= t2;
  }
}

// C (n x d row-major) = Q (n x k col-major) W (k x d row-major); 16 outputs per thread.
typedef double d4_t __attribute__((ext_vector_type(4)));

// C (n x d, row-major, ldc) = Q (n x k, column-major, ldq) W (k x d, row-major)
// on the FP64 matrix cores (v_mfma_f64_16x16x4: per lane l, A[l % 16][l / 16],
// B[l / 16][l % 16], D[l / 16 + 4 v][l % 16]). One wave per 16-row block and
// column groups of up to 8 tiles of 16; the k loop takes 4 steps of 4 per
// batch, loads first. A[i][kk] = Q[r0 + i][k0 + kk] (16 consecutive rows of a
// column per kk), B[kk][j] = W[k0 + kk][c0 + 16 t + j].
__global__ __launch_bounds__(256) void k_combine_mfma(int64_t n, const double *Q, int64_t ldq, int64_t k,
                                                      const double *W, int64_t d, double *C, int64_t ldc) {
  const int lane = threadIdx.x & 63;
  const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (r0 >= n) return;  // the whole wave
  const int i = lane & 15, kk = lane >> 4;
  const int64_t row = r0 + i;
  const bool rok = row < n;
  constexpr int T = 8, S = 4;
  for (int64_t c0 = 0; c0 < d; c0 += 16 * T) {
    d4_t acc[T];
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t] = d4_t{0.0, 0.0, 0.0, 0.0};
    for (int64_t k0 = 0; k0 < k; k0 += 4 * S) {
      double a[S], b[S][T];
#pragma unroll
      for (int st = 0; st < S; ++st) {
        const int64_t kq = k0 + 4 * st + kk;
        const bool kok = kq < k;
        const int64_t kqc = kok ? kq : 0;  // in-range address; the value is dropped below
        const double av = Q[kqc * ldq + (rok ? row : r0)];
        a[st] = kok && rok ? av : 0.0;
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const int64_t col = c0 + 16 * t + i;
          const double bv = W[kqc * d + (col < d ? col : 0)];
          b[st][t] = kok && col < d ? bv : 0.0;
        }
      }
#pragma unroll
      for (int st = 0; st < S; ++st)
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[st], b[st][t], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t orow = r0 + kk + 4 * v, ocol = c0 + 16 * t + i;
        if (orow < n && ocol < d) C[orow * ldc + ocol] = acc[t][v];
      }
  }
}

// Z[r, c] (at r ldz + c zcs, FP64) = beta Z[r, c] + alpha (Y M)[r, c] for a
// row-major panel Y (m x li, FP32 or FP64) and M (li x lo <= 16 T, FP64
// row-major) on the FP64 matrix cores, laid out as k_combine_mfma with
// A[i][kk] = Y[r0 + i][k0 + kk]. A wave reads all of its 16 rows before
// writing any, so Z may be Y (zcs = 1, same rows).
template <typename TY, int T>
__device__ __forceinline__ void apply_mfma_body(int64_t bx, int64_t m, int li, int lo, const TY *Y, int64_t ldy,
                                                const double *M, double *Z, int64_t ldz, int64_t zcs, double alpha,
                                                double beta) {
  const int lane = threadIdx.x & 63;
  const int64_t r0 = (bx * 4 + (threadIdx.x >> 6)) * 16;
  if (r0 >= m) return;  // the whole wave
  const int i = lane & 15, kk = lane >> 4;
  const int64_t row = r0 + i;
  const bool rok = row < m;
  constexpr int S = 4;
  d4_t acc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) acc[t] = d4_t{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < li; k0 += 4 * S) {
    double a[S], b[S][T];
#pragma unroll
    for (int st = 0; st < S; ++st) {
      const int kq = k0 + 4 * st + kk;
      const bool kok = kq < li;
      const int kqc = kok ? kq : 0;  // in-range addresses; the values are dropped below
      const double av = (double)Y[(rok ? row : r0) * ldy + kqc];
      a[st] = kok && rok ? av : 0.0;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int col = 16 * t + i;
        const double bv = M[(int64_t)kqc * lo + (col < lo ? col : 0)];
        b[st][t] = kok && col < lo ? bv : 0.0;
      }
    }
#pragma unroll
    for (int st = 0; st < S; ++st)
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[st], b[st][t], acc[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t orow = r0 + kk + 4 * v;
      const int ocol = 16 * t + i;
      if (orow < m && ocol < lo) {
        const int64_t at = orow * ldz + (int64_t)ocol * zcs;
        Z[at] = (beta != 0.0 ? beta * Z[at] : 0.0) + alpha * acc[t][v];
      }
    }
}

template <typename TY, int T>
__global__ __launch_bounds__(256) void k_apply_mfma(int64_t m, int li, int lo, const TY *Y, int64_t ldy,
                                                    const double *M, double *Z, int64_t ldz, int64_t zcs,
                                                    double alpha, double beta) {
  apply_mfma_body<TY, T>(blockIdx.x, m, li, lo, Y, ldy, M, Z, ldz, zcs, alpha, beta);
}

// Many independent Z_j = Y_j M_j in one launch: job j = blockIdx.y takes row
// blocks 0 .. of the x dimension, each computed exactly as k_apply_mfma does.
struct ApplyJob {
  const void *Y;
  int64_t ldy, m;
  int32_t li, lo;
  const double *M;
  double *Z;
  int64_t ldz, zcs;
};
template <typename TY, int T>
__global__ __launch_bounds__(256) void k_apply_mfma_jobs(const ApplyJob *jobs) {
  const ApplyJob &j = jobs[blockIdx.y];
  apply_mfma_body<TY, T>(blockIdx.x, j.m, j.li, j.lo, static_cast<const TY *>(j.Y), j.ldy, j.M, j.Z, j.ldz, j.zcs,
                         1.0, 0.0);
}

template <typename T> __device__ __forceinline__ double ldd(const T *p, int64_t i);
template <> __device__ __forceinline__ double ldd<double>(const double *p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ double ldd<float>(const float *p, int64_t i) { return (double)p[i]; }
template <> __device__ __forceinline__ double ldd<bf16_t>(const bf16_t *p, int64_t i) {
  return (double)__uint_as_float((uint32_t)p[i] << 16);
}

template <typename T> __device__ __forceinline__ void std_(T *p, int64_t i, double v);
template <> __device__ __forceinline__ void std_<double>(double *p, int64_t i, double v) { p[i] = v; }
template <> __device__ __forceinline__ void std_<float>(float *p, int64_t i, double v) { p[i] = (float)v; }
template <> __device__ __forceinline__ void std_<bf16_t>(bf16_t *p, int64_t i, double v) {
  float f = (float)v;
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  p[i] = (bf16_t)(u >> 16);
}

// Tall-skinny dense kernels work on 64-row tiles of a row-major panel held in
// LDS; each of the 256 threads owns a 4 x 4 block of a 64 x 64 output tile.
constexpr int kTile = 64;

template <typename T, typename C>
__device__ __forceinline__ void load_tile(const T *Y, int64_t ldy, int64_t r0, int rows, int w, C *s, int sld) {
  // s[k * sld + r] = Y[r0 + r, k] (transposed so 4 consecutive rows are one 16-byte read)
  for (int e = threadIdx.x; e < kTile * w; e += kBlock) {
    const int r = e / w, k = e % w;
    s[k * sld + r] = r < rows ? (C)ldd<T>(Y, (r0 + r) * ldy + k) : C(0);
  }
}

// G[e] = sum over the nb partials part[b * ll + e]. 16 outputs per
// workgroup, 16 lanes per output each summing every 16th partial (two
// chains), then the 16 lane sums in order: a fixed order for every launch.
__device__ __forceinline__ void gram_reduce_body(int nb, int64_t ll, const double *part, double *G, int64_t bx) {
  __shared__ double red[16][17];
  const int o = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int64_t e = bx * 16 + o;
  double s0 = 0.0, s1 = 0.0;
  if (e < ll) {
    int b = q;
    for (; b + 16 < nb; b += 32) {
      s0 += part[(int64_t)b * ll + e];
      s1 += part[(int64_t)(b + 16) * ll + e];
    }
    if (b < nb) s0 += part[(int64_t)b * ll + e];
  }
  red[q][o] = s0 + s1;
  __syncthreads();
  if (q == 0 && e < ll) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) t += red[i][o];
    G[e] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_gram_reduce(int nb, int64_t ll, const double *part, double *G) {
  gram_reduce_body(nb, ll, part, G, blockIdx.x);
}

static int reduce_grid(int64_t ll) { return (int)((ll + 15) / 16); }

// Gram partials on the FP64 matrix cores. v_mfma_f64_16x16x4 takes, per lane
// l, A[l % 16][l / 16] and B[l / 16][l % 16] and returns D[l / 16 + 4 v][l % 16]
// in element v (this lane layout was checked on gfx950 with one-hot inputs).
// A 64-column group c0.. is split into 4 strided column blocks
// b = {c0 + 4 t + b : t < 16}, so lane (k, t) = (l / 16, l % 16) reads the 4
// consecutive columns c0 + 4 t .. c0 + 4 t + 3 of row 4 s + k with one vector
// load and feeds all 4 blocks; the MFMA of blocks (b1, b2) adds
// G[c0 + 4 i + b1][d0 + 4 j + b2] += sum_k Y[4 s + k][c0 + 4 i + b1] Y[4 s + k][d0 + 4 j + b2].
// FP32 and BF16 inputs are exact in FP64, so this is the FP64 accumulation of
// exact products. Each wave owns a contiguous range of 4-row steps; the waves
// of a workgroup are summed in LDS in wave order and the workgroup partial is
// written mirrored to part[blockIdx * l * l] (disjoint entries per launch
// when l > 64 needs several (c0, d0) launches).
constexpr int kGramWaves = 8;

template <typename T> struct Quad;
template <> struct Quad<float> {
  float4 v;
  __device__ __forceinline__ void load(const float *p) { v = *reinterpret_cast<const float4 *>(p); }
  __device__ __forceinline__ void zero() { v = make_float4(0.f, 0.f, 0.f, 0.f); }
  __device__ __forceinline__ double get(int b) const { return (double)(b == 0 ? v.x : b == 1 ? v.y : b == 2 ? v.z : v.w); }
};
template <> struct Quad<double> {
  double2 a, b2;
  __device__ __forceinline__ void load(const double *p) {
    a = reinterpret_cast<const double2 *>(p)[0];
    b2 = reinterpret_cast<const double2 *>(p)[1];
  }
  __device__ __forceinline__ void zero() { a = make_double2(0.0, 0.0); b2 = a; }
  __device__ __forceinline__ double get(int b) const { return b == 0 ? a.x : b == 1 ? a.y : b == 2 ? b2.x : b2.y; }
};
template <> struct Quad<bf16_t> {
  uint2 v;
  __device__ __forceinline__ void load(const bf16_t *p) { v = *reinterpret_cast<const uint2 *>(p); }
  __device__ __forceinline__ void zero() { v = make_uint2(0u, 0u); }
  __device__ __forceinline__ double get(int b) const {
    const uint32_t w = b < 2 ? v.x : v.y;
    return (double)__uint_as_float((b & 1) ? (w & 0xffff0000u) : (w << 16));
  }
};

// Two consecutive columns per lane (panels of width <= 32, two strided
// column blocks of 16).
template <typename T> struct Pair;
template <> struct Pair<float> {
  float2 v;
  __device__ __forceinline__ void load(const float *p) { v = *reinterpret_cast<const float2 *>(p); }
  __device__ __forceinline__ void zero() { v = make_float2(0.f, 0.f); }
  __device__ __forceinline__ double get(int b) const { return (double)(b == 0 ? v.x : v.y); }
};
template <> struct Pair<double> {
  double2 v;
  __device__ __forceinline__ void load(const double *p) { v = *reinterpret_cast<const double2 *>(p); }
  __device__ __forceinline__ void zero() { v = make_double2(0.0, 0.0); }
  __device__ __forceinline__ double get(int b) const { return b == 0 ? v.x : v.y; }
};
template <> struct Pair<bf16_t> {
  uint32_t v;
  __device__ __forceinline__ void load(const bf16_t *p) { v = *reinterpret_cast<const uint32_t *>(p); }
  __device__ __forceinline__ void zero() { v = 0u; }
  __device__ __forceinline__ double get(int b) const {
    return (double)__uint_as_float(b ? (v & 0xffff0000u) : (v << 16));
  }
};
template <typename T, int NB> struct GramVec { using type = Quad<T>; };
template <typename T> struct GramVec<T, 2> { using type = Pair<T>; };

// Waves per workgroup: 8 for the diagonal launch (two waves per SIMD), 4 for
// the off-diagonal one, whose 16 accumulators need the whole register file.
template <bool SAME> constexpr int gram_waves() { return SAME ? kGramWaves : 4; }

// Workgroup bid of nbk over the rows of one panel; writes that workgroup's
// partial to out.
template <typename T, bool SAME, int NB = 4>
__device__ __forceinline__ void gram_mfma_body(int64_t bid, int64_t nbk, int64_t m, int l, const T *Y, int64_t ldy,
                                               int l2, const T *Y2, int64_t ld2, int c0, int d0, double *out,
                                               int out_ld, int mirror) {
  constexpr int W = gram_waves<SAME>();
  __shared__ double red[16 * 256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kr = lane >> 4, cc = lane & 15;
  const int64_t nw = nbk * W, gw = bid * W + wv;
  const int64_t steps = (m + 3) / 4;
  const int64_t per = (steps + nw - 1) / nw;
  const int64_t s0 = gw * per, s1 = min<int64_t>(steps, s0 + per);
  const int colc = c0 + NB * cc, cold = d0 + NB * cc;
  // columns at or past l read as zero; a quad (pair) starting below l stays
  // inside the row's leading dimension (a multiple of 4 that is >= l)
  const bool okc = colc < l, okd = cold < l2;
  const int limc = l - colc, limd = l2 - cold;
  d4_t acc[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
  constexpr int G = 4;  // row steps per group; the next group is loaded during this one's MFMAs
  using Vec = typename GramVec<T, NB>::type;
  Vec qc[G], qd[G], nc[G], nd[G];
  // Every load is issued unconditionally at an in-bounds address (rows past
  // the wave's range re-read its last row, columns past l read column 0) and
  // the unused parts are zeroed by selects when the group is consumed: with
  // no load behind a branch the compiler can count the loads in flight and
  // wait only for the group it is about to use, so the next group's loads
  // overlap this group's products.
  const int64_t rlast = min<int64_t>(m, s1 * 4) - 1;
  const int colc_a = okc ? colc : 0, cold_a = okd ? cold : 0;
  auto load = [&](int64_t st0, Vec (&xc)[G], Vec (&xd)[G], bool (&ok)[G]) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int64_t r = (st0 + g) * 4 + kr;
      ok[g] = st0 + g < s1 && r < m;
      const int64_t ra = ok[g] ? r : rlast;
      xc[g].load(Y + ra * ldy + colc_a);
      if (!SAME) xd[g].load(Y2 + ra * ld2 + cold_a);
    }
  };
  auto mfmas = [&](const Vec (&xc)[G], const Vec (&xd)[G], const bool (&ok)[G]) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      double a[NB], b[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        a[q] = ok[g] && q < limc ? xc[g].get(q) : 0.0;
        b[q] = SAME ? a[q] : (ok[g] && q < limd ? xd[g].get(q) : 0.0);
      }
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = (SAME ? i : 0); j < NB; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };
  if (s0 < s1) {
    bool okq[G], okn[G];
    load(s0, qc, qd, okq);
    for (int64_t st = s0;; st += 2 * G) {
      load(st + G, nc, nd, okn);
      mfmas(qc, qd, okq);
      if (st + G >= s1) break;
      load(st + 2 * G, qc, qd, okq);
      mfmas(nc, nd, okn);
      if (st + 2 * G >= s1) break;
    }
  }
  // sum the waves in order through LDS: red[(i * NB + j) * 256 + v * 64 + lane]
  for (int w = 0; w < W; ++w) {
    if (wv == w)
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = (SAME ? i : 0); j < NB; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double *q = red + (i * NB + j) * 256 + v * 64 + lane;
            *q = (w == 0 ? 0.0 : *q) + acc[i][j][v];
          }
    __syncthreads();
  }
  // The tile out[c0 + ri][d0 + ci] (and its mirror out[d0 + ci][c0 + ri]
  // when mirror is set and c0 != d0) in output order, so consecutive threads
  // write consecutive entries. Entry (ri, ci) is column block b1 = ri % NB,
  // b2 = ci % NB, lane 16 ((ri / NB) % 4) + ci / NB, v = ri / (4 NB) of the
  // sums above; with SAME only blocks b1 <= b2 were formed, and (ri, ci) of
  // block b1 > b2 reads its transpose.
  auto raw = [](int ri, int ci) {
    int b1 = ri % NB, b2 = ci % NB;
    if (SAME && b2 < b1) {
      const int t = ri;
      ri = ci;
      ci = t;
      b1 = ri % NB;
      b2 = ci % NB;
    }
    const int rq = ri / NB;
    return (b1 * NB + b2) * 256 + (rq >> 2) * 64 + (rq & 3) * 16 + ci / NB;
  };
  const int rows_t = min(16 * NB, l - c0), cols_t = min(16 * NB, l2 - d0);
  for (int o = threadIdx.x; o < rows_t * cols_t; o += 64 * W) {
    const int ri = o / cols_t, ci = o - ri * cols_t;
    out[(int64_t)(c0 + ri) * out_ld + d0 + ci] = red[raw(ri, ci)];
  }
  if (mirror && c0 != d0)
    for (int o = threadIdx.x; o < rows_t * cols_t; o += 64 * W) {
      const int ci = o / rows_t, ri = o - ci * rows_t;
      out[(int64_t)(d0 + ci) * out_ld + c0 + ri] = red[raw(ri, ci)];
    }
}

template <typename T, bool SAME, int NB = 4>
__global__ __launch_bounds__(64 * gram_waves<SAME>()) void k_gram_mfma(int64_t m, int l, const T *Y, int64_t ldy,
                                                                     int l2, const T *Y2, int64_t ld2, int c0,
                                                                     int d0, double *part, int64_t pstride,
                                                                     int out_ld, int mirror) {
  gram_mfma_body<T, SAME, NB>(blockIdx.x, gridDim.x, m, l, Y, ldy, l2, Y2, ld2, c0, d0,
                              part + (int64_t)blockIdx.x * pstride, out_ld, mirror);
}

// Grams of many panels (width l <= 64) in one launch: panel b = blockIdx.y
// takes workgroups 0 .. nblk - 1 of the x dimension, and its workgroup
// partials go to part[(poff + workgroup) * l * l ...].
struct GramPanel {
  const void *y;
  int64_t rows;
  int64_t poff;
  int64_t nblk;
};

template <typename T, int NB>
__global__ __launch_bounds__(64 * gram_waves<true>()) void k_gram_mfma_batched(const GramPanel *pan, int l,
                                                                             int64_t ldy, double *part) {
  const GramPanel g = pan[blockIdx.y];
  if ((int64_t)blockIdx.x >= g.nblk) return;
  const T *Y = static_cast<const T *>(g.y);
  gram_mfma_body<T, true, NB>(blockIdx.x, g.nblk, g.rows, l, Y, ldy, l, Y, ldy, 0, 0,
                              part + (g.poff + (int64_t)blockIdx.x) * l * l, l, 1);
}

// k_gram_reduce for panel b = blockIdx.y: G + b * l * l = sum of its partials.
__global__ __launch_bounds__(kBlock) void k_gram_reduce_batched(const GramPanel *pan, int64_t ll, const double *part,
                                                                double *G) {
  const GramPanel g = pan[blockIdx.y];
  gram_reduce_body((int)g.nblk, ll, part + g.poff * ll, G + (int64_t)blockIdx.y * ll, blockIdx.x);
}

// Z[r, 0:lo] = Y[r, 0:li] M (li x lo, row-major) on 64-row tiles; the
// tile is read into LDS before any output is written, so Z may alias Y.
// C is the arithmetic type (FP32 for FP32 panels, FP64 otherwise). M is
// staged in LDS too when MLDS is set (li * lo entries of C).
template <typename TY, typename TZ, typename C, bool MLDS>
__global__ __launch_bounds__(kBlock) void k_apply(int64_t m, int li, int lo, const TY *Y, int64_t ldy, const double *M,
                                                 TZ *Z, int64_t ldz, int64_t zcs, double alpha, double beta) {
  extern __shared__ char smem[];
  C *s = reinterpret_cast<C *>(smem);  // li x (kTile + 1)
  C *sm = s + (size_t)li * (kTile + 1);  // li x lo (when MLDS)
  const int sld = kTile + 1;
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  const int64_t r0 = (int64_t)blockIdx.x * kTile;
  const int rows = (int)min<int64_t>(kTile, m - r0);
  load_tile<TY, C>(Y, ldy, r0, rows, li, s, sld);
  if (MLDS)
    for (int e = threadIdx.x; e < li * lo; e += kBlock) sm[e] = (C)M[e];
  __syncthreads();
  for (int c0 = 0; c0 < lo; c0 += kTile) {
    C acc[4][4] = {};
    for (int k = 0; k < li; ++k) {
      C a4[4], b4[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a4[q] = s[k * sld + tr * 4 + q];
        const int c = c0 + tc * 4 + q;
        b4[q] = c < lo ? (MLDS ? sm[k * lo + c] : (C)M[(int64_t)k * lo + c]) : C(0);
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] += a4[p] * b4[q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int r = tr * 4 + p;
      if (r < rows)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c = c0 + tc * 4 + q;
          if (c < lo) {
            const int64_t at = (r0 + r) * ldz + c * zcs;
            const double old = beta != 0.0 ? beta * ldd<TZ>(Z, at) : 0.0;
            std_<TZ>(Z, at, old + alpha * (double)acc[p][q]);
          }
        }
    }
  }
}

// Residual column norms: r = Y1 M1 + Y2 M2 on 64-row tiles (arithmetic in
// C), squared and summed over the tile's rows in FP64; per-workgroup column
// partials (workgroups stride over tiles). M1, M2 (w x k) are staged in LDS
// when MLDS is set.
template <typename T, typename C, bool MLDS>
__global__ __launch_bounds__(kBlock) void k_resid(int64_t m, int w, int w2, int k, const T *Y1, const T *Y2,
                                                 int64_t ld, const double *M1, const double *M2, double *part) {
  // Y1 has w columns, Y2 has w2 <= w; M1 is w x k, M2 is w2 x k (zero rows past w2 are implied)
  extern __shared__ char smem[];
  C *s1 = reinterpret_cast<C *>(smem);
  const int sld = kTile + 1;
  C *s2 = s1 + (size_t)w * sld;
  C *sm1 = s2 + (size_t)w * sld;
  C *sm2 = sm1 + (size_t)w * k;
  __shared__ double red[16][kTile + 4];
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  if (MLDS)
    for (int e = threadIdx.x; e < w * k; e += kBlock) {
      sm1[e] = (C)M1[e];
      sm2[e] = e < w2 * k ? (C)M2[e] : C(0);
    }
  double colsq[2][4] = {};
  for (int64_t r0 = (int64_t)blockIdx.x * kTile; r0 < m; r0 += (int64_t)gridDim.x * kTile) {
    const int rows = (int)min<int64_t>(kTile, m - r0);
    __syncthreads();
    load_tile<T, C>(Y1, ld, r0, rows, w, s1, sld);
    load_tile<T, C>(Y2, ld, r0, rows, w2, s2, sld);
    __syncthreads();
    for (int c0 = 0, pass = 0; c0 < k; c0 += kTile, ++pass) {
      C acc[4][4] = {};
      for (int j = 0; j < w; ++j) {
        C a4[4], b4[4], m1[4], m2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a4[q] = s1[j * sld + tr * 4 + q];
          b4[q] = j < w2 ? s2[j * sld + tr * 4 + q] : C(0);
          const int c = c0 + tc * 4 + q;
          m1[q] = c < k ? (MLDS ? sm1[j * k + c] : (C)M1[(int64_t)j * k + c]) : C(0);
          m2[q] = (c < k && j < w2) ? (MLDS ? sm2[j * k + c] : (C)M2[(int64_t)j * k + c]) : C(0);
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[p][q] += a4[p] * m1[q] + b4[p] * m2[q];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
        if (tr * 4 + p < rows)
#pragma unroll
          for (int q = 0; q < 4; ++q) colsq[pass & 1][q] += (double)acc[p][q] * (double)acc[p][q];
    }
  }
  for (int pass = 0; pass * kTile < k && pass < 2; ++pass) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) red[tr][tc * 4 + q] = colsq[pass][q];
    __syncthreads();
    if (threadIdx.x < kTile) {
      double t = 0.0;
      for (int i = 0; i < 16; ++i) t += red[i][threadIdx.x];
      const int c = pass * kTile + threadIdx.x;
      if (c < k) part[(int64_t)blockIdx.x * k + c] = t;
    }
  }
}

// Column sums of squares: one workgroup per 64-row slab group, FP64.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_colsq_part(int64_t m, int l, const T *Y, int64_t ldy, double *part) {
  double acc = 0.0;
  const int c = threadIdx.x;
  if (c < l)
    for (int64_t r = blockIdx.x; r < m; r += gridDim.x) {
      const double v = ldd<T>(Y, r * ldy + c);
      acc += v * v;
    }
  if (c < l) part[(int64_t)blockIdx.x * l + c] = acc;
}

template <typename T>
__global__ void k_colscale(int64_t m, int l, T *Y, int64_t ldy, const double *scale) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * l; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / l;
    const int c = (int)(e % l);
    std_<T>(Y, r * ldy + c, ldd<T>(Y, r * ldy + c) * scale[c]);
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Standard normals by Box-Muller on two counter-based uniforms; element (i, j)
// of an m x l panel uses counter j*m + i (column-major, the fill order of
// DMatrix::from_fn at src/randomized/mod.rs:310-312).
template <typename T>
__device__ __forceinline__ void gauss_body(int64_t m, int l, uint64_t seed, T *Y, int64_t ldy) {
  const uint64_t phi = 0x9E3779B97F4A7C15ULL;
  const uint64_t key = (seed * phi) ^ 0x4F4D454741303031ULL;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * l; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % m;
    const int64_t j = e / m;
    const uint64_t cnt = (uint64_t)e;
    const uint64_t a = mix64(key + (2 * cnt + 1) * phi);
    const uint64_t b = mix64(key + (2 * cnt + 2) * phi);
    const double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(b >> 11) * 0x1.0p-53;
    const double z = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
    std_<T>(Y, i * ldy + j, z);
  }
}

template <typename T>
__global__ void k_gauss(int64_t m, int l, uint64_t seed, T *Y, int64_t ldy) {
  gauss_body<T>(m, l, seed, Y, ldy);
}

// panel b = blockIdx.y of a batch: rows m[b] at Y[b]
template <typename T>
__global__ void k_gauss_jobs(const int64_t *m, int l, uint64_t seed, T *const *Y, int64_t ldy) {
  gauss_body<T>(m[blockIdx.y], l, seed, Y[blockIdx.y], ldy);
}

template <typename S, typename D>
__global__ void k_pconvert(int64_t m, int l, const S *X, int64_t ldx, D *Y, int64_t ldy) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * l; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / l;
    const int c = (int)(e % l);
    std_<D>(Y, r * ldy + c, ldd<S>(X, r * ldx + c));
  }
}

template <typename T>
__global__ void k_pfill(int64_t m, int l, double v, T *Y, int64_t ldy) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * l; e += (int64_t)gridDim.x * blockDim.x)
    std_<T>(Y, (e / l) * ldy + (e % l), v);
}

int grid_of(int64_t n, int per = kBlock, int cap = 2048) {
  int64_t g = (n + per - 1) / per;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// one element per thread up to 1024 workgroups: the Lanczos vectors are short
// (tens of thousands), and a few elements per thread already leave the kernel
// waiting on one load chain per element
int fused_grid(int64_t n) { return grid_of(n, kBlock, 1024); }

}  // namespace

void fused_axpy_dot(Ctx *c, int64_t n, const Axpy *ups, int nups, const double *u, const double *v,
                    int out_slot, double *copy_dst, const double *copy_src) {
  FusedArgs a;
  std::memset(&a, 0, sizeof a);
  a.n = n;
  a.nups = nups;
  c->slot_parts.ensure((size_t)S_NSLOT * kSlotParts * sizeof(double));
  for (int i = 0; i < nups; ++i) {
    a.up[i].y = ups[i].y;
    a.up[i].x = ups[i].x;
    a.up[i].imm = ups[i].imm;
    a.up[i].slot = ups[i].slot;
    a.up[i].nb = ups[i].slot >= 0 ? c->slot_nb[ups[i].slot] : 0;
    a.up[i].sign = ups[i].sign;
  }
  a.u = u;
  a.v = v;
  a.out_slot = out_slot;
  a.copy_dst = copy_dst;
  a.copy_src = copy_src;
  a.parts = c->slot_parts.as<double>();
  const int g = fused_grid(n);
  if (out_slot >= 0) {
    if (out_slot >= S_NSLOT) fail(SVDH_ERR_ARG, "scalar slot %d out of range", out_slot);
    c->slot_nb[out_slot] = g;
  }
  KTimer t(c, KC_BLAS1, (double)n * 8.0 * (3 * nups + 2 + (copy_dst ? 2 : 0)));
  hipLaunchKernelGGL(k_fused_axpy_dot, g, kBlock, 0, c->stream, a);
  HIP_CHECK(hipGetLastError());
}

bool fused_chain(Ctx *c, int64_t n, const FusedOp *ops, int nops) {
  if (nops <= 0) return false;
  if (nops > kChainMax || n > kChainMaxN) {
    for (int o = 0; o < nops; ++o)
      fused_axpy_dot(c, n, ops[o].nups ? &ops[o].up : nullptr, ops[o].nups, ops[o].u, ops[o].v, ops[o].out_slot,
                     ops[o].copy_dst, ops[o].copy_src);
    return false;
  }
  c->slot_parts.ensure((size_t)S_NSLOT * kSlotParts * sizeof(double));
  ChainArgs a;
  std::memset(&a, 0, sizeof a);
  a.n = n;
  a.nops = nops;
  bool produced[S_NSLOT] = {false};
  double bytes = 0.0;
  for (int o = 0; o < nops; ++o) {
    const FusedOp &f = ops[o];
    ChainOp &d = a.ops[o];
    d.nups = f.nups;
    if (f.nups) {
      d.up.y = f.up.y;
      d.up.x = f.up.x;
      d.up.imm = f.up.imm;
      d.up.slot = f.up.slot;
      d.up.sign = f.up.sign;
      d.in_chain = f.up.slot >= 0 && produced[f.up.slot];
      d.up.nb = f.up.slot >= 0 ? c->slot_nb[f.up.slot] : 0;
    }
    d.u = f.u;
    d.v = f.v;
    d.out_slot = f.out_slot;
    d.copy_dst = f.copy_dst;
    d.copy_src = f.copy_src;
    d.sqrt_slot = f.out_slot >= 0 ? f.sqrt_slot : -1;
    if (f.out_slot >= 0) {
      if (f.out_slot >= S_NSLOT) fail(SVDH_ERR_ARG, "scalar slot %d out of range", f.out_slot);
      produced[f.out_slot] = true;
    }
    if (d.sqrt_slot >= 0) {
      if (d.sqrt_slot >= S_NSLOT) fail(SVDH_ERR_ARG, "scalar slot %d out of range", d.sqrt_slot);
      produced[d.sqrt_slot] = true;
    }
    bytes += (double)n * 8.0 * (3 * f.nups + 2 + (f.copy_dst ? 2 : 0));
  }
  for (int s = 0; s < S_NSLOT; ++s)
    if (produced[s]) c->slot_nb[s] = 1;
  a.parts = c->slot_parts.as<double>();
  a.scal = c->scal.as<double>();
  a.host = c->dh_scal;
  KTimer t(c, KC_BLAS1, bytes);
  hipLaunchKernelGGL(k_fused_chain, 1, kChainBlock, 0, c->stream, a);
  HIP_CHECK(hipGetLastError());
  return true;
}

void slots_finish(Ctx *c, int n) { slots_finish_range(c, 0, n, -1, -1); }

void slots_finish_range(Ctx *c, int first, int n, int sqrt_from, int sqrt_to) {
  if (n <= 0) return;
  if (first < 0 || first + n > S_NSLOT) fail(SVDH_ERR_ARG, "scalar slots %d..%d out of range", first, first + n);
  if (sqrt_from >= 0 && (sqrt_from < first || sqrt_from >= first + n || sqrt_to < 0 || sqrt_to >= S_NSLOT ||
                         (sqrt_to >= first && sqrt_to < first + n)))
    fail(SVDH_ERR_ARG, "square-root slot %d -> %d outside the finished range", sqrt_from, sqrt_to);
  c->slot_parts.ensure((size_t)S_NSLOT * kSlotParts * sizeof(double));
  SlotCounts cnt;
  for (int s = 0; s < S_NSLOT; ++s) cnt.nb[s] = s >= first && s < first + n ? c->slot_nb[s] : 0;
  cnt.first = first;
  cnt.sqrt_from = sqrt_from;
  cnt.sqrt_to = sqrt_to;
  if (sqrt_from >= 0) c->slot_nb[sqrt_to] = 1;
  hipLaunchKernelGGL(k_slot_finish, n, 64, 0, c->stream, c->slot_parts.as<double>(), cnt, c->scal.as<double>(),
                     c->dh_scal);
  HIP_CHECK(hipGetLastError());
}

void abs_sums_to_slots(Ctx *c, int64_t k, const double *a, const double *b, int slot_a, int slot_b) {
  if (slot_a < 0 || slot_a >= S_NSLOT || slot_b < 0 || slot_b >= S_NSLOT || slot_a == slot_b)
    fail(SVDH_ERR_ARG, "scalar slots %d, %d out of range", slot_a, slot_b);
  c->slot_parts.ensure((size_t)S_NSLOT * kSlotParts * sizeof(double));
  double *parts = c->slot_parts.as<double>();
  hipLaunchKernelGGL(k_abs_sums, 1, kBlock, 0, c->stream, k, a, b, parts + (size_t)slot_a * kSlotParts,
                     parts + (size_t)slot_b * kSlotParts);
  HIP_CHECK(hipGetLastError());
  c->slot_nb[slot_a] = 1;
  c->slot_nb[slot_b] = 1;
}

void scale_step(Ctx *c, int64_t n, int rnm_slot, double a, const double *x, double *y, const double *zin,
                double *zout, double *copy_dst, const double *copy_src) {
  const double *rpart = nullptr;
  if (rnm_slot >= 0) {
    if (rnm_slot >= S_NSLOT || c->slot_nb[rnm_slot] != 1)
      fail(SVDH_ERR_ARG, "slot %d does not hold a single published value", rnm_slot);
    rpart = c->slot_parts.as<double>() + (size_t)rnm_slot * kSlotParts;
  }
  KTimer t(c, KC_BLAS1, (double)n * (32.0 + (copy_dst ? 16.0 : 0.0)));
  hipLaunchKernelGGL(k_scale_step, grid_of(n), kBlock, 0, c->stream, n, rpart, a, x, y, zin, zout, copy_dst,
                     copy_src);
  HIP_CHECK(hipGetLastError());
}

void vec_scale2(Ctx *c, int64_t n, double a, const double *x, double *y, double *z, double *copy_dst,
                const double *copy_src) {
  KTimer t(c, KC_BLAS1, (double)n * (32.0 + (copy_dst ? 16.0 : 0.0)));
  hipLaunchKernelGGL(k_scale2, grid_of(n), kBlock, 0, c->stream, n, a, x, y, z, copy_dst, copy_src);
  HIP_CHECK(hipGetLastError());
}

void convert_f64_f32(Ctx *c, int64_t n, const double *x, float *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_f64_to_f32, grid_of(n), kBlock, 0, c->stream, n, x, y);
  HIP_CHECK(hipGetLastError());
}

void vec_copy(Ctx *c, int64_t n, const double *x, double *y) {
  HIP_CHECK(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}

void basis_gemv_t(Ctx *c, int64_t n, const double *Q, int64_t ldq, int64_t k, const double *u, const double *v,
                  double *cu, double *cv) {
  if (k <= 0) return;
  const int ntiles = (int)((n + kGemvRows - 1) / kGemvRows);
  const int ngroups = (int)((k + kGemvCols - 1) / kGemvCols);
  // workspace 12, sized for at least 1,024 columns so that it rarely grows
  double *part = static_cast<double *>(
      c->workspace(12, (size_t)2 * std::max<int64_t>(k, 1024) * ntiles * sizeof(double)));
  KTimer t(c, KC_GEMV, (double)n * k * 8.0 + 16.0 * n);
  hipLaunchKernelGGL(k_gemv_t, dim3(ntiles, ngroups), kBlock, 0, c->stream, n, Q, ldq, k, u, v, part,
                     part + k * ntiles);
  hipLaunchKernelGGL(k_gemv_t_sum, (int)((k + 255) / 256), 256, 0, c->stream, k, ntiles, part, part + k * ntiles, cu,
                     cv);
  HIP_CHECK(hipGetLastError());
}

void basis_gemv_n(Ctx *c, int64_t n, const double *Q, int64_t ldq, int64_t k, const double *c1, double *y1,
                  const double *c2, double *y2) {
  if (k <= 0) return;
  KTimer t(c, KC_GEMV, (double)n * k * 8.0 + 32.0 * n);
  hipLaunchKernelGGL(k_gemv_n, (int)((n + 63) / 64), kBlock, 0, c->stream, n, Q, ldq, k, c1, y1, c2, y2);
  HIP_CHECK(hipGetLastError());
}

void basis_combine(Ctx *c, int64_t n, const double *Q, int64_t ldq, int64_t k, const double *W, int64_t d,
                   double *C, int64_t ldc) {
  if (n <= 0 || d <= 0) return;
  KTimer t(c, KC_DENSE, (double)n * k * 8.0 * ((d + 127) / 128) + (double)n * d * 8.0);
  const int64_t blocks16 = (n + 15) / 16;
  hipLaunchKernelGGL(k_combine_mfma, (int)((blocks16 + 3) / 4), 256, 0, c->stream, n, Q, ldq, k, W, d, C, ldc);
  HIP_CHECK(hipGetLastError());
}

// Panels of width <= 32 take the two-block variant (3 matrix-core products
// per 4-row step instead of 10); wider ones 64-column groups of four blocks.
template <typename T>
static void launch_gram(Ctx *c, bool same, int nb, int64_t m, int l, const T *Y, int64_t ldy, int c0, int d0,
                        double *part) {
  if (same && l <= 32)
    hipLaunchKernelGGL((k_gram_mfma<T, true, 2>), nb, 64 * gram_waves<true>(), 0, c->stream, m, l, Y, ldy, l, Y,
                       ldy, c0, d0, part, (int64_t)l * l, l, 1);
  else if (same)
    hipLaunchKernelGGL((k_gram_mfma<T, true>), nb, 64 * gram_waves<true>(), 0, c->stream, m, l, Y, ldy, l, Y, ldy,
                       c0, d0, part, (int64_t)l * l, l, 1);
  else
    hipLaunchKernelGGL((k_gram_mfma<T, false>), nb, 64 * gram_waves<false>(), 0, c->stream, m, l, Y, ldy, l, Y, ldy,
                       c0, d0, part, (int64_t)l * l, l, 1);
  HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------------
// FP32 panel products on the matrix cores:
//   Z[r, 0:lo] = beta Z[r, 0:lo] + sum_p Y_p[r, 0:li_p] (alpha_p M_p)
// for up to two terms, lo, li_p <= 64. v_mfma_f32_16x16x4 takes A[l % 16][l / 16]
// and B[l / 16][l % 16] per lane and returns D[4 (l / 16) + v][l % 16] (lane
// layout checked on gfx950). A workgroup (eight waves, one per CU) stages
// alpha_p M_p (FP32, zero padded to 64 x 64) in LDS once; each wave then
// strides over 16-row blocks, with the Y_p rows of a block in registers (the
// next block's loaded during this one's products) and the M_p operands read
// from LDS. A wave reads all rows of a block before writing any, so Z may be
// Y_0 when beta = 0.
// ------------------------------------------------------------------------
typedef float f4_t __attribute__((ext_vector_type(4)));
#ifndef SVDH_PANEL_BLOCK
#define SVDH_PANEL_BLOCK 512
#endif
constexpr int kPanelBlock = SVDH_PANEL_BLOCK;  // threads per workgroup of the panel products
constexpr int kMsStride = 68;  // LDS row of the staged factors (64 + 4: rows 4 apart start 16 banks apart)
// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format;
// with stride 0 an access at byte offset >= the resource's size is out of
// range: a load returns 0, a store is dropped).
constexpr int kBufferWord3 = 0x00020000;
constexpr int kBufferDrop = 0x7ffffff0;  // a byte offset past every resource range used here
constexpr int64_t kBufferMaxLd = 1 << 24;  // leading dimensions whose 16-row blocks stay below kBufferDrop

struct PanelTerm {
  const float *y;
  int64_t ld;
  int li;
  const double *m;  // li x lo, row-major, FP64 device
  double alpha;
};

// One wave per 16-row block and all 64 output columns. The contraction index
// k is taken in groups of 16: in step s of group g, lane l supplies
// Y[row l % 16][16 g + 4 (l / 16) + s] (component s of one 16-byte load) and
// alpha M[16 g + 4 (l / 16) + s][16 q + l % 16] (the four q of one 16-byte LDS
// read), so a wave reads whole panel rows and each lane's A values come from
// one vector load. A wave reads all of its rows before writing any,
// so Z may be Y_0 when beta = 0.
template <int NP>
__device__ __forceinline__ void panel_rows_load(const PanelTerm &t0, const PanelTerm &t1, int64_t rows, int64_t row,
                                                int kq, f4_t (&ya)[NP][4]) {
  // whole 16-byte loads at clamped addresses, always in bounds; the parts
  // outside the panel are zeroed by panel_rows_mask when the block is used,
  // so neither a branch nor a select waits on the loads here
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const PanelTerm &t = p == 0 ? t0 : t1;
    const int64_t rr = row < rows ? row : rows - 1;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int kc = 16 * g + 4 * kq;
      const int kk = kc < t.li ? kc : 0;
      const float4 w = *reinterpret_cast<const float4 *>(t.y + rr * t.ld + kk);
      ya[p][g] = f4_t{w.x, w.y, w.z, w.w};
    }
  }
}

template <int NP>
__device__ __forceinline__ void panel_rows_mask(const PanelTerm &t0, const PanelTerm &t1, int64_t rows, int64_t row,
                                                int kq, f4_t (&ya)[NP][4]) {
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const PanelTerm &t = p == 0 ? t0 : t1;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int kc = 16 * g + 4 * kq;
      const bool ok = row < rows && kc < t.li;
      const f4_t w = ya[p][g];
      ya[p][g] = f4_t{ok ? w[0] : 0.f, ok && kc + 1 < t.li ? w[1] : 0.f, ok && kc + 2 < t.li ? w[2] : 0.f,
                      ok && kc + 3 < t.li ? w[3] : 0.f};
    }
  }
}

// RESID: the product is not stored; its column sums of squares (FP64) are
// accumulated instead and each workgroup writes its lo partials to
// part[blockIdx.x * lo ...] (the adaptive residual check).
// GRAM: the FP64 Gram Z^T Z of the stored (FP32) output is accumulated on the
// FP64 matrix cores from the registers that hold it: for row v of each lane's
// four, v_mfma_f64_16x16x4 with A[i][k] = Z[r0 + 4 k + v][16 bi + i] and
// B[k][j] = Z[r0 + 4 k + v][16 bj + j] (lane (k, i) = (l / 16, l % 16) holds
// exactly those) adds the block (bi, bj) of the 16 rows, for bi <= bj. The
// waves are summed in LDS in wave order and each workgroup writes its
// lo x lo partial (both triangles) to part[blockIdx.x * lo * lo ...].
// Workgroup bid of nbk; part is this workgroup's partial (RESID: lo
// entries, GRAM: lo x lo).
// TRI: every M_p is upper triangular (the inverse Cholesky factors and their
// products), so the 16 x 16 factor blocks below the diagonal, which hold
// exact zeros, are not multiplied: 10 of the 16 block products per term.
template <int NP, bool RESID, bool GRAM, bool TRI = false>
__device__ __forceinline__ void panel_mfma_body(int64_t bid, int64_t nbk, int64_t rows, int lo, const PanelTerm &t0,
                                                const PanelTerm &t1, float *z, int64_t ldz, const float *zin,
                                                int64_t ldzin, float beta, double *part) {
  // alpha_p M_p in LDS, row k at k * kMsStride (zero padded to 64 x 64): for
  // contraction index k a lane reads columns 16 q + rl, q = 0..3. The B
  // operands are read from LDS at each step, which leaves registers for more
  // than one wave per SIMD.
  __shared__ float ms[NP][64 * kMsStride];
  // every load of the factors in flight before the first LDS write (a
  // load-then-store loop waits out one memory latency per element)
  constexpr int S = 64 * 64 / kPanelBlock;
  for (int p = 0; p < NP; ++p) {
    const PanelTerm &t = p == 0 ? t0 : t1;
    double mv[S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int e = threadIdx.x + i * kPanelBlock, k = e >> 6, col = e & 63;
      const bool ok = k < t.li && col < lo;
      mv[i] = t.li > 0 ? t.m[ok ? (int64_t)k * lo + col : 0] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int e = threadIdx.x + i * kPanelBlock, k = e >> 6, col = e & 63;
      ms[p][k * kMsStride + col] = (k < t.li && col < lo) ? (float)(t.alpha * mv[i]) : 0.f;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int kq = lane >> 4, rl = lane & 15;
  const int64_t nblk = (rows + 15) / 16;
  // the wave index in a scalar register: the block loop, its bounds and the
  // stores' buffer resource below are wave-uniform
  const int64_t wave = bid * (kPanelBlock / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nwaves = nbk * (kPanelBlock / 64);
  f4_t ya[NP][4], yn[NP][4];
  double csq[4] = {0.0, 0.0, 0.0, 0.0};  // RESID: columns 16 q + rl
  d4_t gacc[4][4];                       // GRAM: blocks (bi, bj), bi <= bj
  if (GRAM)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = i; j < 4; ++j) gacc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
  // the beta term's rows of a block (lane: rows 4 kq + v, column 16 q + rl),
  // loaded as the block starts: they are first used by its stores, after the
  // products
  float zold[4][4];
  auto zin_load = [&](int64_t r0, float (&dst)[4][4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gr = r0 + 4 * kq + v;
        const int c = 16 * q + rl;
        dst[q][v] = zin[(gr < rows ? gr : rows - 1) * ldzin + (c < lo ? c : 0)];
      }
  };
  const bool use_zin = !RESID && beta != 0.f;
  if (wave < nblk) {
    panel_rows_load<NP>(t0, t1, rows, wave * 16 + rl, kq, ya);
    panel_rows_mask<NP>(t0, t1, rows, wave * 16 + rl, kq, ya);
  }
  for (int64_t b = wave; b < nblk; b += nwaves) {
    // an offset of zero the compiler cannot see through: the B reads below
    // stay in the loop instead of being hoisted into registers
    int lds0;
    asm volatile("v_mov_b32 %0, 0" : "=v"(lds0));
    const int64_t r0 = b * 16;
    const bool more = b + nwaves < nblk;
    // this block's beta rows, then the next block in flight (clamped rows
    // past the end: in bounds, unused); in this order the wait for the beta
    // rows leaves the next block's loads outstanding
    if (use_zin) zin_load(r0, zold);
    panel_rows_load<NP>(t0, t1, rows, (b + nwaves) * 16 + rl, kq, yn);
    // the loads stay here, ahead of the products (the scheduler would
    // otherwise sink them to their first use, after the products)
    __builtin_amdgcn_sched_barrier(0);
    f4_t acc[4] = {f4_t{0.f, 0.f, 0.f, 0.f}, f4_t{0.f, 0.f, 0.f, 0.f}, f4_t{0.f, 0.f, 0.f, 0.f},
                   f4_t{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const float *mr = &ms[p][(16 * g + 4 * kq + st) * kMsStride + rl + lds0];
          if (TRI) {
            // column tiles q >= g only (rows 16 g .. of M are zero left of column 16 g)
#pragma unroll
            for (int q = g; q < 4; ++q)
              acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ya[p][g][st], mr[16 * q], acc[q], 0, 0, 0);
          } else {
            const float4 b = make_float4(mr[0], mr[16], mr[32], mr[48]);
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ya[p][g][st], b.x, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ya[p][g][st], b.y, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(ya[p][g][st], b.z, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(ya[p][g][st], b.w, acc[3], 0, 0, 0);
          }
        }
    if (RESID) {
      // rows past the end were loaded as zeros, so they add nothing
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) csq[q] += (double)acc[q][v] * (double)acc[q][v];
    } else {
      // The block's stores go through a buffer resource over its rows of Z:
      // an element outside the panel (a row past the end, a column from lo)
      // gets an offset past the resource's range, and the hardware drops the
      // store. No store sits behind a branch: the memory counter counts loads
      // and stores in issue order, so one store that may or may not have been
      // issued makes the wait for the next block's loads wait for every store
      // of this block too.
      const int64_t nrow = rows - r0 < 16 ? rows - r0 : 16;
      const __amdgpu_buffer_rsrc_t zrs =
          __builtin_amdgcn_make_buffer_rsrc(z + r0 * ldz, (short)0, (int)(nrow * ldz * 4), kBufferWord3);
      float zv[4][4];  // the stored values; zero outside the panel (GRAM)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t gr = r0 + 4 * kq + v;
          const int c = 16 * q + rl;
          const float val = (use_zin ? beta * zold[q][v] : 0.f) + acc[q][v];
          const bool ok = gr < rows && c < lo;
          const int off = ok ? (int)(((4 * kq + v) * ldz + c) * 4) : kBufferDrop;
          // (the builtin's data operand is a 32-bit integer: the float's bits)
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, val), zrs, off, 0, 0);
          zv[q][v] = ok ? val : 0.f;
        }
      if (GRAM)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = i; j < 4; ++j)
              gacc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)zv[i][v], (double)zv[j][v], gacc[i][j], 0, 0,
                                                                0);
    }
    if (more) {
      panel_rows_mask<NP>(t0, t1, rows, (b + nwaves) * 16 + rl, kq, yn);
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int g = 0; g < 4; ++g) ya[p][g] = yn[p][g];
    }
  }
  if (GRAM) {
    // The waves' accumulators summed in a fixed tree through two LDS slots
    // (block (i, j), i <= j, at (t * 4 + v) * 64 + lane, t = i (7 - i) / 2 + j):
    // at each level h = W/2 .. 1, waves h + s and h + s + 1 store into the
    // slots and waves s and s + 1 add them, for s = 0, 2, .. < h; wave 0 then
    // stores the total. Eight waves: ((w0 + w4) + (w2 + w6)) + ((w1 + w5) +
    // (w3 + w7)), four store/add phases instead of eight serial ones.
    constexpr int W = kPanelBlock / 64;
    static_assert((W & (W - 1)) == 0, "a power-of-two number of waves");
    __shared__ double gsl[2][10 * 4 * 64];
    const int wv = threadIdx.x >> 6;
    auto put = [&](double *dst) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) dst[((i * (7 - i) / 2 + j) * 4 + v) * 64 + lane] = gacc[i][j][v];
    };
    auto add = [&](const double *src) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) gacc[i][j][v] += src[((i * (7 - i) / 2 + j) * 4 + v) * 64 + lane];
    };
    for (int h = W / 2; h >= 1; h /= 2)
      for (int s0 = 0; s0 < h; s0 += 2) {
        if (wv >= h + s0 && wv < h + s0 + 2 && wv < 2 * h) put(gsl[wv - h - s0]);
        __syncthreads();
        if (wv >= s0 && wv < s0 + 2 && wv + h < 2 * h && wv < h) add(gsl[wv - s0]);
        __syncthreads();
      }
    if (wv == 0) put(gsl[0]);
    __syncthreads();
    // lo x lo in output order (consecutive threads, consecutive entries):
    // lane ln of block (bi, bj) holds D[ln / 16 + 4 v][ln % 16], so entry
    // (gi, gj) is at (t(bi, bj) * 4 + v) * 64 + 16 ((gi % 16) % 4) + gj % 16
    // with v = (gi % 16) / 4; only blocks bi <= bj were formed, and an entry
    // of block bi > bj reads its transpose
    for (int o = threadIdx.x; o < lo * lo; o += kPanelBlock) {
      int gi = o / lo, gj = o - gi * lo;
      if ((gj >> 4) < (gi >> 4)) {
        const int t = gi;
        gi = gj;
        gj = t;
      }
      const int r = gi & 15, bi = gi >> 4, bj = gj >> 4;
      part[o] = gsl[0][((bi * (7 - bi) / 2 + bj) * 4 + (r >> 2)) * 64 + 16 * (r & 3) + (gj & 15)];
    }
  }
  if (RESID) {
    // fixed-order sum over the four lane groups of each wave, then the waves
    __shared__ double red[kPanelBlock / 64][4][64];
#pragma unroll
    for (int q = 0; q < 4; ++q) red[threadIdx.x >> 6][kq][16 * q + rl] = csq[q];
    __syncthreads();
    if (threadIdx.x < 64 && (int)threadIdx.x < lo) {
      double t = 0.0;
      for (int w = 0; w < kPanelBlock / 64; ++w)
        for (int g = 0; g < 4; ++g) t += red[w][g][threadIdx.x];
      part[threadIdx.x] = t;
    }
  }
}

template <int NP, bool RESID, bool GRAM = false, bool TRI = false>
__global__ __launch_bounds__(kPanelBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_panel_mfma(int64_t rows, int lo, PanelTerm t0, PanelTerm t1, float *z, int64_t ldz, const float *zin,
                  int64_t ldzin, float beta, double *part) {
  double *pw = part ? part + (int64_t)blockIdx.x * (GRAM ? (int64_t)lo * lo : (int64_t)lo) : nullptr;
  panel_mfma_body<NP, RESID, GRAM, TRI>(blockIdx.x, gridDim.x, rows, lo, t0, t1, z, ldz, zin, ldzin, beta, pw);
}

// Z_b = Y_b M_b for many panels in one launch (one term, beta 0; Z_b may be
// Y_b): job b = blockIdx.y takes workgroups 0 .. nblk - 1 of the x dimension.
// Each row block is computed by one wave as in k_panel_mfma, so the result
// does not depend on the grid.
// GRAM: the job's workgroup partials of Z^T Z go to part (nblk x lo x lo) and
// k_gram_reduce_jobs sums them into gram, as k_panel_mfma<1, false, true> and
// k_gram_reduce do for one panel.
struct PanelJob {
  PanelTerm t;
  float *z;
  int64_t ldz;
  int64_t rows;
  int64_t nblk;
  int lo;
  double *part;
  double *gram;
};

template <bool GRAM, bool TRI = false>
__global__ __launch_bounds__(kPanelBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_panel_mfma_batched(const PanelJob *jobs) {
  const PanelJob j = jobs[blockIdx.y];
  if ((int64_t)blockIdx.x >= j.nblk) return;
  double *pw = GRAM ? j.part + (int64_t)blockIdx.x * j.lo * j.lo : nullptr;
  panel_mfma_body<1, false, GRAM, TRI>(blockIdx.x, j.nblk, j.rows, j.lo, j.t, j.t, j.z, j.ldz, j.z, j.ldz, 0.f, pw);
}

// k_gram_reduce for job b = blockIdx.y (grid x: reduce_grid of the widest job).
__global__ __launch_bounds__(kBlock) void k_gram_reduce_jobs(const PanelJob *jobs) {
  const PanelJob j = jobs[blockIdx.y];
  const int64_t ll = (int64_t)j.lo * j.lo;
  if ((int64_t)blockIdx.x * 16 >= ll) return;  // the whole workgroup leaves together
  gram_reduce_body((int)j.nblk, ll, j.part, j.gram, blockIdx.x);
}

bool upper_triangular(const double *M, int64_t rows, int64_t cols) {
  for (int64_t k = 1; k < rows; ++k)
    for (int64_t j = 0; j < std::min(k, cols); ++j)
      if (M[k * cols + j] != 0.0) return false;
  return true;
}

// Workgroups per CU of the panel kernels (SVDH_PANEL_WG, default 1): the
// staged factors are read once per workgroup; a workgroup of eight waves
// gives two waves per SIMD (the register budget of the kernels), so one
// wave's loads overlap the other's products.
static int panel_wg_per_cu() {
  static const int per_cu = [] {
    const char *e = std::getenv("SVDH_PANEL_WG");
    return e ? std::max(1, std::atoi(e)) : 1;
  }();
  return per_cu;
}

// gram_off >= 0: the Gram Z^T Z of the output (lo x lo, FP64) is accumulated by
// the same kernel and summed into the pinned landing area c->h_small +
// gram_off, after which gram_ev (c->small_ev when null) is recorded.
static void panel_mfma_impl(Ctx *c, int64_t rows, int64_t lo, int nterms, const float *const *Y, const int64_t *ldy,
                            const int64_t *li, const double *const *M_dev, const double *alpha, float *Z,
                            int64_t ldz, double beta, const float *Zin, int64_t ldzin, int64_t gram_off,
                            hipEvent_t gram_ev, bool tri) {
  if (!Zin) {
    Zin = Z;
    ldzin = ldz;
  }
  if (lo > 64 || nterms < 1 || nterms > 2) fail(SVDH_ERR_ARG, "panel_mfma: lo <= 64 and one or two terms");
  const bool gram = gram_off >= 0;
  if (gram && (size_t)gram_off + (size_t)lo * lo > Ctx::kSmall) fail(SVDH_ERR_ARG, "Gram landing area too small");
  PanelTerm t[2] = {};
  double bytes = (double)rows * lo * 4.0 * (beta != 0.0 ? 2.0 : 1.0);
  for (int p = 0; p < nterms; ++p) {
    if (li[p] > 64) fail(SVDH_ERR_ARG, "panel_mfma: li <= 64");
    t[p] = PanelTerm{Y[p], ldy[p], (int)li[p], M_dev[p], alpha[p]};
    bytes += (double)rows * li[p] * 4.0;
  }
  for (int p = 0; p < nterms; ++p)
    if (ldy[p] % 4) fail(SVDH_ERR_ARG, "panel_mfma: leading dimensions must be multiples of 4");
  if (ldz >= kBufferMaxLd) fail(SVDH_ERR_ARG, "panel_mfma: output leading dimension too large");
  const int64_t nblk = (rows + 15) / 16;
  constexpr int W = kPanelBlock / 64;
  const int g =
      (int)std::max<int64_t>(1, std::min<int64_t>((nblk + W - 1) / W, (int64_t)panel_wg_per_cu() * c->ncu));
  double *part = nullptr;
  if (gram) {
    c->red_part.ensure(std::max<size_t>((size_t)g * lo * lo * sizeof(double), c->red_part.bytes));
    part = c->red_part.as<double>();
  }
  {
    KTimer tm(c, KC_DENSE, bytes);
    const int sel = (nterms == 2 ? 4 : 0) + (gram ? 2 : 0) + (tri ? 1 : 0);
#define SVDH_PANEL_LAUNCH(NP, GR, TR)                                                                             \
  hipLaunchKernelGGL((k_panel_mfma<NP, false, GR, TR>), g, kPanelBlock, 0, c->stream, rows, (int)lo, t[0],       \
                     t[NP - 1], Z, ldz, Zin, ldzin, (float)beta, GR ? part : nullptr)
    switch (sel) {
      case 0: SVDH_PANEL_LAUNCH(1, false, false); break;
      case 1: SVDH_PANEL_LAUNCH(1, false, true); break;
      case 2: SVDH_PANEL_LAUNCH(1, true, false); break;
      case 3: SVDH_PANEL_LAUNCH(1, true, true); break;
      case 4: SVDH_PANEL_LAUNCH(2, false, false); break;
      case 5: SVDH_PANEL_LAUNCH(2, false, true); break;
      case 6: SVDH_PANEL_LAUNCH(2, true, false); break;
      default: SVDH_PANEL_LAUNCH(2, true, true); break;
    }
#undef SVDH_PANEL_LAUNCH
    HIP_CHECK(hipGetLastError());
    if (gram) {
      const int64_t ll = lo * lo;
      hipLaunchKernelGGL(k_gram_reduce, reduce_grid(ll), kBlock, 0, c->stream, g, ll, (const double *)part,
                         c->d_small + gram_off);
      HIP_CHECK(hipGetLastError());
    }
  }
  if (gram) HIP_CHECK(hipEventRecord(gram_ev ? gram_ev : c->small_ev, c->stream));
}

void panel_mfma(Ctx *c, int64_t rows, int64_t lo, int nterms, const float *const *Y, const int64_t *ldy,
                const int64_t *li, const double *const *M_dev, const double *alpha, float *Z, int64_t ldz,
                double beta, const float *Zin, int64_t ldzin, bool tri) {
  if (rows == 0 || lo == 0) return;
  panel_mfma_impl(c, rows, lo, nterms, Y, ldy, li, M_dev, alpha, Z, ldz, beta, Zin, ldzin, -1, nullptr, tri);
}

void panel_mfma_gram(Ctx *c, int64_t rows, int64_t lo, int nterms, const float *const *Y, const int64_t *ldy,
                     const int64_t *li, const double *const *M_dev, const double *alpha, float *Z, int64_t ldz,
                     double beta, const float *Zin, int64_t ldzin, size_t gram_off, hipEvent_t gram_ev, bool tri) {
  if (rows == 0 || lo == 0) {
    gram_async(c, rows, lo, Z, ldz, SVDH_F32, gram_off, gram_ev);  // the empty Gram, through the usual path
    return;
  }
  panel_mfma_impl(c, rows, lo, nterms, Y, ldy, li, M_dev, alpha, Z, ldz, beta, Zin, ldzin, (int64_t)gram_off,
                  gram_ev, tri);
}

static void panel_mfma_batched_impl(Ctx *c, int64_t count, const int64_t *rows, int64_t li, const int64_t *lo,
                                    float *const *Z, int64_t ld, const std::vector<double> *M,
                                    std::vector<std::vector<double>> *G) {
  if (li > 64 || ld % 4) fail(SVDH_ERR_ARG, "panel_mfma_batched: li <= 64 and ld a multiple of 4");
  if (ld >= kBufferMaxLd) fail(SVDH_ERR_ARG, "panel_mfma_batched: leading dimension too large");
  constexpr int W = kPanelBlock / 64;
  const bool gram = G != nullptr;
  if (gram) G->assign(count, {});
  // groups whose factors fit one upload slot (and, with the Grams, whose
  // Grams fit the landing area), one launch per group
  int64_t b0 = 0;
  while (b0 < count) {
    int64_t b1 = b0;
    size_t doubles = 0, landing = 0;
    while (b1 < count && (doubles + M[b1].size()) * sizeof(double) <= Ctx::kRingSlot &&
           (!gram || landing + (size_t)(lo[b1] * lo[b1]) <= Ctx::kSmall)) {
      doubles += M[b1].size();
      landing += (size_t)(lo[b1] * lo[b1]);
      ++b1;
    }
    if (b1 == b0) fail(SVDH_ERR_ARG, "panel_mfma_batched: factor larger than the upload slot");
    std::vector<double> mall;
    mall.reserve(doubles);
    bool tri = true;  // every factor of the group upper triangular
    for (int64_t b = b0; b < b1; ++b) {
      mall.insert(mall.end(), M[b].begin(), M[b].end());
      tri = tri && upper_triangular(M[b].data(), li, lo[b]);
    }
    const double *dm = static_cast<const double *>(c->upload(mall.data(), mall.size() * sizeof(double)));
    std::vector<PanelJob> jobs(b1 - b0);
    int64_t maxblk = 1, maxll = 1;
    double bytes = 0.0;
    size_t moff = 0, poff = 0, goff = 0;
    std::vector<size_t> part_off(b1 - b0), gram_off(b1 - b0);
    for (int64_t b = b0; b < b1; ++b) {
      if (lo[b] > 64) fail(SVDH_ERR_ARG, "panel_mfma_batched: lo <= 64");
      const int64_t nblk16 = (rows[b] + 15) / 16;
      const int64_t nblk =
          std::max<int64_t>(1, std::min<int64_t>((nblk16 + W - 1) / W, (int64_t)panel_wg_per_cu() * c->ncu));
      jobs[b - b0] = PanelJob{PanelTerm{Z[b], ld, (int)li, dm + moff, 1.0}, Z[b], ld, rows[b], nblk, (int)lo[b],
                              nullptr, nullptr};
      part_off[b - b0] = poff;
      gram_off[b - b0] = goff;
      poff += (size_t)nblk * lo[b] * lo[b];
      goff += (size_t)(lo[b] * lo[b]);
      moff += M[b].size();
      maxblk = std::max(maxblk, nblk);
      maxll = std::max<int64_t>(maxll, lo[b] * lo[b]);
      bytes += (double)rows[b] * (li + lo[b]) * 4.0;
    }
    if (gram) {
      c->red_part.ensure(std::max<size_t>(std::max<size_t>(poff, 1) * sizeof(double), c->red_part.bytes));
      for (int64_t b = b0; b < b1; ++b) {
        jobs[b - b0].part = c->red_part.as<double>() + part_off[b - b0];
        jobs[b - b0].gram = c->d_small + gram_off[b - b0];
      }
    }
    const PanelJob *dj = static_cast<const PanelJob *>(c->upload(jobs.data(), jobs.size() * sizeof(PanelJob)));
    {
      KTimer tm(c, KC_DENSE, bytes);
      const dim3 grid((unsigned)maxblk, (unsigned)(b1 - b0));
      if (gram) {
        if (tri)
          hipLaunchKernelGGL((k_panel_mfma_batched<true, true>), grid, kPanelBlock, 0, c->stream, dj);
        else
          hipLaunchKernelGGL((k_panel_mfma_batched<true, false>), grid, kPanelBlock, 0, c->stream, dj);
        HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_gram_reduce_jobs, dim3((unsigned)reduce_grid(maxll), (unsigned)(b1 - b0)), kBlock, 0,
                           c->stream, dj);
      } else if (tri) {
        hipLaunchKernelGGL((k_panel_mfma_batched<false, true>), grid, kPanelBlock, 0, c->stream, dj);
      } else {
        hipLaunchKernelGGL((k_panel_mfma_batched<false, false>), grid, kPanelBlock, 0, c->stream, dj);
      }
      HIP_CHECK(hipGetLastError());
    }
    if (gram) {
      HIP_CHECK(hipEventRecord(c->small_ev, c->stream));
      HIP_CHECK(hipEventSynchronize(c->small_ev));
      for (int64_t b = b0; b < b1; ++b)
        (*G)[b].assign(c->h_small + gram_off[b - b0], c->h_small + gram_off[b - b0] + lo[b] * lo[b]);
    }
    b0 = b1;
  }
}

void panel_mfma_batched(Ctx *c, int64_t count, const int64_t *rows, int64_t li, const int64_t *lo, float *const *Z,
                        int64_t ld, const std::vector<double> *M) {
  panel_mfma_batched_impl(c, count, rows, li, lo, Z, ld, M, nullptr);
}

void panel_mfma_gram_batched(Ctx *c, int64_t count, const int64_t *rows, int64_t li, const int64_t *lo,
                             float *const *Z, int64_t ld, const std::vector<double> *M,
                             std::vector<std::vector<double>> &G) {
  panel_mfma_batched_impl(c, count, rows, li, lo, Z, ld, M, &G);
}

__global__ void k_fetch(const uint4 *src, uint4 *dst, int64_t n16) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void fetch_from_host(Ctx *c, const void *src, void *dst, size_t n) {
  if (n == 0) return;
  const int64_t n16 = (int64_t)((n + 15) / 16);  // whole 16-byte vectors (slots are 16-byte multiples)
  const int g = (int)std::min<int64_t>(64, (n16 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_fetch, g, kBlock, 0, c->stream, static_cast<const uint4 *>(src), static_cast<uint4 *>(dst),
                     n16);
  HIP_CHECK(hipGetLastError());
}

void gram_async(Ctx *c, int64_t m, int64_t l, const void *Y, int64_t ldy, int ydt, size_t off, hipEvent_t ev) {
  if (l > 128) fail(SVDH_ERR_ARG, "panel width %lld exceeds the Gram kernel limit 128", (long long)l);
  if (off + (size_t)l * l > Ctx::kSmall) fail(SVDH_ERR_ARG, "Gram landing area too small");
  if (ldy % 4 != 0) fail(SVDH_ERR_ARG, "Gram panel leading dimension %lld is not a multiple of 4", (long long)ldy);
  // one workgroup per CU at most; each wave gets at least 8 row steps
  const int nb = grid_of((m + 3) / 4, 8 * kGramWaves, 256);
  const int64_t ll = l * l;
  const size_t need = (size_t)nb * ll * sizeof(double) + ll * sizeof(double);
  c->red_part.ensure(std::max<size_t>(need, c->red_part.bytes));
  double *part = c->red_part.as<double>();
  double *G = part + (int64_t)nb * ll;
  {
    KTimer t(c, KC_DENSE, (double)m * l * dtype_size(ydt));
    for (int c0 = 0; c0 < l; c0 += 64)
      for (int d0 = c0; d0 < l; d0 += 64) {
        if (ydt == SVDH_F64)
          launch_gram<double>(c, c0 == d0, nb, m, (int)l, (const double *)Y, ldy, c0, d0, part);
        else if (ydt == SVDH_F32)
          launch_gram<float>(c, c0 == d0, nb, m, (int)l, (const float *)Y, ldy, c0, d0, part);
        else
          launch_gram<bf16_t>(c, c0 == d0, nb, m, (int)l, (const bf16_t *)Y, ldy, c0, d0, part);
      }
    // the sums go straight to the pinned landing area
    hipLaunchKernelGGL(k_gram_reduce, reduce_grid(ll), kBlock, 0, c->stream, nb, ll, (const double *)part,
                       c->d_small + off);
    HIP_CHECK(hipGetLastError());
  }
  (void)G;
  HIP_CHECK(hipEventRecord(ev ? ev : c->small_ev, c->stream));
}

void gram_batched_async(Ctx *c, int64_t count, const int64_t *rows, int64_t l, const void *const *P, int64_t ldy,
                        int ydt, size_t off, hipEvent_t ev) {
  if (count <= 0) return;
  if (l > 64) {
    fail(SVDH_ERR_ARG, "batched Gram: panel width %lld exceeds 64", (long long)l);
  }
  if (off + (size_t)count * l * l > Ctx::kSmall) fail(SVDH_ERR_ARG, "Gram landing area too small");
  if (ldy % 4 != 0) fail(SVDH_ERR_ARG, "Gram panel leading dimension %lld is not a multiple of 4", (long long)ldy);
  if ((size_t)count * sizeof(GramPanel) > Ctx::kRingSlot) fail(SVDH_ERR_ARG, "batched Gram: too many panels");
  // each panel is cut exactly as gram_async cuts it (same workgroups, so the
  // same partials and sums: a batched Gram equals the single one bit for bit)
  std::vector<GramPanel> pan(count);
  int64_t total = 0, maxblk = 1;
  double bytes = 0.0;
  for (int64_t b = 0; b < count; ++b) {
    const int64_t nblk = grid_of((rows[b] + 3) / 4, 8 * kGramWaves, 256);
    pan[b] = GramPanel{P[b], rows[b], total, nblk};
    total += nblk;
    maxblk = std::max(maxblk, nblk);
    bytes += (double)rows[b] * l * dtype_size(ydt);
  }
  const int64_t ll = l * l;
  c->red_part.ensure(std::max<size_t>((size_t)total * ll * sizeof(double), c->red_part.bytes));
  double *part = c->red_part.as<double>();
  const GramPanel *dpan = static_cast<const GramPanel *>(c->upload(pan.data(), pan.size() * sizeof(GramPanel)));
  {
    KTimer t(c, KC_DENSE, bytes);
    const dim3 grid((unsigned)maxblk, (unsigned)count);
    const int th = 64 * gram_waves<true>();
    const bool narrow = l <= 32;  // as launch_gram chooses
    if (ydt == SVDH_F64 && narrow)
      hipLaunchKernelGGL((k_gram_mfma_batched<double, 2>), grid, th, 0, c->stream, dpan, (int)l, ldy, part);
    else if (ydt == SVDH_F64)
      hipLaunchKernelGGL((k_gram_mfma_batched<double, 4>), grid, th, 0, c->stream, dpan, (int)l, ldy, part);
    else if (ydt == SVDH_F32 && narrow)
      hipLaunchKernelGGL((k_gram_mfma_batched<float, 2>), grid, th, 0, c->stream, dpan, (int)l, ldy, part);
    else if (ydt == SVDH_F32)
      hipLaunchKernelGGL((k_gram_mfma_batched<float, 4>), grid, th, 0, c->stream, dpan, (int)l, ldy, part);
    else if (narrow)
      hipLaunchKernelGGL((k_gram_mfma_batched<bf16_t, 2>), grid, th, 0, c->stream, dpan, (int)l, ldy, part);
    else
      hipLaunchKernelGGL((k_gram_mfma_batched<bf16_t, 4>), grid, th, 0, c->stream, dpan, (int)l, ldy, part);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_gram_reduce_batched, dim3((unsigned)reduce_grid(ll), (unsigned)count), kBlock, 0, c->stream,
                       dpan, ll, (const double *)part, c->d_small + off);
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipEventRecord(ev ? ev : c->small_ev, c->stream));
}

void gram(Ctx *c, int64_t m, int64_t l, const void *Y, int64_t ldy, int ydt, double *G_host) {
  gram_async(c, m, l, Y, ldy, ydt, 0);
  HIP_CHECK(hipEventSynchronize(c->small_ev));
  std::memcpy(G_host, c->h_small, (size_t)l * l * sizeof(double));
}

template <typename TY, typename TZ, typename C>
static void apply_t(Ctx *c, int64_t m, int64_t li, int64_t lo, const void *Y, int64_t ldy, const double *M, void *Z,
                    int64_t ldz, int64_t zcs, double alpha = 1.0, double beta = 0.0) {
  const int64_t blocks = (m + kTile - 1) / kTile;
  const size_t shm = (size_t)li * (kTile + 1) * sizeof(C);
  if (shm > 65536) fail(SVDH_ERR_ARG, "panel width %lld too wide for the apply kernel", (long long)li);
  const size_t shm_m = shm + (size_t)li * lo * sizeof(C);
  if (shm_m <= 65536)
    hipLaunchKernelGGL((k_apply<TY, TZ, C, true>), (int)blocks, kBlock, shm_m, c->stream, m, (int)li, (int)lo,
                       (const TY *)Y, ldy, M, (TZ *)Z, ldz, zcs, alpha, beta);
  else
    hipLaunchKernelGGL((k_apply<TY, TZ, C, false>), (int)blocks, kBlock, shm, c->stream, m, (int)li, (int)lo,
                       (const TY *)Y, ldy, M, (TZ *)Z, ldz, zcs, alpha, beta);
  HIP_CHECK(hipGetLastError());
}

// FP64 outputs of up to 128 columns on the FP64 matrix cores.
template <typename TY>
static void apply_mfma(Ctx *c, int64_t m, int64_t li, int64_t lo, const TY *Y, int64_t ldy, const double *M,
                       double *Z, int64_t ldz, int64_t zcs) {
  const int g = (int)(((m + 15) / 16 + 3) / 4);
  if (lo <= 16)
    hipLaunchKernelGGL((k_apply_mfma<TY, 1>), g, 256, 0, c->stream, m, (int)li, (int)lo, Y, ldy, M, Z, ldz, zcs, 1.0,
                       0.0);
  else if (lo <= 32)
    hipLaunchKernelGGL((k_apply_mfma<TY, 2>), g, 256, 0, c->stream, m, (int)li, (int)lo, Y, ldy, M, Z, ldz, zcs, 1.0,
                       0.0);
  else if (lo <= 64)
    hipLaunchKernelGGL((k_apply_mfma<TY, 4>), g, 256, 0, c->stream, m, (int)li, (int)lo, Y, ldy, M, Z, ldz, zcs, 1.0,
                       0.0);
  else
    hipLaunchKernelGGL((k_apply_mfma<TY, 8>), g, 256, 0, c->stream, m, (int)li, (int)lo, Y, ldy, M, Z, ldz, zcs, 1.0,
                       0.0);
  HIP_CHECK(hipGetLastError());
}

void panel_apply(Ctx *c, int64_t m, int64_t li, int64_t lo, const void *Y, int64_t ldy, int ydt,
                 const double *M_host, void *Z, int64_t ldz, int zdt, int64_t zcs) {
  if (m == 0 || lo == 0) return;
  const size_t mb = (size_t)li * lo * sizeof(double);
  const double *Md = static_cast<const double *>(c->upload(M_host, mb));
  KTimer t(c, KC_DENSE, (double)m * (li * dtype_size(ydt) + lo * dtype_size(zdt)));
  if (zdt == SVDH_F64 && lo <= 128 && ydt == SVDH_F32)
    apply_mfma<float>(c, m, li, lo, (const float *)Y, ldy, Md, (double *)Z, ldz, zcs);
  else if (zdt == SVDH_F64 && lo <= 128 && ydt == SVDH_F64)
    apply_mfma<double>(c, m, li, lo, (const double *)Y, ldy, Md, (double *)Z, ldz, zcs);
  else if (ydt == SVDH_F64 && zdt == SVDH_F64)
    apply_t<double, double, double>(c, m, li, lo, Y, ldy, Md, Z, ldz, zcs);
  else if (ydt == SVDH_F32 && zdt == SVDH_F32)
    apply_t<float, float, float>(c, m, li, lo, Y, ldy, Md, Z, ldz, zcs);
  else if (ydt == SVDH_F32 && zdt == SVDH_F64 && li <= 64)
    apply_t<float, double, double>(c, m, li, lo, Y, ldy, Md, Z, ldz, zcs);
  else if (ydt == SVDH_F32 && zdt == SVDH_F64)  // FP64 tiles of > 64 columns exceed 64 KB of LDS
    apply_t<float, double, float>(c, m, li, lo, Y, ldy, Md, Z, ldz, zcs);
  else if (ydt == SVDH_F64 && zdt == SVDH_F32)
    apply_t<double, float, double>(c, m, li, lo, Y, ldy, Md, Z, ldz, zcs);
  else if (ydt == SVDH_F32 && zdt == SVDH_BF16)
    apply_t<float, bf16_t, float>(c, m, li, lo, Y, ldy, Md, Z, ldz, zcs);
  else
    fail(SVDH_ERR_ARG, "unsupported apply dtypes %d -> %d", ydt, zdt);
}

void panel_apply_batched(Ctx *c, int64_t count, const int64_t *m, const int64_t *li, int64_t lo,
                         const void *const *Y, int64_t ldy, int ydt, const std::vector<double> *M, double *const *Z,
                         const int64_t *ldz, const int64_t *zcs) {
  if (lo <= 0 || count <= 0) return;
  if (lo > 128 || (ydt != SVDH_F32 && ydt != SVDH_F64))
    fail(SVDH_ERR_ARG, "panel_apply_batched: FP32 or FP64 panels, at most 128 output columns");
  // groups whose factors fit one upload slot, one launch per group
  int64_t b0 = 0;
  while (b0 < count) {
    int64_t b1 = b0;
    size_t doubles = 0;
    while (b1 < count && (doubles + M[b1].size()) * sizeof(double) <= Ctx::kRingSlot &&
           (b1 - b0 + 1) * sizeof(ApplyJob) <= Ctx::kRingSlot) {
      doubles += M[b1].size();
      ++b1;
    }
    if (b1 == b0) fail(SVDH_ERR_ARG, "panel_apply_batched: factor larger than the upload slot");
    std::vector<double> mall;
    mall.reserve(doubles);
    for (int64_t b = b0; b < b1; ++b) mall.insert(mall.end(), M[b].begin(), M[b].end());
    const double *dm = static_cast<const double *>(c->upload(mall.data(), std::max<size_t>(mall.size(), 1) * 8));
    std::vector<ApplyJob> jobs(b1 - b0);
    int64_t maxg = 1;
    size_t moff = 0;
    double bytes = 0.0;
    for (int64_t b = b0; b < b1; ++b) {
      if ((int64_t)M[b].size() != li[b] * lo) fail(SVDH_ERR_ARG, "panel_apply_batched: factor %lld is not li x lo",
                                                   (long long)b);
      jobs[b - b0] = ApplyJob{Y[b], ldy, m[b], (int32_t)li[b], (int32_t)lo, dm + moff, Z[b], ldz[b], zcs[b]};
      moff += M[b].size();
      maxg = std::max<int64_t>(maxg, ((m[b] + 15) / 16 + 3) / 4);
      bytes += (double)m[b] * (li[b] * dtype_size(ydt) + lo * 8.0);
    }
    const ApplyJob *dj = static_cast<const ApplyJob *>(c->upload(jobs.data(), jobs.size() * sizeof(ApplyJob)));
    KTimer t(c, KC_DENSE, bytes);
    const dim3 grid((unsigned)maxg, (unsigned)(b1 - b0));
    const bool f32 = ydt == SVDH_F32;
#define SVDH_AJ(TT)                                                                         \
  if (f32)                                                                                  \
    hipLaunchKernelGGL((k_apply_mfma_jobs<float, TT>), grid, 256, 0, c->stream, dj);        \
  else                                                                                      \
    hipLaunchKernelGGL((k_apply_mfma_jobs<double, TT>), grid, 256, 0, c->stream, dj);
    if (lo <= 16) {
      SVDH_AJ(1)
    } else if (lo <= 32) {
      SVDH_AJ(2)
    } else if (lo <= 64) {
      SVDH_AJ(4)
    } else {
      SVDH_AJ(8)
    }
#undef SVDH_AJ
    HIP_CHECK(hipGetLastError());
    b0 = b1;
  }
}

void panel_gemm_dev(Ctx *c, int64_t m, int64_t li, int64_t lo, const void *Y, int64_t ldy, const double *M_dev,
                    double alpha, double beta, void *Z, int64_t ldz, int dt) {
  if (m == 0 || lo == 0) return;
  KTimer t(c, KC_DENSE, (double)m * (li + 2 * lo) * dtype_size(dt));
  if (dt == SVDH_F64)
    apply_t<double, double, double>(c, m, li, lo, Y, ldy, M_dev, Z, ldz, 1, alpha, beta);
  else if (dt == SVDH_F32)
    apply_t<float, float, float>(c, m, li, lo, Y, ldy, M_dev, Z, ldz, 1, alpha, beta);
  else
    fail(SVDH_ERR_ARG, "unsupported panel dtype %d", dt);
}

void gram_cross(Ctx *c, int64_t m, int64_t k1, const void *X1, int64_t ld1, int64_t k2, const void *X2, int64_t ld2,
                int dt, double *C_dev) {
  if (k1 > 64 || k2 > 64) fail(SVDH_ERR_ARG, "cross Gram limited to 64 x 64");
  if (ld1 % 4 || ld2 % 4) fail(SVDH_ERR_ARG, "cross Gram leading dimensions must be multiples of 4");
  const int nb = grid_of((m + 3) / 4, 8 * 4, 256);
  const int64_t kk = k1 * k2;
  c->red_part.ensure(std::max<size_t>(c->red_part.bytes, (size_t)nb * kk * sizeof(double)));
  double *part = c->red_part.as<double>();
  KTimer t(c, KC_DENSE, (double)m * (k1 + k2) * dtype_size(dt));
  if (dt == SVDH_F64)
    hipLaunchKernelGGL((k_gram_mfma<double, false>), nb, 64 * gram_waves<false>(), 0, c->stream, m, (int)k1,
                       (const double *)X1, ld1, (int)k2, (const double *)X2, ld2, 0, 0, part, kk, (int)k2, 0);
  else if (dt == SVDH_F32)
    hipLaunchKernelGGL((k_gram_mfma<float, false>), nb, 64 * gram_waves<false>(), 0, c->stream, m, (int)k1,
                       (const float *)X1, ld1, (int)k2, (const float *)X2, ld2, 0, 0, part, kk, (int)k2, 0);
  else
    fail(SVDH_ERR_ARG, "unsupported panel dtype %d", dt);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gram_reduce, reduce_grid(kk), kBlock, 0, c->stream, nb, kk, (const double *)part, C_dev);
  HIP_CHECK(hipGetLastError());
}

void panel_residual(Ctx *c, int64_t m, int64_t w, int64_t w2, int64_t k, const void *Y1, const void *Y2, int64_t ld,
                    int dt, const double *M1, const double *M2, double *nr2_host) {
  if (w > 128 || k > 128 || w2 > w) fail(SVDH_ERR_ARG, "residual panel wider than 128");
  if (dt == SVDH_F32 && w <= 64 && k <= 64 && ld % 4 == 0 && m > 0) {
    // FP32 panels on the matrix cores: column norms of Y1 M1 + Y2 M2 (FP32
    // products, as the vector kernel below computes them), squares summed in FP64
    std::vector<double> both(M1, M1 + w * k);
    both.insert(both.end(), M2, M2 + w2 * k);
    const double *Md = static_cast<const double *>(c->upload(both.data(), both.size() * sizeof(double)));
    const int64_t nblk = (m + 15) / 16;
    constexpr int W = kPanelBlock / 64;
    const int nb =
        (int)std::max<int64_t>(1, std::min<int64_t>((nblk + W - 1) / W, (int64_t)panel_wg_per_cu() * c->ncu));
    c->red_part.ensure(std::max<size_t>(c->red_part.bytes, (size_t)nb * k * sizeof(double)));
    double *part = c->red_part.as<double>();
    const PanelTerm t0{static_cast<const float *>(Y1), ld, (int)w, Md, 1.0};
    const PanelTerm t1{static_cast<const float *>(Y2), ld, (int)w2, Md + w * k, 1.0};
    {
      KTimer t(c, KC_DENSE, (double)m * (w + w2) * 4.0);
      if (w2 > 0)
        hipLaunchKernelGGL((k_panel_mfma<2, true>), nb, kPanelBlock, 0, c->stream, m, (int)k, t0, t1, nullptr,
                           (int64_t)0, nullptr, (int64_t)0, 0.f, part);
      else
        hipLaunchKernelGGL((k_panel_mfma<1, true>), nb, kPanelBlock, 0, c->stream, m, (int)k, t0, t0, nullptr,
                           (int64_t)0, nullptr, (int64_t)0, 0.f, part);
      HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_gram_reduce, reduce_grid(k), kBlock, 0, c->stream, nb, k, (const double *)part, c->d_small);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(c->small_ev, c->stream));
    if (nr2_host) panel_residual_wait(c, k, nr2_host);
    return;
  }
  const int nb = grid_of(m, kTile, 1024);
  const size_t mb = (size_t)w * k * sizeof(double);
  const double *M1d = static_cast<const double *>(c->upload(M1, mb));
  const double *M2d = static_cast<const double *>(c->upload(M2, (size_t)w2 * k * sizeof(double)));
  c->red_part.ensure(std::max<size_t>(c->red_part.bytes, (size_t)(nb + 1) * k * sizeof(double)));
  double *part = c->red_part.as<double>();
  {
    KTimer t(c, KC_DENSE, (double)m * 2 * w * dtype_size(dt));
    if (dt == SVDH_F64) {
      const size_t shm = 2 * (size_t)w * (kTile + 1) * sizeof(double);
      if (shm > 65536) fail(SVDH_ERR_ARG, "FP64 residual panel too wide");
      const size_t shm_m = shm + 2 * (size_t)w * k * sizeof(double);
      if (shm_m <= 65536)
        hipLaunchKernelGGL((k_resid<double, double, true>), nb, kBlock, shm_m, c->stream, m, (int)w, (int)w2, (int)k,
                           (const double *)Y1, (const double *)Y2, ld, M1d, M2d, part);
      else
        hipLaunchKernelGGL((k_resid<double, double, false>), nb, kBlock, shm, c->stream, m, (int)w, (int)w2, (int)k,
                           (const double *)Y1, (const double *)Y2, ld, M1d, M2d, part);
    } else {
      const size_t shm = 2 * (size_t)w * (kTile + 1) * sizeof(float);
      const size_t shm_m = shm + 2 * (size_t)w * k * sizeof(float);
      if (shm_m <= 65536)
        hipLaunchKernelGGL((k_resid<float, float, true>), nb, kBlock, shm_m, c->stream, m, (int)w, (int)w2, (int)k,
                           (const float *)Y1, (const float *)Y2, ld, M1d, M2d, part);
      else
        hipLaunchKernelGGL((k_resid<float, float, false>), nb, kBlock, shm, c->stream, m, (int)w, (int)w2, (int)k,
                           (const float *)Y1, (const float *)Y2, ld, M1d, M2d, part);
    }
    HIP_CHECK(hipGetLastError());
  }
  // column sums over the workgroup partials, on the device (same kernel as the Gram reduction)
  double *sums = part + (size_t)nb * k;
  (void)sums;
  hipLaunchKernelGGL(k_gram_reduce, reduce_grid(k), kBlock, 0, c->stream, nb, k, (const double *)part, c->d_small);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(c->small_ev, c->stream));
  if (nr2_host) panel_residual_wait(c, k, nr2_host);
}

void panel_residual_wait(Ctx *c, int64_t k, double *nr2_host) {
  HIP_CHECK(hipEventSynchronize(c->small_ev));
  std::memcpy(nr2_host, c->h_small, k * sizeof(double));
}

void panel_normalize_columns(Ctx *c, int64_t m, int64_t l, void *Y, int64_t ldy, int ydt) {
  if (l > kBlock) fail(SVDH_ERR_ARG, "panel width too large for column normalisation");
  const int nb = grid_of(m, 64, 1024);
  c->red_part.ensure(std::max<size_t>(c->red_part.bytes, (size_t)nb * l * sizeof(double)));
  double *part = c->red_part.as<double>();
  if (ydt == SVDH_F64)
    hipLaunchKernelGGL(k_colsq_part<double>, nb, kBlock, 0, c->stream, m, (int)l, (const double *)Y, ldy, part);
  else
    hipLaunchKernelGGL(k_colsq_part<float>, nb, kBlock, 0, c->stream, m, (int)l, (const float *)Y, ldy, part);
  HIP_CHECK(hipGetLastError());
  std::vector<double> h((size_t)nb * l);
  HIP_CHECK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<double> scale(l, 1.0);
  for (int64_t j = 0; j < l; ++j) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += h[(size_t)b * l + j];
    const double nrm = std::sqrt(s);
    if (nrm > 1e-10) scale[j] = 1.0 / nrm;  // src/randomized/mod.rs:279-283
  }
  c->stage.ensure(l * sizeof(double));
  double *sd = c->stage.as<double>();
  HIP_CHECK(hipMemcpyAsync(sd, scale.data(), l * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (ydt == SVDH_F64)
    hipLaunchKernelGGL(k_colscale<double>, grid_of(m * l), kBlock, 0, c->stream, m, (int)l, (double *)Y, ldy, sd);
  else
    hipLaunchKernelGGL(k_colscale<float>, grid_of(m * l), kBlock, 0, c->stream, m, (int)l, (float *)Y, ldy, sd);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

// Weighted column sums s_c = sum_i wt_i X[i, c] (wt == nullptr: all ones),
// per-workgroup FP64 partials; w <= 128, thread t owns column t % 128 and
// rows of phase t / 128, so each wavefront reads 64 consecutive columns.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_wcolsum(int64_t rows, int w, const T *X, int64_t ld, const double *wt,
                                                   double *part) {
  const int col = threadIdx.x & 127, ph = threadIdx.x >> 7;
  double acc = 0.0;
  if (col < w)
    for (int64_t i = (int64_t)blockIdx.x * 2 + ph; i < rows; i += (int64_t)gridDim.x * 2) {
      const double x = ldd<T>(X, i * ld + col);
      acc += wt ? wt[i] * x : x;
    }
  __shared__ double red[kBlock];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 128 && col < w) part[(int64_t)blockIdx.x * w + col] = red[threadIdx.x] + red[threadIdx.x + 128];
}

// Y[i, c] -= scale * a_i * s_c (a == nullptr: all ones).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rank1(int64_t rows, int w, T *Y, int64_t ld, const double *a,
                                                 const double *sv, double scale) {
  const int64_t total = rows * w;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
    const int64_t i = e / w;
    const int col = (int)(e - i * w);
    const double ai = a ? a[i] : 1.0;
    std_<T>(Y, i * ld + col, ldd<T>(Y, i * ld + col) - scale * ai * sv[col]);
  }
}

// Column dots s_c = sum_i X[i, c] Y[i, c] for k <= 64 columns: thread t owns
// column t % 64 and rows of phase t / 64; per-workgroup FP64 partials.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_coldot(int64_t rows, int k, const T *X, const T *Y, int64_t ld,
                                                  double *part) {
  const int col = threadIdx.x & 63, ph = threadIdx.x >> 6;
  double acc = 0.0;
  if (col < k)
    for (int64_t i = (int64_t)blockIdx.x * 4 + ph; i < rows; i += (int64_t)gridDim.x * 4)
      acc += ldd<T>(X, i * ld + col) * ldd<T>(Y, i * ld + col);
  __shared__ double red[kBlock];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64 && col < k)
    part[(int64_t)blockIdx.x * k + col] =
        (red[threadIdx.x] + red[threadIdx.x + 64]) + (red[threadIdx.x + 128] + red[threadIdx.x + 192]);
}

// Conjugate-gradient step on k columns at once (one shift per column):
// W += alpha_c P, R -= alpha_c MP, and the new column norms of R (partials).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(int64_t rows, int k, T *W, T *R, const T *P, const T *MP,
                                                     int64_t ld, const double *alpha, double *part) {
  const int col = threadIdx.x & 63, ph = threadIdx.x >> 6;
  double acc = 0.0;
  if (col < k) {
    const double a = alpha[col];
    for (int64_t i = (int64_t)blockIdx.x * 4 + ph; i < rows; i += (int64_t)gridDim.x * 4) {
      const int64_t at = i * ld + col;
      std_<T>(W, at, ldd<T>(W, at) + a * ldd<T>(P, at));
      const double r = ldd<T>(R, at) - a * ldd<T>(MP, at);
      std_<T>(R, at, r);
      acc += r * r;
    }
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64 && col < k)
    part[(int64_t)blockIdx.x * k + col] =
        (red[threadIdx.x] + red[threadIdx.x + 64]) + (red[threadIdx.x + 128] + red[threadIdx.x + 192]);
}

// Y[i, c] = a_c X[i, c] + b_c Y[i, c] (a or b null: 1).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_colaxpby(int64_t rows, int k, const double *a, const T *X,
                                                    const double *b, T *Y, int64_t ld) {
  const int64_t total = rows * k;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
    const int64_t i = e / k;
    const int col = (int)(e - i * k);
    const int64_t at = i * ld + col;
    std_<T>(Y, at, (a ? a[col] : 1.0) * ldd<T>(X, at) + (b ? b[col] : 1.0) * ldd<T>(Y, at));
  }
}

void panel_rank1_correct(Ctx *c, int64_t rx, int64_t w, const void *X, int64_t ldx, int xdt, const double *wt,
                         int64_t ry, void *Y, int64_t ldy, int ydt, const double *a, double scale) {
  if (w > 128) fail(SVDH_ERR_ARG, "panel wider than 128");
  if (w == 0) return;
  const int nb = grid_of(rx, 2, 512);
  c->red_part.ensure(std::max<size_t>(c->red_part.bytes, (size_t)(nb + 1) * w * sizeof(double)));
  double *part = c->red_part.as<double>();
  double *sums = part + (size_t)nb * w;
  KTimer t(c, KC_DENSE, (double)rx * w * dtype_size(xdt) + 2.0 * ry * w * dtype_size(ydt));
  if (xdt == SVDH_F64)
    hipLaunchKernelGGL(k_wcolsum<double>, nb, kBlock, 0, c->stream, rx, (int)w, (const double *)X, ldx, wt, part);
  else if (xdt == SVDH_F32)
    hipLaunchKernelGGL(k_wcolsum<float>, nb, kBlock, 0, c->stream, rx, (int)w, (const float *)X, ldx, wt, part);
  else
    fail(SVDH_ERR_ARG, "unsupported panel dtype %d", xdt);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gram_reduce, reduce_grid(w), kBlock, 0, c->stream, nb, w, (const double *)part, sums);
  HIP_CHECK(hipGetLastError());
  const int g = grid_of(ry * w);
  if (ydt == SVDH_F64)
    hipLaunchKernelGGL(k_rank1<double>, g, kBlock, 0, c->stream, ry, (int)w, (double *)Y, ldy, a, sums, scale);
  else if (ydt == SVDH_F32)
    hipLaunchKernelGGL(k_rank1<float>, g, kBlock, 0, c->stream, ry, (int)w, (float *)Y, ldy, a, sums, scale);
  else
    fail(SVDH_ERR_ARG, "unsupported panel dtype %d", ydt);
  HIP_CHECK(hipGetLastError());
}

void panel_coldots(Ctx *c, int64_t m, int64_t k, const void *X, const void *Y, int64_t ld, int dt, double *out_dev) {
  if (k > 64) {  // 64 columns per pass
    const size_t es = dtype_size(dt);
    for (int64_t c0 = 0; c0 < k; c0 += 64)
      panel_coldots(c, m, std::min<int64_t>(64, k - c0), static_cast<const char *>(X) + c0 * es,
                    static_cast<const char *>(Y) + c0 * es, ld, dt, out_dev + c0);
    return;
  }
  const int nb = grid_of(m, 4, 512);
  c->red_part.ensure(std::max<size_t>(c->red_part.bytes, (size_t)nb * k * sizeof(double)));
  double *part = c->red_part.as<double>();
  KTimer t(c, KC_DENSE, 2.0 * m * k * dtype_size(dt));
  if (dt == SVDH_F64)
    hipLaunchKernelGGL(k_coldot<double>, nb, kBlock, 0, c->stream, m, (int)k, (const double *)X, (const double *)Y,
                       ld, part);
  else
    hipLaunchKernelGGL(k_coldot<float>, nb, kBlock, 0, c->stream, m, (int)k, (const float *)X, (const float *)Y, ld,
                       part);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gram_reduce, reduce_grid(k), kBlock, 0, c->stream, nb, k, (const double *)part, out_dev);
  HIP_CHECK(hipGetLastError());
}

void panel_cg_update(Ctx *c, int64_t m, int64_t k, void *W, void *R, const void *P, const void *MP, int64_t ld,
                     int dt, const double *alpha_dev, double *rr_dev) {
  if (k > 64) fail(SVDH_ERR_ARG, "CG panels limited to 64 columns");
  const int nb = grid_of(m, 4, 512);
  c->red_part.ensure(std::max<size_t>(c->red_part.bytes, (size_t)nb * k * sizeof(double)));
  double *part = c->red_part.as<double>();
  KTimer t(c, KC_DENSE, 6.0 * m * k * dtype_size(dt));
  if (dt == SVDH_F64)
    hipLaunchKernelGGL(k_cg_update<double>, nb, kBlock, 0, c->stream, m, (int)k, (double *)W, (double *)R,
                       (const double *)P, (const double *)MP, ld, alpha_dev, part);
  else
    hipLaunchKernelGGL(k_cg_update<float>, nb, kBlock, 0, c->stream, m, (int)k, (float *)W, (float *)R,
                       (const float *)P, (const float *)MP, ld, alpha_dev, part);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gram_reduce, reduce_grid(k), kBlock, 0, c->stream, nb, k, (const double *)part, rr_dev);
  HIP_CHECK(hipGetLastError());
}

void panel_colaxpby(Ctx *c, int64_t m, int64_t k, const double *a_dev, const void *X, const double *b_dev, void *Y,
                    int64_t ld, int dt) {
  KTimer t(c, KC_DENSE, 3.0 * m * k * dtype_size(dt));
  const int g = grid_of(m * k);
  if (dt == SVDH_F64)
    hipLaunchKernelGGL(k_colaxpby<double>, g, kBlock, 0, c->stream, m, (int)k, a_dev, (const double *)X, b_dev,
                       (double *)Y, ld);
  else
    hipLaunchKernelGGL(k_colaxpby<float>, g, kBlock, 0, c->stream, m, (int)k, a_dev, (const float *)X, b_dev,
                       (float *)Y, ld);
  HIP_CHECK(hipGetLastError());
}

// out[c * ldo + i] = scale_c * X[i * ldx + c] for i < rows, c < cols: an
// LDS-tiled (64 x 64) transpose with a per-column scale (null = 1).
__global__ __launch_bounds__(kBlock) void k_transpose_scale(int64_t rows, int cols, const double *X, int64_t ldx,
                                                           const double *scale, double *out, int64_t ldo) {
  __shared__ double t[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int rr = ty; rr < 64; rr += 4) {
    const int64_t i = r0 + rr;
    const int cc = c0 + tx;
    t[rr][tx] = (i < rows && cc < cols) ? X[i * ldx + cc] : 0.0;
  }
  __syncthreads();
  for (int cc = ty; cc < 64; cc += 4) {
    const int col = c0 + cc;
    const int64_t i = r0 + tx;
    if (col < cols && i < rows) out[(int64_t)col * ldo + i] = (scale ? scale[col] : 1.0) * t[tx][cc];
  }
}

void panel_transpose_scale(Ctx *c, int64_t rows, int64_t cols, const double *X, int64_t ldx, const double *scale_dev,
                           double *out, int64_t ldo) {
  if (rows == 0 || cols == 0) return;
  KTimer t(c, KC_DENSE, 2.0 * rows * cols * 8.0);
  dim3 g((unsigned)((rows + 63) / 64), (unsigned)((cols + 63) / 64));
  hipLaunchKernelGGL(k_transpose_scale, g, kBlock, 0, c->stream, rows, (int)cols, X, ldx, scale_dev, out, ldo);
  HIP_CHECK(hipGetLastError());
}

void panel_gaussian(Ctx *c, int64_t m, int64_t l, uint64_t seed, void *Y, int64_t ldy, int ydt) {
  if (ydt == SVDH_F64)
    hipLaunchKernelGGL(k_gauss<double>, grid_of(m * l), kBlock, 0, c->stream, m, (int)l, seed, (double *)Y, ldy);
  else if (ydt == SVDH_F32)
    hipLaunchKernelGGL(k_gauss<float>, grid_of(m * l), kBlock, 0, c->stream, m, (int)l, seed, (float *)Y, ldy);
  else
    hipLaunchKernelGGL(k_gauss<bf16_t>, grid_of(m * l), kBlock, 0, c->stream, m, (int)l, seed, (bf16_t *)Y, ldy);
  HIP_CHECK(hipGetLastError());
}

void panel_gaussian_batched(Ctx *c, int64_t count, const int64_t *m, int64_t l, uint64_t seed, void *const *Y,
                            int64_t ldy, int ydt) {
  // one upload slot: count row counts, then count pointers
  const int64_t per = (int64_t)(Ctx::kRingSlot / 16);
  for (int64_t b0 = 0; b0 < count; b0 += per) {
    const int64_t nb = std::min(per, count - b0);
    std::vector<int64_t> buf(2 * nb);
    int64_t mmax = 0;
    for (int64_t b = 0; b < nb; ++b) {
      buf[b] = m[b0 + b];
      buf[nb + b] = (int64_t)(uintptr_t)Y[b0 + b];
      mmax = std::max(mmax, m[b0 + b]);
    }
    const int64_t *dm = static_cast<const int64_t *>(c->upload(buf.data(), buf.size() * sizeof(int64_t)));
    const dim3 grid((unsigned)grid_of(mmax * l), (unsigned)nb);
    if (ydt == SVDH_F64)
      hipLaunchKernelGGL(k_gauss_jobs<double>, grid, kBlock, 0, c->stream, dm, (int)l, seed,
                         reinterpret_cast<double *const *>(dm + nb), ldy);
    else if (ydt == SVDH_F32)
      hipLaunchKernelGGL(k_gauss_jobs<float>, grid, kBlock, 0, c->stream, dm, (int)l, seed,
                         reinterpret_cast<float *const *>(dm + nb), ldy);
    else
      hipLaunchKernelGGL(k_gauss_jobs<bf16_t>, grid, kBlock, 0, c->stream, dm, (int)l, seed,
                         reinterpret_cast<bf16_t *const *>(dm + nb), ldy);
    HIP_CHECK(hipGetLastError());
  }
}

void panel_convert(Ctx *c, int64_t m, int64_t l, const void *X, int64_t ldx, int xdt, void *Y, int64_t ldy,
                   int ydt) {
  const int g = grid_of(m * l);
  if (xdt == SVDH_F64 && ydt == SVDH_F32)
    hipLaunchKernelGGL((k_pconvert<double, float>), g, kBlock, 0, c->stream, m, (int)l, (const double *)X, ldx, (float *)Y, ldy);
  else if (xdt == SVDH_F32 && ydt == SVDH_F64)
    hipLaunchKernelGGL((k_pconvert<float, double>), g, kBlock, 0, c->stream, m, (int)l, (const float *)X, ldx, (double *)Y, ldy);
  else if (xdt == SVDH_F32 && ydt == SVDH_BF16)
    hipLaunchKernelGGL((k_pconvert<float, bf16_t>), g, kBlock, 0, c->stream, m, (int)l, (const float *)X, ldx, (bf16_t *)Y, ldy);
  else if (xdt == SVDH_F64 && ydt == SVDH_BF16)
    hipLaunchKernelGGL((k_pconvert<double, bf16_t>), g, kBlock, 0, c->stream, m, (int)l, (const double *)X, ldx, (bf16_t *)Y, ldy);
  else if (xdt == SVDH_F64 && ydt == SVDH_F64)
    hipLaunchKernelGGL((k_pconvert<double, double>), g, kBlock, 0, c->stream, m, (int)l, (const double *)X, ldx, (double *)Y, ldy);
  else if (xdt == SVDH_F32 && ydt == SVDH_F32)
    hipLaunchKernelGGL((k_pconvert<float, float>), g, kBlock, 0, c->stream, m, (int)l, (const float *)X, ldx, (float *)Y, ldy);
  else
    fail(SVDH_ERR_ARG, "unsupported panel conversion %d -> %d", xdt, ydt);
  HIP_CHECK(hipGetLastError());
}

// Y[c * ldy + r] = X[r * ldx + c] (converted) for r < rows, c < cols: an
// LDS-tiled 64 x 64 transpose.
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_transpose_cvt(int64_t rows, int64_t cols, const TI *X, int64_t ldx,
                                                          TO *Y, int64_t ldy) {
  __shared__ TI t[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int rr = ty; rr < 64; rr += kBlock / 64) {
    const int64_t r = r0 + rr, cc = c0 + tx;
    t[rr][tx] = (r < rows && cc < cols) ? X[r * ldx + cc] : TI(0);
  }
  __syncthreads();
  for (int cc = ty; cc < 64; cc += kBlock / 64) {
    const int64_t col = c0 + cc, r = r0 + tx;
    if (col < cols && r < rows) Y[col * ldy + r] = static_cast<TO>(t[tx][cc]);
  }
}

void panel_transpose_convert(Ctx *c, int64_t rows, int64_t cols, const void *X, int64_t ldx, int xdt, void *Y,
                             int64_t ldy, int ydt) {
  if (rows == 0 || cols == 0) return;
  if (rows > (int64_t)64 * 0x7fffffff || cols > (int64_t)64 * 65535) fail(SVDH_ERR_ARG, "transpose too large");
  dim3 g((unsigned)((rows + 63) / 64), (unsigned)((cols + 63) / 64));
  if (xdt == SVDH_F32 && ydt == SVDH_F32)
    hipLaunchKernelGGL((k_transpose_cvt<float, float>), g, kBlock, 0, c->stream, rows, cols, (const float *)X, ldx,
                       (float *)Y, ldy);
  else if (xdt == SVDH_F64 && ydt == SVDH_F32)
    hipLaunchKernelGGL((k_transpose_cvt<double, float>), g, kBlock, 0, c->stream, rows, cols, (const double *)X, ldx,
                       (float *)Y, ldy);
  else if (xdt == SVDH_F32 && ydt == SVDH_F64)
    hipLaunchKernelGGL((k_transpose_cvt<float, double>), g, kBlock, 0, c->stream, rows, cols, (const float *)X, ldx,
                       (double *)Y, ldy);
  else if (xdt == SVDH_F64 && ydt == SVDH_F64)
    hipLaunchKernelGGL((k_transpose_cvt<double, double>), g, kBlock, 0, c->stream, rows, cols, (const double *)X,
                       ldx, (double *)Y, ldy);
  else
    fail(SVDH_ERR_ARG, "unsupported transpose %d -> %d", xdt, ydt);
  HIP_CHECK(hipGetLastError());
}

void panel_fill(Ctx *c, int64_t m, int64_t l, double v, void *Y, int64_t ldy, int ydt) {
  const int g = grid_of(m * l);
  if (ydt == SVDH_F64)
    hipLaunchKernelGGL(k_pfill<double>, g, kBlock, 0, c->stream, m, (int)l, v, (double *)Y, ldy);
  else if (ydt == SVDH_F32)
    hipLaunchKernelGGL(k_pfill<float>, g, kBlock, 0, c->stream, m, (int)l, v, (float *)Y, ldy);
  else
    hipLaunchKernelGGL(k_pfill<bf16_t>, g, kBlock, 0, c->stream, m, (int)l, v, (bf16_t *)Y, ldy);
  HIP_CHECK(hipGetLastError());
}

}  // namespace svdh
