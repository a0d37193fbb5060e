// Sparse operators of the svdh library: CSR upload, the transpose (CSC) and
// masked views, nnz-sorted row schedules, SpMV and SpMM kernels for gfx950.
//
// Reference semantics:
//   y = A x, y = A^T x          src/lanczos/mod.rs:1520-1603 (CsrMatrix::svd_opa)
//   masked remap                src/lanczos/masked.rs:104-238
//   Y = A X, Y = A^T X          src/lanczos/masked.rs:265-471 (multiply_with_dense)
//
// Design: every direction of an operator is a CSR whose rows are output
// rows (A^T is the CSC built once at upload). Output rows are sorted by nnz
// and cut into bins; a bin assigns 2^lg lanes to a row (SpMV) or a row group
// of lanes across panel columns (SpMM). Rows longer than a threshold are cut
// into fixed nnz chunks, one workgroup each, whose partials a second small
// kernel adds in chunk order, so results are deterministic, no output is
// written with atomics and no workgroup waits on another.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <utility>

#include <sys/stat.h>

#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "common.h"

namespace svdh {

namespace {

constexpr int kSpmvLong = 1024;   // SpMV rows above this nnz are chunked
#ifndef SVDH_SPMV_PER_LANE
#define SVDH_SPMV_PER_LANE 4
#endif
// Binned SpMV (masked views): a row of nnz nonzeros gets the fewest lanes (a
// power of two) that hold it at this many per lane.
constexpr int kSpmvPerLane = SVDH_SPMV_PER_LANE;
// Streamed SpMV: a wave reads kStreamStep consecutive nonzeros per step, one
// per lane and load; rows longer than a step are one task each (a chunk of at
// most kStreamChunk nonzeros per task).
constexpr int kStreamU = 8;
constexpr int kStreamStep = 64 * kStreamU;
constexpr int kStreamChunk = 8 * kStreamStep;
constexpr int kStreamRows = 64;  // rows per task at most
// Source-tiled SpMV: workgroups of kTiledBlock threads stage kTiledSrc source
// entries in LDS; tasks as in the streamed SpMV with kTiledU nonzeros per lane
// per step (the LDS left after the source block holds the waves' products).
constexpr int kTiledBlock = 1024;
constexpr int kTiledSrc = 16384;
constexpr int kTiledU = 4;
constexpr int kTiledStep = 64 * kTiledU;
constexpr int kTiledMaxRow = 64 * kTiledStep;  // longest stacked row (one task, no chunks)
constexpr int64_t kTiledMinNnz = 1 << 20;
constexpr int kSpmvChunk = 4096;  // SpMV chunk nnz (256 threads x 16)

typedef uint16_t bf16_t;

__device__ __forceinline__ float bf16_to_f32(bf16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

__host__ __device__ __forceinline__ bf16_t f32_to_bf16(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t u = __float_as_uint(f);
#else
  uint32_t u;
  std::memcpy(&u, &f, 4);
#endif
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return static_cast<bf16_t>((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<bf16_t>(u >> 16);
}

template <typename T> __device__ __forceinline__ float to_f(T v);
template <> __device__ __forceinline__ float to_f<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f<double>(double v) { return (float)v; }
template <> __device__ __forceinline__ float to_f<bf16_t>(bf16_t v) { return bf16_to_f32(v); }

template <typename Acc, typename T> __device__ __forceinline__ Acc ld_as(const T *p, int64_t i);
template <> __device__ __forceinline__ double ld_as<double, double>(const double *p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ double ld_as<double, float>(const float *p, int64_t i) { return (double)p[i]; }
template <> __device__ __forceinline__ double ld_as<double, bf16_t>(const bf16_t *p, int64_t i) { return (double)bf16_to_f32(p[i]); }
template <> __device__ __forceinline__ float ld_as<float, float>(const float *p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ float ld_as<float, double>(const double *p, int64_t i) { return (float)p[i]; }
template <> __device__ __forceinline__ float ld_as<float, bf16_t>(const bf16_t *p, int64_t i) { return bf16_to_f32(p[i]); }

// Panel vector width in elements: 16 bytes per lane.
template <typename X> struct PanelVec { static constexpr int N = 16 / sizeof(X); };

template <typename Acc, typename X>
__device__ __forceinline__ void ld_panel(const X *p, Acc (&out)[PanelVec<X>::N]);
template <> __device__ __forceinline__ void ld_panel<double, double>(const double *p, double (&o)[2]) {
  double2 v = *reinterpret_cast<const double2 *>(p);
  o[0] = v.x; o[1] = v.y;
}
template <> __device__ __forceinline__ void ld_panel<float, float>(const float *p, float (&o)[4]) {
  float4 v = *reinterpret_cast<const float4 *>(p);
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
template <> __device__ __forceinline__ void ld_panel<double, float>(const float *p, double (&o)[4]) {
  float4 v = *reinterpret_cast<const float4 *>(p);
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
template <> __device__ __forceinline__ void ld_panel<float, bf16_t>(const bf16_t *p, float (&o)[8]) {
  uint4 v = *reinterpret_cast<const uint4 *>(p);
  uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    o[2 * i] = __uint_as_float(w[i] << 16);
    o[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
  }
}

template <typename Y, typename Acc, int N>
__device__ __forceinline__ void st_panel(Y *p, const Acc (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) p[i] = static_cast<Y>(v[i]);
}
template <> __device__ __forceinline__ void st_panel<float, float, 4>(float *p, const float (&v)[4]) {
  *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
template <> __device__ __forceinline__ void st_panel<float, float, 8>(float *p, const float (&v)[8]) {
  reinterpret_cast<float4 *>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
  reinterpret_cast<float4 *>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
template <> __device__ __forceinline__ void st_panel<double, double, 2>(double *p, const double (&v)[2]) {
  *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------------------------------------
// SpMV
// ------------------------------------------------------------------------

template <typename V, typename X, typename Y>
struct SpmvArgs {
  const int32_t *rowptr;
  const int32_t *colidx;
  const V *vals;
  const int32_t *rowmap;
  const int32_t *colmap;
  const int32_t *rows;
  const int32_t *chunks;
  double *partials;
  const X *x;
  Y *y;
  int32_t nbins;
  int32_t long_block_begin;
  BinDesc bins[kMaxBins];
};

// One row per group of 2^LG lanes. The bin of a row is chosen so that a lane
// holds at most kSpmvPerLane of its nonzeros, and the loads of one lane's
// nonzeros are issued together, predicated on the row end, so that a row
// costs one round of independent gathers rather than a chain of dependent
// ones. Each lane adds its nonzeros in index order, then the lanes are
// reduced by butterfly.
template <int LG, bool MC, typename V, typename X, typename Y>
__device__ __forceinline__ void spmv_group(const SpmvArgs<V, X, Y> &a, const BinDesc &d, int lb) {
  constexpr int G = 1 << LG;
  constexpr int U = kSpmvPerLane;
  constexpr int RPB = kBlock >> LG;
  const int gid = threadIdx.x >> LG;
  const int lane = threadIdx.x & (G - 1);
  const int idx = lb * RPB + gid;
  const bool active = idx < d.row_count;
  double acc = 0.0;
  int r = 0;
  if (active) {
    r = a.rows[d.row_begin + idx];
    const int src = a.rowmap ? a.rowmap[r] : r;
    const int e = a.rowptr[src + 1];
    for (int k = a.rowptr[src] + lane; k < e; k += U * G) {
      int c[U];
      V v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        // unconditional loads (entry 0 stands in past the row end), so that
        // all of them are in flight before the first is used
        const bool ok = k + u * G < e;
        const int ks = ok ? k + u * G : 0;
        c[u] = a.colidx[ks];
        v[u] = a.vals[ks];
        if (!ok) c[u] = -1;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int cc = c[u];
        if (MC && cc >= 0) cc = a.colmap[cc];
        // unconditional gather (entry 0 stands in for a dropped one), then a select
        const double xv = ld_as<double>(a.x, cc >= 0 ? cc : 0);
        acc += cc >= 0 ? ld_as<double>(v, u) * xv : 0.0;
      }
    }
  }
  if (G > 1) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
  }
  if (active && lane == 0) a.y[r] = static_cast<Y>(acc);
}

template <bool MC, typename V, typename X, typename Y>
__device__ __forceinline__ void spmv_long(const SpmvArgs<V, X, Y> &a, int ch) {
  __shared__ double red[kBlock / 64];
  const int *cd = a.chunks + 6 * ch;
  const int r = cd[0], beg = cd[1], end = cd[2], nch = cd[4];
  double acc = 0.0;
  int k = beg + threadIdx.x;
  for (; k + 3 * kBlock < end; k += 4 * kBlock) {
    int c[4];
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c[u] = a.colidx[k + u * kBlock];
      v[u] = ld_as<double>(a.vals, k + u * kBlock);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int cc = c[u];
      if (MC) cc = a.colmap[cc];
      if (!MC || cc >= 0) acc += v[u] * ld_as<double>(a.x, cc);
    }
  }
  for (; k < end; k += kBlock) {
    int cc = a.colidx[k];
    if (MC) cc = a.colmap[cc];
    if (!MC || cc >= 0) acc += ld_as<double>(a.vals, k) * ld_as<double>(a.x, cc);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += red[w];
    if (nch == 1)
      a.y[r] = static_cast<Y>(t);
    else
      a.partials[ch] = t;
  }
}

// Second pass for SpMV rows cut into several chunks: partials added in chunk
// order, one thread per such row (the kernel boundary orders the passes).
template <typename Y>
__global__ void k_spmv_fin(int nrows, const int32_t *multi, const double *partials, Y *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int r = multi[3 * i], first = multi[3 * i + 1], nch = multi[3 * i + 2];
  double t = 0.0;
  for (int k = 0; k < nch; ++k) t += partials[first + k];
  y[r] = static_cast<Y>(t);
}

// Streamed SpMV over the rows in storage order. Task t = (ra, rb, kb, ke):
//  - rb > ra: rows ra .. rb - 1, whose nonzeros kb .. ke - 1 fit in one step.
//    Lane l reads nonzeros kb + l + 64 j (each load instruction covers 64
//    consecutive entries), the products go to the wave's LDS row, and the
//    rows are summed from there: with nr rows, G = 64 / 2^ceil(log2 nr) lanes
//    per row, lane s of a group adds entries s, s + G, ... of its row and the
//    group is reduced by butterfly.
//  - rb <= 0: one row ra (longer than a step), nonzeros kb .. ke - 1 in steps;
//    each lane adds its entries in order, then the wave is reduced. rb = 0:
//    the task is the whole row (written to y); rb < 0: a chunk, written to
//    partials[-rb - 1] and added in chunk order by k_spmv_fin.
// No workgroup-level wait: a wave's LDS row is its own.
template <typename V, typename X, typename Y>
struct StreamArgs {
  const int32_t *rowptr;
  const int32_t *colidx;
  const V *vals;
  const int32_t *colmap;
  const int4 *tasks;
  double *partials;
  const X *x;
  Y *y;
  int32_t ntasks;
};

template <bool MC, typename V, typename X, typename Y>
__global__ __launch_bounds__(kBlock) void k_spmv_stream(StreamArgs<V, X, Y> a) {
  __shared__ double prod[kBlock / 64][kStreamStep];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = blockIdx.x * (kBlock / 64) + w;
  if (t >= a.ntasks) return;
  const int4 tk = a.tasks[t];
  const int ra = tk.x, rb = tk.y, kb = tk.z, ke = tk.w;
  if (rb > ra) {
    const int nr = rb - ra;
    const int lg = nr > 1 ? 32 - __clz(nr - 1) : 0;  // ceil(log2 nr)
    const int sh = 6 - lg;                            // G = 2^sh lanes per row
    const int gid = lane >> sh, sub = lane & ((1 << sh) - 1);
    // lanes past the last row load row ra's bounds and then take an empty range
    const int q = gid < nr ? ra + gid : ra;
    const int p0 = a.rowptr[q], p1 = a.rowptr[q + 1];
    const int rs = gid < nr ? p0 - kb : 0, re = gid < nr ? p1 - kb : 0;
    int c[kStreamU];
    V v[kStreamU];
#pragma unroll
    for (int j = 0; j < kStreamU; ++j) {
      const int k = kb + lane + 64 * j;
      const bool ok = k < ke;
      const int ks = ok ? k : 0;  // unconditional loads, all in flight together
      c[j] = a.colidx[ks];
      v[j] = a.vals[ks];
      if (!ok) c[j] = -1;
    }
#pragma unroll
    for (int j = 0; j < kStreamU; ++j) {
      int cc = c[j];
      if (MC && cc >= 0) cc = a.colmap[cc];
      const double xv = ld_as<double>(a.x, cc >= 0 ? cc : 0);  // unconditional gather, then a select
      prod[w][lane + 64 * j] = cc >= 0 ? ld_as<double>(v, j) * xv : 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double acc = 0.0;
    for (int k = rs + sub; k < re; k += 1 << sh) acc += prod[w][k];
    for (int off = 32 >> lg; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (gid < nr && sub == 0) a.y[ra + gid] = static_cast<Y>(acc);
    return;
  }
  double acc = 0.0;
  for (int k0 = kb; k0 < ke; k0 += kStreamStep) {
    int c[kStreamU];
    V v[kStreamU];
#pragma unroll
    for (int j = 0; j < kStreamU; ++j) {
      const int k = k0 + lane + 64 * j;
      const bool ok = k < ke;
      const int ks = ok ? k : 0;
      c[j] = a.colidx[ks];
      v[j] = a.vals[ks];
      if (!ok) c[j] = -1;
    }
#pragma unroll
    for (int j = 0; j < kStreamU; ++j) {
      int cc = c[j];
      if (MC && cc >= 0) cc = a.colmap[cc];
      const double xv = ld_as<double>(a.x, cc >= 0 ? cc : 0);
      acc += cc >= 0 ? ld_as<double>(v, j) * xv : 0.0;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) {
    if (rb == 0)
      a.y[ra] = static_cast<Y>(acc);
    else
      a.partials[-rb - 1] = acc;
  }
}

// Source-tiled SpMV. Workgroup g takes range g = (block b, tasks t0 .. t1):
// it copies source entries [b sb, b sb + sb) into LDS, then its waves take the
// range's tasks in turn, each as in k_spmv_stream with the source read from
// LDS through the 16-bit in-block column index. Stacked row q = b nout + r
// is written to part[q] (nb > 1, added in block order by k_spmv_tiled_fin) or
// to y[r] (nb = 1).
template <typename V, typename X, typename Y>
struct TiledArgs {
  const int32_t *rowptr;  // stacked, nb * nout + 1
  const uint16_t *col;
  const V *vals;
  const int4 *tasks;
  const int4 *ranges;
  const X *x;
  double *part;
  Y *y;
  int64_t nsrc;
  int32_t sb;
  int32_t direct;
};

template <typename V, typename X, typename Y>
__global__ __launch_bounds__(kTiledBlock) void k_spmv_tiled(TiledArgs<V, X, Y> a) {
  __shared__ X xs[kTiledSrc];
  __shared__ double prod[kTiledBlock / 64][kTiledStep];
  const int4 rg = a.ranges[blockIdx.x];
  const int64_t s0 = (int64_t)rg.x * a.sb;
  const int slen = (int)min<int64_t>(a.sb, a.nsrc - s0);
  // 8 loads per thread in flight, then the LDS writes (addresses past the
  // block end read its last entry and are not written)
  for (int i0 = threadIdx.x; i0 < slen; i0 += 8 * kTiledBlock) {
    X t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = a.x[s0 + min(i0 + u * kTiledBlock, slen - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i0 + u * kTiledBlock < slen) xs[i0 + u * kTiledBlock] = t[u];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int WV = kTiledBlock / 64;
  // A wave takes tasks t, t + WV, ... of the range. While it forms one task's
  // products, the next task's row bounds and first step of entries are in
  // flight (their descriptor was loaded one task earlier), so a task costs
  // its arithmetic rather than a memory round trip. Every load is
  // unconditional at a clamped index (the descriptor of the range's last
  // task, entry 0), and entries past a task's end are dropped where they are
  // used.
  auto ld_task = [&](int t) { return a.tasks[t < rg.z ? t : rg.z - 1]; };
  auto ld_first = [&](const int4 &tk, int &p0, int &p1, int (&c)[kTiledU], V (&v)[kTiledU]) {
    const int nr = tk.y - tk.x;
    const int lg = nr > 1 ? 32 - __clz(nr - 1) : 0;
    const int gid = lane >> (6 - lg);
    const int q = gid < nr ? tk.x + gid : tk.x;
    p0 = a.rowptr[q];
    p1 = a.rowptr[q + 1];
#pragma unroll
    for (int j = 0; j < kTiledU; ++j) {
      const int k = tk.z + lane + 64 * j;
      const int ks = k < tk.w ? k : 0;
      c[j] = a.col[ks];
      v[j] = a.vals[ks];
    }
  };
  int t = rg.y + w;
  int4 tk = ld_task(t), tkn = ld_task(t + WV);
  int p0, p1, c[kTiledU];
  V v[kTiledU];
  ld_first(tk, p0, p1, c, v);
  for (; t < rg.z; t += WV) {
    // the next task's first step (and the descriptor after it) in flight
    int q0n, q1n, cn[kTiledU];
    V vn[kTiledU];
    ld_first(tkn, q0n, q1n, cn, vn);
    const int4 tknn = ld_task(t + 2 * WV);
    const int qa = tk.x, qb = tk.y, kb = tk.z, ke = tk.w;
    if (qb > qa) {
      const int nr = qb - qa;
      const int lg = nr > 1 ? 32 - __clz(nr - 1) : 0;
      const int sh = 6 - lg;
      const int gid = lane >> sh, sub = lane & ((1 << sh) - 1);
      // lanes past the last row loaded row qa's bounds and take an empty range
      const int rs = gid < nr ? p0 - kb : 0, re = gid < nr ? p1 - kb : 0;
      // the previous task's reads of this wave's products come first
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int j = 0; j < kTiledU; ++j) {
        // entries past the task end land in slots no row reads: no select
        const int cj = kb + lane + 64 * j < ke ? c[j] : 0;
        prod[w][lane + 64 * j] = ld_as<double>(v, j) * (double)xs[cj];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double acc = 0.0;
      for (int k = rs + sub; k < re; k += 1 << sh) acc += prod[w][k];
      for (int off = 32 >> lg; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if (gid < nr && sub == 0) {
        if (a.direct)
          a.y[qa + gid] = static_cast<Y>(acc);
        else
          a.part[qa + gid] = acc;
      }
    } else {
      // one long row: the first step from the prefetched entries, then the
      // rest, each step's loads all in flight before the first is used
      double acc = 0.0;
      for (int k0 = kb; k0 < ke; k0 += kTiledStep) {
        if (k0 > kb) {
#pragma unroll
          for (int j = 0; j < kTiledU; ++j) {
            const int k = k0 + lane + 64 * j;
            const int ks = k < ke ? k : 0;
            c[j] = a.col[ks];
            v[j] = a.vals[ks];
          }
        }
#pragma unroll
        for (int j = 0; j < kTiledU; ++j) {
          // unconditional LDS read at a clamped index, then a select
          const bool ok = k0 + lane + 64 * j < ke;
          const double p = ld_as<double>(v, j) * (double)xs[ok ? c[j] : 0];
          acc += ok ? p : 0.0;
        }
      }
      acc = wave_sum(acc);
      if (lane == 0) {
        if (a.direct)
          a.y[qa] = static_cast<Y>(acc);
        else
          a.part[qa] = acc;
      }
    }
    tk = tkn;
    tkn = tknn;
    p0 = q0n;
    p1 = q1n;
#pragma unroll
    for (int j = 0; j < kTiledU; ++j) {
      c[j] = cn[j];
      v[j] = vn[j];
    }
  }
}

// y[r] = sum over source blocks b of part[b nout + r]: P lanes per row, lane
// t adds blocks [t nb / P, (t + 1) nb / P) in order (8 loads in flight), then
// the P lane sums are added by butterfly (a fixed order).
template <int P, typename Y>
__global__ void k_spmv_tiled_fin(int64_t nout, int nb, const double *part, Y *y) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t r = g / P;
  const int t = (int)(g % P);
  double s = 0.0;
  if (r < nout) {
    const int b1 = (t + 1) * nb / P;
    int b = t * nb / P;
    for (; b + 8 <= b1; b += 8) {
      double p[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) p[u] = part[(int64_t)(b + u) * nout + r];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += p[u];
    }
    for (; b < b1; ++b) s += part[(int64_t)b * nout + r];
  }
#pragma unroll
  for (int off = 1; off < P; off <<= 1) s += __shfl_xor(s, off, 64);
  if (r < nout && t == 0) y[r] = static_cast<Y>(s);
}

template <bool MC, typename V, typename X, typename Y>
__global__ __launch_bounds__(kBlock) void k_spmv(SpmvArgs<V, X, Y> a) {
  const int b = blockIdx.x;
  if (b >= a.long_block_begin) {
    spmv_long<MC>(a, b - a.long_block_begin);
    return;
  }
  BinDesc d = a.bins[0];
#pragma unroll
  for (int i = 1; i < kMaxBins; ++i)
    if (i < a.nbins && b >= a.bins[i].block_begin) d = a.bins[i];
  const int lb = b - d.block_begin;
  switch (d.lg) {
    case 0: spmv_group<0, MC>(a, d, lb); break;
    case 1: spmv_group<1, MC>(a, d, lb); break;
    case 2: spmv_group<2, MC>(a, d, lb); break;
    case 3: spmv_group<3, MC>(a, d, lb); break;
    case 4: spmv_group<4, MC>(a, d, lb); break;
    case 5: spmv_group<5, MC>(a, d, lb); break;
    default: spmv_group<6, MC>(a, d, lb); break;
  }
}

// ------------------------------------------------------------------------
// SpMM: Y (rows x l, row-major, ldy) = A X (X row-major, ldx)
//
// Two kinds of work in one launch:
//  - chunks of long rows (more than SpmmTune::long_nnz), one workgroup each: the
//    row groups of the workgroup take slices of the chunk and are reduced in
//    LDS in group order. When the operand panel X is larger than one XCD's
//    L2, long rows are first cut at 8 source-block boundaries and workgroup
//    w takes chunks of block w % 8 only: workgroups are dispatched
//    round-robin over the 8 XCDs, so each XCD gathers from one eighth of X.
//    A row with several chunks writes one partial per chunk; a second kernel
//    (k_spmm_fin) adds them in chunk order, so results do not depend on
//    timing and no workgroup waits on or signals another;
//  - short rows, one row group of R = 2^LGR lanes each, rows sorted by nnz.
// Lane t of a row group covers panel columns [c0 + t*NV, c0 + (t+1)*NV) of
// each column pass.
// ------------------------------------------------------------------------

constexpr int kChunkInts = 8;  // r, b, e, slot, first_slot, nchunks, row_id (unused), pad
constexpr int kSegInts = 4;    // r, b, e, slot (-1: the row's only segment, written to y)
#ifndef SVDH_SPMM_UNROLL
#define SVDH_SPMM_UNROLL 4  // panel gathers in flight per row group
#endif

template <typename V, typename X, typename Y, typename Acc>
struct SpmmArgs {
  const int32_t *rowptr;
  const int32_t *colidx;
  const V *vals;
  const int32_t *rowmap;
  const int32_t *colmap;
  const int32_t *rows;    // short rows, sorted by nnz (descending)
  const int32_t *chunks;  // kChunkInts per chunk, grouped by source block
  Acc *partials;  // slot x ldp
  const X *x;
  Y *y;
  int64_t ldx, ldy, ldp;
  int32_t l;
  int32_t nshort;
  int32_t nblk;
  int32_t chunk_blocks;  // nblk * max chunks per block
  int32_t seg_blocks;    // nblk * max segment workgroups per block
  int32_t total_blocks;  // chunk blocks + segment blocks + short-row blocks
  const int32_t *segs;   // kSegInts per segment, grouped by source block
  int32_t seg_off[8];
  int32_t seg_cnt[8];
  int32_t blk_off[8];
  int32_t blk_cnt[8];
  int32_t off32;  // every panel row's byte offset fits 31 bits (the 16-lane DPP path)
};

// One fused multiply-add per entry, written out: every entry is added the
// same way whether it falls in a full batch or in the last one, so the
// batch size (gathers in flight) and the lanes per row leave the sums
// unchanged.
// Lane n of each 16-lane row of the wave, to every lane of that row (DPP
// row_newbcast, a VALU move instead of an LDS permute); n must fold to a
// constant once the caller's loops are unrolled.
__device__ __forceinline__ int row_bcast16(int x, int n) {
#define SVDH_RB(N) \
  case N: return __builtin_amdgcn_update_dpp(0, x, 0x150 + N, 0xf, 0xf, false);
  switch (n) {
    SVDH_RB(0) SVDH_RB(1) SVDH_RB(2) SVDH_RB(3) SVDH_RB(4) SVDH_RB(5) SVDH_RB(6) SVDH_RB(7)
    SVDH_RB(8) SVDH_RB(9) SVDH_RB(10) SVDH_RB(11) SVDH_RB(12) SVDH_RB(13) SVDH_RB(14)
    default: return __builtin_amdgcn_update_dpp(0, x, 0x15f, 0xf, 0xf, false);
  }
#undef SVDH_RB
}
__device__ __forceinline__ float row_bcast16(float x, int n) { return __int_as_float(row_bcast16(__float_as_int(x), n)); }
// The same broadcast as a plain DPP move (every lane is written, so no prior
// value is kept: the compiler can fold it into the instruction that uses it).
__device__ __forceinline__ int row_bcast16m(int x, int n) {
#define SVDH_RB(N) \
  case N: return __builtin_amdgcn_mov_dpp(x, 0x150 + N, 0xf, 0xf, false);
  switch (n) {
    SVDH_RB(0) SVDH_RB(1) SVDH_RB(2) SVDH_RB(3) SVDH_RB(4) SVDH_RB(5) SVDH_RB(6) SVDH_RB(7)
    SVDH_RB(8) SVDH_RB(9) SVDH_RB(10) SVDH_RB(11) SVDH_RB(12) SVDH_RB(13) SVDH_RB(14)
    default: return __builtin_amdgcn_mov_dpp(x, 0x15f, 0xf, 0xf, false);
  }
#undef SVDH_RB
}
__device__ __forceinline__ float row_bcast16m(float x, int n) {
  return __int_as_float(row_bcast16m(__float_as_int(x), n));
}
__device__ __forceinline__ double row_bcast16m(double x, int n) {
  const long long b = __double_as_longlong(x);
  const int lo = row_bcast16m((int)(b & 0xffffffffll), n), hi = row_bcast16m((int)(b >> 32), n);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double row_bcast16(double x, int n) {
  const long long b = __double_as_longlong(x);
  const int lo = row_bcast16((int)(b & 0xffffffffll), n), hi = row_bcast16((int)(b >> 32), n);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// DPP row broadcasts folded into the instructions that use them (the
// compiler keeps a broadcast with several uses as a separate move). Each
// asm block opens with s_nop 1: a DPP read of a VGPR needs two wait states
// after a VALU write of it, and the compiler's hazard checks do not look
// inside inline assembly. The source lane must be active (EXEC): a DPP read
// of an inactive lane does not deliver its value.
// x of lane N of this lane's 16-lane row, plus y
template <int N>
__device__ __forceinline__ uint32_t add_bcast16(uint32_t x, uint32_t y);
// acc[i] = fma(v of lane N of this lane's 16-lane row, x[i], acc[i]), i < 4
template <int N>
__device__ __forceinline__ void fmac_bcast16(float (&acc)[4], float v, const float (&x)[4]);
#define SVDH_BCAST16(N)                                                                                  \
  template <>                                                                                            \
  __device__ __forceinline__ uint32_t add_bcast16<N>(uint32_t x, uint32_t y) {                           \
    uint32_t r;                                                                                          \
    asm("s_nop 1\n\tv_add_u32_dpp %0, %1, %2 row_newbcast:" #N " row_mask:0xf bank_mask:0xf"             \
        : "=v"(r)                                                                                        \
        : "v"(x), "v"(y));                                                                               \
    return r;                                                                                            \
  }                                                                                                      \
  template <>                                                                                            \
  __device__ __forceinline__ void fmac_bcast16<N>(float (&acc)[4], float v, const float (&x)[4]) {       \
    asm("s_nop 1\n\t"                                                                                    \
        "v_fmac_f32_dpp %0, %4, %5 row_newbcast:" #N " row_mask:0xf bank_mask:0xf\n\t"                   \
        "v_fmac_f32_dpp %1, %4, %6 row_newbcast:" #N " row_mask:0xf bank_mask:0xf\n\t"                   \
        "v_fmac_f32_dpp %2, %4, %7 row_newbcast:" #N " row_mask:0xf bank_mask:0xf\n\t"                   \
        "v_fmac_f32_dpp %3, %4, %8 row_newbcast:" #N " row_mask:0xf bank_mask:0xf"                         \
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])                                         \
        : "v"(v), "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]));                                           \
  }
SVDH_BCAST16(0) SVDH_BCAST16(1) SVDH_BCAST16(2) SVDH_BCAST16(3) SVDH_BCAST16(4) SVDH_BCAST16(5)
SVDH_BCAST16(6) SVDH_BCAST16(7) SVDH_BCAST16(8) SVDH_BCAST16(9) SVDH_BCAST16(10) SVDH_BCAST16(11)
SVDH_BCAST16(12) SVDH_BCAST16(13) SVDH_BCAST16(14) SVDH_BCAST16(15)
#undef SVDH_BCAST16

// One batch of U entries of a 16-lane row group, FP32 values and panel: lane
// n of the row holds entry jb U + n's panel-row byte offset `off` and value
// `v`; every lane of the row gathers the 16 bytes at src + offset + colb and
// adds its products. All 16 lanes run it (the DPP reads need the entry
// lanes active): lanes whose columns are past the panel pass colb = 0 and
// their sums are not stored.
// FP32 values and panel: the broadcasts folded into the adds and FMAs
template <typename Acc, typename X>
constexpr bool kBcastFold = std::is_same<Acc, float>::value && std::is_same<X, float>::value;

template <int JB, int U, int... I>
__device__ __forceinline__ void batch_bcast16(std::integer_sequence<int, I...>, const char *src, uint32_t off,
                                              uint32_t colb, float v, float (&acc)[4]) {
  uint32_t o[U];
  float xv[U][4];
  ((o[I] = add_bcast16<JB * U + I>(off, colb)), ...);
  (ld_panel<float, float>(reinterpret_cast<const float *>(src + o[I]), xv[I]), ...);
  (fmac_bcast16<JB * U + I>(acc, v, xv[I]), ...);
}
// Batches of U while the row's entries last (n, the same on the row's
// lanes); lanes past the row's end hold offset 0 and value 0, whose fused
// products add nothing.
template <int U, int... J>
__device__ __forceinline__ void entries_bcast16(std::integer_sequence<int, J...>, const char *src, uint32_t off,
                                                uint32_t colb, float v, int n, float (&acc)[4]) {
  (void)((J * U < n ? (batch_bcast16<J, U>(std::make_integer_sequence<int, U>{}, src, off, colb, v, acc), true)
                    : false) &&
         ...);
}

__device__ __forceinline__ float fma_acc(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_acc(double a, double b, double c) { return __builtin_fma(a, b, c); }

// acc += sum over nnz k in [kb, ke) of A_k * X[col_k, cols of this lane]
template <int R, bool MC, int UU = SVDH_SPMM_UNROLL, typename V, typename X, typename Y, typename Acc>
__device__ __forceinline__ void spmm_accum(const SpmmArgs<V, X, Y, Acc> &a, int kb, int ke, int lane, int col,
                                           bool colok, Acc (&acc)[PanelVec<X>::N]) {
  constexpr int NV = PanelVec<X>::N;
  if constexpr (R == 16) {
    if (a.off32) {
      // 16 lanes per row: the lane that loads an entry computes the byte
      // offset of its panel row, and DPP row broadcasts (VALU moves) hand
      // offset and value to the row's lanes; the same products in the same
      // order as the general path below, so the results are the same bits.
      const uint32_t rowb = (uint32_t)(a.ldx * (int64_t)sizeof(X));
      const uint32_t colb = (uint32_t)(col * (int)sizeof(X));
      const char *xb = reinterpret_cast<const char *>(a.x);
      constexpr int U = UU < R ? UU : R;
      for (int base = kb; base < ke; base += R) {
        int c = 0;
        Acc v = Acc(0);
        const int kk = base + lane;
        if (kk < ke) {
          c = __builtin_nontemporal_load(&a.colidx[kk]);
          v = ld_as<Acc>(a.vals, kk);
          if (MC) {
            c = a.colmap[c];
            if (c < 0) {
              c = 0;
              v = Acc(0);
            }
          }
        }
        const int n = min(R, ke - base);
        const int off = (int)((uint32_t)c * rowb);
        if constexpr (kBcastFold<Acc, X>) {
          entries_bcast16<U>(std::make_integer_sequence<int, R / U>{}, xb, (uint32_t)off, colok ? colb : 0u, v, n,
                             acc);
          continue;
        }
#pragma unroll
        for (int jb = 0; jb < R / U; ++jb) {
          if (jb * U >= n) break;
          uint32_t oj[U];
          Acc vj[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            oj[u] = (uint32_t)row_bcast16m(off, jb * U + u);
            vj[u] = row_bcast16m(v, jb * U + u);
          }
          if (colok) {
            Acc xv[U][NV];
#pragma unroll
            for (int u = 0; u < U; ++u)
              ld_panel<Acc, X>(reinterpret_cast<const X *>(xb + (size_t)(oj[u] + colb)), xv[u]);
            // lanes past the end hold offset 0 and value 0: their fused
            // products add exactly nothing
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
              for (int i = 0; i < NV; ++i) acc[i] = fma_acc(vj[u], xv[u][i], acc[i]);
          }
        }
      }
      return;
    }
  }
  for (int base = kb; base < ke; base += R) {
    int c = 0;
    Acc v = Acc(0);
    const int kk = base + lane;
    if (kk < ke) {
      c = __builtin_nontemporal_load(&a.colidx[kk]);
      v = ld_as<Acc>(a.vals, kk);
      if (MC) {
        c = a.colmap[c];
        if (c < 0) {
          c = 0;
          v = Acc(0);
        }
      }
    }
    const int n = min(R, ke - base);
    constexpr int U = UU < R ? UU : R;
    int j = 0;
    for (; j + U <= n; j += U) {
      int cj[U];
      Acc vj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cj[u] = __shfl(c, j + u, R);
        vj[u] = __shfl(v, j + u, R);
      }
      if (colok) {
        Acc xv[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) ld_panel<Acc, X>(a.x + (int64_t)cj[u] * a.ldx + col, xv[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int i = 0; i < NV; ++i) acc[i] = fma_acc(vj[u], xv[u][i], acc[i]);
      }
    }
    if (j < n) {
      // the last n - j (< U) entries as one batch: lanes past the end hold
      // column 0 and value 0, so their fused products add exactly nothing
      int cj[U];
      Acc vj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cj[u] = __shfl(c, j + u, R);
        vj[u] = __shfl(v, j + u, R);
      }
      if (colok) {
        Acc xv[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) ld_panel<Acc, X>(a.x + (int64_t)cj[u] * a.ldx + col, xv[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int i = 0; i < NV; ++i) acc[i] = fma_acc(vj[u], xv[u][i], acc[i]);
      }
    }
  }
}

template <int LGR, bool MC, int UU = SVDH_SPMM_UNROLL, typename V, typename X, typename Y, typename Acc>
__device__ __forceinline__ void spmm_short(const SpmmArgs<V, X, Y, Acc> &a, int b) {
  constexpr int R = 1 << LGR;
  constexpr int NV = PanelVec<X>::N;
  constexpr int RPB = kBlock >> LGR;
  const int gid = threadIdx.x >> LGR;
  const int lane = threadIdx.x & (R - 1);
  const int idx = b * RPB + gid;
  if (idx >= a.nshort) return;  // the whole row group leaves together
  const int r = a.rows[idx];
  const int src = a.rowmap ? a.rowmap[r] : r;
  const int kb = a.rowptr[src], ke = a.rowptr[src + 1];
  for (int c0 = 0; c0 < a.l; c0 += R * NV) {
    const int col = c0 + lane * NV;
    const bool colok = col < a.l;
    Acc acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = Acc(0);
    spmm_accum<R, MC, UU>(a, kb, ke, lane, col, colok, acc);
    if (colok) st_panel<Y, Acc, NV>(a.y + (int64_t)r * a.ldy + col, acc);
  }
}

// Segments: the part of a short row that falls in one source block, one row
// group each; workgroup w takes segments of block w % nblk only. A row with
// several segments writes one partial per segment (summed by k_spmm_fin).
template <int LGR, bool MC, int UU = SVDH_SPMM_UNROLL, typename V, typename X, typename Y, typename Acc>
__device__ __forceinline__ void spmm_segment(const SpmmArgs<V, X, Y, Acc> &a, int w) {
  constexpr int R = 1 << LGR;
  constexpr int NV = PanelVec<X>::N;
  constexpr int RPB = kBlock >> LGR;
  const int xb = w % a.nblk, kb = w / a.nblk;
  int off = 0, cnt = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (i == xb) {
      off = a.seg_off[i];
      cnt = a.seg_cnt[i];
    }
  const int gid = threadIdx.x >> LGR;
  const int lane = threadIdx.x & (R - 1);
  const int idx = kb * RPB + gid;
  if (idx >= cnt) return;  // the whole row group leaves together
  const int *sd = a.segs + (int64_t)(off + idx) * kSegInts;
  const int r = sd[0], kb0 = sd[1], ke = sd[2], slot = sd[3];
  for (int c0 = 0; c0 < a.l; c0 += R * NV) {
    const int col = c0 + lane * NV;
    const bool colok = col < a.l;
    Acc acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = Acc(0);
    spmm_accum<R, MC, UU>(a, kb0, ke, lane, col, colok, acc);
    if (colok) {
      if (slot < 0)
        st_panel<Y, Acc, NV>(a.y + (int64_t)r * a.ldy + col, acc);
      else
        st_panel<Acc, Acc, NV>(a.partials + (int64_t)slot * a.ldp + col, acc);
    }
  }
}

template <int LGR, bool MC, int UU = SVDH_SPMM_UNROLL, typename V, typename X, typename Y, typename Acc>
__device__ __forceinline__ void spmm_chunk(const SpmmArgs<V, X, Y, Acc> &a, int ci) {
  constexpr int R = 1 << LGR;
  constexpr int NV = PanelVec<X>::N;
  constexpr int NG = kBlock >> LGR;
  __shared__ Acc red[kBlock * PanelVec<X>::N];
  const int *cd = a.chunks + (int64_t)ci * kChunkInts;
  const int r = cd[0], cb = cd[1], ce = cd[2], slot = cd[3], nch = cd[5];
  const int gid = threadIdx.x >> LGR;
  const int lane = threadIdx.x & (R - 1);
  const int per = (ce - cb + NG - 1) / NG;
  const int gb = min(ce, cb + gid * per);
  const int ge = min(ce, gb + per);
  for (int c0 = 0; c0 < a.l; c0 += R * NV) {
    const int col = c0 + lane * NV;
    const bool colok = col < a.l;
    Acc acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = Acc(0);
    spmm_accum<R, MC, UU>(a, gb, ge, lane, col, colok, acc);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) red[gid * (R * NV) + lane * NV + i] = acc[i];
    __syncthreads();
    for (int t = threadIdx.x; t < R * NV; t += kBlock) {
      const int cc = c0 + t;
      if (cc < a.l) {
        Acc sum = Acc(0);
        for (int g = 0; g < NG; ++g) sum += red[g * (R * NV) + t];
        if (nch == 1)
          a.y[(int64_t)r * a.ldy + cc] = static_cast<Y>(sum);
        else
          a.partials[(int64_t)slot * a.ldp + cc] = sum;
      }
    }
  }
}

// Second pass for rows cut into several chunks: one workgroup per such row
// adds the chunk partials in chunk order (4 interleaved chains per column,
// combined in a fixed order). The kernel boundary makes the partials of the
// first pass visible, so no fences or atomics are needed.
// In a workgroup of more than kBlock threads the others only take part in
// the barriers (the sums are the same bits as with kBlock threads).
template <typename Y, typename Acc>
__device__ __forceinline__ void spmm_fin_row(const int32_t *multi, const Acc *partials, int64_t ldp, Y *y,
                                             int64_t ldy, int l, int row) {
  __shared__ Acc red[kBlock];
  const int *md = multi + 3 * (int64_t)row;
  const int r = md[0], first = md[1], nch = md[2];
  const int q = threadIdx.x >> 6, o = threadIdx.x & 63;
  const bool act = threadIdx.x < kBlock;
  for (int c0 = 0; c0 < l; c0 += 64) {
    const int cc = c0 + o;
    // thread q sums pieces q, q + 4, ... in four interleaved chains
    Acc s0 = Acc(0), s1 = Acc(0), s2 = Acc(0), s3 = Acc(0);
    if (act && cc < l) {
      const Acc *pp = partials + (int64_t)first * ldp + cc;
      int i = q;
      for (; i + 12 < nch; i += 16) {
        s0 += pp[(int64_t)i * ldp];
        s1 += pp[(int64_t)(i + 4) * ldp];
        s2 += pp[(int64_t)(i + 8) * ldp];
        s3 += pp[(int64_t)(i + 12) * ldp];
      }
      for (; i < nch; i += 4) s0 += pp[(int64_t)i * ldp];
    }
    const Acc sum = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (act) red[q * 64 + o] = sum;
    __syncthreads();
    if (q == 0 && cc < l) y[(int64_t)r * ldy + cc] = static_cast<Y>(((red[o] + red[64 + o]) + red[128 + o]) + red[192 + o]);
  }
}

template <typename Y, typename Acc>
__global__ __launch_bounds__(kBlock) void k_spmm_fin(const int32_t *multi, const Acc *partials, int64_t ldp, Y *y,
                                                    int64_t ldy, int l) {
  spmm_fin_row<Y, Acc>(multi, partials, ldp, y, ldy, l, blockIdx.x);
}

// The same sum for panels of width l <= 64, four rows per workgroup: one
// thread per column adds the row's pieces in four interleaved chains (piece
// i goes to chain i mod 4), combined in a fixed order.
template <typename Y, typename Acc, int RPB = 4>
__device__ __forceinline__ void spmm_fin4_rows(const int32_t *multi, int nrows, const Acc *partials, int64_t ldp,
                                               Y *y, int64_t ldy, int l, int blk) {
  const int row = blk * RPB + (threadIdx.x >> 6), o = threadIdx.x & 63;
  if (row >= nrows || o >= l) return;
  const int *md = multi + 3 * (int64_t)row;
  const int r = md[0], first = md[1], nch = md[2];
  const Acc *pp = partials + (int64_t)first * ldp + o;
  Acc s0 = Acc(0), s1 = Acc(0), s2 = Acc(0), s3 = Acc(0);
  int i = 0;
  for (; i + 3 < nch; i += 4) {
    s0 += pp[(int64_t)i * ldp];
    s1 += pp[(int64_t)(i + 1) * ldp];
    s2 += pp[(int64_t)(i + 2) * ldp];
    s3 += pp[(int64_t)(i + 3) * ldp];
  }
  for (; i < nch; ++i) s0 += pp[(int64_t)i * ldp];
  y[(int64_t)r * ldy + o] = static_cast<Y>((s0 + s1) + (s2 + s3));
}

template <typename Y, typename Acc>
__global__ __launch_bounds__(kBlock) void k_spmm_fin4(const int32_t *multi, int nrows, const Acc *partials,
                                                     int64_t ldp, Y *y, int64_t ldy, int l) {
  spmm_fin4_rows<Y, Acc>(multi, nrows, partials, ldp, y, ldy, l, blockIdx.x);
}

// The same sum for rows with hundreds of pieces (l <= 64): 1024 threads per
// row, thread q of a column summing pieces q, q + 16, ... in four chains,
// the 16 thread sums then added in q order. red holds 1024 sums.
template <typename Y, typename Acc>
__device__ __forceinline__ void spmm_fin_wide_row(const int32_t *md, const Acc *partials, int64_t ldp, Y *y,
                                                  int64_t ldy, int l, Acc *red) {
  const int r = md[0], first = md[1], nch = md[2];
  const int q = threadIdx.x >> 6, o = threadIdx.x & 63;
  Acc s0 = Acc(0), s1 = Acc(0), s2 = Acc(0), s3 = Acc(0);
  if (o < l) {
    const Acc *pp = partials + (int64_t)first * ldp + o;
    int i = q;
    for (; i + 48 < nch; i += 64) {
      s0 += pp[(int64_t)i * ldp];
      s1 += pp[(int64_t)(i + 16) * ldp];
      s2 += pp[(int64_t)(i + 32) * ldp];
      s3 += pp[(int64_t)(i + 48) * ldp];
    }
    for (; i < nch; i += 16) s0 += pp[(int64_t)i * ldp];
  }
  red[threadIdx.x] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (q == 0 && o < l) {
    Acc t = Acc(0);
    for (int k = 0; k < 16; ++k) t += red[k * 64 + o];
    y[(int64_t)r * ldy + o] = static_cast<Y>(t);
  }
}

template <typename Y, typename Acc>
__global__ __launch_bounds__(1024) void k_spmm_fin_wide(const int32_t *multi, const Acc *partials, int64_t ldp, Y *y,
                                                       int64_t ldy, int l) {
  __shared__ Acc red[1024];
  spmm_fin_wide_row<Y, Acc>(multi + 3 * (int64_t)blockIdx.x, partials, ldp, y, ldy, l, red);
}

// All piece sums of one A^T Y with staged columns in one launch of
// 1024-thread workgroups (l <= 64), each sum the same bits as in the
// separate kernels: the table holds nch chunked rows, then n4 rows of a few
// segments, then nwide staged rows (one piece per matrix-core workgroup).
// Workgroups [0, nwide) take one staged row each as k_spmm_fin_wide does
// (the staged rows have the most pieces: 16 threads per column),
// [nwide, nwide + nch) one chunked row each as k_spmm_fin does (its first
// kBlock threads), the rest 16 rows of a few segments each as k_spmm_fin4.
template <typename Y, typename Acc>
__global__ __launch_bounds__(1024) void k_spmm_fin_merged(const int32_t *multi, int nch, int n4, int nwide,
                                                         const Acc *partials, int64_t ldp, Y *y, int64_t ldy,
                                                         int l) {
  __shared__ Acc red[1024];
  const int w = blockIdx.x;
  if (w < nwide) {
    spmm_fin_wide_row<Y, Acc>(multi + 3 * (int64_t)(nch + n4 + w), partials, ldp, y, ldy, l, red);
  } else if (w < nwide + nch) {
    spmm_fin_row<Y, Acc>(multi, partials, ldp, y, ldy, l, w - nwide);
  } else {
    spmm_fin4_rows<Y, Acc, 16>(multi + 3 * (int64_t)nch, n4, partials, ldp, y, ldy, l, w - nwide - nch);
  }
}

// Batched forms: workgroup w belongs to operator b with boff[b] <= w < boff[b + 1].
__device__ __forceinline__ int batch_of(const int32_t *boff, int nb, int w) {
  int lo = 0, hi = nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (boff[mid] <= w) lo = mid; else hi = mid;
  }
  return lo;
}

template <typename Y, typename Acc>
struct FinArgs {
  const int32_t *multi;
  const Acc *partials;
  int64_t ldp;
  Y *y;
  int64_t ldy;
  int32_t l;
  int32_t nchunk;  // entries [0, nchunk) summed as k_spmm_fin does, the rest as k_spmm_fin4
};

// One workgroup per multi-row entry of every operator, each summed in the
// order of the single-operator pass (launch_fin): chunked rows as k_spmm_fin,
// the others in k_spmm_fin4's four chains, so a batch rounds as separate
// products do.
template <typename Y, typename Acc>
__global__ __launch_bounds__(kBlock) void k_spmm_fin_batched(const FinArgs<Y, Acc> *args, const int32_t *boff,
                                                            int nb) {
  const int b = batch_of(boff, nb, blockIdx.x);
  const FinArgs<Y, Acc> &f = args[b];
  const int row = (int)blockIdx.x - boff[b];
  if (row < f.nchunk) {
    spmm_fin_row<Y, Acc>(f.multi, f.partials, f.ldp, f.y, f.ldy, f.l, row);
    return;
  }
  const int o = threadIdx.x;
  if (o >= f.l) return;
  const int *md = f.multi + 3 * (int64_t)row;
  const int r = md[0], first = md[1], nch = md[2];
  const Acc *pp = f.partials + (int64_t)first * f.ldp + o;
  Acc s0 = Acc(0), s1 = Acc(0), s2 = Acc(0), s3 = Acc(0);
  int i = 0;
  for (; i + 3 < nch; i += 4) {
    s0 += pp[(int64_t)i * f.ldp];
    s1 += pp[(int64_t)(i + 1) * f.ldp];
    s2 += pp[(int64_t)(i + 2) * f.ldp];
    s3 += pp[(int64_t)(i + 3) * f.ldp];
  }
  for (; i < nch; ++i) s0 += pp[(int64_t)i * f.ldp];
  f.y[(int64_t)r * f.ldy + o] = static_cast<Y>((s0 + s1) + (s2 + s3));
}

// ------------------------------------------------------------------------
// A X with the panel rows of the most popular columns staged in LDS
// (power-law column popularity: a few hundred columns hold half of the
// nonzeros). hcol is the column index array with those columns encoded as
// -(1 + slot); a persistent grid of 1024-thread workgroups (two per CU, eight
// waves per SIMD) copies the h staged rows once and then strides over the
// short rows. Gathers of staged rows read LDS, the others global memory.
// ------------------------------------------------------------------------
constexpr int kHotBlock = 1024;

// acc += sum over k in [kb, ke) of vals[k] * panel row of hcol[k]: staged
// rows from LDS (LDS_SRC, hcol = -(1 + slot)) or panel rows from global memory.
// With 16 lanes per row the entries reach the lanes by row broadcasts, with
// fewer by LDS permutes.
template <int R, bool LDS_SRC, typename V, typename X, typename Y, typename Acc>
__device__ __forceinline__ void spmm_accum_hot(const SpmmArgs<V, X, Y, Acc> &a, const int32_t *hcol,
                                               const V *hval, const X *xs, int kb, int ke, int lane, int col,
                                               bool colok, Acc (&acc)[PanelVec<X>::N]) {
  constexpr int NV = PanelVec<X>::N;
  const X *src = LDS_SRC ? xs : a.x;
  // 16 lanes per row: byte offsets of the entries' panel rows are computed by
  // the lane that loaded the entry (32-bit: the launch checks that the panel
  // fits) and reach the row's lanes by DPP broadcasts, so an entry costs two
  // moves, one add, one load and the FMAs.
  const uint32_t rowb = (uint32_t)(a.ldx * (int64_t)sizeof(X));
  const uint32_t colb = (uint32_t)(col * (int)sizeof(X));
  const char *srcb = reinterpret_cast<const char *>(src);
  for (int base = kb; base < ke; base += R) {
    int c = 0;
    Acc v = Acc(0);
    const int kk = base + lane;
    if (kk < ke) {
      c = __builtin_nontemporal_load(&hcol[kk]);
      if (LDS_SRC) c = -1 - c;
      v = ld_as<Acc>(hval, kk);
    }
    const int n = min(R, ke - base);
    constexpr int U = SVDH_SPMM_UNROLL < R ? SVDH_SPMM_UNROLL : R;
    int j = 0;
    // The staged rows take the folded broadcasts; the global gathers keep the
    // separate moves below, which measure faster here (100,000 x 20,000 Zipf,
    // l = 60: 106.5 us against 115.1 us with the fold), unlike the gathers
    // of the other SpMM kernels (95.2 -> 92.7 us).
    if constexpr (R == 16 && kBcastFold<Acc, X> && LDS_SRC) {
      entries_bcast16<U>(std::make_integer_sequence<int, R / U>{}, srcb, (uint32_t)c * rowb, colok ? colb : 0u, v, n,
                         acc);
      j = n;
    } else if constexpr (R == 16) {
      const int off = (int)((uint32_t)c * rowb);
#pragma unroll
      for (int jb = 0; jb < R / U; ++jb) {
        if (jb * U >= n) break;
        uint32_t oj[U];
        Acc vj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          oj[u] = (uint32_t)row_bcast16m(off, jb * U + u);
          vj[u] = row_bcast16m(v, jb * U + u);
        }
        if (colok) {
          Acc xv[U][NV];
#pragma unroll
          for (int u = 0; u < U; ++u)
            ld_panel<Acc, X>(reinterpret_cast<const X *>(srcb + (size_t)(oj[u] + colb)), xv[u]);
          if (jb * U + U <= n) {
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
              for (int i = 0; i < NV; ++i) acc[i] = fma_acc(vj[u], xv[u][i], acc[i]);
          } else {
            // the last entries of the row: products past the end dropped by a select
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
              for (int i = 0; i < NV; ++i) {
                const Acc t = fma_acc(vj[u], xv[u][i], acc[i]);
                acc[i] = jb * U + u < n ? t : acc[i];
              }
          }
        }
      }
      j = n;
    } else {
      for (; j + U <= n; j += U) {
        int cj[U];
        Acc vj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          cj[u] = __shfl(c, j + u, R);
          vj[u] = __shfl(v, j + u, R);
        }
        if (colok) {
          Acc xv[U][NV];
#pragma unroll
          for (int u = 0; u < U; ++u) ld_panel<Acc, X>(src + (int64_t)cj[u] * a.ldx + col, xv[u]);
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < NV; ++i) acc[i] += vj[u] * xv[u][i];
        }
      }
    }
    if (j < n) {
      // the last n - j (< U) entries as one batch; lanes past the end hold
      // slot/column 0 and value 0, and their products are dropped by a select
      int cj[U];
      Acc vj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cj[u] = __shfl(c, j + u, R);
        vj[u] = __shfl(v, j + u, R);
      }
      if (colok) {
        Acc xv[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) ld_panel<Acc, X>(src + (int64_t)cj[u] * a.ldx + col, xv[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int i = 0; i < NV; ++i) acc[i] += j + u < n ? vj[u] * xv[u][i] : Acc(0);
      }
    }
  }
}

// Rows in `order` (nrows of them); row r's entries are reordered so that its
// split[r] staged entries come first: one loop over LDS rows, then one over
// global rows, with no per-entry branch between the two sources.
template <int LGR, typename V, typename X, typename Y, typename Acc>
__global__ __launch_bounds__(kHotBlock) void k_spmm_hot(SpmmArgs<V, X, Y, Acc> a, const int32_t *hcol,
                                                        const V *hval, const int32_t *split, const int32_t *order,
                                                        int nrows, const int32_t *hot_rows, int h) {
  extern __shared__ uint4 xs_raw[];
  X *xs = reinterpret_cast<X *>(xs_raw);
  constexpr int R = 1 << LGR;
  constexpr int NV = PanelVec<X>::N;
  constexpr int RPB = kHotBlock >> LGR;
  const int vpr = (int)(a.ldx / NV);  // 16-byte vectors per panel row
  for (int e = threadIdx.x; e < h * vpr; e += kHotBlock) {
    const int r = e / vpr, q = e - r * vpr;
    xs_raw[e] = reinterpret_cast<const uint4 *>(a.x + (int64_t)hot_rows[r] * a.ldx)[q];
  }
  __syncthreads();
  const int gid = threadIdx.x >> LGR;
  const int lane = threadIdx.x & (R - 1);
  for (int64_t b0 = (int64_t)blockIdx.x * RPB; b0 < nrows; b0 += (int64_t)gridDim.x * RPB) {
    const int64_t idx = b0 + gid;
    if (idx >= nrows) break;
    const int r = order[idx];
    const int kb = a.rowptr[r], ke = a.rowptr[r + 1];
    const int kh = kb + split[r];
    for (int c0 = 0; c0 < a.l; c0 += R * NV) {
      const int col = c0 + lane * NV;
      const bool colok = col < a.l;
      Acc acc[NV];
#pragma unroll
      for (int i = 0; i < NV; ++i) acc[i] = Acc(0);
      spmm_accum_hot<R, true>(a, hcol, hval, xs, kb, kh, lane, col, colok, acc);
      spmm_accum_hot<R, false>(a, hcol, hval, xs, kh, ke, lane, col, colok, acc);
      if (colok) st_panel<Y, Acc, NV>(a.y + (int64_t)r * a.ldy + col, acc);
    }
  }
}

// ------------------------------------------------------------------------
// A^T Y for the h = 256 staged (most popular) columns of A on the matrix
// cores. Those columns are kept as a dense m x 256 FP32 block Ad (built once;
// about half of their entries are nonzero in the power-law matrices this
// serves), so the product is a plain tall GEMM Z_h = Ad^T Y that streams Ad
// and Y once instead of gathering one panel row per nonzero. Workgroup w
// takes the 32-row blocks w, w + G, ...: the block's Ad rows and Y rows are
// copied into LDS (the next block's are loaded into registers meanwhile) and
// Z_h += Ad_blk^T Y_blk runs on v_mfma_f32_16x16x4 (A[l % 16][l / 16],
// B[l / 16][l % 16] per lane, D[4 (l / 16) + v][l % 16]). Wave q owns slot
// tiles 2q, 2q + 1 and all four 16-column tiles. At the end the workgroup
// writes its 256 x l partial into slot base + i G + w of staged row i; the
// fin pass sums those in workgroup order. LDS rows are 272 (tile) and 80 (Y)
// floats apart, 16 banks, so the operand reads are free of bank conflicts.
// ------------------------------------------------------------------------
constexpr int kTHotRows = 32, kTHotYStride = 80, kTHotBlock = 512;
#ifndef SVDH_THOT_DEFAULT_H
#define SVDH_THOT_DEFAULT_H 128
#endif
constexpr int kTHotDefaultH = SVDH_THOT_DEFAULT_H;
typedef float f4_t __attribute__((ext_vector_type(4)));

// Ad[r][slot] += staged entries of row r, in CSR order (one thread per row);
// Ad is m x h.
__global__ void k_hot_dense(int64_t m, const int32_t *rowptr, const int32_t *hcol, const float *hval,
                            const int32_t *split, float *Ad, int h) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m; r += (int64_t)gridDim.x * blockDim.x) {
    const int kb = rowptr[r], ke = kb + split[r];
    for (int k = kb; k < ke; ++k) Ad[r * h + (-1 - hcol[k])] += hval[k];
  }
}

// T slot tiles of 16 per wave: h = 128 T staged columns.
template <int T>
__global__ __launch_bounds__(kTHotBlock) void k_spmm_t_hot(int64_t m, const float *Ad, const float *Y, int64_t ldy,
                                                          int l, float *partials, int64_t ldp, int32_t slot_base) {
  constexpr int H = 128 * T, STRIDE = H + 16, VPR = H / 4, NA = 32 * VPR / kTHotBlock;
  __shared__ float tile[kTHotRows * STRIDE];
  __shared__ float ytile[kTHotRows * kTHotYStride];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kq = lane >> 4, rl = lane & 15;
  f4_t acc[T][4];
#pragma unroll
  for (int a = 0; a < T; ++a)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[a][q] = f4_t{0.f, 0.f, 0.f, 0.f};
  const int64_t nblocks = (m + kTHotRows - 1) / kTHotRows;
  // per thread: NA float4 of the 32 x H Ad block (row e / VPR, vector e % VPR)
  // and one float4 of the 32 x 64 Y block (row tid / 16, vector tid % 16)
  float4 ra[NA], ry;
  const int yr = tid >> 4, yq = tid & 15;
  auto fetch = [&](int64_t r0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = tid + kTHotBlock * i;
      const int64_t r = r0 + e / VPR;
      ra[i] = r < m ? reinterpret_cast<const float4 *>(Ad + r * H)[e % VPR] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int64_t r = r0 + yr;
    const int c = 4 * yq;
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
    if (r < m && c < l) w = *reinterpret_cast<const float4 *>(Y + r * ldy + c);
    ry = make_float4(w.x, c + 1 < l ? w.y : 0.f, c + 2 < l ? w.z : 0.f, c + 3 < l ? w.w : 0.f);
  };
  if ((int64_t)blockIdx.x < nblocks) fetch((int64_t)blockIdx.x * kTHotRows);
  for (int64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    __syncthreads();  // the previous block's operands are consumed
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = tid + kTHotBlock * i;
      *reinterpret_cast<float4 *>(tile + (e / VPR) * STRIDE + 4 * (e % VPR)) = ra[i];
    }
    *reinterpret_cast<float4 *>(ytile + yr * kTHotYStride + 4 * yq) = ry;
    __syncthreads();
    if (b + gridDim.x < nblocks) fetch((b + gridDim.x) * kTHotRows);  // in flight during the products
#pragma unroll
    for (int s = 0; s < kTHotRows / 4; ++s) {
      const int row = 4 * s + kq;
      float av[T];
#pragma unroll
      for (int a = 0; a < T; ++a) av[a] = tile[row * STRIDE + 16 * (T * wave + a) + rl];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float bv = ytile[row * kTHotYStride + 16 * q + rl];
#pragma unroll
        for (int a = 0; a < T; ++a) acc[a][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv, acc[a][q], 0, 0, 0);
      }
    }
  }
  const int G = gridDim.x;
#pragma unroll
  for (int a = 0; a < T; ++a)
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int slot = 16 * (T * wave + a) + 4 * kq + v;
        const int col = 16 * q + rl;
        if (col < l) partials[((int64_t)slot_base + (int64_t)slot * G + blockIdx.x) * ldp + col] = acc[a][q][v];
      }
}

template <int LGR, bool MC, typename V, typename X, typename Y, typename Acc>
__device__ __forceinline__ void spmm_block(const SpmmArgs<V, X, Y, Acc> &a, int w) {
  if (w < a.chunk_blocks) {
    const int xb = w % a.nblk;
    const int kb = w / a.nblk;
    int off = 0, cnt = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (i == xb) {
        off = a.blk_off[i];
        cnt = a.blk_cnt[i];
      }
    if (kb < cnt) spmm_chunk<LGR, MC>(a, off + kb);
  } else if (w < a.chunk_blocks + a.seg_blocks) {
    spmm_segment<LGR, MC>(a, w - a.chunk_blocks);
  } else {
    spmm_short<LGR, MC>(a, w - a.chunk_blocks - a.seg_blocks);
  }
}

template <int LGR, bool MC, typename V, typename X, typename Y, typename Acc>
__global__ __launch_bounds__(kBlock) void k_spmm(SpmmArgs<V, X, Y, Acc> a) {
  spmm_block<LGR, MC>(a, blockIdx.x);
}

// Several operators in one launch. Each operator's first workgroup is a
// multiple of 8, so w mod 8 (its XCD under round-robin dispatch) is the same
// as in a launch of its own.
template <int LGR, bool MC, typename V, typename X, typename Y, typename Acc>
__global__ __launch_bounds__(kBlock) void k_spmm_batched(const SpmmArgs<V, X, Y, Acc> *args, const int32_t *boff,
                                                        int nb) {
  const int b = batch_of(boff, nb, blockIdx.x);
  const SpmmArgs<V, X, Y, Acc> &a = args[b];
  const int w = (int)blockIdx.x - boff[b];
  if (w < a.total_blocks) spmm_block<LGR, MC>(a, w);
}

// ------------------------------------------------------------------------
// Value conversion and masked nnz
// ------------------------------------------------------------------------

template <typename S, typename D> __device__ __forceinline__ D cvt(S v) { return static_cast<D>(v); }
template <> __device__ __forceinline__ bf16_t cvt<double, bf16_t>(double v) { return f32_to_bf16((float)v); }
template <> __device__ __forceinline__ bf16_t cvt<float, bf16_t>(float v) { return f32_to_bf16(v); }

template <typename S, typename D>
__global__ void k_convert(const S *s, D *d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = cvt<S, D>(s[i]);
}

int grid_for(int64_t n, int per_block = kBlock) {
  int64_t g = (n + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 4096));
}

}  // namespace

size_t dtype_size(int dt) {
  switch (dt) {
    case SVDH_F64: return 8;
    case SVDH_F32: return 4;
    case SVDH_BF16: return 2;
  }
  fail(SVDH_ERR_ARG, "unknown dtype %d", dt);
}

// ------------------------------------------------------------------------
// Host side: operator construction
// ------------------------------------------------------------------------

static void convert_values(Ctx *c, const void *src, int sdt, void *dst, int ddt, int64_t n) {
  if (n == 0) return;
  const int g = grid_for(n);
  if (sdt == SVDH_F64 && ddt == SVDH_F32)
    hipLaunchKernelGGL((k_convert<double, float>), g, kBlock, 0, c->stream, (const double *)src, (float *)dst, n);
  else if (sdt == SVDH_F32 && ddt == SVDH_F64)
    hipLaunchKernelGGL((k_convert<float, double>), g, kBlock, 0, c->stream, (const float *)src, (double *)dst, n);
  else if (sdt == SVDH_F64 && ddt == SVDH_BF16)
    hipLaunchKernelGGL((k_convert<double, bf16_t>), g, kBlock, 0, c->stream, (const double *)src, (bf16_t *)dst, n);
  else if (sdt == SVDH_F32 && ddt == SVDH_BF16)
    hipLaunchKernelGGL((k_convert<float, bf16_t>), g, kBlock, 0, c->stream, (const float *)src, (bf16_t *)dst, n);
  else
    fail(SVDH_ERR_ARG, "unsupported value conversion %d -> %d", sdt, ddt);
  HIP_CHECK(hipGetLastError());
}

static void gather_compact_values(svdh_op *op, int dtype);

void ensure_values(svdh_op *op, int dtype) {
  svdh_op *owner = const_cast<svdh_op *>(op->base ? op->base : op);
  if (op->base && op->compact && !op->c_vals[dtype].p) {
    if (!owner->vals[dtype].p) ensure_values(owner, dtype);
    gather_compact_values(op, dtype);
  }
  if (owner->vals[dtype].p) return;
  if (owner->base_dtype == SVDH_BF16)
    fail(SVDH_ERR_ARG, "operator values are BF16; wider copies are not derivable");
  const int64_t nnz = owner->nnz;
  const size_t sz = dtype_size(dtype);
  owner->vals[dtype].alloc(std::max<size_t>(nnz * sz, 16));
  owner->t_vals[dtype].alloc(std::max<size_t>(nnz * sz, 16));
  Ctx *c = owner->ctx;
  convert_values(c, owner->vals[owner->base_dtype].p, owner->base_dtype, owner->vals[dtype].p, dtype, nnz);
  convert_values(c, owner->t_vals[owner->base_dtype].p, owner->base_dtype, owner->t_vals[dtype].p, dtype, nnz);
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace svdh

using namespace svdh;

svdh::CsrView svdh_op::view(int direction, bool pattern) const {
  CsrView v;
  const svdh_op *own = base ? base : this;
  if (direction == 0 && base && compact && !pattern) {
    v.nrows = nrows;
    v.ncols = ncols;
    v.rowptr = c_rowptr.as<int32_t>();
    v.colidx = c_colidx.as<int32_t>();
    for (int d = 0; d < 3; ++d) v.vals[d] = c_vals[d].p;
    v.h_rowptr = const_cast<std::vector<int32_t> *>(&h_c_rowptr);
    v.nnz = c_nnz;
    return v;
  }
  if (direction == 0) {
    v.nrows = nrows;
    v.ncols = ncols;
    v.rowptr = own->rowptr.as<int32_t>();
    v.colidx = own->colidx.as<int32_t>();
    for (int d = 0; d < 3; ++d) v.vals[d] = own->vals[d].p;
    v.colmap = base ? o2m.as<int32_t>() : nullptr;
    v.h_rowptr = const_cast<std::vector<int32_t> *>(&own->h_rowptr);
  } else {
    v.nrows = ncols;
    v.ncols = nrows;
    v.rowptr = own->t_rowptr.as<int32_t>();
    v.colidx = own->t_colidx.as<int32_t>();
    for (int d = 0; d < 3; ++d) v.vals[d] = own->t_vals[d].p;
    v.rowmap = base ? m2o.as<int32_t>() : nullptr;
    v.h_rowptr = const_cast<std::vector<int32_t> *>(&own->h_t_rowptr);
    v.h_rowmap = base ? &h_m2o : nullptr;
  }
  v.nnz = nnz;
  return v;
}

namespace svdh {

template <typename T>
__global__ void k_gather(int64_t n, const int32_t *perm, const T *src, T *dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

template <typename T> struct DtypeOf;
template <> struct DtypeOf<double> { static constexpr int v = SVDH_F64; };
template <> struct DtypeOf<float> { static constexpr int v = SVDH_F32; };
template <> struct DtypeOf<bf16_t> { static constexpr int v = SVDH_BF16; };

// Build (once) the schedule of one operator direction.
//   kind 0 (SpMV): bins of 2^lg lanes per row,
//                  lg = ceil(log2(ceil(nnz / kSpmvPerLane))) clamped to [0, 6];
//                  rows above kSpmvLong nnz are chunked.
static bool build_stream_schedule(svdh_op *op, const CsrView &v, Schedule &s);
static bool build_tiled_schedule(svdh_op *op, int direction, const CsrView &v, Schedule &s);

static Schedule &get_schedule(svdh_op *op, int direction, int kind) {
  Schedule &s = op->sched[direction][kind];
  if (s.built) return s;
  CsrView v = op->view(direction);
  if (kind == 0 && (build_tiled_schedule(op, direction, v, s) || build_stream_schedule(op, v, s))) return s;
  const int64_t nout = v.nrows;
  const std::vector<int32_t> &rp = *v.h_rowptr;
  std::vector<int32_t> len(nout);
  for (int64_t r = 0; r < nout; ++r) {
    const int64_t src = v.h_rowmap ? (*v.h_rowmap)[r] : r;
    len[r] = rp[src + 1] - rp[src];
  }
  std::vector<int32_t> order(nout);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return len[a] < len[b]; });
  const int long_t = kSpmvLong;
  const int chunk = kSpmvChunk;
  int64_t nshort = 0;
  while (nshort < nout && len[order[nshort]] <= long_t) ++nshort;
  // Long-row chunks.
  std::vector<int32_t> ch;
  int32_t nlong = 0;
  for (int64_t i = nshort; i < nout; ++i) {
    const int32_t r = order[i];
    const int64_t src = v.h_rowmap ? (*v.h_rowmap)[r] : r;
    const int32_t b = rp[src], e = rp[src + 1];
    const int32_t n = (e - b + chunk - 1) / chunk;
    const int32_t first = (int32_t)(ch.size() / 6);
    for (int32_t k = 0; k < n; ++k) {
      ch.push_back(r);
      ch.push_back(b + k * chunk);
      ch.push_back(std::min(e, b + (k + 1) * chunk));
      ch.push_back(first);
      ch.push_back(n);
      ch.push_back(nlong);
    }
    ++nlong;
  }
  s.nchunks = (int32_t)(ch.size() / 6);
  s.nlong_rows = nlong;
  s.nbins = 0;
  int32_t blocks = 0;
  if (kind == 0) {
    int64_t i = 0;
    for (int lg = 0; lg <= 6 && i < nshort; ++lg) {
      const int64_t lim = lg == 6 ? long_t : ((int64_t)kSpmvPerLane << lg);
      int64_t j = i;
      while (j < nshort && len[order[j]] <= lim) ++j;
      if (j > i) {
        BinDesc &d = s.bins[s.nbins++];
        d.block_begin = blocks;
        d.row_begin = (int32_t)i;
        d.row_count = (int32_t)(j - i);
        d.lg = lg;
        blocks += (int32_t)((d.row_count + (kBlock >> lg) - 1) / (kBlock >> lg));
      }
      i = j;
    }
  } else {
    BinDesc &d = s.bins[s.nbins++];
    d.block_begin = 0;
    d.row_begin = 0;
    d.row_count = (int32_t)nshort;
    d.lg = 0;
  }
  s.long_block_begin = blocks;
  s.total_blocks = blocks + s.nchunks;
  Ctx *c = op->ctx;
  s.rows.alloc(std::max<int64_t>(nout, 1) * sizeof(int32_t));
  if (nout) HIP_CHECK(hipMemcpy(s.rows.p, order.data(), nout * sizeof(int32_t), hipMemcpyHostToDevice));
  s.chunks.alloc(std::max<size_t>(ch.size(), 6) * sizeof(int32_t));
  if (!ch.empty()) HIP_CHECK(hipMemcpy(s.chunks.p, ch.data(), ch.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  // rows with several chunks: r, first chunk, nchunks (for the second pass)
  std::vector<int32_t> multi;
  for (size_t k = 0; k < ch.size(); k += 6)
    if (ch[k + 4] > 1 && (int32_t)(k / 6) == ch[k + 3]) {
      multi.push_back(ch[k]);
      multi.push_back(ch[k + 3]);
      multi.push_back(ch[k + 4]);
    }
  s.nslots = (int32_t)(multi.size() / 3);
  s.counters.alloc(std::max<size_t>(multi.size(), 3) * sizeof(int32_t));
  if (!multi.empty())
    HIP_CHECK(hipMemcpy(s.counters.p, multi.data(), multi.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  (void)c;
  s.built = true;
  return s;
}

// Source-tiled SpMV schedule (see k_spmv_tiled), for plain operators with at
// least kTiledMinNnz nonzeros (SVDH_SPMV_TILED=0 turns it off). Source blocks
// of equal width (a multiple of 64, at most kTiledSrc); each output row's
// entries are split by block, keeping their order, and the pieces stacked
// block-major. Tasks over the stacked rows of each block are cut as in the
// streamed SpMV (runs of rows within one step, longer rows alone), and the
// tasks of each block are dealt to workgroup ranges of about
// nnz / (SVDH_SPMV_TILED_WG x CUs) nonzeros (default 2).
static bool build_tiled_schedule(svdh_op *op, int direction, const CsrView &v, Schedule &s) {
  if (op->base || v.rowmap || v.colmap || v.nnz < kTiledMinNnz) return false;
  const char *env = std::getenv("SVDH_SPMV_TILED");
  if (env && std::atoi(env) == 0) return false;
  const std::vector<int32_t> &ci = direction == 0 ? op->h_colidx : op->h_t_colidx;
  const std::vector<int32_t> &rp = *v.h_rowptr;
  const int64_t nout = v.nrows, nsrc = v.ncols, nnz = v.nnz;
  if ((int64_t)ci.size() != nnz || nsrc < 1 || nout < 1) return false;
  const int64_t nb = (nsrc + kTiledSrc - 1) / kTiledSrc;
  const int64_t sb = ((nsrc + nb - 1) / nb + 63) / 64 * 64;
  if (nb * nout + 1 > INT32_MAX) return false;
  // stacked row pointers
  std::vector<int32_t> srp(nb * nout + 1, 0);
  for (int64_t r = 0; r < nout; ++r)
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) ++srp[(ci[k] / sb) * nout + r + 1];
  for (int64_t q = 0; q < nb * nout; ++q) {
    if (srp[q + 1] > kTiledMaxRow) return false;
    srp[q + 1] += srp[q];
  }
  std::vector<int32_t> next(srp.begin(), srp.end() - 1);
  std::vector<uint16_t> col(nnz);
  std::vector<int32_t> perm(nnz);
  for (int64_t r = 0; r < nout; ++r)
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
      const int64_t b = ci[k] / sb;
      const int32_t p = next[b * nout + r]++;
      col[p] = (uint16_t)(ci[k] - b * sb);
      perm[p] = k;
    }
  // tasks per block, then workgroup ranges
  std::vector<int4> tasks;
  std::vector<int4> ranges;
  const char *wenv = std::getenv("SVDH_SPMV_TILED_WG");
  const double per_cu = wenv ? std::max(0.25, std::atof(wenv)) : 2.0;
  const int64_t target = std::max<int64_t>(kTiledStep, (int64_t)(nnz / (per_cu * op->ctx->ncu)));
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t q0 = b * nout, q1 = q0 + nout;
    const int32_t tfirst = (int32_t)tasks.size();
    int64_t q = q0;
    while (q < q1) {
      const int32_t len = srp[q + 1] - srp[q];
      if (len > kTiledStep) {
        tasks.push_back(make_int4((int)q, 0, srp[q], srp[q + 1]));
        ++q;
        continue;
      }
      const int64_t qa = q;
      const int32_t kb = srp[q];
      while (q < q1 && q - qa < kStreamRows && srp[q + 1] - kb <= kTiledStep) ++q;
      tasks.push_back(make_int4((int)qa, (int)q, kb, srp[q]));
    }
    // ranges of about `target` nonzeros, never crossing a block
    int32_t t0 = tfirst;
    int64_t acc = 0;
    for (int32_t t = tfirst; t < (int32_t)tasks.size(); ++t) {
      acc += tasks[t].w - tasks[t].z;
      if (acc >= target) {
        ranges.push_back(make_int4((int)b, t0, t + 1, 0));
        t0 = t + 1;
        acc = 0;
      }
    }
    if (t0 < (int32_t)tasks.size()) ranges.push_back(make_int4((int)b, t0, (int)tasks.size(), 0));
  }
  Schedule::Tiled &T = s.tiled;
  T.nb = (int32_t)nb;
  T.sb = (int32_t)sb;
  T.ntasks = (int32_t)tasks.size();
  T.nranges = (int32_t)ranges.size();
  auto up = [](DevBuf &d, const void *src, size_t bytes) {
    d.alloc(std::max<size_t>(bytes, 16));
    if (bytes) HIP_CHECK(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
  };
  try {  // the layout is a second copy of the matrix: without room for it, the streamed SpMV runs instead
    up(T.rowptr, srp.data(), srp.size() * sizeof(int32_t));
    up(T.col, col.data(), col.size() * sizeof(uint16_t));
    up(T.perm, perm.data(), perm.size() * sizeof(int32_t));
    up(T.tasks, tasks.data(), tasks.size() * sizeof(int4));
    up(T.ranges, ranges.data(), ranges.size() * sizeof(int4));
    if (nb > 1) T.part.alloc((size_t)(nb * nout) * sizeof(double));
    // the values of the operator's own dtype in the stacked order (others on first use)
    const int bdt = op->base_dtype;
    if (v.vals[bdt]) {
      DevBuf &d = T.vals[bdt];
      const size_t es = bdt == SVDH_F64 ? 8 : bdt == SVDH_F32 ? 4 : 2;
      d.alloc(std::max<size_t>((size_t)nnz * es, 16));
      const int g = grid_for(nnz);
      if (bdt == SVDH_F64)
        hipLaunchKernelGGL(k_gather<double>, g, kBlock, 0, op->ctx->stream, nnz, T.perm.as<int32_t>(),
                           static_cast<const double *>(v.vals[bdt]), d.as<double>());
      else if (bdt == SVDH_F32)
        hipLaunchKernelGGL(k_gather<float>, g, kBlock, 0, op->ctx->stream, nnz, T.perm.as<int32_t>(),
                           static_cast<const float *>(v.vals[bdt]), d.as<float>());
      else
        hipLaunchKernelGGL(k_gather<bf16_t>, g, kBlock, 0, op->ctx->stream, nnz, T.perm.as<int32_t>(),
                           static_cast<const bf16_t *>(v.vals[bdt]), d.as<bf16_t>());
      HIP_CHECK(hipGetLastError());
    }
  } catch (const Error &) {
    T = Schedule::Tiled();
    return false;
  }
  T.on = true;
  s.built = true;
  return true;
}

// Values of dtype V in the stacked order of the tiled schedule.
template <typename V>
static const V *tiled_values(Ctx *c, Schedule::Tiled &T, const V *vals, int64_t nnz) {
  DevBuf &d = T.vals[DtypeOf<V>::v];
  if (!d.p) {
    d.alloc(std::max<size_t>(nnz * sizeof(V), 16));
    hipLaunchKernelGGL(k_gather<V>, grid_for(nnz), kBlock, 0, c->stream, nnz, T.perm.as<int32_t>(), vals,
                       d.as<V>());
    HIP_CHECK(hipGetLastError());
  }
  return d.as<V>();
}

// Streamed SpMV tasks (rows in storage order; see k_spmv_stream): runs of
// consecutive rows whose nonzeros fit in one step (at most kStreamRows rows),
// and rows longer than a step as tasks of their own, cut into chunks of
// kStreamChunk nonzeros. Not used for views with a row map (a masked
// operator's transposed direction) or when SVDH_SPMV_STREAM=0.
static bool build_stream_schedule(svdh_op *op, const CsrView &v, Schedule &s) {
  if (v.h_rowmap) return false;
  const char *env = std::getenv("SVDH_SPMV_STREAM");
  if (env && std::atoi(env) == 0) return false;
  const int64_t nout = v.nrows;
  const std::vector<int32_t> &rp = *v.h_rowptr;
  std::vector<int4> tasks;
  std::vector<int32_t> multi;
  int32_t nslots = 0;
  int64_t r = 0;
  while (r < nout) {
    const int32_t len = rp[r + 1] - rp[r];
    if (len > kStreamStep) {
      const int32_t nch = (len + kStreamChunk - 1) / kStreamChunk;
      if (nch == 1) {
        tasks.push_back(make_int4((int)r, 0, rp[r], rp[r + 1]));
      } else {
        multi.push_back((int32_t)r);
        multi.push_back(nslots);
        multi.push_back(nch);
        for (int32_t k = 0; k < nch; ++k)
          tasks.push_back(make_int4((int)r, -1 - (nslots++), rp[r] + k * kStreamChunk,
                                    std::min(rp[r + 1], rp[r] + (k + 1) * kStreamChunk)));
      }
      ++r;
      continue;
    }
    const int64_t ra = r;
    const int32_t kb = rp[r];
    while (r < nout && r - ra < kStreamRows && rp[r + 1] - kb <= kStreamStep) ++r;
    tasks.push_back(make_int4((int)ra, (int)r, kb, rp[r]));
  }
  s.stream = true;
  s.ntasks = (int32_t)tasks.size();
  s.tasks.alloc(std::max<size_t>(tasks.size(), 1) * sizeof(int4));
  if (!tasks.empty())
    HIP_CHECK(hipMemcpy(s.tasks.p, tasks.data(), tasks.size() * sizeof(int4), hipMemcpyHostToDevice));
  s.nchunks = nslots;
  s.nslots = (int32_t)(multi.size() / 3);
  s.counters.alloc(std::max<size_t>(multi.size(), 3) * sizeof(int32_t));
  if (!multi.empty())
    HIP_CHECK(hipMemcpy(s.counters.p, multi.data(), multi.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  (void)op;
  s.built = true;
  return true;
}

// SpMM schedule parameters (defaults; SVDH_SPMM_LONG / _CHUNK / _SPLIT /
// _NBLK override them for tuning).
// Defaults measured on the 100,000 x 20,000 Zipf workload (width 60, FP32):
// A^T Y 195 us with these, 356 us without source blocks, 660 us when chunk
// partials were combined inside the kernel.
struct SpmmTune {
  int long_nnz = 256;      // rows above this nnz are cut into chunks
  int chunk = 1024;        // nnz per chunk (one workgroup)
  int split_min = 512;     // rows with at least this nnz are cut at source blocks
  int nblk = 8;            // source blocks when the panel exceeds one XCD's L2
};

static SpmmTune spmm_tune() {
  SpmmTune t;
  auto get = [](const char *k, int d) {
    const char *e = std::getenv(k);
    return e ? std::atoi(e) : d;
  };
  t.long_nnz = std::max(1, get("SVDH_SPMM_LONG", t.long_nnz));
  t.chunk = std::max(64, get("SVDH_SPMM_CHUNK", t.chunk));
  t.split_min = std::max(1, get("SVDH_SPMM_SPLIT", t.split_min));
  t.nblk = std::max(1, std::min(8, get("SVDH_SPMM_NBLK", t.nblk)));
  return t;
}

// Build (once) the SpMM schedule of one operator direction (see k_spmm).
// Source blocks: 8 when the operand panel has at least 65,536 rows and the
// source indices of every long row are sorted, otherwise 1.
// hot (optional): output rows left to the matrix-core pass of the staged
// columns; they get no chunks or segments here, only hot_g partial slots each
// (consecutive, in workgroup order) and an entry in the multi-row table, so
// the fin pass sums them.
static void build_spmm_schedule(svdh_op *op, int direction, Schedule &s, const std::vector<int32_t> *hot, int hot_g) {
  CsrView v = op->view(direction);
  const svdh_op *own = op->base ? op->base : op;
  const std::vector<int32_t> &rp = *v.h_rowptr;
  const std::vector<int32_t> &ci = direction == 0 ? own->h_colidx : own->h_t_colidx;
  const int64_t nout = v.nrows, nsrc = v.ncols;
  auto src_of = [&](int64_t r) { return v.h_rowmap ? (int64_t)(*v.h_rowmap)[r] : r; };
  const SpmmTune tune = spmm_tune();
  int nblk = (nsrc >= 65536 && !(direction == 0 && op->base)) ? tune.nblk : 1;
  // every row's source indices must be sorted for the block cuts
  for (int64_t r = 0; r < nout && nblk > 1; ++r) {
    const int64_t sr = src_of(r);
    for (int32_t k = rp[sr] + 1; k < rp[sr + 1]; ++k)
      if (ci[k - 1] > ci[k]) {
        nblk = 1;
        break;
      }
  }
  // With source blocks, rows below split_min become per-block segments (one
  // row group each) and longer rows chunks cut at the block bounds; without
  // them, rows above long_nnz are chunked and the rest are short rows.
  std::vector<int32_t> shortr, longr, segr;
  std::vector<char> skip(hot ? nout : 0, 0);
  if (hot)
    for (int32_t r : *hot) skip[r] = 1;
  for (int64_t r = 0; r < nout; ++r) {
    if (hot && skip[r]) continue;
    const int64_t sr = src_of(r);
    const int32_t len = rp[sr + 1] - rp[sr];
    if (nblk > 1)
      (len >= tune.split_min ? longr : segr).push_back((int32_t)r);
    else
      (len > tune.long_nnz ? longr : shortr).push_back((int32_t)r);
  }
  std::stable_sort(shortr.begin(), shortr.end(), [&](int32_t a, int32_t b) {
    return rp[src_of(a) + 1] - rp[src_of(a)] > rp[src_of(b) + 1] - rp[src_of(b)];
  });
  struct Ch { int32_t r, b, e, slot, first, n, rid, blk; };
  std::vector<std::vector<Ch>> per(nblk);
  std::vector<Ch> row;
  int32_t nslots = 0, nmulti = 0;
  int64_t rr = 0;  // round-robin list of chunks of rows that are not cut at source blocks
  for (int32_t r : longr) {
    const int64_t sr = src_of(r);
    const int32_t b0 = rp[sr], e0 = rp[sr + 1];
    row.clear();
    int32_t lo = b0;
    const int nb = (e0 - b0 >= tune.split_min) ? nblk : 1;
    for (int x = 0; x < nb; ++x) {
      int32_t hi = e0;
      if (x + 1 < nb) {
        const int32_t bound = (int32_t)(nsrc * (x + 1) / nb);
        hi = (int32_t)(std::lower_bound(ci.begin() + lo, ci.begin() + e0, bound) - ci.begin());
      }
      const int32_t len = hi - lo;
      const int32_t pieces = (len + tune.chunk - 1) / tune.chunk;
      for (int32_t p = 0; p < pieces; ++p) {
        const int32_t pb = lo + (int32_t)((int64_t)len * p / pieces);
        const int32_t pe = lo + (int32_t)((int64_t)len * (p + 1) / pieces);
        row.push_back(Ch{r, pb, pe, 0, 0, 0, 0, nb > 1 ? x : (int)(rr++ % nblk)});
      }
      lo = hi;
    }
    const int32_t n = (int32_t)row.size();
    const int32_t first = n > 1 ? nslots : 0;
    const int32_t rid = n > 1 ? nmulti++ : 0;
    for (int32_t i = 0; i < n; ++i) {
      Ch g = row[i];
      g.n = n;
      g.first = first;
      g.slot = n > 1 ? first + i : 0;
      g.rid = rid;
      per[g.blk].push_back(g);
    }
    if (n > 1) nslots += n;
  }
  std::vector<int32_t> multi;  // r, first_slot, nchunks of every row with several chunks
  for (auto &L : per)
    for (const Ch &g : L)
      if (g.n > 1 && g.slot == g.first) {
        multi.push_back(g.r);
        multi.push_back(g.first);
        multi.push_back(g.n);
      }
  s.chunk_multi = (int32_t)(multi.size() / 3);  // rows cut into chunks come first in the table
  // segments of the short rows, one per non-empty source block
  std::vector<std::vector<int32_t>> segper(nblk);
  for (int32_t r : segr) {
    const int64_t sr = src_of(r);
    const int32_t b0 = rp[sr], e0 = rp[sr + 1];
    int32_t cuts[9];
    cuts[0] = b0;
    for (int x = 0; x < nblk; ++x) {
      int32_t hi = e0;
      if (x + 1 < nblk) {
        const int32_t bound = (int32_t)(nsrc * (x + 1) / nblk);
        hi = (int32_t)(std::lower_bound(ci.begin() + cuts[x], ci.begin() + e0, bound) - ci.begin());
      }
      cuts[x + 1] = hi;
    }
    int nseg = 0;
    for (int x = 0; x < nblk; ++x) nseg += cuts[x + 1] > cuts[x];
    if (nseg == 0) {  // empty row: one segment writes its zeros
      segper[0].insert(segper[0].end(), {r, b0, e0, -1});
      continue;
    }
    const int32_t first = nslots;
    for (int x = 0; x < nblk; ++x)
      if (cuts[x + 1] > cuts[x]) segper[x].insert(segper[x].end(), {r, cuts[x], cuts[x + 1], nseg > 1 ? nslots++ : -1});
    if (nseg > 1) {
      multi.push_back(r);
      multi.push_back(first);
      multi.push_back(nseg);
      ++nmulti;
    }
  }
  s.hot_slot_base = nslots;
  s.hot_multi = hot ? (int32_t)hot->size() : 0;
  if (hot)
    for (int32_t r : *hot) {
      multi.push_back(r);
      multi.push_back(nslots);
      multi.push_back(hot_g);
      nslots += hot_g;
      ++nmulti;
    }
  // longest segments first within each block, so the row groups of a wave
  // (consecutive segments) have nearly equal trip counts
  for (auto &v : segper) {
    const size_t ns = v.size() / kSegInts;
    std::vector<int32_t> idx(ns);
    for (size_t i = 0; i < ns; ++i) idx[i] = (int32_t)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) {
      return v[x * kSegInts + 2] - v[x * kSegInts + 1] > v[y * kSegInts + 2] - v[y * kSegInts + 1];
    });
    std::vector<int32_t> sorted(v.size());
    for (size_t i = 0; i < ns; ++i)
      std::copy(v.begin() + (size_t)idx[i] * kSegInts, v.begin() + (size_t)(idx[i] + 1) * kSegInts,
                sorted.begin() + i * kSegInts);
    v.swap(sorted);
  }
  {
    std::vector<int32_t> sflat;
    int32_t soff = 0;
    for (int x = 0; x < 8; ++x) {
      s.seg_off[x] = 0;
      s.seg_cnt[x] = 0;
    }
    for (int x = 0; x < nblk; ++x) {
      s.seg_off[x] = soff;
      s.seg_cnt[x] = (int32_t)(segper[x].size() / kSegInts);
      soff += s.seg_cnt[x];
      sflat.insert(sflat.end(), segper[x].begin(), segper[x].end());
    }
    s.segs.alloc(std::max<size_t>(sflat.size(), kSegInts) * sizeof(int32_t));
    if (!sflat.empty())
      HIP_CHECK(hipMemcpy(s.segs.p, sflat.data(), sflat.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  std::vector<int32_t> flat;
  int32_t off = 0, maxcnt = 0;
  s.nblk = nblk;
  for (int x = 0; x < 8; ++x) {
    s.blk_off[x] = 0;
    s.blk_cnt[x] = 0;
  }
  for (int x = 0; x < nblk; ++x) {
    auto &L = per[x];
    std::stable_sort(L.begin(), L.end(), [](const Ch &a, const Ch &b) { return (a.e - a.b) > (b.e - b.b); });
    s.blk_off[x] = off;
    s.blk_cnt[x] = (int32_t)L.size();
    maxcnt = std::max(maxcnt, (int32_t)L.size());
    for (const Ch &g : L) {
      const int32_t w[kChunkInts] = {g.r, g.b, g.e, g.slot, g.first, g.n, g.rid, 0};
      flat.insert(flat.end(), w, w + kChunkInts);
    }
    off += (int32_t)L.size();
  }
  s.nchunks = off;
  s.total_blocks = maxcnt * nblk;  // chunk blocks; short-row blocks depend on l
  s.nslots = nslots;
  s.nlong_rows = nmulti;
  s.bins[0].row_count = (int32_t)shortr.size();
  s.rows.alloc(std::max<size_t>(shortr.size(), 1) * sizeof(int32_t));
  if (!shortr.empty())
    HIP_CHECK(hipMemcpy(s.rows.p, shortr.data(), shortr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  s.chunks.alloc(std::max<size_t>(flat.size(), kChunkInts) * sizeof(int32_t));
  if (!flat.empty()) HIP_CHECK(hipMemcpy(s.chunks.p, flat.data(), flat.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  s.counters.alloc(std::max<size_t>(multi.size(), 3) * sizeof(int32_t));
  if (!multi.empty())
    HIP_CHECK(hipMemcpy(s.counters.p, multi.data(), multi.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  s.built = true;
}

static Schedule &get_spmm_schedule(svdh_op *op, int direction) {
  Schedule &s = op->sched[direction][1];
  if (!s.built) build_spmm_schedule(op, direction, s, nullptr, 0);
  return s;
}

// scr (allocation-free calls, spmv_noalloc): the workgroup partials go to
// this device buffer instead of the schedule's own, the schedule and value
// layouts must exist already (checked by the caller), and no timer events
// are taken.
template <typename V, typename X, typename Y>
static void launch_spmv(svdh_op *op, int direction, const void *vals, const void *x, void *y,
                        double *scr = nullptr) {
  Ctx *c = op->ctx;
  CsrView v = op->view(direction);
  Schedule &s = get_schedule(op, direction, 0);
  if (v.nrows == 0) return;
  if (!scr) s.partials.ensure(std::max<int32_t>(s.nchunks, 1) * sizeof(double));
  double *partials = scr ? scr : s.partials.as<double>();
  Ctx *tc = scr ? nullptr : c;  // timer context
  const double bytes = (double)v.nnz * (sizeof(V) + 4) + (double)(v.nrows + 1) * 4 +
                       (double)v.ncols * sizeof(X) + (double)v.nrows * sizeof(Y);
  if (s.tiled.on) {
    Schedule::Tiled &T = s.tiled;
    double *part = scr ? scr : T.part.as<double>();
    TiledArgs<V, X, Y> a;
    a.rowptr = T.rowptr.as<int32_t>();
    a.col = T.col.as<uint16_t>();
    a.vals = scr ? static_cast<const V *>(T.vals[DtypeOf<V>::v].p)
                 : tiled_values<V>(c, T, static_cast<const V *>(vals), v.nnz);
    a.tasks = T.tasks.as<int4>();
    a.ranges = T.ranges.as<int4>();
    a.x = static_cast<const X *>(x);
    a.part = part;
    a.y = static_cast<Y *>(y);
    a.nsrc = v.ncols;
    a.sb = T.sb;
    a.direct = T.nb == 1;
    KTimer t(tc, KC_SPMV, bytes);
    hipLaunchKernelGGL((k_spmv_tiled<V, X, Y>), T.nranges, kTiledBlock, 0, c->stream, a);
    HIP_CHECK(hipGetLastError());
    if (T.nb > 1) {
      if (T.nb >= 8)
        hipLaunchKernelGGL((k_spmv_tiled_fin<4, Y>), (int)((4 * v.nrows + kBlock - 1) / kBlock), kBlock, 0,
                           c->stream, v.nrows, (int)T.nb, (const double *)part, static_cast<Y *>(y));
      else
        hipLaunchKernelGGL((k_spmv_tiled_fin<1, Y>), (int)((v.nrows + kBlock - 1) / kBlock), kBlock, 0,
                           c->stream, v.nrows, (int)T.nb, (const double *)part, static_cast<Y *>(y));
      HIP_CHECK(hipGetLastError());
    }
    return;
  }
  if (s.stream) {
    StreamArgs<V, X, Y> a;
    a.rowptr = v.rowptr;
    a.colidx = v.colidx;
    a.vals = static_cast<const V *>(vals);
    a.colmap = v.colmap;
    a.tasks = s.tasks.as<int4>();
    a.partials = partials;
    a.x = static_cast<const X *>(x);
    a.y = static_cast<Y *>(y);
    a.ntasks = s.ntasks;
    const int blocks = (s.ntasks + kBlock / 64 - 1) / (kBlock / 64);
    KTimer t(tc, KC_SPMV, bytes);
    if (v.colmap)
      hipLaunchKernelGGL((k_spmv_stream<true, V, X, Y>), blocks, kBlock, 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_spmv_stream<false, V, X, Y>), blocks, kBlock, 0, c->stream, a);
    HIP_CHECK(hipGetLastError());
    if (s.nslots > 0) {
      hipLaunchKernelGGL((k_spmv_fin<Y>), (s.nslots + kBlock - 1) / kBlock, kBlock, 0, c->stream, s.nslots,
                         s.counters.as<int32_t>(), (const double *)partials, static_cast<Y *>(y));
      HIP_CHECK(hipGetLastError());
    }
    return;
  }
  SpmvArgs<V, X, Y> a;
  a.rowptr = v.rowptr;
  a.colidx = v.colidx;
  a.vals = static_cast<const V *>(vals);
  a.rowmap = v.rowmap;
  a.colmap = v.colmap;
  a.rows = s.rows.as<int32_t>();
  a.chunks = s.chunks.as<int32_t>();
  a.partials = partials;
  a.x = static_cast<const X *>(x);
  a.y = static_cast<Y *>(y);
  a.nbins = s.nbins;
  a.long_block_begin = s.long_block_begin;
  for (int i = 0; i < kMaxBins; ++i) a.bins[i] = i < s.nbins ? s.bins[i] : BinDesc{INT32_MAX, 0, 0, 0};
  if (s.nbins == 0) a.bins[0] = BinDesc{INT32_MAX, 0, 0, 0};
  KTimer t(tc, KC_SPMV, bytes);
  if (v.colmap)
    hipLaunchKernelGGL((k_spmv<true, V, X, Y>), s.total_blocks, kBlock, 0, c->stream, a);
  else
    hipLaunchKernelGGL((k_spmv<false, V, X, Y>), s.total_blocks, kBlock, 0, c->stream, a);
  HIP_CHECK(hipGetLastError());
  if (s.nslots > 0) {
    hipLaunchKernelGGL((k_spmv_fin<Y>), (s.nslots + kBlock - 1) / kBlock, kBlock, 0, c->stream, s.nslots,
                       s.counters.as<int32_t>(), (const double *)partials, static_cast<Y *>(y));
    HIP_CHECK(hipGetLastError());
  }
}

// ------------------------------------------------------------------------
// Allocation-free SpMV (svdh_spmv_noalloc). Its workgroup partials live in a
// file-scope device array of fixed size; each GPU holds its own copy, whose
// device address is looked up once per device (scratch_init, at context
// creation with the device current) and passed to the kernels. Every partial
// a launch reads is written earlier in the same launch sequence (each
// stacked row / chunk task writes its slot), so nothing needs clearing
// between calls; calls that overlap on one GPU race on the array.
// ------------------------------------------------------------------------
constexpr size_t kScratchDoubles = (size_t)2 << 20;  // 16 MiB of FP64 partials
static __device__ __align__(256) double scratch[kScratchDoubles];

void scratch_init(Ctx *c) {
  constexpr int kMaxDev = 64;
  static std::mutex mu;
  static void *addr[kMaxDev] = {};
  if (c->device < 0 || c->device >= kMaxDev) fail(SVDH_ERR_ARG, "device %d out of range", c->device);
  std::lock_guard<std::mutex> g(mu);
  if (!addr[c->device]) {
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(SVDH_ERR_HIP, "no device address for the SpMV scratch");
    addr[c->device] = p;
  }
  c->scratch = static_cast<double *>(addr[c->device]);
}

void spmv_noalloc(svdh_op *op, int direction, int vdt, const void *x, void *y, int xdt) {
  Ctx *c = op->ctx;
  if (!c->scratch) fail(SVDH_ERR_HIP, "the SpMV scratch is not set up");
  const CsrView v = op->view(direction);
  if (v.nrows == 0) return;
  const Schedule &s = op->sched[direction][0];
  const void *vals = op->values(direction, vdt);
  const char *how = "call svdh_spmv once with the same operator, direction and vector dtype first";
  if (!s.built || !vals) fail(SVDH_ERR_ARG, "operator not prepared for this product: %s", how);
  size_t need = (size_t)std::max<int32_t>(s.nchunks, s.nslots);
  if (s.tiled.on) {
    if (!s.tiled.vals[vdt].p) fail(SVDH_ERR_ARG, "operator not prepared for this product: %s", how);
    need = s.tiled.nb > 1 ? (size_t)s.tiled.nb * (size_t)v.nrows : 0;
  }
  if (need > kScratchDoubles)
    fail(SVDH_ERR_ARG, "this product needs %zu FP64 partials; the allocation-free SpMV holds at most %zu (%zu MiB)",
         need, kScratchDoubles, kScratchDoubles * sizeof(double) >> 20);
  if (vdt == SVDH_F64 && xdt == SVDH_F64)
    launch_spmv<double, double, double>(op, direction, vals, x, y, c->scratch);
  else if (vdt == SVDH_F32 && xdt == SVDH_F64)
    launch_spmv<float, double, double>(op, direction, vals, x, y, c->scratch);
  else if (vdt == SVDH_F32 && xdt == SVDH_F32)
    launch_spmv<float, float, float>(op, direction, vals, x, y, c->scratch);
  else if (vdt == SVDH_F64 && xdt == SVDH_F32)
    launch_spmv<double, float, float>(op, direction, vals, x, y, c->scratch);
  else
    fail(SVDH_ERR_ARG, "unsupported SpMV dtypes values=%d vectors=%d", vdt, xdt);
}

void spmv(svdh_op *op, int direction, int vdt, const void *x, void *y, int xdt) {
  ensure_values(op, vdt);
  const void *vals = op->values(direction, vdt);
  if (vdt == SVDH_F64 && xdt == SVDH_F64)
    launch_spmv<double, double, double>(op, direction, vals, x, y);
  else if (vdt == SVDH_F32 && xdt == SVDH_F64)
    launch_spmv<float, double, double>(op, direction, vals, x, y);
  else if (vdt == SVDH_F32 && xdt == SVDH_F32)
    launch_spmv<float, float, float>(op, direction, vals, x, y);
  else if (vdt == SVDH_F64 && xdt == SVDH_F32)
    launch_spmv<double, float, float>(op, direction, vals, x, y);
  else
    fail(SVDH_ERR_ARG, "unsupported SpMV dtypes values=%d vectors=%d", vdt, xdt);
}

template <int LGR, bool MC, typename V, typename X, typename Y, typename Acc>
static void spmm_launch_one(Ctx *c, const SpmmArgs<V, X, Y, Acc> &a, int blocks) {
  hipLaunchKernelGGL((k_spmm<LGR, MC, V, X, Y, Acc>), blocks, kBlock, 0, c->stream, a);
}

template <bool MC, typename V, typename X, typename Y, typename Acc>
static void spmm_dispatch(Ctx *c, const SpmmArgs<V, X, Y, Acc> &a, int lgr, int blocks) {
  switch (lgr) {
    case 1: spmm_launch_one<1, MC>(c, a, blocks); break;
    case 2: spmm_launch_one<2, MC>(c, a, blocks); break;
    case 3: spmm_launch_one<3, MC>(c, a, blocks); break;
    case 4: spmm_launch_one<4, MC>(c, a, blocks); break;
    case 5: spmm_launch_one<5, MC>(c, a, blocks); break;
    default: spmm_launch_one<6, MC>(c, a, blocks); break;
  }
}

template <typename V, typename X, typename Y, typename Acc>
struct SpmmPlan {
  SpmmArgs<V, X, Y, Acc> a;
  Schedule *s = nullptr;
  int lgr = 1, blocks = 0;
  bool mc = false;
  double bytes = 0.0;
};

// Arguments and launch shape of one SpMM (false: nothing to compute).
template <typename V, typename X, typename Y, typename Acc>
static bool plan_spmm(svdh_op *op, int direction, const void *vals, const void *X_, int64_t ldx, void *Y_,
                      int64_t ldy, int64_t l, SpmmPlan<V, X, Y, Acc> &p, Schedule *sched = nullptr) {
  CsrView v = op->view(direction);
  Schedule &s = sched ? *sched : get_spmm_schedule(op, direction);
  if (v.nrows == 0 || l == 0) return false;
  constexpr int NV = PanelVec<X>::N;
  if (ldx % NV || ldy % NV) fail(SVDH_ERR_ARG, "panel leading dimensions must be multiples of %d", NV);
  // lanes per row group: enough to cover l with NV columns each, 2..64
  int lgr = 1;
  while ((1 << lgr) * NV < l && lgr < 6) ++lgr;
  const int rpb = kBlock >> lgr;
  const int nshort = s.bins[0].row_count;
  int seg_max = 0;
  for (int x = 0; x < s.nblk; ++x) seg_max = std::max(seg_max, (s.seg_cnt[x] + rpb - 1) / rpb);
  const int seg_blocks = seg_max * s.nblk;
  const int blocks = s.total_blocks + seg_blocks + (nshort + rpb - 1) / rpb;
  const int64_t ldp = (l + 7) / 8 * 8;
  s.partials.ensure((size_t)std::max<int32_t>(s.nslots, 1) * ldp * sizeof(Acc));
  SpmmArgs<V, X, Y, Acc> &a = p.a;
  a.segs = s.segs.as<int32_t>();
  a.seg_blocks = seg_blocks;
  for (int x = 0; x < 8; ++x) {
    a.seg_off[x] = s.seg_off[x];
    a.seg_cnt[x] = x < s.nblk ? s.seg_cnt[x] : 0;
  }
  a.rowptr = v.rowptr;
  a.colidx = v.colidx;
  a.vals = static_cast<const V *>(vals);
  a.rowmap = v.rowmap;
  a.colmap = v.colmap;
  a.rows = s.rows.as<int32_t>();
  a.chunks = s.chunks.as<int32_t>();
  a.partials = s.partials.as<Acc>();
  a.x = static_cast<const X *>(X_);
  a.y = static_cast<Y *>(Y_);
  a.ldx = ldx;
  a.ldy = ldy;
  a.ldp = ldp;
  a.l = (int32_t)l;
  a.nshort = nshort;
  a.nblk = s.nblk;
  a.chunk_blocks = s.total_blocks;
  a.total_blocks = blocks;
  for (int x = 0; x < 8; ++x) {
    a.blk_off[x] = s.blk_off[x];
    a.blk_cnt[x] = x < s.nblk ? s.blk_cnt[x] : 0;
  }
  a.off32 = (uint64_t)std::max<int64_t>(v.ncols, 1) * (uint64_t)ldx * sizeof(X) < (1ull << 31) ? 1 : 0;
  p.s = &s;
  p.lgr = lgr;
  p.blocks = blocks;
  p.mc = v.colmap != nullptr;
  p.bytes = (double)v.nnz * (sizeof(V) + 4) + (double)(v.nrows + 1) * 4 + (double)v.ncols * l * sizeof(X) +
            (double)v.nrows * l * sizeof(Y);
  return blocks > 0;
}

// Staged hot columns for A x of a plain operator whose rows are all short:
// h = as many panel rows as fit 64 KB of LDS, used when the h most popular
// columns hold at least a fifth of the nonzeros. Returns the encoding for h,
// or null when staging is not used. The popularity order is computed once per
// operator; encodings for two values of h (FP32 and BF16 panels) are kept.
static Schedule::HotEnc *hot_columns(svdh_op *op, Schedule &s, int h) {
  if (op->base || s.total_blocks > 0 || s.nblk > 1 || h < 16 || op->ncols < 4 * (int64_t)h) return nullptr;
  const char *env = std::getenv("SVDH_SPMM_HOT");
  if (env && std::atoi(env) == 0) return nullptr;
  if (s.hot_state == 0) return nullptr;
  const int64_t n = op->ncols, nnz = op->nnz;
  if (s.hot_state < 0) {
    std::vector<int64_t> cnt(n, 0);
    for (int32_t cidx : op->h_colidx) ++cnt[cidx];
    std::vector<int32_t> order(n);
    for (int64_t j = 0; j < n; ++j) order[j] = (int32_t)j;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cnt[x] > cnt[y]; });
    const int64_t keep = std::min<int64_t>(n, 1024);
    s.hot_cover.assign(keep, 0);
    int64_t cover = 0;
    for (int64_t i = 0; i < keep; ++i) s.hot_cover[i] = cover += cnt[order[i]];
    order.resize(keep);
    s.hot_order = std::move(order);
    s.hot_state = 1;
  }
  if (h > (int)s.hot_cover.size() || 5 * s.hot_cover[h - 1] < nnz) return nullptr;
  for (int i = 0; i < Schedule::kHotEnc; ++i)
    if (s.hot[i].h == h) {
      s.hot_use[i] = ++s.hot_clock;
      return &s.hot[i];
    }
  int victim = 0;  // an unused entry, else the least recently used
  for (int i = 1; i < Schedule::kHotEnc; ++i)
    if ((s.hot[i].h == 0) != (s.hot[victim].h == 0) ? s.hot[i].h == 0 : s.hot_use[i] < s.hot_use[victim])
      victim = i;
  s.hot_use[victim] = ++s.hot_clock;
  Schedule::HotEnc &e = s.hot[victim];
  for (auto &v : e.vals) v.release();
  e.dense.release();
  Ctx *c = op->ctx;
  const int64_t m = op->nrows;
  std::vector<int32_t> slot(n, -1);
  for (int i = 0; i < h; ++i) slot[s.hot_order[i]] = i;
  // staged-first reordering of every row (entry order otherwise kept)
  const std::vector<int32_t> &rp = op->h_rowptr, &ci = op->h_colidx;
  std::vector<int32_t> enc(std::max<int64_t>(nnz, 1)), perm(std::max<int64_t>(nnz, 1)), split(std::max<int64_t>(m, 1));
  std::vector<int32_t> cold(std::max<int64_t>(m, 1));
  for (int64_t r = 0; r < m; ++r) {
    int32_t pos = rp[r];
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
      if (slot[ci[k]] >= 0) {
        enc[pos] = -1 - slot[ci[k]];
        perm[pos++] = k;
      }
    split[r] = pos - rp[r];
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
      if (slot[ci[k]] < 0) {
        enc[pos] = ci[k];
        perm[pos++] = k;
      }
    cold[r] = rp[r + 1] - rp[r] - split[r];
  }
  std::vector<int32_t> order(std::max<int64_t>(m, 1));
  for (int64_t r = 0; r < m; ++r) order[r] = (int32_t)r;
  std::stable_sort(order.begin(), order.begin() + m, [&](int32_t x, int32_t y) {
    return cold[x] != cold[y] ? cold[x] > cold[y] : split[x] > split[y];
  });
  e.rows.alloc((size_t)h * sizeof(int32_t));
  HIP_CHECK(hipMemcpyAsync(e.rows.p, s.hot_order.data(), (size_t)h * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
  e.colidx.alloc(enc.size() * sizeof(int32_t));
  HIP_CHECK(hipMemcpyAsync(e.colidx.p, enc.data(), enc.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  e.perm.alloc(perm.size() * sizeof(int32_t));
  HIP_CHECK(hipMemcpyAsync(e.perm.p, perm.data(), perm.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  e.split.alloc(split.size() * sizeof(int32_t));
  HIP_CHECK(hipMemcpyAsync(e.split.p, split.data(), split.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
  e.order.alloc(order.size() * sizeof(int32_t));
  HIP_CHECK(hipMemcpyAsync(e.order.p, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  e.h = h;
  return &e;
}

// Values of dtype V in the staged-first layout of e (gathered on first use).
template <typename V>
static const V *hot_values(Ctx *c, Schedule::HotEnc &e, const V *vals, int64_t nnz) {
  DevBuf &d = e.vals[DtypeOf<V>::v];
  if (!d.p) {
    d.alloc(std::max<size_t>(nnz * sizeof(V), 16));
    if (nnz > 0) {
      hipLaunchKernelGGL(k_gather<V>, grid_for(nnz), kBlock, 0, c->stream, nnz, e.perm.as<int32_t>(), vals,
                         d.as<V>());
      HIP_CHECK(hipGetLastError());
    }
  }
  return d.as<V>();
}

template <int LGR, typename V, typename X, typename Y, typename Acc>
static void spmm_hot_launch(Ctx *c, const SpmmArgs<V, X, Y, Acc> &a, const Schedule::HotEnc &e, const V *hval,
                            int nrows) {
  constexpr int RPB = kHotBlock >> LGR;
  const int64_t need = ((int64_t)nrows + RPB - 1) / RPB;
  const size_t shm = (size_t)e.h * a.ldx * sizeof(X);
  // two workgroups per CU while two stagings fit the 160 KB of LDS, else one
  const int per_cu = shm <= 80 * 1024 ? 2 : 1;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(per_cu * (int64_t)c->ncu, need));
  hipLaunchKernelGGL((k_spmm_hot<LGR, V, X, Y, Acc>), blocks, kHotBlock, shm, c->stream, a, e.colidx.as<int32_t>(),
                     hval, e.split.as<int32_t>(), e.order.as<int32_t>(), nrows, e.rows.as<int32_t>(), e.h);
}

// Sums the pieces of multi-row entries [0, n) of schedule s into Y: rows cut
// into chunks (the first s.chunk_multi entries) one workgroup each, rows with
// a few segments four per workgroup (l <= 64).
template <typename Y, typename Acc>
static void launch_fin(Ctx *c, const Schedule &s, int n, int64_t ldp, Y *y, int64_t ldy, int l) {
  const int nch = l <= 64 ? std::min<int>(n, s.chunk_multi) : n;
  if (nch > 0)
    hipLaunchKernelGGL((k_spmm_fin<Y, Acc>), nch, kBlock, 0, c->stream, s.counters.as<int32_t>(),
                       (const Acc *)s.partials.p, ldp, y, ldy, l);
  if (n > nch)
    hipLaunchKernelGGL((k_spmm_fin4<Y, Acc>), (n - nch + 3) / 4, kBlock, 0, c->stream,
                       s.counters.as<int32_t>() + 3 * (int64_t)nch, n - nch, (const Acc *)s.partials.p, ldp, y, ldy,
                       l);
  HIP_CHECK(hipGetLastError());
}

// A^T Y of a plain FP32 operator with Zipf columns: the 256 staged columns
// (output rows of A^T) on the matrix cores (k_spmm_t_hot), the other rows by
// k_spmm on a schedule without them, and the fin pass for both. Returns false
// (nothing launched) when the split does not apply.
static bool spmm_t_hot(svdh_op *op, const void *vals_t, const void *X_, int64_t ldx, void *Y_, int64_t ldy,
                       int64_t l) {
  (void)vals_t;
  if (op->base || l > 64 || ldx % 4 || op->thot_state == 0) return false;
  const char *env = std::getenv("SVDH_SPMM_THOT");
  if (env && std::atoi(env) == 0) return false;
  Ctx *c = op->ctx;
  const void *vals0 = op->values(0, SVDH_F32);
  if (!vals0) return false;
  if (op->thot_h == 0) {
    // staged columns on the matrix cores: 128 or 256 (the MFMA work and the
    // dense block grow with h, the gathered entries shrink)
    op->thot_h = kTHotDefaultH;
    if (const char *eh = std::getenv("SVDH_SPMM_THOT_H")) op->thot_h = std::atoi(eh) == 256 ? 256 : 128;
  }
  const int kH = op->thot_h;
  Schedule::HotEnc *e = hot_columns(op, get_spmm_schedule(op, 0), kH);
  if (!e || e->h != kH) return false;
  if (!op->sched_thot.built) {
    const Schedule &s0 = get_spmm_schedule(op, 0);
    std::vector<int32_t> hot(s0.hot_order.begin(), s0.hot_order.begin() + kH);
    op->thot_g = 2 * c->ncu;
    if (const char *eg = std::getenv("SVDH_SPMM_THOT_G")) op->thot_g = std::max(1, std::atoi(eg));
    build_spmm_schedule(op, 1, op->sched_thot, &hot, op->thot_g);
    op->thot_state = 1;
  }
  SpmmPlan<float, float, float, float> p;
  if (!plan_spmm(op, 1, vals_t, X_, ldx, Y_, ldy, l, p, &op->sched_thot)) return false;
  if (!e->dense.p) {
    const float *hval = hot_values<float>(c, *e, static_cast<const float *>(vals0), op->nnz);
    e->dense.alloc(std::max<size_t>((size_t)op->nrows * kH * sizeof(float), 16));
    HIP_CHECK(hipMemsetAsync(e->dense.p, 0, e->dense.bytes, c->stream));
    hipLaunchKernelGGL(k_hot_dense, grid_for(op->nrows), kBlock, 0, c->stream, op->nrows, op->rowptr.as<int32_t>(),
                       e->colidx.as<int32_t>(), hval, e->split.as<int32_t>(), e->dense.as<float>(), kH);
    HIP_CHECK(hipGetLastError());
  }
  KTimer t(c, KC_SPMM_T, p.bytes);
  if (kH == 256)
    hipLaunchKernelGGL(k_spmm_t_hot<2>, op->thot_g, kTHotBlock, 0, c->stream, op->nrows, e->dense.as<float>(),
                       static_cast<const float *>(X_), ldx, (int)l, p.a.partials, p.a.ldp,
                       op->sched_thot.hot_slot_base);
  else
    hipLaunchKernelGGL(k_spmm_t_hot<1>, op->thot_g, kTHotBlock, 0, c->stream, op->nrows, e->dense.as<float>(),
                       static_cast<const float *>(X_), ldx, (int)l, p.a.partials, p.a.ldp,
                       op->sched_thot.hot_slot_base);
  HIP_CHECK(hipGetLastError());
  spmm_dispatch<false>(c, p.a, p.lgr, p.blocks);
  HIP_CHECK(hipGetLastError());
  const Schedule &sp = op->sched_thot;
  const int ncold = sp.nlong_rows - sp.hot_multi;  // the staged rows come last in the table
  static const bool merged = [] {
    const char *e = std::getenv("SVDH_SPMM_FIN_MERGE");
    return !e || std::atoi(e) != 0;
  }();
  if (merged) {
    const int nch = std::min<int>(ncold, sp.chunk_multi), n4 = ncold - nch;
    hipLaunchKernelGGL((k_spmm_fin_merged<float, float>), sp.hot_multi + nch + (n4 + 15) / 16, 1024, 0, c->stream,
                       sp.counters.as<int32_t>(), nch, n4, sp.hot_multi, (const float *)sp.partials.p, p.a.ldp,
                       static_cast<float *>(Y_), ldy, (int)l);
  } else {
    if (ncold > 0) launch_fin<float, float>(c, sp, ncold, p.a.ldp, static_cast<float *>(Y_), ldy, (int)l);
    hipLaunchKernelGGL((k_spmm_fin_wide<float, float>), sp.hot_multi, 1024, 0, c->stream,
                       sp.counters.as<int32_t>() + 3 * (int64_t)ncold, (const float *)sp.partials.p, p.a.ldp,
                       static_cast<float *>(Y_), ldy, (int)l);
  }
  HIP_CHECK(hipGetLastError());
  return true;
}

template <typename V, typename X, typename Y, typename Acc>
static void launch_spmm(svdh_op *op, int direction, const void *vals, const void *X_, int64_t ldx,
                        void *Y_, int64_t ldy, int64_t l) {
  Ctx *c = op->ctx;
  SpmmPlan<V, X, Y, Acc> p;
  if (!plan_spmm(op, direction, vals, X_, ldx, Y_, ldy, l, p)) return;
  // (the staged path addresses panel rows by 32-bit byte offsets)
  if (direction == 0 && !p.mc && (uint64_t)op->ncols * (uint64_t)ldx * sizeof(X) < (1ull << 31)) {
    static const int64_t hot_lds = [] {  // LDS bytes for the staged panel rows (SVDH_SPMM_HOT_LDS)
      const char *e = std::getenv("SVDH_SPMM_HOT_LDS");
      // 80 KB: two workgroups still fit a CU's 160 KB (320 rows of l = 60 FP32: A X 110.1 -> 106.9 us
      // against 64 KB at 100,000 x 20,000; one workgroup per CU, 96 KB and more, is slower)
      return e ? std::max<int64_t>(4096, std::min<int64_t>(155648, std::atoll(e))) : 81920;
    }();
    Schedule::HotEnc *e = hot_columns(op, *p.s, std::min<int>(1024, (int)(hot_lds / (ldx * (int64_t)sizeof(X)))));
    if (e) {
      const V *hval = hot_values<V>(c, *e, static_cast<const V *>(vals), op->nnz);
      const int nrows = (int)op->nrows;
      KTimer t(c, KC_SPMM, p.bytes);
      switch (p.lgr) {
        case 1: spmm_hot_launch<1>(c, p.a, *e, hval, nrows); break;
        case 2: spmm_hot_launch<2>(c, p.a, *e, hval, nrows); break;
        case 3: spmm_hot_launch<3>(c, p.a, *e, hval, nrows); break;
        case 4: spmm_hot_launch<4>(c, p.a, *e, hval, nrows); break;
        case 5: spmm_hot_launch<5>(c, p.a, *e, hval, nrows); break;
        default: spmm_hot_launch<6>(c, p.a, *e, hval, nrows); break;
      }
      HIP_CHECK(hipGetLastError());
      return;
    }
  }
  if constexpr (std::is_same<V, float>::value && std::is_same<X, float>::value && std::is_same<Y, float>::value &&
                std::is_same<Acc, float>::value) {
    if (direction == 1 && !p.mc && spmm_t_hot(op, vals, X_, ldx, Y_, ldy, l)) return;
  }
  KTimer t(c, direction == 0 ? KC_SPMM : KC_SPMM_T, p.bytes);
  if (p.mc)
    spmm_dispatch<true>(c, p.a, p.lgr, p.blocks);
  else
    spmm_dispatch<false>(c, p.a, p.lgr, p.blocks);
  HIP_CHECK(hipGetLastError());
  const Schedule *sp = p.s;
  if (sp->nlong_rows > 0) launch_fin<Y, Acc>(c, *sp, sp->nlong_rows, p.a.ldp, static_cast<Y *>(Y_), ldy, (int)l);
}

// ------------------------------------------------------------------------
// Persistent batched SpMM: one launch of a fixed grid of workgroups that take
// their work from eight queues, one per XCD (workgroup w serves queue w mod 8,
// the XCD round-robin dispatch puts it on), until the queues run dry. A work
// item is a row group of one operator's schedule: a chunk of a long row, a
// group of segments (rows cut at source blocks) or a group of short rows.
// Chunks and segments of source block x go to queue x, so each XCD still
// gathers from one block of the panel; the other items are dealt round robin.
// The queues list the operators by decreasing nonzero count (large jobs
// first). Each item computes its rows exactly as the per-operator launch does
// (one lane sums a row's entries for its columns in storage order), so the
// grid, the queue order, the gathers in flight (UU) and the lanes per short
// row (LGRS) change the time, not the result; chunks keep the lane layout
// LGRC of the panel width, which fixes their split of a row.
// ------------------------------------------------------------------------
struct PqRange {
  int32_t b, kind, first, stride;  // kind 0: chunk index, 1: segment workgroup, 2: short-row group
  int32_t count, start, pad0, pad1;  // items first + k * stride, k < count; queue positions start ..
};

// Two ways of taking items: claim = 0, workgroup j of the queue's W takes
// items j, j + W, j + 2W, ... (no synchronisation; the queue order spreads
// large and small items over the workgroups); claim > 0, a workgroup claims
// `claim` consecutive items at a time from the queue's counter, the next
// claim issued before the current items are processed.
// The queue's ranges are copied into LDS first (kPqMaxRanges per queue at
// most), and a workgroup walks them forward, since the items it takes come
// in increasing order.
constexpr int kPqMaxRanges = 384;

template <int LGRC, int LGRS, int UU, bool MC, typename V, typename X, typename Y, typename Acc>
__global__ __launch_bounds__(kBlock) void k_spmm_pq(const SpmmArgs<V, X, Y, Acc> *args, const PqRange *ranges,
                                                   const int32_t *qoff, const int32_t *qlen, int32_t *counters,
                                                   int claim) {
  __shared__ int s_it;
  __shared__ int4 s_rg[kPqMaxRanges];  // (b, kind, first, stride) of each range
  __shared__ int s_end[kPqMaxRanges];  // start + count
  const int x = blockIdx.x & 7;
  const int r0 = qoff[x], nr = qoff[x + 1] - r0;
  const int total = qlen[x];
  for (int i = threadIdx.x; i < nr; i += kBlock) {
    const PqRange g = ranges[r0 + i];
    s_rg[i] = make_int4(g.b, g.kind, g.first, g.stride);
    s_end[i] = g.start + g.count;
  }
  __syncthreads();
  int cur = 0, cur_start = 0;  // the range holding the current item, and its first position
  const int nwg = (int)(gridDim.x >> 3), jw = (int)(blockIdx.x >> 3);
  int32_t *ctr = counters + 64 * x;  // one counter per 256-byte line
  int it0, step, span;
  if (claim > 0) {
    if (threadIdx.x == 0) s_it = atomicAdd(ctr, claim);
    __syncthreads();
    it0 = s_it;
    __syncthreads();
    span = claim;
    step = 0;
  } else {
    it0 = jw;
    span = 1;
    step = nwg;
  }
  for (;;) {
    if (it0 >= total) break;
    int nxt = 0;
    if (claim > 0 && threadIdx.x == 0) nxt = atomicAdd(ctr, claim);  // in flight during the items
    const int e = min(total, it0 + span);
    for (int it = it0; it < e; ++it) {
      while (s_end[cur] <= it) cur_start = s_end[cur++];  // items only increase
      const int4 rg = s_rg[cur];
      const int idx = rg.z + (it - cur_start) * rg.w;
      const SpmmArgs<V, X, Y, Acc> &a = args[rg.x];
      if (rg.y == 0)
        spmm_chunk<LGRC, MC, UU>(a, idx);
      else if (rg.y == 1)
        spmm_segment<LGRS, MC, UU>(a, idx);
      else
        spmm_short<LGRS, MC, UU>(a, idx);
    }
    if (claim > 0) {
      __syncthreads();  // the previous value of s_it has been read by every thread
      if (threadIdx.x == 0) s_it = nxt;
      __syncthreads();
      it0 = s_it;
    } else {
      it0 += step;
    }
  }
}

__global__ void k_zero_counters(int32_t *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

template <int LGR, bool MC, typename V, typename X, typename Y, typename Acc>
static void spmm_batched_one(Ctx *c, const SpmmArgs<V, X, Y, Acc> *args, const int32_t *boff, int nb, int blocks) {
  hipLaunchKernelGGL((k_spmm_batched<LGR, MC, V, X, Y, Acc>), blocks, kBlock, 0, c->stream, args, boff, nb);
}

template <typename V, typename X, typename Y, typename Acc>
static void launch_spmm_static(Ctx *c, int direction, bool mc, int lgr, const std::vector<SpmmArgs<V, X, Y, Acc>> &args,
                               const std::vector<int32_t> &boff, double bytes);

// ---- autotuner of the persistent batched SpMM ----------------------------
// Candidates: lanes per short row (the panel width's layout, or half of it:
// twice the rows per wavefront in two column passes), gathers in flight per
// row group (4 or 8), workgroups per CU of the persistent grid (4, 8, 16) and
// how workgroups take items (static interleave, or claims of 16 items).
// The first batched SpMM of a workload runs every candidate once (each
// computes the same result), keeps the fastest, and appends the choice to a
// tuning file keyed by the direction, value type, panel width, operator count,
// rows, nonzeros and row-length histogram (SVDH_TUNE_FILE, default
// ~/.cache/svdh/tune.txt), so later processes start tuned. SVDH_BATCH_TUNE=0
// uses the default candidate without tuning.
struct PqConfig {
  int dl;     // lanes per short row = 2^(LGRC - dl)
  int uu;     // gathers in flight per row group
  int wpc;    // workgroups (256 threads) per CU
  int claim;  // 0: static interleave; > 0: items claimed per counter update
};
// wpc = 0 stands for the static launch (one workgroup per schedule block of
// every operator, dispatched by the hardware), which competes with the
// persistent candidates.
constexpr PqConfig kPqDefault{0, 4, 8, 0};
constexpr PqConfig kPqCandidates[] = {
    {0, 4, 0, 0}, {0, 4, 8, 0}, {0, 8, 8, 0}, {1, 4, 8, 0}, {1, 8, 8, 0}, {0, 4, 4, 0}, {0, 8, 4, 0},
    {1, 4, 4, 0}, {1, 8, 4, 0}, {0, 4, 16, 0}, {0, 8, 16, 0}, {1, 4, 16, 0}, {1, 8, 16, 0}, {0, 4, 8, 16},
    {0, 8, 8, 16}, {1, 4, 8, 16}, {1, 8, 8, 16}, {0, 4, 16, 16}, {1, 8, 16, 16}};

struct PqTuner {
  std::mutex mu;
  bool loaded = false;
  std::map<std::string, PqConfig> best;
  std::string path() const {
    if (const char *e = std::getenv("SVDH_TUNE_FILE")) return e;
    const char *h = std::getenv("HOME");
    return std::string(h ? h : ".") + "/.cache/svdh/tune.txt";
  }
  void load() {
    if (loaded) return;
    loaded = true;
    std::ifstream f(path());
    std::string line;
    while (std::getline(f, line)) {
      const size_t t = line.find('\t');
      if (t == std::string::npos) continue;
      std::istringstream is(line.substr(t + 1));
      PqConfig c{};
      if (is >> c.dl >> c.uu >> c.wpc >> c.claim) best[line.substr(0, t)] = c;
    }
  }
  void store(const std::string &key, const PqConfig &c, double ms) {
    best[key] = c;
    const std::string p = path();
    const size_t slash = p.rfind('/');
    if (slash != std::string::npos) {  // create the directory (and its parent) if missing
      const std::string dir = p.substr(0, slash);
      const size_t s2 = dir.rfind('/');
      if (s2 != std::string::npos && s2 > 0) (void)::mkdir(dir.substr(0, s2).c_str(), 0755);
      (void)::mkdir(dir.c_str(), 0755);
    }
    std::ofstream f(p, std::ios::app);
    if (f) f << key << '\t' << c.dl << ' ' << c.uu << ' ' << c.wpc << ' ' << c.claim << '\t' << ms << '\n';
  }
};

static PqTuner &pq_tuner() {
  static PqTuner t;
  return t;
}

// Row-length histogram of an operator direction (log2 buckets), cached.
static const std::vector<int64_t> &row_histogram(svdh_op *op, int direction) {
  std::vector<int64_t> &h = op->row_hist[direction];
  if (!h.empty()) return h;
  h.assign(16, 0);
  const CsrView v = op->view(direction);
  const std::vector<int32_t> &rp = *v.h_rowptr;
  for (int64_t r = 0; r < v.nrows; ++r) {
    const int64_t src = v.h_rowmap ? (*v.h_rowmap)[r] : r;
    const int32_t len = rp[src + 1] - rp[src];
    int bkt = 0;
    while (bkt < 15 && (1 << bkt) <= len) ++bkt;
    ++h[bkt];
  }
  return h;
}

// Work ranges of the eight queues for the short-row layout LGRS: per
// operator (largest first) its chunks and segments by source block, its
// short-row groups round robin.
template <typename V, typename X, typename Y, typename Acc>
static void pq_ranges(const std::vector<const Schedule *> &sched, const std::vector<int64_t> &weight, int lgrs,
                      std::vector<PqRange> &ranges, int32_t qoff[9], int32_t qlen[8]) {
  const int rpbs = kBlock >> lgrs;
  std::vector<int> order(sched.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return weight[a] > weight[b]; });
  std::vector<PqRange> q[8];
  auto add = [&](int x, int b, int kind, int first, int stride, int count) {
    if (count > 0) q[x].push_back(PqRange{b, kind, first, stride, count, 0, 0, 0});
  };
  for (int b : order) {
    const Schedule &s = *sched[b];
    const int nblk = std::max(1, s.nblk);
    for (int x = 0; x < nblk; ++x) {  // chunks of source block x
      if (nblk > 1) {
        add(x % 8, b, 0, s.blk_off[x], 1, s.blk_cnt[x]);
      } else {
        for (int y = 0; y < 8; ++y)
          add(y, b, 0, s.blk_off[x] + y, 8, s.blk_cnt[x] > y ? (s.blk_cnt[x] - y + 7) / 8 : 0);
      }
    }
    for (int x = 0; x < nblk; ++x) {  // segment groups kb of block x: workgroup index kb * nblk + x
      const int groups = (s.seg_cnt[x] + rpbs - 1) / rpbs;
      if (nblk > 1)
        add(x % 8, b, 1, x, nblk, groups);
      else
        for (int y = 0; y < 8; ++y) add(y, b, 1, y, 8, groups > y ? (groups - y + 7) / 8 : 0);
    }
    const int ngroups = (s.bins[0].row_count + rpbs - 1) / rpbs;  // short rows
    for (int y = 0; y < 8; ++y) add(y, b, 2, y, 8, ngroups > y ? (ngroups - y + 7) / 8 : 0);
  }
  ranges.clear();
  qoff[0] = 0;
  for (int x = 0; x < 8; ++x) {
    int32_t pos = 0;
    for (PqRange r : q[x]) {
      r.start = pos;
      pos += r.count;
      ranges.push_back(r);
    }
    qlen[x] = pos;
    qoff[x + 1] = (int32_t)ranges.size();
  }
}

template <int LGRC, int LGRS, int UU, bool MC, typename V, typename X, typename Y, typename Acc>
static void pq_launch_one(Ctx *c, int grid, int claim, const SpmmArgs<V, X, Y, Acc> *args, const PqRange *rg,
                          const int32_t *qoff, const int32_t *qlen) {
  if (claim > 0)
    hipLaunchKernelGGL(k_zero_counters, 1, 512, 0, c->stream, c->pq_counters.as<int32_t>(), 512);
  hipLaunchKernelGGL((k_spmm_pq<LGRC, LGRS, UU, MC, V, X, Y, Acc>), grid, kBlock, 0, c->stream, args, rg, qoff, qlen,
                     c->pq_counters.as<int32_t>(), claim);
}

template <int LGRC, bool MC, typename V, typename X, typename Y, typename Acc>
static void pq_launch_lgr(Ctx *c, const PqConfig &cf, int grid, const SpmmArgs<V, X, Y, Acc> *args, const PqRange *rg,
                          const int32_t *qoff, const int32_t *qlen) {
  if (cf.dl == 0 && cf.uu == 4) pq_launch_one<LGRC, LGRC, 4, MC>(c, grid, cf.claim, args, rg, qoff, qlen);
  else if (cf.dl == 0) pq_launch_one<LGRC, LGRC, 8, MC>(c, grid, cf.claim, args, rg, qoff, qlen);
  else if (cf.uu == 4) pq_launch_one<LGRC, LGRC - 1, 4, MC>(c, grid, cf.claim, args, rg, qoff, qlen);
  else pq_launch_one<LGRC, LGRC - 1, 8, MC>(c, grid, cf.claim, args, rg, qoff, qlen);
}

// The persistent launch of one group of plans (all with the chunk layout lgr,
// which must be 3, 4 or 5); returns false when the layout is not covered.
template <typename V, typename X, typename Y, typename Acc, typename StaticRun>
static bool launch_spmm_pq(Ctx *c, int direction, bool mc, int lgr, int64_t l,
                           const std::vector<SpmmArgs<V, X, Y, Acc>> &args, const std::vector<const Schedule *> &sched,
                           const std::vector<svdh_op *> &gops, double bytes, StaticRun static_run) {
  if (lgr < 3 || lgr > 5) return false;
  if (const char *e = std::getenv("SVDH_BATCH_PERSISTENT"))
    if (std::atoi(e) == 0) return false;
  c->pq_counters.ensure(512 * sizeof(int32_t));
  std::vector<int64_t> weight(gops.size());
  int64_t rows = 0, nnz = 0;
  std::vector<int64_t> hist(16, 0);
  for (size_t b = 0; b < gops.size(); ++b) {
    const CsrView v = gops[b]->view(direction);
    weight[b] = v.nnz;
    rows += v.nrows;
    nnz += v.nnz;
    const std::vector<int64_t> &h = row_histogram(gops[b], direction);
    for (int i = 0; i < 16; ++i) hist[i] += h[i];
  }
  std::ostringstream key;
  key << "spmm_batched dir=" << direction << " mc=" << (int)mc << " val=" << sizeof(V) << " l=" << l
      << " ops=" << gops.size() << " rows=" << rows << " nnz=" << nnz << " hist=";
  for (int i = 0; i < 16; ++i) key << (i ? "," : "") << (rows ? (hist[i] * 1000 + rows / 2) / rows : 0);
  const std::string k = key.str();
  const char *te = std::getenv("SVDH_BATCH_TUNE");
  const bool tune = !(te && std::atoi(te) == 0);
  PqTuner &tn = pq_tuner();
  PqConfig cfg = kPqDefault;
  bool known = false;
  {
    std::lock_guard<std::mutex> g(tn.mu);
    if (tune) {
      tn.load();
      auto it = tn.best.find(k);
      if (it != tn.best.end()) {
        cfg = it->second;
        known = true;
      }
    }
  }
  // Each run uploads its own arguments: the upload ring's slots are reused
  // in turn, and a slot's contents are valid only for the kernels launched
  // before the next upload into it.
  auto run = [&](const PqConfig &cf) {
    if (cf.wpc == 0) {
      static_run();
      return;
    }
    const auto *dargs =
        static_cast<const SpmmArgs<V, X, Y, Acc> *>(c->upload(args.data(), args.size() * sizeof(args[0])));
    std::vector<PqRange> ranges;
    int32_t qo[9], ql[8];
    pq_ranges<V, X, Y, Acc>(sched, weight, lgr - cf.dl, ranges, qo, ql);
    std::vector<int32_t> tab(ranges.size() * (sizeof(PqRange) / 4) + 17);
    std::memcpy(tab.data(), ranges.data(), ranges.size() * sizeof(PqRange));
    int32_t *tq = tab.data() + ranges.size() * (sizeof(PqRange) / 4);
    std::memcpy(tq, qo, sizeof qo);
    std::memcpy(tq + 9, ql, sizeof ql);
    if (tab.size() * 4 > Ctx::kRingSlot) fail(SVDH_ERR_ARG, "batched SpMM: too many work ranges");
    for (int x = 0; x < 8; ++x)
      if (qo[x + 1] - qo[x] > kPqMaxRanges) fail(SVDH_ERR_ARG, "batched SpMM: too many work ranges in one queue");
    const int32_t *dt = static_cast<const int32_t *>(c->upload(tab.data(), tab.size() * 4));
    const PqRange *drg = reinterpret_cast<const PqRange *>(dt);
    const int32_t *dqo = dt + ranges.size() * (sizeof(PqRange) / 4);
    const int grid = std::max(8, cf.wpc * c->ncu / 8 * 8);
    switch (lgr) {
      case 3: if (mc) pq_launch_lgr<3, true>(c, cf, grid, dargs, drg, dqo, dqo + 9);
              else pq_launch_lgr<3, false>(c, cf, grid, dargs, drg, dqo, dqo + 9); break;
      case 4: if (mc) pq_launch_lgr<4, true>(c, cf, grid, dargs, drg, dqo, dqo + 9);
              else pq_launch_lgr<4, false>(c, cf, grid, dargs, drg, dqo, dqo + 9); break;
      default: if (mc) pq_launch_lgr<5, true>(c, cf, grid, dargs, drg, dqo, dqo + 9);
               else pq_launch_lgr<5, false>(c, cf, grid, dargs, drg, dqo, dqo + 9); break;
    }
    HIP_CHECK(hipGetLastError());
  };
  if (tune && !known) {
    // every candidate once, timed with events; all compute the same result
    constexpr int nc = (int)(sizeof(kPqCandidates) / sizeof(kPqCandidates[0]));
    hipEvent_t ev[nc + 1];
    for (int i = 0; i <= nc; ++i) HIP_CHECK(hipEventCreate(&ev[i]));
    HIP_CHECK(hipEventRecord(ev[0], c->stream));
    for (int i = 0; i < nc; ++i) {
      run(kPqCandidates[i]);
      HIP_CHECK(hipEventRecord(ev[i + 1], c->stream));
    }
    HIP_CHECK(hipEventSynchronize(ev[nc]));
    int bi = 0;
    float bt = 0.f;
    for (int i = 0; i < nc; ++i) {
      float t = 0.f;
      HIP_CHECK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      if (i == 0 || t < bt) {
        bt = t;
        bi = i;
      }
    }
    for (int i = 0; i <= nc; ++i) (void)hipEventDestroy(ev[i]);
    std::lock_guard<std::mutex> g(tn.mu);
    tn.store(k, kPqCandidates[bi], bt);
    return true;
  }
  if (cfg.wpc == 0) return false;  // the static launch (timed by the caller)
  KTimer t(c, direction == 0 ? KC_SPMM : KC_SPMM_T, bytes);
  run(cfg);
  return true;
}

// One launch for the SpMMs of several operators of the same direction and
// panel width (plus one launch of the chunk sums). Operators with and without
// a column remap go to separate launches.
template <typename V, typename X, typename Y, typename Acc>
static void launch_spmm_batched(Ctx *c, int nops, svdh_op *const *ops, int direction, int vdt, const void *const *Xs,
                                int64_t ldx, void *const *Ys, int64_t ldy, int64_t l) {
  std::vector<SpmmPlan<V, X, Y, Acc>> plans(nops);
  std::vector<char> live(nops, 0);
  for (int b = 0; b < nops; ++b)
    live[b] = plan_spmm(ops[b], direction, ops[b]->values(direction, vdt), Xs[b], ldx, Ys[b], ldy, l, plans[b]);
  for (int g = 0; g < 2; ++g) {
    const bool mc = g == 1;
    std::vector<SpmmArgs<V, X, Y, Acc>> args;
    std::vector<int32_t> boff(1, 0);
    std::vector<FinArgs<Y, Acc>> fargs;
    std::vector<int32_t> foff(1, 0);
    std::vector<const Schedule *> sched;
    std::vector<svdh_op *> gops;
    double bytes = 0.0;
    int lgr = 0;
    for (int b = 0; b < nops; ++b) {
      if (!live[b] || plans[b].mc != mc) continue;
      if (lgr && plans[b].lgr != lgr) fail(SVDH_ERR_ARG, "batched SpMM needs one lane layout");
      lgr = plans[b].lgr;
      args.push_back(plans[b].a);
      sched.push_back(plans[b].s);
      gops.push_back(ops[b]);
      boff.push_back(boff.back() + (plans[b].blocks + 7) / 8 * 8);
      bytes += plans[b].bytes;
      const Schedule &s = *plans[b].s;
      if (s.nlong_rows > 0) {
        const int32_t nchunk = l <= 64 ? std::min<int32_t>(s.nlong_rows, s.chunk_multi) : s.nlong_rows;
        fargs.push_back(FinArgs<Y, Acc>{s.counters.as<int32_t>(), (const Acc *)s.partials.p, plans[b].a.ldp,
                                        plans[b].a.y, ldy, (int32_t)l, nchunk});
        foff.push_back(foff.back() + s.nlong_rows);
      }
    }
    if (args.empty()) continue;
    auto static_run = [&] { launch_spmm_static<V, X, Y, Acc>(c, direction, mc, lgr, args, boff, bytes); };
    if (!launch_spmm_pq<V, X, Y, Acc>(c, direction, mc, lgr, l, args, sched, gops, bytes, static_run)) static_run();
    if (!fargs.empty()) {
      const auto *dfargs = static_cast<const FinArgs<Y, Acc> *>(c->upload(fargs.data(), fargs.size() * sizeof(fargs[0])));
      const auto *dfoff = static_cast<const int32_t *>(c->upload(foff.data(), foff.size() * sizeof(int32_t)));
      hipLaunchKernelGGL((k_spmm_fin_batched<Y, Acc>), foff.back(), kBlock, 0, c->stream, dfargs, dfoff,
                         (int)fargs.size());
      HIP_CHECK(hipGetLastError());
    }
  }
}

// The static launch of one group: one workgroup per schedule block of every
// operator (each operator's blocks start at a multiple of 8, so block w of an
// operator lands on XCD w mod 8 as in a launch of its own).
template <typename V, typename X, typename Y, typename Acc>
static void launch_spmm_static(Ctx *c, int direction, bool mc, int lgr, const std::vector<SpmmArgs<V, X, Y, Acc>> &args,
                               const std::vector<int32_t> &boff, double bytes) {
  {
    const int nb = (int)args.size();
    const auto *dargs = static_cast<const SpmmArgs<V, X, Y, Acc> *>(c->upload(args.data(), args.size() * sizeof(args[0])));
    const auto *dboff = static_cast<const int32_t *>(c->upload(boff.data(), boff.size() * sizeof(int32_t)));
    KTimer t(c, direction == 0 ? KC_SPMM : KC_SPMM_T, bytes);
    const int blocks = boff.back();
#define SVDH_BATCH_CASE(L)                                                                   \
  case L:                                                                                   \
    if (mc) spmm_batched_one<L, true>(c, dargs, dboff, nb, blocks);                          \
    else spmm_batched_one<L, false>(c, dargs, dboff, nb, blocks);                            \
    break;
    switch (lgr) {
      SVDH_BATCH_CASE(1)
      SVDH_BATCH_CASE(2)
      SVDH_BATCH_CASE(3)
      SVDH_BATCH_CASE(4)
      SVDH_BATCH_CASE(5)
      default:
        if (mc) spmm_batched_one<6, true>(c, dargs, dboff, nb, blocks);
        else spmm_batched_one<6, false>(c, dargs, dboff, nb, blocks);
    }
#undef SVDH_BATCH_CASE
    HIP_CHECK(hipGetLastError());
  }
}

void spmm(svdh_op *op, int direction, int vdt, const void *X, int64_t ldx, int xdt, void *Y,
          int64_t ldy, int ydt, int64_t l) {
  ensure_values(op, vdt);
  const void *vals = op->values(direction, vdt);
  if (vdt == SVDH_F64 && xdt == SVDH_F64 && ydt == SVDH_F64)
    launch_spmm<double, double, double, double>(op, direction, vals, X, ldx, Y, ldy, l);
  else if (vdt == SVDH_F32 && xdt == SVDH_F64 && ydt == SVDH_F64)
    launch_spmm<float, double, double, double>(op, direction, vals, X, ldx, Y, ldy, l);
  else if (vdt == SVDH_F32 && xdt == SVDH_F32 && ydt == SVDH_F32)
    launch_spmm<float, float, float, float>(op, direction, vals, X, ldx, Y, ldy, l);
  else if (vdt == SVDH_BF16 && xdt == SVDH_BF16 && ydt == SVDH_F32)
    launch_spmm<bf16_t, bf16_t, float, float>(op, direction, vals, X, ldx, Y, ldy, l);
  else if (vdt == SVDH_F32 && xdt == SVDH_BF16 && ydt == SVDH_F32)
    launch_spmm<float, bf16_t, float, float>(op, direction, vals, X, ldx, Y, ldy, l);
  else
    fail(SVDH_ERR_ARG, "unsupported SpMM dtypes values=%d in=%d out=%d", vdt, xdt, ydt);
}

// ------------------------------------------------------------------------
// SDDMM over the CSR pattern (the adjoint of the SVD):
//   grad[e] = sum_c s1_c L1[r_e, c] R1[col_e, c] + s2_c L2[r_e, c] R2[col_e, c]
// for k <= 64 factor columns. One row group of R = 2^LGR lanes per row, each
// lane owning NV consecutive factor columns; the left rows (scaled) stay in
// registers, each nonzero gathers one right row, and the R lane partials are
// combined with a butterfly so lane t keeps the value of the row's t-th
// nonzero of the batch and the batch is written coalesced. A masked view
// maps the column through colmap; dropped columns get 0.
// ------------------------------------------------------------------------
template <typename T, typename G>
struct SddmmArgs {
  const int32_t *rowptr;
  const int32_t *colidx;
  const int32_t *colmap;
  const T *l1, *r1, *l2, *r2;
  const double *s1, *s2;
  int64_t ldl, ldr;
  int32_t k;
  int64_t nrows;
  G *grad;
};

// T is the storage type of the factor panels (FP64, FP32 or BF16); the
// products are accumulated in FP64 for FP64 factors, FP32 otherwise.
template <typename T> struct SddmmAcc { using type = float; };
template <> struct SddmmAcc<double> { using type = double; };

template <int LGR, bool MC, typename T, typename G>
__global__ __launch_bounds__(kBlock) void k_sddmm(SddmmArgs<T, G> a) {
  using A = typename SddmmAcc<T>::type;
  constexpr int R = 1 << LGR;
  constexpr int NV = PanelVec<T>::N;
  constexpr int RPB = kBlock >> LGR;
  const int lane = threadIdx.x & (R - 1);
  const int col = lane * NV;
  const bool colok = col < a.k;
  const bool two = a.l2 != nullptr;
  for (int64_t r = (int64_t)blockIdx.x * RPB + (threadIdx.x >> LGR); r < a.nrows; r += (int64_t)gridDim.x * RPB) {
    A lv1[NV], lv2[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      lv1[v] = A(0);
      lv2[v] = A(0);
    }
    if (colok) {
      ld_panel<A, T>(a.l1 + r * a.ldl + col, lv1);
      if (two) ld_panel<A, T>(a.l2 + r * a.ldl + col, lv2);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const bool in = col + v < a.k;
        lv1[v] = in ? (A)(a.s1 ? a.s1[col + v] * (double)lv1[v] : (double)lv1[v]) : A(0);
        lv2[v] = in && two ? (A)(a.s2 ? a.s2[col + v] * (double)lv2[v] : (double)lv2[v]) : A(0);
      }
    }
    const int kb = a.rowptr[r], ke = a.rowptr[r + 1];
    for (int base = kb; base < ke; base += R) {
      int c = -1;
      if (base + lane < ke) {
        c = __builtin_nontemporal_load(&a.colidx[base + lane]);
        if (MC) c = a.colmap[c];
      }
      const int n = min(R, ke - base);
      A mine = A(0);
      // U nonzeros at a time: their right rows are gathered together (clamped
      // to row 0 and column 0 where the entry or the lane's columns are
      // absent, so no load waits behind a branch), then each product is
      // formed and reduced as for a single nonzero (the same bits)
      constexpr int U = 4;
      for (int j0 = 0; j0 < n; j0 += U) {
        int cj[U];
        A rv1[U][NV], rv2[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) cj[u] = __shfl(c, (j0 + u) & (R - 1), R);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t rr = cj[u] >= 0 ? (int64_t)cj[u] : 0;
          const int cc = colok ? col : 0;
          ld_panel<A, T>(a.r1 + rr * a.ldr + cc, rv1[u]);
          if (two) ld_panel<A, T>(a.r2 + rr * a.ldr + cc, rv2[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          A part = A(0);
          if (colok && cj[u] >= 0 && j0 + u < n) {
#pragma unroll
            for (int v = 0; v < NV; ++v) part += lv1[v] * rv1[u][v];
            if (two)
#pragma unroll
              for (int v = 0; v < NV; ++v) part += lv2[v] * rv2[u][v];
          }
#pragma unroll
          for (int off = R / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, R);
          if (lane == j0 + u) mine = part;
        }
      }
      if (lane < n) a.grad[base + lane] = static_cast<G>(mine);
    }
  }
}

template <int LGR, bool MC, typename T, typename G>
static void sddmm_launch(Ctx *c, const SddmmArgs<T, G> &a) {
  constexpr int RPB = kBlock >> LGR;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((a.nrows + RPB - 1) / RPB, 8192));
  hipLaunchKernelGGL((k_sddmm<LGR, MC, T, G>), blocks, kBlock, 0, c->stream, a);
}

template <typename T, typename G>
static void sddmm_t(svdh_op *op, int64_t k, const void *L1, const void *R1, const double *s1, const void *L2,
                    const void *R2, const double *s2, int64_t ldl, int64_t ldr, void *grad) {
  Ctx *c = op->ctx;
  CsrView v = op->view(0, true);
  constexpr int NV = PanelVec<T>::N;
  if (ldl % NV || ldr % NV) fail(SVDH_ERR_ARG, "factor leading dimensions must be multiples of %d", NV);
  SddmmArgs<T, G> a;
  a.rowptr = v.rowptr;
  a.colidx = v.colidx;
  a.colmap = v.colmap;
  a.l1 = static_cast<const T *>(L1);
  a.r1 = static_cast<const T *>(R1);
  a.l2 = static_cast<const T *>(L2);
  a.r2 = static_cast<const T *>(R2);
  a.s1 = s1;
  a.s2 = s2;
  a.ldl = ldl;
  a.ldr = ldr;
  a.k = (int32_t)k;
  a.nrows = v.nrows;
  a.grad = static_cast<G *>(grad);
  const int pairs = L2 ? 2 : 1;
  const double bytes = (double)v.nnz * (4 + sizeof(G)) + (double)(v.nrows + 1) * 4 +
                       (double)pairs * k * sizeof(T) * (double)(v.nrows + op->ncols);
  KTimer t(c, KC_SDDMM, bytes);
  int lgr = 1;
  while ((1 << lgr) * NV < k && lgr < 6) ++lgr;
  const bool mc = v.colmap != nullptr;
#define SVDH_SDDMM_CASE(L)                                   \
  case L:                                                    \
    if (mc) sddmm_launch<L, true>(c, a); else sddmm_launch<L, false>(c, a); \
    break;
  switch (lgr) {
    SVDH_SDDMM_CASE(1)
    SVDH_SDDMM_CASE(2)
    SVDH_SDDMM_CASE(3)
    SVDH_SDDMM_CASE(4)
    SVDH_SDDMM_CASE(5)
    default:
      if (mc) sddmm_launch<6, true>(c, a); else sddmm_launch<6, false>(c, a);
  }
#undef SVDH_SDDMM_CASE
  HIP_CHECK(hipGetLastError());
}

void sddmm(svdh_op *op, int64_t k, const void *L1, const void *R1, const double *s1, const void *L2, const void *R2,
           const double *s2, int64_t ldl, int64_t ldr, int dt, void *grad, int gdt) {
  if (k <= 0 || k > 64) fail(SVDH_ERR_ARG, "SDDMM factor width %lld outside 1..64", (long long)k);
  if (dt == SVDH_F64 && gdt == SVDH_F64)
    sddmm_t<double, double>(op, k, L1, R1, s1, L2, R2, s2, ldl, ldr, grad);
  else if (dt == SVDH_F32 && gdt == SVDH_F32)
    sddmm_t<float, float>(op, k, L1, R1, s1, L2, R2, s2, ldl, ldr, grad);
  else if (dt == SVDH_F32 && gdt == SVDH_F64)
    sddmm_t<float, double>(op, k, L1, R1, s1, L2, R2, s2, ldl, ldr, grad);
  else if (dt == SVDH_BF16 && gdt == SVDH_F32)
    sddmm_t<bf16_t, float>(op, k, L1, R1, s1, L2, R2, s2, ldl, ldr, grad);
  else if (dt == SVDH_BF16 && gdt == SVDH_F64)
    sddmm_t<bf16_t, double>(op, k, L1, R1, s1, L2, R2, s2, ldl, ldr, grad);
  else
    fail(SVDH_ERR_ARG, "unsupported SDDMM dtypes factors=%d grad=%d", dt, gdt);
}

void spmm_batched(Ctx *c, int nops, svdh_op *const *ops, int direction, int vdt, const void *const *X, int64_t ldx,
                  int xdt, void *const *Y, int64_t ldy, int ydt, int64_t l) {
  for (int b = 0; b < nops; ++b) ensure_values(ops[b], vdt);
  if (vdt == SVDH_F64 && xdt == SVDH_F64 && ydt == SVDH_F64)
    launch_spmm_batched<double, double, double, double>(c, nops, ops, direction, vdt, X, ldx, Y, ldy, l);
  else if (vdt == SVDH_F32 && xdt == SVDH_F32 && ydt == SVDH_F32)
    launch_spmm_batched<float, float, float, float>(c, nops, ops, direction, vdt, X, ldx, Y, ldy, l);
  else
    fail(SVDH_ERR_ARG, "unsupported batched SpMM dtypes values=%d in=%d out=%d", vdt, xdt, ydt);
}

void panel_to_bf16(Ctx *c, const float *src, uint16_t *dst, int64_t n) {
  convert_values(c, src, SVDH_F32, dst, SVDH_BF16, n);
}

}  // namespace svdh

const void *svdh_op::values(int direction, int dtype) const {
  if (direction == 0 && base && compact) return c_vals[dtype].p;
  const svdh_op *own = base ? base : this;
  return direction == 0 ? own->vals[dtype].p : own->t_vals[dtype].p;
}

// ------------------------------------------------------------------------
// Operator construction (called from api.cpp)
// ------------------------------------------------------------------------

namespace svdh {

svdh_op *csr_create(Ctx *c, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rowptr,
                    const int64_t *colidx, const void *vals, int32_t vdt, uint32_t flags) {
  if (nrows < 0 || ncols < 0 || nnz < 0) fail(SVDH_ERR_ARG, "negative shape");
  if (nnz >= (int64_t)INT32_MAX || nrows >= (int64_t)INT32_MAX || ncols >= (int64_t)INT32_MAX)
    fail(SVDH_ERR_ARG, "nnz=%lld or a dimension does not fit 32-bit indices", (long long)nnz);
  if (!rowptr || (nnz > 0 && (!colidx || !vals))) fail(SVDH_ERR_ARG, "null CSR array");
  if (vdt != SVDH_F64 && vdt != SVDH_F32 && vdt != SVDH_BF16) fail(SVDH_ERR_ARG, "bad value dtype");
  if (rowptr[0] != 0 || rowptr[nrows] != nnz) fail(SVDH_ERR_ARG, "row pointers must start at 0 and end at nnz");
  auto op = new svdh_op();
  try {
    op->ctx = static_cast<svdh_ctx *>(c);
    op->nrows = nrows;
    op->ncols = ncols;
    op->nnz = nnz;
    op->base_dtype = vdt;
    const size_t vs = dtype_size(vdt);
    op->h_rowptr.resize(nrows + 1);
    std::vector<int32_t> ci(nnz);
    std::vector<int32_t> tcount(ncols + 1, 0);
    for (int64_t i = 0; i < nrows; ++i) {
      if (rowptr[i + 1] < rowptr[i]) fail(SVDH_ERR_ARG, "row pointers decrease at row %lld", (long long)i);
      op->h_rowptr[i] = (int32_t)rowptr[i];
    }
    op->h_rowptr[nrows] = (int32_t)nnz;
    for (int64_t k = 0; k < nnz; ++k) {
      const int64_t cc = colidx[k];
      if (cc < 0 || cc >= ncols) fail(SVDH_ERR_ARG, "column index %lld out of range", (long long)cc);
      ci[k] = (int32_t)cc;
      ++tcount[cc + 1];
    }
    // Transpose by a stable counting sort on the column index.
    op->h_t_rowptr.assign(ncols + 1, 0);
    for (int64_t j = 0; j < ncols; ++j) op->h_t_rowptr[j + 1] = op->h_t_rowptr[j] + tcount[j + 1];
    std::vector<int32_t> next(op->h_t_rowptr.begin(), op->h_t_rowptr.end() - 1);
    std::vector<int32_t> tci(nnz);
    std::vector<unsigned char> tv(nnz * vs);
    const unsigned char *vb = static_cast<const unsigned char *>(vals);
    for (int64_t i = 0; i < nrows; ++i)
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int32_t pos = next[ci[k]]++;
        tci[pos] = (int32_t)i;
        std::memcpy(&tv[pos * vs], vb + k * vs, vs);
      }
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
      b.alloc(std::max<size_t>(bytes, 16));
      if (bytes) HIP_CHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    op->h_colidx = ci;
    op->h_t_colidx = tci;
    up(op->rowptr, op->h_rowptr.data(), (nrows + 1) * sizeof(int32_t));
    up(op->colidx, ci.data(), nnz * sizeof(int32_t));
    up(op->vals[vdt], vals, nnz * vs);
    up(op->t_rowptr, op->h_t_rowptr.data(), (ncols + 1) * sizeof(int32_t));
    up(op->t_colidx, tci.data(), nnz * sizeof(int32_t));
    up(op->t_vals[vdt], tv.data(), nnz * vs);
    if (flags & SVDH_CSR_KEEP_F32) ensure_values(op, SVDH_F32);
    if (flags & SVDH_CSR_KEEP_BF16) ensure_values(op, SVDH_BF16);
  } catch (...) {
    delete op;
    throw;
  }
  return op;
}

// Compaction of a masked view's forward CSR: one wave per row counts the kept
// entries, the host scans the counts, and one wave per row writes them in
// order (ballot + popcount give each kept entry its slot).
__global__ __launch_bounds__(kBlock) void k_mask_count(int64_t m, const int32_t *rowptr, const int32_t *colidx,
                                                      const int32_t *o2m, int32_t *cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t r = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (r >= m) return;
  int n = 0;
  for (int k = rowptr[r] + lane; k < rowptr[r + 1]; k += 64) n += o2m[colidx[k]] >= 0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if (lane == 0) cnt[r] = n;
}

__global__ __launch_bounds__(kBlock) void k_mask_fill(int64_t m, const int32_t *rowptr, const int32_t *colidx,
                                                     const int32_t *o2m, const int32_t *c_rowptr, int32_t *c_colidx,
                                                     int32_t *perm) {
  const int lane = threadIdx.x & 63;
  const int64_t r = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (r >= m) return;
  int out = c_rowptr[r];
  const int e = rowptr[r + 1];
  for (int k0 = rowptr[r]; k0 < e; k0 += 64) {
    const int k = k0 + lane;
    const int mc = k < e ? o2m[colidx[k]] : -1;
    const uint64_t keep = __ballot(mc >= 0);
    if (mc >= 0) {
      const int pos = out + __popcll(keep & ((1ull << lane) - 1ull));
      c_colidx[pos] = mc;
      perm[pos] = k;
    }
    out += __popcll(keep);
  }
}


static void gather_compact_values(svdh_op *op, int dtype) {
  const svdh_op *own = op->base;
  Ctx *c = op->ctx;
  const int64_t n = op->c_nnz;
  op->c_vals[dtype].alloc(std::max<size_t>(n * dtype_size(dtype), 16));
  if (n > 0) {
    const int g = grid_for(n);
    const int32_t *perm = op->c_perm.as<int32_t>();
    if (dtype == SVDH_F64)
      hipLaunchKernelGGL(k_gather<double>, g, kBlock, 0, c->stream, n, perm, own->vals[dtype].as<double>(),
                         op->c_vals[dtype].as<double>());
    else if (dtype == SVDH_F32)
      hipLaunchKernelGGL(k_gather<float>, g, kBlock, 0, c->stream, n, perm, own->vals[dtype].as<float>(),
                         op->c_vals[dtype].as<float>());
    else
      hipLaunchKernelGGL(k_gather<bf16_t>, g, kBlock, 0, c->stream, n, perm, own->vals[dtype].as<bf16_t>(),
                         op->c_vals[dtype].as<bf16_t>());
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

static void build_compact(Ctx *c, svdh_op *op) {
  const svdh_op *own = op->base;
  const int64_t m = own->nrows;
  DevBuf cnt(std::max<size_t>(m * 4, 16));
  const int g = (int)std::max<int64_t>(1, (m * 64 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_mask_count, g, kBlock, 0, c->stream, m, own->rowptr.as<int32_t>(), own->colidx.as<int32_t>(),
                     op->o2m.as<int32_t>(), cnt.as<int32_t>());
  HIP_CHECK(hipGetLastError());
  std::vector<int32_t> h(m);
  if (m) HIP_CHECK(hipMemcpyAsync(h.data(), cnt.p, m * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  op->h_c_rowptr.assign(m + 1, 0);
  for (int64_t r = 0; r < m; ++r) op->h_c_rowptr[r + 1] = op->h_c_rowptr[r] + h[r];
  op->c_nnz = op->h_c_rowptr[m];
  op->c_rowptr.alloc((m + 1) * 4);
  HIP_CHECK(hipMemcpyAsync(op->c_rowptr.p, op->h_c_rowptr.data(), (m + 1) * 4, hipMemcpyHostToDevice, c->stream));
  op->c_colidx.alloc(std::max<size_t>(op->c_nnz * 4, 16));
  op->c_perm.alloc(std::max<size_t>(op->c_nnz * 4, 16));
  hipLaunchKernelGGL(k_mask_fill, g, kBlock, 0, c->stream, m, own->rowptr.as<int32_t>(), own->colidx.as<int32_t>(),
                     op->o2m.as<int32_t>(), op->c_rowptr.as<int32_t>(), op->c_colidx.as<int32_t>(),
                     op->c_perm.as<int32_t>());
  HIP_CHECK(hipGetLastError());
  op->compact = true;
  for (int d = 0; d < 3; ++d)
    if (own->vals[d].p) gather_compact_values(op, d);
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

svdh_op *mask_create(Ctx *c, const svdh_op *base, const uint8_t *mask) {
  if (!base || !mask) fail(SVDH_ERR_ARG, "null base or mask");
  if (base->base) fail(SVDH_ERR_ARG, "a masked view cannot be masked again");
  auto op = new svdh_op();
  try {
    op->ctx = static_cast<svdh_ctx *>(c);
    op->base = base;
    op->nrows = base->nrows;
    op->base_dtype = base->base_dtype;
    std::vector<int32_t> o2m(base->ncols, -1);
    for (int64_t j = 0; j < base->ncols; ++j)
      if (mask[j]) {
        o2m[j] = (int32_t)op->h_m2o.size();
        op->h_m2o.push_back((int32_t)j);
      }
    op->ncols = (int64_t)op->h_m2o.size();
    int64_t nnz = 0;
    for (int64_t j = 0; j < base->ncols; ++j)
      if (mask[j]) nnz += base->h_t_rowptr[j + 1] - base->h_t_rowptr[j];
    op->nnz = nnz;
    op->o2m.alloc(std::max<size_t>(o2m.size() * 4, 16));
    if (!o2m.empty()) HIP_CHECK(hipMemcpy(op->o2m.p, o2m.data(), o2m.size() * 4, hipMemcpyHostToDevice));
    op->m2o.alloc(std::max<size_t>(op->h_m2o.size() * 4, 16));
    if (!op->h_m2o.empty())
      HIP_CHECK(hipMemcpy(op->m2o.p, op->h_m2o.data(), op->h_m2o.size() * 4, hipMemcpyHostToDevice));
    // Forward products of a mask that drops a quarter of the nonzeros or more
    // read a compacted copy instead of skipping entries of the full rows.
    if (4 * nnz <= 3 * base->nnz) build_compact(c, op);
  } catch (...) {
    delete op;
    throw;
  }
  return op;
}

}  // namespace svdh
