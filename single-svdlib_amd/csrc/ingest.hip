// Input formats on the device: COO triplets and CSC arrays to the CSR
// operator (the CooMatrix and CscMatrix operands of SMat, src/lanczos/
// mod.rs:1415-1504 and 1657-1740, whose products add every stored entry,
// repeated entries included).
//
// Entries get the key (row << 32 | column); a stable LSD radix sort over the
// key bits that can be nonzero (4-bit digits, column bits first) carries the
// values along and orders the entries by (row, column) with repeated pairs
// in input order; each run of equal keys is summed left to right in the value
// type; row pointers come from the sorted rows. Everything here is integer
// and byte traffic bound by HBM: per radix pass one read and one write of the
// (key, value) pairs, staged through LDS so global loads are coalesced. The CSR arrays then take the ordinary operator construction
// (indices narrowed to int32, the CSC copy).
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"

namespace svdh {

svdh_op *csr_create(Ctx *c, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rowptr, const int64_t *colidx,
                    const void *vals, int32_t vdt, uint32_t flags);

namespace {

constexpr int kIngBlock = 256;
constexpr int kTileItems = 2048;  // radix tile: 8 consecutive items per thread
constexpr int kScanTile = 4096;   // scan tile: 16 per thread
constexpr int kDigits = 16;

int ing_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kIngBlock - 1) / kIngBlock, 65536)); }

// key = row << 32 | column (the values travel with the keys through the
// sort); bad = 1 on an index out of range (checked on the host before use).
__global__ void k_coo_pack(int64_t nnz, const int64_t *rows, const int64_t *cols, int64_t m, int64_t n,
                           uint64_t *key, int32_t *bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rows[i], c = cols[i];
    const bool ok = r >= 0 && r < m && c >= 0 && c < n;
    if (!ok) bad[0] = 1;
    key[i] = ok ? ((uint64_t)r << 32 | (uint64_t)c) : 0;
  }
}

// CSC: the column of entry i is the last c with colptr[c] <= i (binary
// search over the validated column pointers).
__global__ void k_csc_pack(int64_t nnz, const int64_t *colptr, int64_t n, const int64_t *rows, int64_t m,
                           uint64_t *key, int32_t *bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = n;  // colptr[lo] <= i < colptr[hi]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (colptr[mid] <= i)
        lo = mid;
      else
        hi = mid;
    }
    const int64_t r = rows[i];
    const bool ok = r >= 0 && r < m;
    if (!ok) bad[0] = 1;
    key[i] = ok ? ((uint64_t)r << 32 | (uint64_t)lo) : 0;
  }
}

// Digit counts of one radix tile: hist[d * ntiles + tile].
__global__ __launch_bounds__(kIngBlock) void k_radix_count(int64_t n, const uint64_t *key, int shift, int64_t ntiles,
                                                           int32_t *hist) {
  __shared__ int32_t h[kDigits];
  if (threadIdx.x < kDigits) h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTileItems;
  for (int j = threadIdx.x; j < kTileItems; j += kIngBlock) {
    const int64_t i = base + j;
    if (i < n) atomicAdd(&h[(int)((key[i] >> shift) & (kDigits - 1))], 1);
  }
  __syncthreads();
  if (threadIdx.x < kDigits) hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one radix tile (payload P carried along). Thread t owns
// items 8t .. 8t + 7 of the
// tile (input order = thread order, then item order), so an item's rank among
// the tile's items of its digit is: (items of that digit in threads < t) +
// (earlier items of that digit in thread t). off = exclusive scan of hist.
template <typename P>
__global__ __launch_bounds__(kIngBlock) void k_radix_scatter(int64_t n, const uint64_t *key, const P *val, int shift,
                                                             int64_t ntiles, const int32_t *off, uint64_t *key_out,
                                                             P *val_out) {
  constexpr int IPT = kTileItems / kIngBlock;  // 8
  __shared__ uint64_t sk[kTileItems];
  __shared__ P sp[kTileItems];
  __shared__ int32_t cnt[kDigits][kIngBlock];
  __shared__ int32_t gsum[kDigits][kIngBlock / 16];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kTileItems;
  const int nt = (int)std::min<int64_t>(kTileItems, n - base);
  for (int j = t; j < kTileItems; j += kIngBlock) {  // coalesced staging
    if (j < nt) {
      sk[j] = key[base + j];
      sp[j] = val[base + j];
    }
  }
#pragma unroll
  for (int d = 0; d < kDigits; ++d) cnt[d][t] = 0;
  __syncthreads();
  for (int q = 0; q < IPT; ++q) {
    const int j = t * IPT + q;
    if (j < nt) ++cnt[(int)((sk[j] >> shift) & (kDigits - 1))][t];
  }
  __syncthreads();
  {
    // level 1: thread (d, g) turns cnt[d][16g .. 16g + 15] into prefixes
    const int d = t & (kDigits - 1), g = t >> 4;
    int32_t s = 0;
    for (int k = 0; k < 16; ++k) {
      const int32_t v = cnt[d][16 * g + k];
      cnt[d][16 * g + k] = s;
      s += v;
    }
    gsum[d][g] = s;
  }
  __syncthreads();
  if (t < kDigits) {  // level 2: prefixes over the 16 groups, plus the tile's global offset
    int32_t s = off[(int64_t)t * ntiles + blockIdx.x];
    for (int g = 0; g < kIngBlock / 16; ++g) {
      const int32_t v = gsum[t][g];
      gsum[t][g] = s;
      s += v;
    }
  }
  __syncthreads();
  for (int q = 0; q < IPT; ++q) {
    const int j = t * IPT + q;
    if (j < nt) {
      const int d = (int)((sk[j] >> shift) & (kDigits - 1));
      const int32_t dst = gsum[d][t >> 4] + cnt[d][t]++;
      key_out[dst] = sk[j];
      val_out[dst] = sp[j];
    }
  }
}

// Exclusive scan, three passes: tile scans with tile totals, a scan of the
// totals in one workgroup, then the tile offsets added.
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t *sh, int32_t &total) {
  // kIngBlock threads; returns the exclusive prefix of v over the block
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int32_t wo = 0, tot = 0;
  for (int k = 0; k < kIngBlock / 64; ++k) {
    if (k < w) wo += sh[k];
    tot += sh[k];
  }
  __syncthreads();
  total = tot;
  return wo + x - v;
}

__global__ __launch_bounds__(kIngBlock) void k_scan_tiles(int64_t n, const int32_t *in, int32_t *out,
                                                          int32_t *tile_sum) {
  constexpr int IPT = kScanTile / kIngBlock;  // 16
  __shared__ int32_t sh[kIngBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * IPT;
  int32_t v[IPT], s = 0;
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    v[q] = base + q < n ? in[base + q] : 0;
    s += v[q];
  }
  int32_t total;
  int32_t run = block_excl_scan(s, sh, total);
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the tile totals in place; total[0] = sum.
__global__ __launch_bounds__(kIngBlock) void k_scan_top(int64_t ntiles, int32_t *tile_sum, int32_t *total) {
  __shared__ int32_t sh[kIngBlock / 64];
  int32_t carry = 0;
  for (int64_t b = 0; b < ntiles; b += kIngBlock) {
    const int64_t i = b + threadIdx.x;
    const int32_t v = i < ntiles ? tile_sum[i] : 0;
    int32_t tot;
    const int32_t ex = block_excl_scan(v, sh, tot);
    if (i < ntiles) tile_sum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) total[0] = carry;
}

__global__ void k_scan_add(int64_t n, int32_t *out, const int32_t *tile_off) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] += tile_off[i / kScanTile];
}

__global__ void k_run_heads(int64_t n, const uint64_t *key, int32_t *head) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// One thread per run head: the run's values (sorted along with the keys, so
// in input order within the run) summed left to right, written at the
// head's rank with its row and column.
template <typename T>
__global__ void k_run_sum(int64_t n, const uint64_t *key, const int32_t *head, const int32_t *rank, const T *vals,
                          int32_t *out_row, int32_t *out_col, T *out_val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (!head[i]) continue;
    const uint64_t k = key[i];
    T s = vals[i];
    for (int64_t j = i + 1; j < n && key[j] == k; ++j) s += vals[j];
    const int32_t p = rank[i];
    out_row[p] = (int32_t)(k >> 32);
    out_col[p] = (int32_t)(k & 0xffffffffu);
    out_val[p] = s;
  }
}

// rowptr[r] = first output entry of row r (rows sorted): entry p sets the
// pointers of the rows after its predecessor's row up to its own; the last
// entry also sets those after it.
__global__ void k_row_pointers(int64_t u, int64_t m, const int32_t *row, int32_t *rowptr) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < u; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = row[p], prev = p > 0 ? row[p - 1] : -1;
    for (int32_t q = prev + 1; q <= r; ++q) rowptr[q] = (int32_t)p;
    if (p == u - 1)
      for (int64_t q = r + 1; q <= m; ++q) rowptr[q] = (int32_t)u;
  }
}

int bits_for(int64_t extent) {  // bits of the largest index extent - 1
  int b = 0;
  while (b < 31 && ((int64_t)1 << b) < extent) ++b;
  return b;
}

struct DevScan {
  Ctx *c;
  DevBuf tiles, total;
  // out = exclusive scan of in (n int32); with want_total, returns the sum
  // (and waits for it), otherwise 0 with the kernels only enqueued
  int64_t run(const int32_t *in, int32_t *out, int64_t n, bool want_total = true) {
    if (n == 0) return 0;
    const int64_t nt = (n + kScanTile - 1) / kScanTile;
    tiles.ensure(std::max<size_t>(nt * sizeof(int32_t), 16));
    total.ensure(16);
    hipLaunchKernelGGL(k_scan_tiles, (int)nt, kIngBlock, 0, c->stream, n, in, out, tiles.as<int32_t>());
    hipLaunchKernelGGL(k_scan_top, 1, kIngBlock, 0, c->stream, nt, tiles.as<int32_t>(), total.as<int32_t>());
    hipLaunchKernelGGL(k_scan_add, ing_grid(n), kIngBlock, 0, c->stream, n, out, tiles.as<int32_t>());
    HIP_CHECK(hipGetLastError());
    if (!want_total) return 0;
    int32_t h = 0;
    HIP_CHECK(hipMemcpyAsync(&h, total.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return h;
  }
};

// Sorted, summed CSR arrays of the (key, value) pairs on the device.
template <typename T>
svdh_op *finish_csr(Ctx *c, int64_t m, int64_t n, int64_t nnz, DevBuf &key, DevBuf &dvals, int32_t vdt,
                    uint32_t flags) {
  std::vector<int32_t> rp32(m + 1, 0), ci32, rows32;
  std::vector<T> v;
  int64_t u = 0;
  if (nnz > 0) {
    DevBuf key2(nnz * sizeof(uint64_t)), val2(nnz * sizeof(T));
    const int64_t ntiles = (nnz + kTileItems - 1) / kTileItems;
    DevBuf hist(std::max<size_t>(kDigits * ntiles * sizeof(int32_t), 16));
    DevBuf off(std::max<size_t>(kDigits * ntiles * sizeof(int32_t), 16));
    DevScan scan{c};
    DevBuf orow, ocol, oval, drp((m + 1) * sizeof(int32_t));
    {
      KTimer tm(c, KC_INGEST, 0.0);
      uint64_t *k0 = key.as<uint64_t>(), *k1 = key2.as<uint64_t>();
      T *p0 = dvals.as<T>(), *p1 = val2.as<T>();
      std::vector<int> shifts;
      for (int s = 0; s < bits_for(n); s += 4) shifts.push_back(s);
      for (int s = 0; s < bits_for(m); s += 4) shifts.push_back(32 + s);
      for (int sh : shifts) {
        hipLaunchKernelGGL(k_radix_count, (int)ntiles, kIngBlock, 0, c->stream, nnz, k0, sh, ntiles,
                           hist.as<int32_t>());
        scan.run(hist.as<int32_t>(), off.as<int32_t>(), kDigits * ntiles, false);
        hipLaunchKernelGGL(k_radix_scatter<T>, (int)ntiles, kIngBlock, 0, c->stream, nnz, k0, p0, sh, ntiles,
                           off.as<int32_t>(), k1, p1);
        HIP_CHECK(hipGetLastError());
        std::swap(k0, k1);
        std::swap(p0, p1);
      }
      DevBuf head(nnz * sizeof(int32_t)), rank(nnz * sizeof(int32_t));
      hipLaunchKernelGGL(k_run_heads, ing_grid(nnz), kIngBlock, 0, c->stream, nnz, k0, head.as<int32_t>());
      u = scan.run(head.as<int32_t>(), rank.as<int32_t>(), nnz);
      orow.alloc(std::max<size_t>(u * sizeof(int32_t), 16));
      ocol.alloc(std::max<size_t>(u * sizeof(int32_t), 16));
      oval.alloc(std::max<size_t>(u * sizeof(T), 16));
      hipLaunchKernelGGL(k_run_sum<T>, ing_grid(nnz), kIngBlock, 0, c->stream, nnz, k0, head.as<int32_t>(),
                         rank.as<int32_t>(), p0, orow.as<int32_t>(), ocol.as<int32_t>(), oval.as<T>());
      hipLaunchKernelGGL(k_row_pointers, ing_grid(u), kIngBlock, 0, c->stream, u, m, orow.as<int32_t>(),
                         drp.as<int32_t>());
      HIP_CHECK(hipGetLastError());
    }
    // the CSR arrays to the host for the operator construction (schedules
    // are planned from host copies of the indices)
    ci32.resize(u);
    v.resize(u);
    HIP_CHECK(hipMemcpyAsync(rp32.data(), drp.p, (m + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(ci32.data(), ocol.p, u * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(v.data(), oval.p, u * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  std::vector<int64_t> rp(m + 1), ci(u);
  for (int64_t i = 0; i <= m; ++i) rp[i] = nnz > 0 ? rp32[i] : 0;
  for (int64_t k = 0; k < u; ++k) ci[k] = ci32[k];
  return csr_create(c, m, n, u, rp.data(), ci.data(), v.data(), vdt, flags);
}

void check_shape(int64_t m, int64_t n, int64_t nnz, int32_t vdt) {
  if (m < 0 || n < 0 || nnz < 0) fail(SVDH_ERR_ARG, "negative shape");
  if (nnz >= (int64_t)INT32_MAX || m >= (int64_t)INT32_MAX || n >= (int64_t)INT32_MAX)
    fail(SVDH_ERR_ARG, "nnz=%lld or a dimension does not fit 32-bit indices", (long long)nnz);
  if (vdt != SVDH_F64 && vdt != SVDH_F32) fail(SVDH_ERR_ARG, "COO and CSC values must be F64 or F32");
}

template <typename PackFn>
svdh_op *ingest(Ctx *c, int64_t m, int64_t n, int64_t nnz, const void *vals, int32_t vdt, uint32_t flags,
                PackFn pack) {
  const size_t vs = dtype_size(vdt);
  DevBuf key(std::max<size_t>(nnz * sizeof(uint64_t), 16));
  DevBuf dvals(std::max<size_t>(nnz * vs, 16)), bad(16);
  if (nnz > 0) {
    HIP_CHECK(hipMemsetAsync(bad.p, 0, 16, c->stream));
    HIP_CHECK(hipMemcpyAsync(dvals.p, vals, nnz * vs, hipMemcpyHostToDevice, c->stream));
    pack(key, bad.as<int32_t>());
    HIP_CHECK(hipGetLastError());
    int32_t hb = 0;
    HIP_CHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (hb) fail(SVDH_ERR_ARG, "an index is out of range");
  }
  if (vdt == SVDH_F64) return finish_csr<double>(c, m, n, nnz, key, dvals, vdt, flags);
  return finish_csr<float>(c, m, n, nnz, key, dvals, vdt, flags);
}

}  // namespace

svdh_op *coo_create(Ctx *c, int64_t m, int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols,
                    const void *vals, int32_t vdt, uint32_t flags) {
  check_shape(m, n, nnz, vdt);
  if (nnz > 0 && (!rows || !cols || !vals)) fail(SVDH_ERR_ARG, "null COO array");
  return ingest(c, m, n, nnz, vals, vdt, flags, [&](DevBuf &key, int32_t *bad) {
    DevBuf dr(nnz * sizeof(int64_t)), dc(nnz * sizeof(int64_t));
    HIP_CHECK(hipMemcpyAsync(dr.p, rows, nnz * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(dc.p, cols, nnz * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    KTimer tm(c, KC_INGEST, 0.0);
    hipLaunchKernelGGL(k_coo_pack, ing_grid(nnz), kIngBlock, 0, c->stream, nnz, dr.as<int64_t>(), dc.as<int64_t>(),
                       m, n, key.as<uint64_t>(), bad);
    HIP_CHECK(hipStreamSynchronize(c->stream));  // dr, dc are freed on return
  });
}

svdh_op *csc_create(Ctx *c, int64_t m, int64_t n, int64_t nnz, const int64_t *colptr, const int64_t *rows,
                    const void *vals, int32_t vdt, uint32_t flags) {
  check_shape(m, n, nnz, vdt);
  if (!colptr || (nnz > 0 && (!rows || !vals))) fail(SVDH_ERR_ARG, "null CSC array");
  if (colptr[0] != 0 || colptr[n] != nnz) fail(SVDH_ERR_ARG, "column pointers must start at 0 and end at nnz");
  for (int64_t j = 0; j < n; ++j)
    if (colptr[j + 1] < colptr[j]) fail(SVDH_ERR_ARG, "column pointers decrease at column %lld", (long long)j);
  return ingest(c, m, n, nnz, vals, vdt, flags, [&](DevBuf &key, int32_t *bad) {
    DevBuf dp((n + 1) * sizeof(int64_t)), dr(nnz * sizeof(int64_t));
    HIP_CHECK(hipMemcpyAsync(dp.p, colptr, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(dr.p, rows, nnz * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    KTimer tm(c, KC_INGEST, 0.0);
    hipLaunchKernelGGL(k_csc_pack, ing_grid(nnz), kIngBlock, 0, c->stream, nnz, dp.as<int64_t>(), n,
                       dr.as<int64_t>(), m, key.as<uint64_t>(), bad);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

}  // namespace svdh
