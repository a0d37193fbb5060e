// Gather-ceiling probe for the SpMM (measurement only; built as
// libsvdh_probe.so beside libsvdh.so and called by bench.py). It runs the
// SpMM's memory access without the arithmetic or the output: for every entry
// of a column-index stream (the column indices of a CSR matrix in storage
// order), one 16-lane group loads the panel row X[col, 0:l] (16 bytes per
// lane, as the SpMM kernels do) and folds it into a register sum. Its rate is
// the measured ceiling of the SpMM's gathers on that stream: the bytes the
// SpMM must gather (entries x row bytes) over the probe's time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>

namespace {

constexpr int kBlock = 256;

// Each wave takes a contiguous range of the stream; per step it loads 64
// indices with one coalesced load, then its four 16-lane groups gather U
// rows each per round (entries j + 4 u + g), all U loads in flight before
// the first is used.
// Source-blocked mode (nblk > 1): the stream is grouped by source block
// (blk_off[x] .. blk_off[x + 1]) and workgroup w takes only block w mod nblk,
// as the SpMM's segment schedule does: workgroups are dispatched round-robin
// over the XCDs, so with nblk = 8 each XCD gathers from one eighth of the
// panel, which its L2 can hold.
template <int U>
__global__ __launch_bounds__(kBlock) void k_gather(const int32_t *stream, int64_t n, const uint4 *panel,
                                                   int64_t ld16, int vecs, int64_t per_wave, float *sink,
                                                   const int64_t *blk_off, int nblk) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, q = lane & 15;
  int64_t b0, b1;
  if (nblk > 1) {
    const int x = blockIdx.x % nblk;
    const int64_t slots = (int64_t)(gridDim.x / nblk) * (kBlock / 64);
    const int64_t slot = (int64_t)(blockIdx.x / nblk) * (kBlock / 64) + (threadIdx.x >> 6);
    const int64_t lo = blk_off[x], hi = blk_off[x + 1];
    const int64_t per = ((hi - lo + slots - 1) / slots + 63) / 64 * 64;
    b0 = lo + slot * per;
    b1 = min(hi, b0 + per);
    if ((int64_t)blockIdx.x >= (int64_t)(gridDim.x / nblk) * nblk) b0 = b1 = 0;  // the remainder takes nothing
  } else {
    const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    b0 = wid * per_wave;
    b1 = min(n, b0 + per_wave);
  }
  const bool qok = q < vecs;
  float acc = 0.f;
  for (int64_t base = b0; base < b1; base += 64) {
    const int64_t k = base + lane;
    const int c = k < b1 ? stream[k] : -1;
#pragma unroll
    for (int j = 0; j < 64; j += 4 * U) {
      uint4 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int cj = __shfl(c, j + 4 * u + g, 64);
        v[u] = (cj >= 0 && qok) ? panel[(int64_t)cj * ld16 + q] : make_uint4(0u, 0u, 0u, 0u);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        acc += __uint_as_float(v[u].x) + __uint_as_float(v[u].y) + __uint_as_float(v[u].z) +
               __uint_as_float(v[u].w);
    }
  }
  sink[blockIdx.x * (int64_t)blockDim.x + threadIdx.x] = acc;
}

template <int U>
static void launch(const int32_t *stream, int64_t n, const uint4 *panel, int64_t ld16, int vecs, int waves,
                   float *sink, const int64_t *blk_off, int nblk, hipStream_t s) {
  const int64_t per_wave = ((n + waves - 1) / waves + 63) / 64 * 64;
  const int blocks = (int)((waves * 64 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_gather<U>, blocks, kBlock, 0, s, stream, n, panel, ld16, vecs, per_wave, sink, blk_off,
                     nblk);
}

}  // namespace

extern "C" {

// Gathers row_bytes of every panel row named by stream (n entries, indices
// below nsrc; rows are ld_bytes apart, ld_bytes a multiple of 16) from a
// device panel filled with small values; `waves` waves share the stream in
// contiguous ranges, U (1, 2, 4 or 8) loads per 16-lane group in flight.
// Returns the average kernel time over reps launches (after one warm-up) in
// *ms. Status 0 on success, else a HIP error code.
// blk_off_host (nullable, nblk + 1 offsets with nblk <= 8): the stream is
// grouped by source block and workgroup w gathers only from block w mod nblk.
int32_t svdh_probe_gather(int32_t device, const int32_t *stream_host, int64_t n, int64_t nsrc, int64_t ld_bytes,
                          int64_t row_bytes, int32_t u, int32_t waves, int32_t reps, const int64_t *blk_off_host,
                          int32_t nblk, double *ms) {
  if (ld_bytes % 16 || row_bytes > ld_bytes || row_bytes > 256 || n < 0 || nsrc <= 0 || reps <= 0 || waves <= 0)
    return -1;
  if (nblk < 1 || nblk > 8 || (nblk > 1 && !blk_off_host)) return -1;
  if (nblk > 1) {
    if (blk_off_host[0] != 0 || blk_off_host[nblk] != n) return -1;
    for (int x = 0; x < nblk; ++x)
      if (blk_off_host[x + 1] < blk_off_host[x]) return -1;
    if (waves * 64 / kBlock < nblk) return -1;
  }
  for (int64_t i = 0; i < n; ++i)
    if (stream_host[i] < 0 || stream_host[i] >= nsrc) return -2;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (int32_t)e;
  int32_t *d_stream = nullptr;
  uint4 *d_panel = nullptr;
  float *d_sink = nullptr;
  int64_t *d_off = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int32_t st = 0;
  const int vecs = (int)((row_bytes + 15) / 16);
  const int64_t ld16 = ld_bytes / 16;
  do {
    if ((e = hipStreamCreate(&s)) != hipSuccess) break;
    if ((e = hipMalloc(&d_stream, std::max<int64_t>(n, 1) * sizeof(int32_t))) != hipSuccess) break;
    if ((e = hipMalloc(&d_panel, nsrc * ld_bytes)) != hipSuccess) break;
    if ((e = hipMalloc(&d_sink, (int64_t)(waves + 4) * 64 * sizeof(float))) != hipSuccess) break;
    if ((e = hipMemcpy(d_stream, stream_host, n * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemset(d_panel, 0x3c, nsrc * ld_bytes)) != hipSuccess) break;  // small finite floats
    if ((e = hipMalloc(&d_off, 9 * sizeof(int64_t))) != hipSuccess) break;
    if (nblk > 1 &&
        (e = hipMemcpy(d_off, blk_off_host, (nblk + 1) * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess)
      break;
    if ((e = hipEventCreate(&a)) != hipSuccess) break;
    if ((e = hipEventCreate(&b)) != hipSuccess) break;
    auto run = [&]() {
      switch (u) {
        case 1: launch<1>(d_stream, n, d_panel, ld16, vecs, waves, d_sink, d_off, nblk, s); break;
        case 2: launch<2>(d_stream, n, d_panel, ld16, vecs, waves, d_sink, d_off, nblk, s); break;
        case 8: launch<8>(d_stream, n, d_panel, ld16, vecs, waves, d_sink, d_off, nblk, s); break;
        default: launch<4>(d_stream, n, d_panel, ld16, vecs, waves, d_sink, d_off, nblk, s); break;
      }
    };
    run();
    if ((e = hipEventRecord(a, s)) != hipSuccess) break;
    for (int i = 0; i < reps; ++i) run();
    if ((e = hipEventRecord(b, s)) != hipSuccess) break;
    if ((e = hipEventSynchronize(b)) != hipSuccess) break;
    float t = 0.f;
    if ((e = hipEventElapsedTime(&t, a, b)) != hipSuccess) break;
    *ms = (double)t / reps;
  } while (false);
  if (e != hipSuccess) st = (int32_t)e;
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (d_stream) (void)hipFree(d_stream);
  if (d_panel) (void)hipFree(d_panel);
  if (d_sink) (void)hipFree(d_sink);
  if (d_off) (void)hipFree(d_off);
  if (s) (void)hipStreamDestroy(s);
  return st;
}

}  // extern "C"
