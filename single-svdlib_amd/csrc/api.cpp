// C ABI of the svdh library (include/svdh.h): contexts, operators, the SMat
// entry points and the whole-solve entry points. Every entry point converts
// exceptions into a status code and a thread-local message; none aborts.
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace svdh {

static thread_local std::string g_last_error;

// Host result arrays (SvdRec u, s, vt) come from pinned memory that
// svdh_result_free returns to a process-wide pool instead of releasing: the
// next result of a similar size reuses it, so the hand-over of U and V^T is
// a DMA copy into memory that is already mapped (no page faults). Free
// buffers beyond kPoolKeep bytes are unpinned.
namespace {
std::mutex g_pool_mu;
std::unordered_map<void *, size_t> g_pool_live;  // pinned buffers handed out: bytes
std::multimap<size_t, void *> g_pool_free;        // pinned buffers to reuse, by size
size_t g_pool_free_bytes = 0;
constexpr size_t kPoolKeep = size_t(4) << 30;
}  // namespace

double *result_alloc(int64_t count) { return static_cast<double *>(result_alloc_bytes((size_t)count * sizeof(double))); }
float *result_alloc_f32(int64_t count) { return static_cast<float *>(result_alloc_bytes((size_t)count * sizeof(float))); }

void *result_alloc_bytes(size_t want) {
  const size_t bytes = std::max<size_t>(std::max<size_t>(want, 8), 64);
  {
    std::lock_guard<std::mutex> g(g_pool_mu);
    auto it = g_pool_free.lower_bound(bytes);
    if (it != g_pool_free.end() && it->first <= 2 * bytes) {
      void *p = it->second;
      g_pool_free_bytes -= it->first;
      g_pool_live[p] = it->first;
      g_pool_free.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
    (void)hipGetLastError();
    p = std::malloc(bytes);  // unpinned: still correct, the copy is slower
    if (!p) fail(SVDH_ERR_ALLOC, "result allocation of %zu bytes failed", bytes);
    return p;
  }
  std::lock_guard<std::mutex> g(g_pool_mu);
  g_pool_live[p] = bytes;
  return p;
}

void result_release(void *p) {
  if (!p) return;
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> g(g_pool_mu);
    auto it = g_pool_live.find(p);
    if (it == g_pool_live.end()) {
      std::free(p);
      return;
    }
    g_pool_free.emplace(it->second, p);
    g_pool_free_bytes += it->second;
    g_pool_live.erase(it);
    while (g_pool_free_bytes > kPoolKeep && !g_pool_free.empty()) {
      auto big = std::prev(g_pool_free.end());
      g_pool_free_bytes -= big->first;
      drop.push_back(big->second);
      g_pool_free.erase(big);
    }
  }
  for (void *q : drop) (void)hipHostFree(q);
}

void fail(int32_t code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

void DevBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 16;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    fail(SVDH_ERR_ALLOC, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
  }
  bytes = n;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void Ctx::sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

void *Ctx::upload(const void *host, size_t n) {
  if (n > kRingSlot) fail(SVDH_ERR_ARG, "upload of %zu bytes exceeds the staging slot", n);
  const int i = ring_next;
  ring_next = (ring_next + 1) % kRing;
  HIP_CHECK(hipEventSynchronize(ring_ev[i]));  // the slot's previous copy has left the host buffer
  char *h = h_ring + (size_t)i * kRingSlot;
  std::memcpy(h, host, n);
  char *d = d_ring.as<char>() + (size_t)i * kRingSlot;
  fetch_from_host(this, dh_ring + (size_t)i * kRingSlot, d, n);
  HIP_CHECK(hipEventRecord(ring_ev[i], stream));
  return d;
}

hipEvent_t Ctx::get_event() {
  if (!free_events.empty()) {
    hipEvent_t e = free_events.back();
    free_events.pop_back();
    return e;
  }
  hipEvent_t e;
  HIP_CHECK(hipEventCreate(&e));
  return e;
}

void Ctx::drain_profile() {
  if (pending.empty()) return;
  HIP_CHECK(hipStreamSynchronize(stream));
  for (auto &p : pending) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    prof_ms[p.cls] += ms;
    prof_n[p.cls] += 1;
    prof_bytes[p.cls] += p.bytes;
    free_events.push_back(p.a);
    free_events.push_back(p.b);
  }
  pending.clear();
}

KTimer::KTimer(Ctx *ctx, int k, double algorithmic_bytes) : c(ctx), cls(k), bytes(algorithmic_bytes) {
  if (!c || !c->profile) return;
  a = c->get_event();
  HIP_CHECK(hipEventRecord(a, c->stream));
}

KTimer::~KTimer() {
  if (!a) return;
  hipEvent_t b = c->get_event();
  (void)hipEventRecord(b, c->stream);
  c->pending.push_back(Ctx::Pending{a, b, cls, bytes});
  if (c->pending.size() > 4096) c->drain_profile();
}

svdh_op *csr_create(Ctx *c, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rowptr, const int64_t *colidx,
                    const void *vals, int32_t vdt, uint32_t flags);
svdh_op *mask_create(Ctx *c, const svdh_op *base, const uint8_t *mask);
svdh_op *coo_create(Ctx *c, int64_t m, int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols,
                    const void *vals, int32_t vdt, uint32_t flags);
svdh_op *csc_create(Ctx *c, int64_t m, int64_t n, int64_t nnz, const int64_t *colptr, const int64_t *rows,
                    const void *vals, int32_t vdt, uint32_t flags);
int32_t las2_solve(svdh_ctx *ctx, svdh_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
                   double kappa, uint32_t random_seed, const double *start, int32_t precision, svdh_result *out,
                   const RsvdDeviceOut *dev, bool legacy = false);
int32_t rsvd_solve(svdh_ctx *ctx, svdh_op *op, int64_t target_rank, int64_t n_oversamples, int64_t n_power_iters,
                   int32_t normalizer, int32_t mean_center, uint64_t seed, int32_t has_seed, const double *omega,
                   int32_t precision, double tol, int64_t max_power_iters, svdh_result *out,
                   const RsvdDeviceOut *dev);

int32_t rsvd_batched_solve(svdh_ctx *ctx, int64_t nops, svdh_op *const *ops, int64_t target_rank,
                           int64_t n_oversamples, int64_t n_power_iters, int32_t normalizer, uint64_t seed,
                           int32_t has_seed, const double *const *omegas, int32_t precision, svdh_result *outs,
                           const BatchedDeviceOut *dev);
int32_t backward_solve(svdh_ctx *ctx, svdh_op *op, const svdh_result *res, int32_t loss, int64_t k, const double *y,
                       int32_t precision, double tol, int64_t max_iters, void *grad, int32_t grad_dtype,
                       svdh_backward_info *info);

template <typename F>
static int32_t guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return SVDH_OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return SVDH_ERR_ALLOC;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return SVDH_ERR_ARG;
  }
}

static void check_ctx(svdh_ctx *ctx) {
  if (!ctx) fail(SVDH_ERR_ARG, "null context");
  HIP_CHECK(hipSetDevice(ctx->device));
}

static void check_op(svdh_ctx *ctx, const svdh_op *op) {
  if (!op) fail(SVDH_ERR_ARG, "null operator");
  if (op->ctx != ctx) fail(SVDH_ERR_ARG, "operator belongs to another context");
}

}  // namespace svdh

using namespace svdh;

extern "C" {

const char *svdh_last_error(void) { return g_last_error.c_str(); }

#ifndef SVDH_BUILD_ID
#define SVDH_BUILD_ID "unknown"
#endif
const char *svdh_build_id(void) { return SVDH_BUILD_ID; }

int32_t svdh_ctx_create(int32_t device, svdh_ctx **out) {
  return guarded([&] {
    if (!out) fail(SVDH_ERR_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
      (void)hipGetLastError();
      fail(SVDH_ERR_NO_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) fail(SVDH_ERR_ARG, "device %d out of range (%d devices)", device, count);
    HIP_CHECK(hipSetDevice(device));
    auto c = new svdh_ctx();
    try {
      c->device = device;
      HIP_CHECK(hipStreamCreate(&c->stream));
      int ncu = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        c->ncu = ncu;
      (void)hipGetLastError();
      c->scal.alloc(S_NSLOT * sizeof(double));
      HIP_CHECK(hipMemset(c->scal.p, 0, S_NSLOT * sizeof(double)));
      // coherent and mapped: k_slot_finish writes the step's scalars straight into it
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c->h_scal), S_NSLOT * sizeof(double),
                              hipHostMallocCoherent | hipHostMallocMapped));
      HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->dh_scal), c->h_scal, 0));
      c->red_part.alloc(1 << 20);
      c->stage.alloc(1 << 16);
      // fine-grained (coherent) pinned memory: kernels read uploads from it
      // and write small results into it, with no copy engine round trip
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c->h_ring), Ctx::kRing * Ctx::kRingSlot,
                              hipHostMallocCoherent | hipHostMallocMapped));
      HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->dh_ring), c->h_ring, 0));
      c->d_ring.alloc(Ctx::kRing * Ctx::kRingSlot);
      for (int i = 0; i < Ctx::kRing; ++i) {
        HIP_CHECK(hipEventCreateWithFlags(&c->ring_ev[i], hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(c->ring_ev[i], c->stream));
      }
      scratch_init(c);  // the device is current here
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c->h_small), Ctx::kSmall * sizeof(double),
                              hipHostMallocCoherent | hipHostMallocMapped));
      HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->d_small), c->h_small, 0));
      HIP_CHECK(hipEventCreateWithFlags(&c->small_ev, hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&c->small_ev2, hipEventDisableTiming));
    } catch (...) {
      if (c->h_ring) (void)hipHostFree(c->h_ring);
      if (c->h_small) (void)hipHostFree(c->h_small);
      if (c->h_scal) (void)hipHostFree(c->h_scal);
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
      throw;
    }
    *out = c;
  });
}

int32_t svdh_ctx_destroy(svdh_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    for (int i = 0; i < Ctx::kRing; ++i)
      if (ctx->ring_ev[i]) (void)hipEventDestroy(ctx->ring_ev[i]);
    if (ctx->small_ev) (void)hipEventDestroy(ctx->small_ev);
    if (ctx->small_ev2) (void)hipEventDestroy(ctx->small_ev2);
    if (ctx->h_ring) (void)hipHostFree(ctx->h_ring);
    if (ctx->h_small) (void)hipHostFree(ctx->h_small);
    if (ctx->h_scal) (void)hipHostFree(ctx->h_scal);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int32_t svdh_ctx_synchronize(svdh_ctx *ctx) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->sync();
  });
}

int32_t svdh_ctx_profile(svdh_ctx *ctx, int32_t enable) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->drain_profile();
    for (int i = 0; i < KC_COUNT; ++i) {
      ctx->prof_ms[i] = 0.0;
      ctx->prof_n[i] = 0;
      ctx->prof_bytes[i] = 0.0;
    }
    ctx->profile = enable != 0;
  });
}

int32_t svdh_ctx_kernel_time(svdh_ctx *ctx, int32_t kernel_class, double *total_ms, int64_t *launches,
                             double *algorithmic_bytes) {
  return guarded([&] {
    check_ctx(ctx);
    if (kernel_class < 0 || kernel_class >= KC_COUNT) fail(SVDH_ERR_ARG, "bad kernel class");
    ctx->drain_profile();
    if (total_ms) *total_ms = ctx->prof_ms[kernel_class];
    if (launches) *launches = ctx->prof_n[kernel_class];
    if (algorithmic_bytes) *algorithmic_bytes = ctx->prof_bytes[kernel_class];
  });
}

int32_t svdh_csr_create(svdh_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rowptr,
                        const int64_t *colidx, const void *vals, int32_t val_dtype, uint32_t flags, svdh_op **out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) fail(SVDH_ERR_ARG, "null output pointer");
    *out = csr_create(ctx, nrows, ncols, nnz, rowptr, colidx, vals, val_dtype, flags);
  });
}

int32_t svdh_coo_create(svdh_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rows,
                        const int64_t *cols, const void *vals, int32_t val_dtype, uint32_t flags, svdh_op **out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) fail(SVDH_ERR_ARG, "null output pointer");
    *out = coo_create(ctx, nrows, ncols, nnz, rows, cols, vals, val_dtype, flags);
  });
}

int32_t svdh_csc_create(svdh_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *colptr,
                        const int64_t *rowidx, const void *vals, int32_t val_dtype, uint32_t flags, svdh_op **out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) fail(SVDH_ERR_ARG, "null output pointer");
    *out = csc_create(ctx, nrows, ncols, nnz, colptr, rowidx, vals, val_dtype, flags);
  });
}

int32_t svdh_mask_create(svdh_ctx *ctx, const svdh_op *base, const uint8_t *mask, svdh_op **out) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, base);
    if (!out) fail(SVDH_ERR_ARG, "null output pointer");
    *out = mask_create(ctx, base, mask);
  });
}

int32_t svdh_op_destroy(svdh_op *op) {
  return guarded([&] {
    if (!op) return;
    (void)hipSetDevice(op->ctx->device);
    (void)hipStreamSynchronize(op->ctx->stream);
    delete op;
  });
}

int32_t svdh_op_shape(const svdh_op *op, int64_t *nrows, int64_t *ncols, int64_t *nnz) {
  return guarded([&] {
    if (!op) fail(SVDH_ERR_ARG, "null operator");
    if (nrows) *nrows = op->nrows;
    if (ncols) *ncols = op->ncols;
    if (nnz) *nnz = op->nnz;
  });
}

int32_t svdh_spmv(svdh_ctx *ctx, const svdh_op *op_, int32_t transposed, const void *x, void *y, int32_t vec_dtype,
                  int32_t on_device) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op_);
    svdh_op *op = const_cast<svdh_op *>(op_);
    if (vec_dtype != SVDH_F64 && vec_dtype != SVDH_F32) fail(SVDH_ERR_ARG, "vector dtype must be F64 or F32");
    if (!x || !y) fail(SVDH_ERR_ARG, "null vector");
    const int dir = transposed ? 1 : 0;
    const int64_t nin = dir ? op->nrows : op->ncols;
    const int64_t nout = dir ? op->ncols : op->nrows;
    const size_t es = dtype_size(vec_dtype);
    // Values: the operator's own dtype for F64 vectors; FP32 values for FP32
    // vectors unless the matrix is stored in FP64 only (then FP64 values).
    const svdh_op *own = op->base ? op->base : op;
    int vdt = own->base_dtype;
    if (vec_dtype == SVDH_F64 && vdt == SVDH_BF16) fail(SVDH_ERR_ARG, "BF16 operators take F32 vectors");
    if (vdt == SVDH_BF16) fail(SVDH_ERR_ARG, "SpMV on BF16 values is not provided");
    if (on_device) {
      spmv(op, dir, vdt, x, y, vec_dtype);
      ctx->sync();
      return;
    }
    DevBuf dx(std::max<size_t>(nin * es, 16)), dy(std::max<size_t>(nout * es, 16));
    HIP_CHECK(hipMemcpyAsync(dx.p, x, nin * es, hipMemcpyHostToDevice, ctx->stream));
    spmv(op, dir, vdt, dx.p, dy.p, vec_dtype);
    HIP_CHECK(hipMemcpyAsync(y, dy.p, nout * es, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

int32_t svdh_spmv_noalloc(svdh_ctx *ctx, const svdh_op *op_, int32_t transposed, const void *x, void *y,
                          int32_t vec_dtype) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op_);
    svdh_op *op = const_cast<svdh_op *>(op_);
    if (vec_dtype != SVDH_F64 && vec_dtype != SVDH_F32) fail(SVDH_ERR_ARG, "vector dtype must be F64 or F32");
    if (!x || !y) fail(SVDH_ERR_ARG, "null vector");
    // the value dtype svdh_spmv picks for these vectors
    const svdh_op *own = op->base ? op->base : op;
    const int vdt = own->base_dtype;
    if (vdt == SVDH_BF16) fail(SVDH_ERR_ARG, "SpMV on BF16 values is not provided");
    spmv_noalloc(op, transposed ? 1 : 0, vdt, x, y, vec_dtype);
  });
}

// Column-major host panel (rows x l, ld) -> row-major device panel (ld8) and back.
static void to_rowmajor(const void *src, int64_t rows, int64_t l, int64_t ld, int dt, std::vector<unsigned char> &dst,
                        int64_t ldr) {
  const size_t es = dtype_size(dt);
  dst.assign((size_t)rows * ldr * es, 0);
  const unsigned char *s = static_cast<const unsigned char *>(src);
  for (int64_t j = 0; j < l; ++j)
    for (int64_t i = 0; i < rows; ++i) std::memcpy(&dst[(i * ldr + j) * es], s + (i + j * ld) * es, es);
}

static void to_colmajor(const std::vector<unsigned char> &src, int64_t rows, int64_t l, int64_t ldr, int dt, void *dst,
                        int64_t ld) {
  const size_t es = dtype_size(dt);
  unsigned char *d = static_cast<unsigned char *>(dst);
  for (int64_t j = 0; j < l; ++j)
    for (int64_t i = 0; i < rows; ++i) std::memcpy(d + (i + j * ld) * es, &src[(i * ldr + j) * es], es);
}

// means (nullable): the centred products of the implicit A - 1 mu^T, mu the
// ncols host means of dtype dt, as exact rank-one corrections on the device:
//   dir 0: (A - 1 mu^T) X   = A X   - 1 (mu^T X)
//   dir 1: (A - 1 mu^T)^T X = A^T X - mu (1^T X)
static void host_spmm(svdh_ctx *ctx, svdh_op *op, int dir, const void *X, int64_t ldx, int64_t l, void *Y, int64_t ldy,
                      int dt, bool transpose_out, const void *means = nullptr) {
  const int64_t nin = dir ? op->nrows : op->ncols;
  const int64_t nout = dir ? op->ncols : op->nrows;
  if (dt != SVDH_F64 && dt != SVDH_F32) fail(SVDH_ERR_ARG, "panel dtype must be F64 or F32");
  if (l < 0 || ldx < std::max<int64_t>(nin, 1) || (!transpose_out && ldy < std::max<int64_t>(nout, 1)) ||
      (transpose_out && ldy < std::max<int64_t>(l, 1)))
    fail(SVDH_ERR_SHAPE, "incompatible panel leading dimension");
  if (!X || !Y) fail(SVDH_ERR_ARG, "null panel");
  const svdh_op *own = op->base ? op->base : op;
  if (own->base_dtype == SVDH_BF16) fail(SVDH_ERR_ARG, "BF16 operators are solver-only");
  // FP64 panels use the stored values as they are; FP32 panels use FP32 values.
  const int vuse = dt == SVDH_F64 ? own->base_dtype : SVDH_F32;
  const int64_t ldr = (l + 7) / 8 * 8;
  const size_t es = dtype_size(dt);
  std::vector<unsigned char> hx, hy((size_t)nout * ldr * es);
  to_rowmajor(X, nin, l, ldx, dt, hx, ldr);
  DevBuf dx(std::max<size_t>(hx.size(), 16)), dy(std::max<size_t>(hy.size(), 16));
  if (!hx.empty()) HIP_CHECK(hipMemcpyAsync(dx.p, hx.data(), hx.size(), hipMemcpyHostToDevice, ctx->stream));
  if (l > 0) spmm(op, dir, vuse, dx.p, ldr, dt, dy.p, ldr, dt, l);
  DevBuf dmu;
  if (means && l > 0) {
    std::vector<double> mu(op->ncols);
    for (int64_t j = 0; j < op->ncols; ++j)
      mu[j] = dt == SVDH_F64 ? static_cast<const double *>(means)[j] : (double)static_cast<const float *>(means)[j];
    dmu.alloc(std::max<size_t>(mu.size() * sizeof(double), 16));
    HIP_CHECK(hipMemcpyAsync(dmu.p, mu.data(), mu.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (dir == 0)
      panel_rank1_correct(ctx, nin, l, dx.p, ldr, dt, dmu.as<double>(), nout, dy.p, ldr, dt, nullptr, 1.0);
    else
      panel_rank1_correct(ctx, nin, l, dx.p, ldr, dt, nullptr, nout, dy.p, ldr, dt, dmu.as<double>(), 1.0);
  }
  if (!hy.empty()) HIP_CHECK(hipMemcpyAsync(hy.data(), dy.p, hy.size(), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  if (!transpose_out) {
    to_colmajor(hy, nout, l, ldr, dt, Y, ldy);
  } else {
    // B (l x nout, column-major ldy) = Y^T: element (c, i) at c + i*ldy.
    unsigned char *d = static_cast<unsigned char *>(Y);
    for (int64_t i = 0; i < nout; ++i)
      for (int64_t c = 0; c < l; ++c) std::memcpy(d + (c + i * ldy) * es, &hy[(i * ldr + c) * es], es);
  }
}

int32_t svdh_spmm(svdh_ctx *ctx, const svdh_op *op, int32_t transpose_self, const void *X, int64_t ldx, int64_t l,
                  void *Y, int64_t ldy, int32_t dtype) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    host_spmm(ctx, const_cast<svdh_op *>(op), transpose_self ? 1 : 0, X, ldx, l, Y, ldy, dtype, false);
  });
}

int32_t svdh_spmm_qta(svdh_ctx *ctx, const svdh_op *op, const void *Q, int64_t ldq, int64_t l, void *B, int64_t ldb,
                      int32_t dtype) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    // Q^T A = (A^T Q)^T: the transposed SpMM, written transposed.
    host_spmm(ctx, const_cast<svdh_op *>(op), 1, Q, ldq, l, B, ldb, dtype, true);
  });
}

int32_t svdh_spmm_centered(svdh_ctx *ctx, const svdh_op *op, int32_t transpose_self, const void *X, int64_t ldx,
                           int64_t l, void *Y, int64_t ldy, const void *means, int32_t dtype) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!means) fail(SVDH_ERR_ARG, "null means");
    host_spmm(ctx, const_cast<svdh_op *>(op), transpose_self ? 1 : 0, X, ldx, l, Y, ldy, dtype, false, means);
  });
}

int32_t svdh_spmm_qta_centered(svdh_ctx *ctx, const svdh_op *op, const void *Q, int64_t ldq, int64_t l, void *B,
                               int64_t ldb, const void *means, int32_t dtype) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!means) fail(SVDH_ERR_ARG, "null means");
    // Q^T (A - 1 mu^T) = ((A - 1 mu^T)^T Q)^T
    host_spmm(ctx, const_cast<svdh_op *>(op), 1, Q, ldq, l, B, ldb, dtype, true, means);
  });
}

int32_t svdh_column_means(svdh_ctx *ctx, const svdh_op *op_, void *means, int32_t dtype) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op_);
    if (!means) fail(SVDH_ERR_ARG, "null output");
    if (dtype != SVDH_F64 && dtype != SVDH_F32) fail(SVDH_ERR_ARG, "dtype must be F64 or F32");
    svdh_op *op = const_cast<svdh_op *>(op_);
    // Column means = A^T 1 / nrows (masked.rs:240-263).
    const svdh_op *own = op->base ? op->base : op;
    if (own->base_dtype == SVDH_BF16) fail(SVDH_ERR_ARG, "column means need F32 or F64 values");
    std::vector<double> ones(op->nrows, 1.0), sums(op->ncols, 0.0);
    DevBuf dx(std::max<size_t>(op->nrows * 8, 16)), dy(std::max<size_t>(op->ncols * 8, 16));
    HIP_CHECK(hipMemcpyAsync(dx.p, ones.data(), op->nrows * 8, hipMemcpyHostToDevice, ctx->stream));
    spmv(op, 1, own->base_dtype, dx.p, dy.p, SVDH_F64);
    HIP_CHECK(hipMemcpyAsync(sums.data(), dy.p, op->ncols * 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    const double r = op->nrows ? 1.0 / (double)op->nrows : 0.0;
    for (int64_t j = 0; j < op->ncols; ++j) {
      if (dtype == SVDH_F64)
        static_cast<double *>(means)[j] = sums[j] * r;
      else
        static_cast<float *>(means)[j] = (float)(sums[j] * r);
    }
  });
}

int32_t svdh_las2(svdh_ctx *ctx, const svdh_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
                  double kappa, uint32_t random_seed, const double *start, int32_t precision, svdh_result *out) {
  if (out) std::memset(out, 0, sizeof *out);
  int32_t st = guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!out) fail(SVDH_ERR_ARG, "null result");
    if (precision == SVDH_PREC_F32) ensure_values(const_cast<svdh_op *>(op), SVDH_F32);
    las2_solve(ctx, const_cast<svdh_op *>(op), dimensions, iterations, endl, endr, kappa, random_seed, start,
               precision, out, nullptr);
  });
  if (st != SVDH_OK && out) svdh_result_free(out);
  return st;
}

int32_t svdh_las2_legacy(svdh_ctx *ctx, const svdh_op *op, int64_t dimensions, int64_t iterations, double endl,
                         double endr, double kappa, uint32_t random_seed, const double *start, svdh_result *out) {
  if (out) std::memset(out, 0, sizeof *out);
  int32_t st = guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!out) fail(SVDH_ERR_ARG, "null result");
    las2_solve(ctx, const_cast<svdh_op *>(op), dimensions, iterations, endl, endr, kappa, random_seed, start,
               SVDH_PREC_F64, out, nullptr, true);
  });
  if (st != SVDH_OK && out) svdh_result_free(out);
  return st;
}

int32_t svdh_las2_device(svdh_ctx *ctx, const svdh_op *op, int64_t dimensions, int64_t iterations, double endl,
                         double endr, double kappa, uint32_t random_seed, const double *start, int32_t precision,
                         double *u, int64_t ldu, double *s, double *vt, int64_t ldvt, svdh_result *meta) {
  if (meta) std::memset(meta, 0, sizeof *meta);
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!meta || !u || !s || !vt) fail(SVDH_ERR_ARG, "null output");
    // the written rows are d <= dims; the leading dimensions must cover a row
    const int64_t dims_max = std::min(op->nrows, op->ncols);
    const int64_t dims = dimensions <= 0 || dimensions > dims_max ? dims_max : dimensions;
    (void)dims;
    if (ldu < op->nrows || ldvt < op->ncols)
      fail(SVDH_ERR_ARG, "ldu >= nrows and ldvt >= ncols are required (%lld, %lld)", (long long)ldu,
           (long long)ldvt);
    if (precision == SVDH_PREC_F32) ensure_values(const_cast<svdh_op *>(op), SVDH_F32);
    const RsvdDeviceOut dev{u, ldu, s, vt, ldvt};
    las2_solve(ctx, const_cast<svdh_op *>(op), dimensions, iterations, endl, endr, kappa, random_seed, start,
               precision, meta, &dev);
  });
}

int32_t svdh_rsvd(svdh_ctx *ctx, const svdh_op *op, int64_t target_rank, int64_t n_oversamples,
                  int64_t n_power_iters, int32_t normalizer, int32_t mean_center, uint64_t seed, int32_t has_seed,
                  const double *omega, int32_t precision, double tol, int64_t max_power_iters, svdh_result *out) {
  if (out) std::memset(out, 0, sizeof *out);
  int32_t st = guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!out) fail(SVDH_ERR_ARG, "null result");
    rsvd_solve(ctx, const_cast<svdh_op *>(op), target_rank, n_oversamples, n_power_iters, normalizer, mean_center,
               seed, has_seed, omega, precision, tol, max_power_iters, out, nullptr);
  });
  if (st != SVDH_OK && out) svdh_result_free(out);
  return st;
}

int32_t svdh_rsvd_device(svdh_ctx *ctx, const svdh_op *op, int64_t target_rank, int64_t n_oversamples,
                         int64_t n_power_iters, int32_t normalizer, int32_t mean_center, uint64_t seed,
                         int32_t has_seed, const double *omega, int32_t precision, double tol,
                         int64_t max_power_iters, double *u, int64_t ldu, double *s, double *vt, int64_t ldvt,
                         svdh_result *meta) {
  if (meta) std::memset(meta, 0, sizeof *meta);
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!meta) fail(SVDH_ERR_ARG, "null result");
    const int64_t k = std::max<int64_t>(0, std::min(target_rank, std::min(op->nrows, op->ncols)));
    if (k > 0 && (!u || !s || !vt)) fail(SVDH_ERR_ARG, "null output buffer");
    if (ldu < k || ldvt < op->ncols) fail(SVDH_ERR_ARG, "leading dimension too small (ldu %lld < %lld or ldvt %lld < %lld)",
                                          (long long)ldu, (long long)k, (long long)ldvt, (long long)op->ncols);
    RsvdDeviceOut dev{u, ldu, s, vt, ldvt};
    rsvd_solve(ctx, const_cast<svdh_op *>(op), target_rank, n_oversamples, n_power_iters, normalizer, mean_center,
               seed, has_seed, omega, precision, tol, max_power_iters, meta, &dev);
  });
}

int32_t svdh_rsvd_batched(svdh_ctx *ctx, int64_t nops, const svdh_op *const *ops, int64_t target_rank,
                          int64_t n_oversamples, int64_t n_power_iters, int32_t normalizer, uint64_t seed,
                          int32_t has_seed, const double *const *omegas, int32_t precision, svdh_result *outs) {
  if (outs && nops > 0) std::memset(outs, 0, sizeof(svdh_result) * (size_t)nops);
  int32_t st = guarded([&] {
    check_ctx(ctx);
    if (nops < 0 || (nops > 0 && (!ops || !outs))) fail(SVDH_ERR_ARG, "null argument");
    std::vector<svdh_op *> v(nops);
    for (int64_t b = 0; b < nops; ++b) {
      check_op(ctx, ops[b]);
      v[b] = const_cast<svdh_op *>(ops[b]);
    }
    rsvd_batched_solve(ctx, nops, v.data(), target_rank, n_oversamples, n_power_iters, normalizer, seed, has_seed,
                       omegas, precision, outs, nullptr);
  });
  if (st != SVDH_OK && outs)
    for (int64_t b = 0; b < nops; ++b) svdh_result_free(&outs[b]);
  return st;
}

int32_t svdh_rsvd_batched_device(svdh_ctx *ctx, int64_t nops, const svdh_op *const *ops, int64_t target_rank,
                                 int64_t n_oversamples, int64_t n_power_iters, int32_t normalizer, uint64_t seed,
                                 int32_t has_seed, const double *const *omegas, int32_t precision,
                                 double *const *u, const int64_t *ldu,
                                 double *const *vt, const int64_t *ldvt, svdh_result *metas) {
  if (metas && nops > 0) std::memset(metas, 0, sizeof(svdh_result) * (size_t)nops);
  int32_t st = guarded([&] {
    check_ctx(ctx);
    if (nops < 0 || (nops > 0 && (!ops || !metas || !u || !ldu || !vt || !ldvt))) fail(SVDH_ERR_ARG, "null argument");
    std::vector<svdh_op *> v(nops);
    for (int64_t b = 0; b < nops; ++b) {
      check_op(ctx, ops[b]);
      v[b] = const_cast<svdh_op *>(ops[b]);
      if (!u[b] || !vt[b]) fail(SVDH_ERR_ARG, "null output buffer for operator %lld", (long long)b);
    }
    BatchedDeviceOut dev{u, ldu, vt, ldvt};
    rsvd_batched_solve(ctx, nops, v.data(), target_rank, n_oversamples, n_power_iters, normalizer, seed, has_seed,
                       omegas, precision, metas, &dev);
  });
  if (st != SVDH_OK && metas)
    for (int64_t b = 0; b < nops; ++b) svdh_result_free(&metas[b]);
  return st;
}

int32_t svdh_backward(svdh_ctx *ctx, const svdh_op *op, const svdh_result *res, int32_t loss, int64_t k,
                      const double *y, int32_t precision, double tol, int64_t max_iters, void *grad,
                      int32_t grad_dtype, svdh_backward_info *info) {
  return guarded([&] {
    check_ctx(ctx);
    check_op(ctx, op);
    if (!res || !grad || !info) fail(SVDH_ERR_ARG, "null argument");
    backward_solve(ctx, const_cast<svdh_op *>(op), res, loss, k, y, precision, tol, max_iters, grad, grad_dtype,
                   info);
  });
}

void svdh_result_free(svdh_result *res) {
  if (!res) return;
  result_release(res->u);
  result_release(res->s);
  result_release(res->vt);
  result_release(res->u32);
  result_release(res->s32);
  result_release(res->vt32);
  res->u = res->s = res->vt = nullptr;
  res->u32 = res->s32 = res->vt32 = nullptr;
}

}  // extern "C"
