// Internal declarations shared by the svdh library sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/svdh.h"

namespace svdh {

// Error carried from the point of failure to the C ABI, where it becomes a
// status code plus the thread-local message.
struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int32_t code, const char *fmt, ...);

#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      ::svdh::fail(SVDH_ERR_HIP, "%s failed at %s:%d: %s", #expr, __FILE__, __LINE__, \
                   hipGetErrorString(e_));                                           \
  } while (0)

// Owning device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n);
  void ensure(size_t n) { if (n > bytes) alloc(n); }
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Kernel classes for event timing (svdh_ctx_kernel_time).
enum KernelClass {
  KC_SPMV = 0, KC_SPMM = 1, KC_DENSE = 2, KC_BLAS1 = 3, KC_GEMV = 4, KC_SPMM_T = 5, KC_SDDMM = 6, KC_INGEST = 7,
  KC_COUNT = 8
};

struct Ctx {
  int ncu = 256;  // compute units of the device
  int device = 0;
  hipStream_t stream = nullptr;
  // Device scalar block written by reduction kernels, and its pinned mirror.
  DevBuf scal;
  double *h_scal = nullptr;
  double *dh_scal = nullptr;  // device address of h_scal
  // Reduction partials.
  DevBuf red_part;
  // Workgroup partials of the fused BLAS-1 dot products, per scalar slot, and
  // how many each slot's last writer left.
  DevBuf slot_parts;
  int slot_nb[64] = {0};
  // Small host-to-device matrices (staged, then read by one kernel).
  DevBuf stage;
  // Device workspaces kept for the context's lifetime (panels of the
  // randomized solver), grown on demand; slot meanings belong to the caller.
  static constexpr int kWorkspaces = 14;  // 0-11: randomized solver panels, 12: basis GEMV partials,
                                          // 13: single-precision copies of host results
  DevBuf ws[kWorkspaces];
  void *workspace(int slot, size_t bytes) {
    if (bytes > ws[slot].bytes) {
      sync();  // queued kernels may still use the old buffer
      ws[slot].alloc(bytes);
    }
    return ws[slot].p;
  }
  // Panels of the batched randomized solver (4 per operator), grown on demand.
  // batch_pad[i]: rows [0, rows) of batch_ws[i] laid out with leading
  // dimension ld hold zeros in columns [l, ld) (no kernel writes past a
  // panel's width); l = -1 when nothing is known.
  struct PadState {
    int64_t l = -1, ld = 0, rows = 0;
  };
  std::vector<DevBuf> batch_ws;
  std::vector<PadState> batch_pad;
  // Pinned ring for small host-to-device uploads: slot i is reused only after
  // the event of its previous copy has completed.
  static constexpr int kRing = 8;
  static constexpr size_t kRingSlot = 2 * 128 * 128 * sizeof(double);
  char *h_ring = nullptr;
  DevBuf d_ring;
  hipEvent_t ring_ev[kRing] = {};
  int ring_next = 0;
  // Queue counters of the persistent batched SpMM (8 per launch).
  DevBuf pq_counters;
  // Pinned landing area for small device-to-host results (Gram, norms).
  double *h_small = nullptr;
  double *d_small = nullptr;  // device address of h_small (kernels write results there)
  char *dh_ring = nullptr;    // device address of h_ring (a kernel fetches uploads from it)
  static constexpr size_t kSmall = 4 * 128 * 128;  // doubles
  hipEvent_t small_ev = nullptr;
  hipEvent_t small_ev2 = nullptr;  // a second landing event (two Grams in flight)
  // Copy n bytes (<= kRingSlot) to the device, ordered on the stream; returns
  // the device address, valid for kernels launched after this call.
  void *upload(const void *host, size_t n);
  // Profiling.
  bool profile = false;
  struct Pending { hipEvent_t a, b; int cls; double bytes; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  double prof_ms[KC_COUNT] = {0};
  int64_t prof_n[KC_COUNT] = {0};
  double prof_bytes[KC_COUNT] = {0};
  // Device address of this GPU's copy of the allocation-free SpMV scratch.
  double *scratch = nullptr;

  void sync();
  hipEvent_t get_event();
  void drain_profile();
};

// RAII helper: records events around the kernels launched in its scope when
// profiling is on (a null context records nothing).
struct KTimer {
  Ctx *c;
  int cls;
  double bytes;
  hipEvent_t a = nullptr;
  KTimer(Ctx *ctx, int k, double algorithmic_bytes);
  ~KTimer();
};

// Row schedule of one operator direction: output rows sorted by nnz, split
// into groups-per-row bins, plus block-level chunks of long rows.
struct BinDesc {
  int32_t block_begin;  // first workgroup of the bin
  int32_t row_begin;    // offset in the sorted row list
  int32_t row_count;
  int32_t lg;           // log2 lanes per row (SpMV bins)
};

constexpr int kMaxBins = 8;
constexpr int kBlock = 256;

struct Schedule {
  bool built = false;
  int nbins = 0;
  BinDesc bins[kMaxBins];
  int32_t long_block_begin = 0;  // first workgroup of the long-row chunks
  int32_t nchunks = 0;           // number of long-row chunks (one workgroup each)
  int32_t nlong_rows = 0;
  int32_t total_blocks = 0;
  DevBuf rows;      // int32 sorted output rows
  DevBuf chunks;    // int32 x 5 per chunk: row, begin, end, first_slot, nchunks_of_row
  DevBuf counters;  // int32 x 3 per row with several chunks: row, first chunk, chunk count
  DevBuf partials;  // chunk partial results (grown on demand)
  // Streamed SpMV (views without a row map): int4 tasks (row begin, row end or
  // chunk code, first nonzero, end nonzero) over the rows in storage order.
  bool stream = false;
  int32_t ntasks = 0;
  DevBuf tasks;
  // Source-tiled SpMV (plain operators with many nonzeros): the matrix cut at
  // source blocks of width sb and stacked, block-major, as nb * nrows rows with
  // 16-bit in-block column indices; perm[p] = CSR position of stacked entry p,
  // vals[dt] = the values in that order (gathered on first use). Workgroup
  // ranges (block, first task, end task) of int4 tasks over the stacked rows;
  // part holds one FP64 partial per stacked row when nb > 1.
  struct Tiled {
    bool on = false;
    int32_t nb = 0, sb = 0, ntasks = 0, nranges = 0;
    DevBuf rowptr, col, perm, vals[3], tasks, ranges, part;
  } tiled;
  // SpMM segment schedule: per source block offsets/counts into `chunks`.
  int nblk = 1;
  int32_t blk_off[8] = {0};
  int32_t blk_cnt[8] = {0};
  int32_t nslots = 0;
  // A x with staged hot columns. hot_state: -1 not decided yet, 0 not used,
  // 1 used; hot_cover[i] = nonzeros in the i + 1 most popular columns,
  // hot_order = columns by popularity. Each encoding stages the h most
  // popular columns (h depends on the panel element size): colidx = column
  // indices with the staged ones encoded as -(1 + slot), rows = the staged
  // columns by slot.
  struct HotEnc {
    int h = 0;
    DevBuf colidx, rows;
    // each row's entries reordered staged-first: perm[k] = CSR position of
    // entry k, split[r] = staged entries of row r, order = rows by descending
    // count of unstaged entries, vals[dt] = values in the reordered layout
    DevBuf perm, split, order, vals[3];
    DevBuf dense;  // m x h FP32: the staged columns as a dense block (A^T panels)
  };
  int32_t hot_slot_base = 0;  // first partial slot of the matrix-core rows
  int32_t hot_multi = 0;      // matrix-core rows (the last entries of the multi-row table)
  int32_t chunk_multi = 0;    // rows cut into chunks (the first entries of the table)
  int hot_state = -1;
  std::vector<int64_t> hot_cover;
  std::vector<int32_t> hot_order;
  // Encodings for up to four values of h (panel widths and element sizes of
  // the callers), replaced least recently used first.
  static constexpr int kHotEnc = 4;
  HotEnc hot[kHotEnc];
  uint64_t hot_use[kHotEnc] = {0, 0, 0, 0};
  uint64_t hot_clock = 0;
  // Short rows cut at source blocks: kSegInts per segment, grouped by block.
  DevBuf segs;
  int32_t seg_off[8] = {0};
  int32_t seg_cnt[8] = {0};
};

// One direction of an operator as a CSR: rows of this matrix are output rows.
struct CsrView {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  const int32_t *rowptr = nullptr;
  const int32_t *colidx = nullptr;
  const void *vals[3] = {nullptr, nullptr, nullptr};  // by svdh_dtype
  const int32_t *rowmap = nullptr;  // output row -> source row (masked A^T)
  const int32_t *colmap = nullptr;  // source column -> operand index or -1
  std::vector<int32_t> *h_rowptr = nullptr;  // host row pointers (for schedules)
  const std::vector<int32_t> *h_rowmap = nullptr;
};

}  // namespace svdh

// The opaque ABI types.
struct svdh_ctx : svdh::Ctx {};

struct svdh_op {
  svdh_ctx *ctx = nullptr;
  const svdh_op *base = nullptr;  // masked view: the borrowed matrix
  int64_t nrows = 0, ncols = 0, nnz = 0;
  int32_t base_dtype = SVDH_F64;
  // Owned storage (plain CSR only).
  svdh::DevBuf rowptr, colidx, vals[3];
  svdh::DevBuf t_rowptr, t_colidx, t_vals[3];  // the transpose (CSC of A)
  std::vector<int32_t> h_rowptr, h_t_rowptr;
  std::vector<int32_t> h_colidx, h_t_colidx;  // host indices, for SpMM segment schedules
  // Masked view storage.
  svdh::DevBuf o2m, m2o;
  std::vector<int32_t> h_m2o;
  // Compacted forward CSR of a masked view (built when the mask drops a
  // quarter of the nonzeros or more): only the kept entries, column indices
  // already in masked numbering, perm[e] = position in the base CSR.
  bool compact = false;
  int64_t c_nnz = 0;
  svdh::DevBuf c_rowptr, c_colidx, c_perm, c_vals[3];
  std::vector<int32_t> h_c_rowptr;
  // Schedules: [direction][kind], direction 0 = A x, 1 = A^T x; kind 0 = SpMV, 1 = SpMM.
  svdh::Schedule sched[2][2];
  // A^T panels with the staged columns on the matrix cores: the schedule of
  // the other output rows (thot_state -1: not decided, 0: not used, 1: used)
  svdh::Schedule sched_thot;
  // Row-length histogram per direction (log2 buckets of the output rows'
  // stored entries), built on first use: part of the batched SpMM tuning key.
  std::vector<int64_t> row_hist[2];
  int thot_state = -1;
  int thot_g = 0;
  int thot_h = 0;  // staged columns of that product (0: not decided)
  // pattern = true: the base matrix's own pattern (with the column remap),
  // never the compacted one (the adjoint writes one value per base nonzero).
  svdh::CsrView view(int direction, bool pattern = false) const;
  const void *values(int direction, int dtype) const;
};

namespace svdh {

// ---- sparse kernels (sparse.hip) -----------------------------------------
// y = op_dir x for vectors; dtypes: values vdt, vectors xdt (F64 or F32).
// Accumulation is FP64 when the vectors are FP64, FP32 otherwise.
void spmv(svdh_op *op, int direction, int vdt, const void *x, void *y, int xdt);
// spmv with the workgroup partials in the static device scratch: allocates
// nothing; fails unless the operator's schedule and value layouts for this
// direction and value dtype exist already (an earlier spmv builds them).
void spmv_noalloc(svdh_op *op, int direction, int vdt, const void *x, void *y, int xdt);
// Looks up the device address of the scratch for c's device (the device
// must be current); once per device, at context creation.
void scratch_init(Ctx *c);
// Y = op_dir X for row-major panels of width l; X has leading dimension ldx
// (elements) and dtype xdt (F64, F32 or BF16); Y has ldy and dtype ydt (F64 or
// F32). Accumulation is FP64 for FP64 panels, FP32 otherwise.
void spmm(svdh_op *op, int direction, int vdt, const void *X, int64_t ldx, int xdt, void *Y,
          int64_t ldy, int ydt, int64_t l);
void ensure_values(svdh_op *op, int dtype);
// spmm for nops operators in one launch (same direction, dtypes and width l;
// panel b is X[b] / Y[b] with the shared leading dimensions).
void spmm_batched(Ctx *c, int nops, svdh_op *const *ops, int direction, int vdt, const void *const *X, int64_t ldx,
                  int xdt, void *const *Y, int64_t ldy, int ydt, int64_t l);
// grad[e] = sum_c s1_c L1[r_e, c] R1[col_e, c] (+ the same with L2, R2, s2 when
// L2 is not null) over the pattern of op (CSR order of the underlying matrix;
// columns a mask drops get 0). L panels have op->nrows rows, R panels
// op->ncols rows; k <= 64; s1, s2 device vectors or null (all ones).
void sddmm(svdh_op *op, int64_t k, const void *L1, const void *R1, const double *s1, const void *L2, const void *R2,
           const double *s2, int64_t ldl, int64_t ldr, int dt, void *grad, int gdt);

// Device-resident outputs of the randomized solver: u (m x k, row-major,
// leading dimension ldu) and vt (k x n, row-major, ldvt) in device memory,
// s (k entries) in host memory.
struct RsvdDeviceOut {
  double *u;
  int64_t ldu;
  double *s;
  double *vt;
  int64_t ldvt;
};

// Device-resident outputs of the batched randomized solver, per operator b:
// u[b] (m x k, ldu[b]) and vt[b] (k x n_b, ldvt[b]), FP64 device memory.
struct BatchedDeviceOut {
  double *const *u;
  const int64_t *ldu;
  double *const *vt;
  const int64_t *ldvt;
};

// ---- vector / dense kernels (dense.hip) -----------------------------------
// Scalar slots of Ctx::scal.
enum Slot { S_A = 0, S_B = 1, S_C = 2, S_D = 3, S_E = 4, S_F = 5, S_NSLOT = 64 };
// y[i] += coef * x[i] for all listed updates, then dot(u, v) -> slot out_slot.
// coef is  sign * (value of slot)  when slot >= 0, else the immediate imm.
// A slot's value stays on the device as workgroup partials; slots_finish(n)
// writes the values of slots 0 .. n-1 to Ctx::scal and to the host mirror
// Ctx::h_scal (coherent memory: valid on the host once the stream has passed).
struct Axpy { double *y; const double *x; double imm; int slot; double sign; };
void fused_axpy_dot(Ctx *c, int64_t n, const Axpy *ups, int nups, const double *u,
                    const double *v, int out_slot, double *copy_dst, const double *copy_src);
void slots_finish(Ctx *c, int n);
// Publishes slots first .. first + n - 1; with sqrt_from >= 0 (a slot of that
// range) also sqrt(value of sqrt_from) as the single partial of sqrt_to.
void slots_finish_range(Ctx *c, int first, int n, int sqrt_from, int sqrt_to);
// Slots slot_a and slot_b get sum |a_i| and sum |b_i| over i < k (device
// vectors), each summed in one fixed order.
void abs_sums_to_slots(Ctx *c, int64_t k, const double *a, const double *b, int slot_a, int slot_b);
// y = s x, zout = s zin (and copy_dst = copy_src) with s = 1 / (value of
// rnm_slot, a slot holding one published value) or, with rnm_slot < 0, s = a.
void scale_step(Ctx *c, int64_t n, int rnm_slot, double a, const double *x, double *y, const double *zin,
                double *zout, double *copy_dst = nullptr, const double *copy_src = nullptr);
// A chain of fused ops (each as fused_axpy_dot with at most one update). On
// vectors of up to 8,192 entries the whole chain is one single-workgroup
// launch that also publishes the produced slots to Ctx::scal and the host
// mirror (returns true); longer vectors take one fused_axpy_dot launch per op
// (returns false: slots_finish publishes them).
struct FusedOp {
  Axpy up;
  int nups;
  const double *u;
  const double *v;
  int out_slot;
  double *copy_dst;
  const double *copy_src;
  int sqrt_slot = -1;  // chain only: also publish sqrt(dot) to this slot
};
bool fused_chain(Ctx *c, int64_t n, const FusedOp *ops, int nops);
// y = a x, z *= a, and (when copy_dst is set) copy_dst = copy_src, in one pass.
void vec_scale2(Ctx *c, int64_t n, double a, const double *x, double *y, double *z, double *copy_dst = nullptr,
                const double *copy_src = nullptr);
void vec_copy(Ctx *c, int64_t n, const double *x, double *y);
// y = x rounded to single precision (round to nearest), on the stream.
void convert_f64_f32(Ctx *c, int64_t n, const double *x, float *y);
// Basis GEMV: cout[2k] = Q[:, c0:c0+k]^T [u v]  (k <= 4096)
void basis_gemv_t(Ctx *c, int64_t n, const double *Q, int64_t ldq, int64_t k, const double *u,
                  const double *v, double *cu, double *cv);
// y1 -= Q c1, y2 -= Q c2 (Q is n x k column-major)
void basis_gemv_n(Ctx *c, int64_t n, const double *Q, int64_t ldq, int64_t k, const double *c1,
                  double *y1, const double *c2, double *y2);
// C (n x d row-major, ldc) = Q (n x k col-major, ldq) * W (k x d row-major)
void basis_combine(Ctx *c, int64_t n, const double *Q, int64_t ldq, int64_t k, const double *W,
                   int64_t d, double *C, int64_t ldc);
// Gram G (l x l, FP64, row-major) = Y^T Y for row-major Y (m x l, ldy) of dtype ydt.
void gram(Ctx *c, int64_t m, int64_t l, const void *Y, int64_t ldy, int ydt, double *G_host);
// Copies n bytes from pinned host memory (device address src) to device
// memory with a kernel: a small copy costs a dispatch, not a DMA round trip.
void fetch_from_host(Ctx *c, const void *src, void *dst, size_t n);
// Asynchronous Gram: enqueues the kernels and the copy of G into the pinned
// area c->h_small + off (doubles) and records ev (c->small_ev when null); the
// caller waits on the event before reading.
void gram_async(Ctx *c, int64_t m, int64_t l, const void *Y, int64_t ldy, int ydt, size_t off,
                hipEvent_t ev = nullptr);
// Grams of count panels P[b] (rows[b] x l, l <= 64, shared ldy and dtype) in
// one launch: Gram b lands at c->h_small + off + b * l * l; ev (c->small_ev
// when null) is recorded after the last.
void gram_batched_async(Ctx *c, int64_t count, const int64_t *rows, int64_t l, const void *const *P, int64_t ldy,
                        int ydt, size_t off, hipEvent_t ev = nullptr);
// Z (m x lo) = Y (m x li) * M (li x lo, FP64 row-major, host), in place or to
// another panel. Element (i, c) of Z is stored at Z[i * ldz + c * zcs]
// (zcs = 1: row-major; ldz = 1, zcs = ld: the transposed, column-major store).
// Z_b = Y_b M_b (FP64 outputs, lo <= 128 columns, Z_b element (r, c) at
// r ldz[b] + c zcs[b]) for many panels in as few launches as the upload
// slots allow; each result equals panel_apply's for the same panel.
void panel_apply_batched(Ctx *c, int64_t count, const int64_t *m, const int64_t *li, int64_t lo,
                         const void *const *Y, int64_t ldy, int ydt, const std::vector<double> *M, double *const *Z,
                         const int64_t *ldz, const int64_t *zcs);
void panel_apply(Ctx *c, int64_t m, int64_t li, int64_t lo, const void *Y, int64_t ldy, int ydt,
                 const double *M_host, void *Z, int64_t ldz, int zdt, int64_t zcs = 1);
// nr2[c] = |Y1 M1[:, c] + Y2 M2[:, c]|^2 for c < k (FP64); Y1 (m x w) and
// Y2 (m x w2, w2 <= w) are row-major panels with leading dimension ld,
// M1 (w x k) and M2 (w2 x k) are host matrices.
// With nr2_host null the kernels are only enqueued; panel_residual_wait then
// waits for them and copies the k norms out.
void panel_residual(Ctx *c, int64_t m, int64_t w, int64_t w2, int64_t k, const void *Y1, const void *Y2, int64_t ld,
                    int dt, const double *M1, const double *M2, double *nr2_host);
void panel_residual_wait(Ctx *c, int64_t k, double *nr2_host);
// Column L2 normalisation, columns with norm <= 1e-10 left as they are.
void panel_normalize_columns(Ctx *c, int64_t m, int64_t l, void *Y, int64_t ldy, int ydt);
// Standard normal panel (m x l row-major) from a counter-based generator.
void panel_gaussian(Ctx *c, int64_t m, int64_t l, uint64_t seed, void *Y, int64_t ldy, int ydt);
// panel_gaussian for count panels (rows m[b], pointer Y[b]) in one launch;
// panel b holds the values panel_gaussian(c, m[b], l, seed, Y[b], ...) gives.
void panel_gaussian_batched(Ctx *c, int64_t count, const int64_t *m, int64_t l, uint64_t seed, void *const *Y,
                            int64_t ldy, int ydt);
// Convert a panel between dtypes (same shape).
void panel_convert(Ctx *c, int64_t m, int64_t l, const void *X, int64_t ldx, int xdt, void *Y,
                   int64_t ldy, int ydt);
// Transpose-copy between a host column-major panel and a device row-major panel.
void panel_fill(Ctx *c, int64_t m, int64_t l, double v, void *Y, int64_t ldy, int ydt);
// Z (m x lo) = alpha * Y (m x li) M + beta * Z with M (li x lo, FP64 row-major)
// already in device memory; Y and Z share the dtype dt (F64 or F32).
void panel_gemm_dev(Ctx *c, int64_t m, int64_t li, int64_t lo, const void *Y, int64_t ldy, const double *M_dev,
                    double alpha, double beta, void *Z, int64_t ldz, int dt);
// FP32 panels on the matrix cores: Z = beta Zin + sum_p alpha_p Y_p M_p for
// one or two terms (M_p li_p x lo FP64 row-major in device memory, lo, li_p
// <= 64); Zin is Z unless given (leading dimension ldzin). Z may be Y[0] when
// beta = 0.
// tri: every M_p is upper triangular (zero below the diagonal), so the
// products skip the factor blocks that hold only zeros (same result).
void panel_mfma(Ctx *c, int64_t rows, int64_t lo, int nterms, const float *const *Y, const int64_t *ldy,
                const int64_t *li, const double *const *M_dev, const double *alpha, float *Z, int64_t ldz,
                double beta, const float *Zin = nullptr, int64_t ldzin = 0, bool tri = false);
// M (rows x cols, row-major) is zero below the diagonal.
bool upper_triangular(const double *M, int64_t rows, int64_t cols);
// Z_b (rows[b] x lo[b]) = Z_b (rows[b] x li) M_b in place for count FP32
// panels (shared leading dimension ld), M_b li x lo[b] FP64 row-major on the
// host: the factors travel in upload slots and each group that fits one slot
// is one launch. Equal to panel_mfma on each panel, bit for bit.
void panel_mfma_batched(Ctx *c, int64_t count, const int64_t *rows, int64_t li, const int64_t *lo, float *const *Z,
                        int64_t ld, const std::vector<double> *M);
// panel_mfma_batched plus each panel's FP64 Gram Z_b^T Z_b accumulated in the
// same kernel (as panel_mfma_gram does for one panel, bit for bit); waits for
// them and returns G[b] (lo[b] x lo[b], row-major) on the host.
void panel_mfma_gram_batched(Ctx *c, int64_t count, const int64_t *rows, int64_t li, const int64_t *lo,
                             float *const *Z, int64_t ld, const std::vector<double> *M,
                             std::vector<std::vector<double>> &G);
// panel_mfma plus the FP64 Gram Z^T Z of its output, accumulated in the same
// kernel: the equivalent of panel_mfma followed by gram_async(Z, gram_off,
// gram_ev) without reading Z again.
void panel_mfma_gram(Ctx *c, int64_t rows, int64_t lo, int nterms, const float *const *Y, const int64_t *ldy,
                     const int64_t *li, const double *const *M_dev, const double *alpha, float *Z, int64_t ldz,
                     double beta, const float *Zin, int64_t ldzin, size_t gram_off, hipEvent_t gram_ev = nullptr,
                     bool tri = false);
// C (k1 x k2, FP64 row-major, device) = X1^T X2 for row-major panels (k1, k2 <= 64).
void gram_cross(Ctx *c, int64_t m, int64_t k1, const void *X1, int64_t ld1, int64_t k2, const void *X2, int64_t ld2,
                int dt, double *C_dev);
// out (cols x rows, leading dimension ldo) = (X diag(scale))^T for a
// row-major FP64 panel X (rows x cols); scale is a device vector or null.
void panel_transpose_scale(Ctx *c, int64_t rows, int64_t cols, const double *X, int64_t ldx, const double *scale_dev,
                           double *out, int64_t ldo);
// Y (cols x rows, leading dimension ldy, dtype ydt) = X^T for a row-major
// device matrix X (rows x cols, dtype xdt), converted; FP32 / FP64.
void panel_transpose_convert(Ctx *c, int64_t rows, int64_t cols, const void *X, int64_t ldx, int xdt, void *Y,
                             int64_t ldy, int ydt);
// out[c] = sum_i X[i, c] Y[i, c] for k columns (device FP64 vector).
void panel_coldots(Ctx *c, int64_t m, int64_t k, const void *X, const void *Y, int64_t ld, int dt, double *out_dev);
// W += alpha_c P, R -= alpha_c MP per column; rr[c] = |R[:, c]|^2 afterwards (device).
void panel_cg_update(Ctx *c, int64_t m, int64_t k, void *W, void *R, const void *P, const void *MP, int64_t ld,
                     int dt, const double *alpha_dev, double *rr_dev);
// Y[:, c] = a_c X[:, c] + b_c Y[:, c] (device coefficient vectors; null = 1).
void panel_colaxpby(Ctx *c, int64_t m, int64_t k, const double *a_dev, const void *X, const double *b_dev, void *Y,
                    int64_t ld, int dt);
// Rank-one correction Y (ry x w) -= scale * a s^T with s = X^T wt, X (rx x w).
// wt and a are device FP64 vectors; nullptr stands for all ones.
void panel_rank1_correct(Ctx *c, int64_t rx, int64_t w, const void *X, int64_t ldx, int xdt, const double *wt,
                         int64_t ry, void *Y, int64_t ldy, int ydt, const double *a, double scale);

// ---- host dense algebra (hostla.cpp) --------------------------------------
double pythag(double a, double b, double eps);
int imtqlb(int64_t n, double *d, double *e, double *bnd, int64_t max_iter, double eps);
int imtqlb_legacy(int64_t n, double *d, double *e, double *bnd, double eps);
int imtql2(int64_t nm, int64_t n, double *d, double *e, double *z, int64_t max_iter, double eps);
// Symmetric eigen-decomposition of an l x l row-major matrix: eigenvalues
// ascending in w, eigenvectors as columns of V (row-major l x l); V null:
// eigenvalues only (the same values).
void sym_eig(int64_t l, const double *A, double *w, double *V);
// Smallest eigenvalue of a symmetric l x l row-major matrix.
double sym_eig_min(int64_t l, const double *A);
// Cholesky A = R^T R (upper R, row-major). Returns false on a non-positive pivot.
bool cholesky_upper(int64_t l, const double *A, double *R);
// Inverse of an upper-triangular matrix.
void upper_inverse(int64_t l, const double *R, double *Rinv);

size_t dtype_size(int dt);

// Host arrays of a result (SvdRec u, s, vt): pinned buffers from a pool that
// result_release (svdh_result_free) refills; contents are not cleared.
double *result_alloc(int64_t count);
float *result_alloc_f32(int64_t count);
void *result_alloc_bytes(size_t bytes);
void result_release(void *p);

}  // namespace svdh
