// LAS2 (Lanczos with selective reorthogonalisation) on one MI355X.
//
// Control flow follows src/lanczos/mod.rs:104-1406 step for step (the
// reference's sparsity-adaptive tolerances and iteration caps included);
// numerical defects are corrected as src/legacy.rs does (pythag,
// imtqlb's p/f refresh, Ritz-vector indexing s[(js-1-i)*js + k], largest-d
// descending selection).
//
// Data placement: the working vectors w0..w4, the operator temp vector and
// the Lanczos basis (Store, mod.rs:193-225: column i+2 holds q_i, columns 0
// and 1 hold p_0, p_1) live in HBM as FP64. Every dot product of a step is a
// fused "apply the pending axpy, then reduce" kernel whose result stays in a
// device scalar slot, so the step's scalars travel to the host once per step
// (one 32-byte read) for the host-side recurrences (ortbnd, error bounds,
// tridiagonal QL), which are O(j) / O(j^2) scalar work.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "common.h"

namespace svdh {

namespace {

constexpr int64_t kMaxLL = 2;

struct Las2 {
  Ctx *c;
  svdh_op *op;
  int vdt;          // sparse value dtype
  double eps;       // epsilon of the reference's T
  bool legacy = false;  // legacy::svdLAS2 control flow (src/legacy.rs:569-2014)
  bool transposed;
  int fwd, bwd;     // operator directions for "A x" and "A^T y" in working terms
  int64_t nrows, ncols;  // working shape
  int64_t iterations;
  // device vectors: the five working vectors w[] and free buffers that a
  // step writes its outputs into (vbufs owns them all)
  std::vector<DevBuf> vbufs;
  std::vector<double *> vfree;
  DevBuf tempb, cbuf;
  double *w[5];
  double *temp;
  bool pipeline = true;  // enqueue step j + 1 before reading step j's scalars
  hipEvent_t ev = nullptr;  // recorded after a step that has one enqueued behind it
  ~Las2() {
    if (ev) (void)hipEventDestroy(ev);
  }
  double *take() {
    if (vfree.empty()) fail(SVDH_ERR_ARG, "LAS2 vector pool exhausted");
    double *p = vfree.back();
    vfree.pop_back();
    return p;
  }
  void give(double *p) { vfree.push_back(p); }
  // basis
  DevBuf basis;
  int64_t ldq = 0, cap = 0;
  // host recurrences
  std::vector<double> alf, eta, oldeta, bet, ritz, bnd, w5;

  double *col(int64_t slot) const { return basis.as<double>() + slot * ldq; }
  double *q(int64_t i) const { return col(i + kMaxLL); }
  double *p(int64_t i) const { return col(i); }

  void ensure_slots(int64_t slots) {
    if (slots <= cap) return;
    int64_t ncap = std::max<int64_t>(slots, cap * 2);
    ncap = std::min<int64_t>(std::max(ncap, slots), iterations + kMaxLL + 2);
    DevBuf nb((size_t)ncap * ldq * sizeof(double));
    HIP_CHECK(hipMemsetAsync(nb.p, 0, (size_t)ncap * ldq * sizeof(double), c->stream));
    if (cap) HIP_CHECK(hipMemcpyAsync(nb.p, basis.p, (size_t)cap * ldq * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    basis = std::move(nb);
    cap = ncap;
  }
  void storq(int64_t i, const double *v) {
    ensure_slots(i + kMaxLL + 1);
    vec_copy(c, ncols, v, q(i));
  }
  void storp(int64_t i, const double *v) {
    ensure_slots(i + 1);
    vec_copy(c, ncols, v, p(i));
  }

  bool zero(double x) const { return std::fabs(x) < eps; }  // compare(x, 0), utils.rs:110-112

  void opb(const double *x, double *y) {
    spmv(op, fwd, vdt, x, temp, SVDH_F64);
    spmv(op, bwd, vdt, temp, y, SVDH_F64);
  }

  // Read n scalar slots (synchronises the stream).
  void read_slots(int n) {
    slots_finish(c, n);  // writes the values into the host mirror too
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  double dot_now(const double *u, const double *v) {
    fused_axpy_dot(c, ncols, nullptr, 0, u, v, S_A, nullptr, nullptr);
    read_slots(1);
    return c->h_scal[S_A];
  }

  double start_vec(uint32_t seed) {
    std::vector<double> h(ncols);
    svdh_las2_start_vector(seed, ncols, h.data());
    HIP_CHECK(hipMemcpyAsync(w[0], h.data(), ncols * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0.0;
  }

  // Classical Gram-Schmidt of y against basis columns [slot0, slot0+k):
  // coefficients from u (u unchanged), y -= Q (Q^T u).  c1/c2 on the device.
  void cgs(int64_t slot0, int64_t k, const double *u, double *y) {
    if (k <= 0) return;
    cbuf.ensure((size_t)2 * k * sizeof(double));
    double *cu = cbuf.as<double>();
    double *cv = cu + k;
    basis_gemv_t(c, ncols, col(slot0), ldq, k, u, u, cu, cv);
    basis_gemv_n(c, ncols, col(slot0), ldq, k, cu, y, cv, temp_scratch());
  }
  // basis_gemv_n always updates two outputs; the second goes to a scratch
  // vector when only one is wanted.
  DevBuf scratchb;
  double *temp_scratch() {
    scratchb.ensure(ncols * sizeof(double));
    return scratchb.as<double>();
  }

  // mod.rs:587-640
  int startv(int64_t step, uint32_t seed, double *rnm_out) {
    double rnm2 = dot_now(w[0], w[0]);
    for (int id = 0; id < 3; ++id) {
      if (id > 0 || step > 0 || zero(rnm2)) start_vec(seed);
      vec_copy(c, ncols, w[0], w[3]);
      opb(w[3], w[0]);
      vec_copy(c, ncols, w[0], w[3]);
      rnm2 = dot_now(w[3], w[3]);
      if (rnm2 > 0.0) break;
    }
    if (rnm2 <= 0.0) fail(SVDH_ERR_STARTV, "rnm2 <= 0.0, rnm2 = %g", rnm2);
    if (step > 0) {
      cgs(kMaxLL, step, w[3], w[0]);
      const double t = dot_now(w[4], w[0]);
      Axpy up{w[0], w[2], -t, -1, 1.0};
      fused_axpy_dot(c, ncols, &up, 1, w[0], w[0], S_A, w[3], w[0]);
      read_slots(1);
      const double d = c->h_scal[S_A];
      rnm2 = d <= eps * rnm2 ? 0.0 : d;
    }
    *rnm_out = std::sqrt(rnm2);
    return SVDH_OK;
  }

  // mod.rs:643-670
  void stpone(uint32_t seed, double *rnm, double *tol) {
    double r;
    startv(0, seed, &r);
    if (zero(r)) fail(SVDH_ERR_STPONE, "rnm == 0.0");
    vec_scale2(c, ncols, 1.0 / r, w[0], w[1], w[3]);
    opb(w[3], w[0]);
    fused_axpy_dot(c, ncols, nullptr, 0, w[0], w[3], S_A, nullptr, nullptr);
    Axpy u1{w[0], w[1], 0.0, S_A, -1.0};
    fused_axpy_dot(c, ncols, &u1, 1, w[0], w[3], S_B, nullptr, nullptr);
    Axpy u2{w[0], w[1], 0.0, S_B, -1.0};
    fused_axpy_dot(c, ncols, &u2, 1, w[0], w[0], S_C, w[4], w[0]);
    read_slots(3);
    alf[0] = c->h_scal[S_A] + c->h_scal[S_B];
    r = std::sqrt(c->h_scal[S_C]);
    *rnm = r;
    *tol = std::sqrt(eps) * (r + std::fabs(alf[0]));
  }

  // mod.rs:815-838
  void ortbnd(int64_t step, double rnm, double eps1) {
    if (step < 1) return;
    if (!zero(rnm) && step > 1) {
      oldeta[0] = (bet[1] * eta[1] + (alf[0] - alf[step]) * eta[0] - bet[step] * oldeta[0]) / rnm + eps1;
      for (int64_t i = 1; i + 2 <= step; ++i)
        oldeta[i] = (bet[i + 1] * eta[i + 1] + (alf[i] - alf[step]) * eta[i] + bet[i] * eta[i - 1] -
                     bet[step] * oldeta[i]) /
                        rnm +
                    eps1;
    }
    oldeta[step - 1] = eps1;
    std::swap(oldeta, eta);
    eta[step] = eps1;
  }

  // mod.rs:760-813: up to two classical Gram-Schmidt passes of q (w1) and
  // r (w0) against q_ll .. q_{step-1}, done as two block GEMVs per pass.
  // Returns whether the reorthogonalisation ran (it then changed w1, w0 and
  // rnm, or left rnm <= tol).
  bool purge(int64_t ll, int64_t step, double *rnm, double tol) {
    if (step < ll + 2) return false;
    const double reps = std::sqrt(eps);
    const double eps1 = eps * std::sqrt((double)ncols);
    int64_t k = 0;
    for (int64_t i = 1; i < step - (ll + 1); ++i)
      if (std::fabs(eta[i]) > std::fabs(eta[k])) k = i;
    k += ll;
    if (!(std::fabs(eta[k]) > reps)) return false;
    const double reps1 = eps1 / reps;
    const int64_t cnt = step - ll;
    bool flag = true;
    for (int it = 0; it < 2 && flag; ++it) {
      if (!(*rnm > tol)) continue;
      cbuf.ensure((size_t)2 * cnt * sizeof(double));
      double *cq = cbuf.as<double>();
      double *cr = cq + cnt;
      basis_gemv_t(c, ncols, q(ll), ldq, cnt, w[3], w[4], cq, cr);
      basis_gemv_n(c, ncols, q(ll), ldq, cnt, cq, w[1], cr, w[0]);
      fused_axpy_dot(c, ncols, nullptr, 0, w[0], w[1], S_A, w[3], w[1]);
      Axpy up{w[0], w[1], 0.0, S_A, -1.0};
      fused_axpy_dot(c, ncols, &up, 1, w[0], w[0], S_B, w[4], w[0]);
      // sum |coefficients| on the device
      abs_sums_to_slots(c, cnt, cq, cr, S_C, S_D);
      read_slots(4);
      const double tq = c->h_scal[S_C];
      const double tr = c->h_scal[S_D] + std::fabs(c->h_scal[S_A]);
      *rnm = std::sqrt(c->h_scal[S_B]);
      if (tq <= reps1 && tr <= *rnm * reps1) flag = false;
    }
    for (int64_t i = ll; i <= step; ++i) {
      eta[i] = eps1;
      oldeta[i] = eps1;
    }
    return true;
  }

  // One Lanczos step's device work (mod.rs:690-745) after the buffer
  // rotation W = (r, q_{j-1}, q_{j-2}, old r copy, p_{j-1}) and any restart:
  // q_j = r / rnm, p = w3 / rnm, the store of q_{j-1} (unless done), A^T A p,
  // the fused updates and dots, with the step's scalars in the slots at base
  // (alternating between steps, so that a step enqueued ahead leaves the
  // previous step's values alone) and rnm = sqrt(r.r) published beside them.
  // Every output goes to a buffer taken from the pool, never to a vector the
  // step reads: a step enqueued ahead is dropped by returning its outputs.
  // rnm comes from the host (rnm_slot < 0) or, for a step enqueued ahead,
  // from the previous step's published slot. Steps j <= kMaxLL wait for alpha
  // and may move ll (mod.rs:719-731).
  struct Step {
    int64_t j = -1;
    int base = 0;
    double *w[5];
    double *outs[4];   // buffers taken by the step
    double *freed[4];  // buffers of W that the new state no longer holds
  };
  static int slot_base(int64_t j) { return (j & 1) ? 32 : 0; }
  static constexpr int S_R = 6;  // rnm of the step, next to its dot slots

  void enqueue_body(int64_t j, double *const W[5], double rnm, int rnm_slot, bool stored, int64_t *ll,
                    Step *st) {
    const int base = slot_base(j);
    double *F1 = take(), *F2 = take(), *F3 = take(), *F4 = take();
    scale_step(c, ncols, rnm_slot, 1.0 / rnm, W[0], F1, W[3], F2, stored ? nullptr : q(j - 1), W[2]);
    opb(F2, F3);
    std::vector<FusedOp> ch;
    const Axpy first = rnm_slot >= 0 ? Axpy{F3, W[2], 0.0, rnm_slot, -1.0} : Axpy{F3, W[2], -rnm, -1, 1.0};
    ch.push_back(FusedOp{first, 1, F3, F2, base + S_A, nullptr, nullptr});
    if (j <= kMaxLL) {
      if (!fused_chain(c, ncols, ch.data(), (int)ch.size())) slots_finish_range(c, base + S_A, 1, -1, -1);
      ch.clear();
      HIP_CHECK(hipStreamSynchronize(c->stream));
      if (std::fabs(alf[j - 1]) > 4.0 * std::fabs(c->h_scal[base + S_A])) *ll = j;
    }
    // Pending update: r -= alpha_j q_j, applied by the next fused op.
    Axpy pend{F3, F1, 0.0, base + S_A, -1.0};
    const int64_t lim = std::min(j - 1, *ll);
    for (int64_t i = 0; i < lim; ++i) {
      const int slot = base + 8 + (int)i;
      ch.push_back(FusedOp{pend, 1, p(i), F3, slot, nullptr, nullptr});
      pend = Axpy{F3, q(i), 0.0, slot, -1.0};
    }
    ch.push_back(FusedOp{pend, 1, F3, W[4], base + S_B, nullptr, nullptr});
    ch.push_back(FusedOp{Axpy{F3, W[2], 0.0, base + S_B, -1.0}, 1, F3, F2, base + S_C, nullptr, nullptr});
    FusedOp last{Axpy{F3, F1, 0.0, base + S_C, -1.0}, 1, F3, F3, base + S_E, F4, F3};
    last.sqrt_slot = base + S_R;
    ch.push_back(last);
    if (!fused_chain(c, ncols, ch.data(), (int)ch.size()))
      slots_finish_range(c, base, S_E + 1, base + S_E, base + S_R);
    st->j = j;
    st->base = base;
    double *nw[5] = {F3, F1, W[2], F2, F4};
    double *fr[4] = {W[0], W[1], W[3], W[4]};
    double *ou[4] = {F1, F2, F3, F4};
    std::copy(nw, nw + 5, st->w);
    std::copy(fr, fr + 4, st->freed);
    std::copy(ou, ou + 4, st->outs);
  }
  void commit(const Step &st) {
    std::copy(st.w, st.w + 5, w);
    for (double *f : st.freed) give(f);
  }
  void drop(const Step &st) {
    for (double *o : st.outs) give(o);
  }

  // Whether step j's reorthogonalisation (mod.rs:777-779) is likely, judged
  // before step j's scalars arrive: ortbnd's recurrence (mod.rs:815-838) with
  // alpha_j and rnm_j taken as alpha_{j-1} and beta_j = rnm_{j-1}, against half
  // the threshold. Only the choice of enqueuing step j + 1 ahead depends on it.
  std::vector<double> ne;  // purge_likely's predicted eta
  bool purge_likely(int64_t j, int64_t ll, double eps1) {
    if (j < ll + 2) return false;
    const double rg = bet[j];
    if (zero(rg)) return true;
    const double ag = alf[j - 1];
    const int64_t cnt = j - (ll + 1);
    ne.assign(j + 1, eps1);
    if (j > 1) {
      ne[0] = (bet[1] * eta[1] + (alf[0] - ag) * eta[0] - rg * oldeta[0]) / rg + eps1;
      for (int64_t i = 1; i + 2 <= j; ++i)
        ne[i] = (bet[i + 1] * eta[i + 1] + (alf[i] - ag) * eta[i] + bet[i] * eta[i - 1] - rg * oldeta[i]) / rg +
                eps1;
    }
    int64_t k = 0;
    for (int64_t i = 1; i < cnt; ++i)
      if (std::fabs(ne[i]) > std::fabs(ne[k])) k = i;
    return std::fabs(ne[k + ll]) > 0.5 * std::sqrt(eps);
  }

  // mod.rs:674-758. With the pipeline on, step j + 1 is enqueued on the
  // device (its rnm taken from step j's published slot) before the host
  // waits for step j's scalars; it is kept when step j's host side leaves
  // the device state and rnm as they were (no reorthogonalisation, rnm above
  // tol), else dropped and enqueued again from the host's values. Either way
  // every step runs the same kernels on the same values as without the
  // pipeline.
  int64_t lanczos_step(int64_t first, int64_t last, int64_t *ll, bool *enough, double *rnm, double *tol) {
    const double eps1 = eps * std::sqrt((double)ncols);
    int64_t j = first;
    Step ahead;
    bool have_ahead = false;
    for (; j < last; ++j) {
      Step cur;
      bet[j] = *rnm;
      if (have_ahead) {
        cur = ahead;
        have_ahead = false;
        commit(cur);
      } else {
        double *W[5] = {w[0], w[2], w[1], w[4], w[3]};
        ensure_slots(j - 1 + kMaxLL + 1);
        bool stored = false;
        if (j - 1 < kMaxLL) storp(j - 1, W[4]);
        if (zero(*rnm)) {
          storq(j - 1, W[2]);
          stored = true;
          std::copy(W, W + 5, w);
          startv(j, 0u, rnm);
          if (zero(*rnm)) *enough = true;
          std::copy(w, w + 5, W);
        }
        if (*enough) {
          if (!stored) storq(j - 1, W[2]);
          w[0] = W[0];
          w[1] = W[2];
          w[2] = W[1];
          w[3] = W[3];
          w[4] = W[4];
          break;
        }
        enqueue_body(j, W, *rnm, -1, stored, ll, &cur);
        commit(cur);
      }
      const int64_t lim = std::min(j - 1, *ll);
      for (int64_t i = 0; i < lim; ++i) {
        eta[i] = eps1;
        oldeta[i] = eps1;
      }
      if (pipeline && j + 1 < last && j + 1 > kMaxLL && !purge_likely(j, *ll, eps1)) {
        ensure_slots(j + kMaxLL + 1);
        HIP_CHECK(hipEventRecord(ev, c->stream));
        double *W[5] = {w[0], w[2], w[1], w[4], w[3]};
        enqueue_body(j + 1, W, 0.0, cur.base + S_R, false, ll, &ahead);
        have_ahead = true;
      }
      // with a step behind it, wait for this one's event (a stream
      // synchronisation would wait for both); else for the stream, which
      // measured faster than an event wait
      if (have_ahead)
        HIP_CHECK(hipEventSynchronize(ev));
      else
        HIP_CHECK(hipStreamSynchronize(c->stream));
      const double *h = c->h_scal + cur.base;
      const double a = h[S_A], t1 = h[S_B], t2 = h[S_C];
      alf[j] = a;
      if (bet[j] > 0.0) bet[j] += t1;
      alf[j] += t2;
      *rnm = h[S_R];
      const double anorm = bet[j] + std::fabs(alf[j]) + *rnm;
      *tol = std::sqrt(eps) * anorm;
      ortbnd(j, *rnm, eps1);
      const bool purged = purge(*ll, j, rnm, *tol);
      if (*rnm <= *tol) *rnm = 0.0;
      if (have_ahead && (purged || *rnm == 0.0)) {
        drop(ahead);
        have_ahead = false;
      }
    }
    if (have_ahead) drop(ahead);
    return j;
  }

  double sparsity() const {
    return 1.0 - (double)op->nnz / ((double)op->nrows * (double)op->ncols);
  }

  // mod.rs:840-899
  int64_t error_bound(bool *enough, double endl, double endr, int64_t step, double tol) {
    const double eps34 = std::pow(eps, 0.75);
    int64_t mid = 0;
    for (int64_t i = 1; i <= step; ++i)
      if (std::fabs(bnd[i]) > std::fabs(bnd[mid])) mid = i;
    for (int64_t i = ((step + 1) + (step - 1)) / 2; i > mid + 1; --i)
      if (std::fabs(ritz[i - 1] - ritz[i]) < eps34 * std::fabs(ritz[i]) && bnd[i] > tol && bnd[i - 1] > tol) {
        bnd[i - 1] = std::sqrt(bnd[i] * bnd[i] + bnd[i - 1] * bnd[i - 1]);
        bnd[i] = 0.0;
      }
    for (int64_t i = ((step + 1) - (step - 1)) / 2; i + 1 < mid; ++i)
      if (std::fabs(ritz[i + 1] - ritz[i]) < eps34 * std::fabs(ritz[i]) && bnd[i] > tol && bnd[i + 1] > tol) {
        bnd[i + 1] = std::sqrt(bnd[i] * bnd[i] + bnd[i + 1] * bnd[i + 1]);
        bnd[i] = 0.0;
      }
    int64_t neig = 0;
    double gapl = ritz[step] - ritz[0];
    for (int64_t i = 0; i <= step; ++i) {
      double gap = gapl;
      if (i < step) gapl = ritz[i + 1] - ritz[i];
      gap = std::min(gap, gapl);
      if (gap > bnd[i]) bnd[i] *= bnd[i] / gap;
      if (bnd[i] <= 16.0 * eps * std::fabs(ritz[i])) {
        ++neig;
        if (!*enough) *enough = endl < ritz[i] && ritz[i] < endr;
      }
    }
    return neig;
  }

  static void copy_reversed(int64_t n, int64_t off, const double *x, double *y) {
    for (int64_t i = 0; i < n; ++i) y[off + n - 1 - i] = x[off + i];
  }

  // mod.rs:1263-1406; legacy.rs:1924-2014 in legacy mode (plain eps, a split
  // where compare(bet, 0), legacy imtqlb, no extra steps)
  int64_t lanso(int64_t dim, double endl, double endr, uint32_t seed, int64_t *neig) {
    const double sp = sparsity();
    const int64_t max_imtqlb = sp > 0.999 ? 500 : sp > 0.99 ? 300 : sp > 0.9 ? 100 : 50;
    const double aeps = legacy ? eps : sp > 0.99 ? eps * 100.0 : sp > 0.9 ? eps * 10.0 : eps;
    double rnm, tol;
    stpone(seed, &rnm, &tol);
    const double eps1 = aeps * std::sqrt((double)ncols);
    eta[0] = eps1;
    oldeta[0] = eps1;
    int64_t ll = 0, first = 1, j = 0, intro = 0;
    int64_t last = std::min(iterations, std::max<int64_t>(dim, 8) + dim);
    bool enough = false;
    while (!enough) {
      if (rnm <= tol) rnm = 0.0;
      const int64_t steps = lanczos_step(first, last, &ll, &enough, &rnm, &tol);
      j = enough ? steps - 1 : last - 1;
      first = j + 1;
      bet[first] = rnm;
      int64_t l = 0;
      for (int64_t it = 0; it < j && l <= j; ++it) {
        int64_t i = l;
        while (i <= j && !(legacy ? zero(bet[i + 1]) : std::fabs(bet[i + 1]) <= aeps)) ++i;
        i = std::min(i, j);
        const int64_t sz = i - l;
        copy_reversed(sz + 1, l, alf.data(), ritz.data());
        copy_reversed(sz, l + 1, bet.data(), w5.data());
        if (legacy)
          imtqlb_legacy(sz + 1, ritz.data() + l, w5.data() + l, bnd.data() + l, eps);
        else
          imtqlb(sz + 1, ritz.data() + l, w5.data() + l, bnd.data() + l, max_imtqlb, eps);
        for (int64_t m = l; m <= i; ++m) bnd[m] = rnm * std::fabs(bnd[m]);
        l = i + 1;
      }
      // insertion sort of ritz (carrying bnd), mod.rs:297-307
      for (int64_t i = 1; i <= j; ++i)
        for (int64_t k = i; k >= 1 && ritz[k - 1] > ritz[k]; --k) {
          std::swap(ritz[k - 1], ritz[k]);
          std::swap(bnd[k - 1], bnd[k]);
        }
      if (j < 1) fail(SVDH_ERR_LAS2, "error_bound: expected 'step' to be non-zero");
      *neig = error_bound(&enough, endl, endr, j, tol);
      if (*neig < dim) {
        if (*neig == 0) {
          last = first + 9;
          intro = first;
        } else {
          const int64_t extra = !legacy && sp > 0.99 ? 5 : 0;
          last = first + std::max<int64_t>(3, 1 + ((j - intro) * (dim - *neig)) / *neig) + extra;
        }
        last = std::min(last, iterations);
      } else {
        enough = true;
      }
      enough = enough || first >= iterations;
    }
    storq(j, w[1]);
    return j;
  }

  // mod.rs:1056-1259 with the Ritz indexing and ordering of legacy.rs:1796-1835;
  // in legacy mode legacy.rs:1765-1870 (imtql2 capped at 30 iterations, the
  // significance test bnd <= kappa |ritz|, every non-zero coefficient, u = A v
  // / sigma).
  // dev (nullable): U (d x m) and V^T (d x n) go to these device buffers
  // instead of host memory, s to dev->s.
  void ritvec(int64_t dims, double kappa, int64_t steps, int64_t neig, svdh_result *out, int64_t *nsig_out,
              const RsvdDeviceOut *dev) {
    const int64_t js = steps + 1;
    const double sp = sparsity();
    const double aeps = legacy ? eps : sp > 0.99 ? eps * 100.0 : sp > 0.9 ? eps * 10.0 : eps;
    const int64_t max_imtql2 = legacy ? 30 : sp > 0.999 ? 500 : sp > 0.99 ? 300 : sp > 0.9 ? 200 : 50;
    std::vector<double> s(js * js, 0.0), ev(js), e(std::max<int64_t>(js, 1) + 1, 0.0);
    for (int64_t i = 0; i < js; ++i) s[i * js + i] = 1.0;
    copy_reversed(js, 0, alf.data(), ev.data());
    copy_reversed(steps, 1, bet.data(), e.data());
    imtql2(js, js, ev.data(), e.data(), s.data(), max_imtql2, eps);
    double max_ev = 0.0;
    for (double v : ev) max_ev = std::max(max_ev, std::fabs(v));
    const double akappa = !legacy && sp > 0.99 ? kappa * 10.0 : kappa;
    std::vector<int64_t> sig;
    for (int64_t k = 0; k < js; ++k) {
      const double rel = legacy ? kappa * std::fabs(ritz[k]) : akappa * std::max(std::fabs(ritz[k]), max_ev * aeps);
      if (bnd[k] <= rel && k + 1 > js - neig) sig.push_back(k);
    }
    const int64_t nsig = (int64_t)sig.size();
    const int64_t d = std::min(dims, nsig);
    *nsig_out = nsig;
    out->d = d;
    out->u_rows = d;
    out->u_cols = transposed ? ncols : nrows;
    out->vt_rows = d;
    out->vt_cols = transposed ? nrows : ncols;
    // FP32-value solves hand over single-precision host arrays (SvdRec<f32>)
    const bool single = vdt != SVDH_F64;
    if (!dev && !single) {
      out->s = result_alloc(d);
      out->u = result_alloc(d * out->u_cols);
      out->vt = result_alloc(d * out->vt_cols);
    } else if (!dev) {
      out->s32 = result_alloc_f32(d);
      out->u32 = result_alloc_f32(d * out->u_cols);
      out->vt32 = result_alloc_f32(d * out->vt_cols);
    }
    if (d == 0) return;
    // Coefficients W (js x d): column r is the eigenvector of T for the r-th
    // largest significant Ritz value, in the basis q_0..q_{js-1}.
    std::vector<double> W(js * d, 0.0);
    for (int64_t r = 0; r < d; ++r) {
      const int64_t k = sig[nsig - 1 - r];
      for (int64_t i = 0; i < js; ++i) {
        const double coef = s[(js - 1 - i) * js + k];
        if (legacy ? coef != 0.0 : std::fabs(coef) > aeps) W[i * d + r] = coef;
      }
    }
    const int64_t ldv = (d + 1) / 2 * 2;
    DevBuf Wd(W.size() * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(Wd.p, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    DevBuf V((size_t)ncols * ldv * sizeof(double)), AV((size_t)nrows * ldv * sizeof(double));
    HIP_CHECK(hipMemsetAsync(V.p, 0, V.bytes, c->stream));
    basis_combine(c, ncols, q(0), ldq, js, Wd.as<double>(), d, V.as<double>(), ldv);
    spmm(op, fwd, vdt, V.p, ldv, SVDH_F64, AV.p, ldv, SVDH_F64, d);
    // sigma_r = sqrt(max(v_r^T B v_r, 0)) with B = A^T A (mod.rs:1212-1225),
    // taken as ||A v_r||^2 = v_r^T A^T A v_r from the A V needed for U anyway
    // (one sparse product instead of two); u_r = A v_r / max(sigma_r, eps')
    // (mod.rs:1228-1243): dots, scaling and the transposes to the d x n row
    // layout of SvdRec run on the device; only U^T, V^T and sigma come back.
    DevBuf sc(2 * (size_t)std::max<int64_t>(d, 1) * sizeof(double));
    double *dots = sc.as<double>(), *scale = dots + d;
    panel_coldots(c, nrows, d, AV.p, AV.p, ldv, SVDH_F64, dots);
    std::vector<double> hd;
    hd.resize(d);
    HIP_CHECK(hipMemcpyAsync(hd.data(), dots, d * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    std::vector<double> hs(d);
    for (int64_t r = 0; r < d; ++r) {
      const double sval = std::sqrt(std::max(hd[r], 0.0));
      if (dev)
        dev->s[r] = sval;
      else if (single)
        out->s32[r] = (float)sval;
      else
        out->s[r] = sval;
      hs[r] = legacy || sval > aeps ? 1.0 / sval : 1.0 / std::max(sval, aeps);
    }
    HIP_CHECK(hipMemcpyAsync(scale, hs.data(), d * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (dev) {
      // the working (possibly transposed) layout written straight into the
      // caller's device buffers: Ut (d x nrows) and Vt (d x ncols)
      double *Ut = transposed ? dev->vt : dev->u;
      double *Vt = transposed ? dev->u : dev->vt;
      const int64_t ldut = transposed ? dev->ldvt : dev->ldu, ldvt = transposed ? dev->ldu : dev->ldvt;
      panel_transpose_scale(c, nrows, d, AV.as<double>(), ldv, scale, Ut, ldut);
      panel_transpose_scale(c, ncols, d, V.as<double>(), ldv, nullptr, Vt, ldvt);
      HIP_CHECK(hipStreamSynchronize(c->stream));
      return;
    }
    // The working (possibly transposed) layout: Ut is d x nrows, Vt is d x ncols.
    void *Ut = transposed ? (single ? (void *)out->vt32 : (void *)out->vt) : (single ? (void *)out->u32 : (void *)out->u);
    void *Vt = transposed ? (single ? (void *)out->u32 : (void *)out->u) : (single ? (void *)out->vt32 : (void *)out->vt);
    const size_t esz = single ? sizeof(float) : sizeof(double);
    const int64_t big = std::max(nrows, ncols);
    DevBuf T((size_t)d * big * sizeof(double) + (single ? (size_t)d * big * sizeof(float) : 0));
    float *Tf = single ? reinterpret_cast<float *>(T.as<double>() + d * big) : nullptr;
    auto hand_over = [&](int64_t len, void *dst) {
      if (single) {
        convert_f64_f32(c, d * len, T.as<double>(), Tf);
        HIP_CHECK(hipMemcpyAsync(dst, Tf, (size_t)d * len * esz, hipMemcpyDeviceToHost, c->stream));
      } else {
        HIP_CHECK(hipMemcpyAsync(dst, T.p, (size_t)d * len * esz, hipMemcpyDeviceToHost, c->stream));
      }
      HIP_CHECK(hipStreamSynchronize(c->stream));
    };
    panel_transpose_scale(c, nrows, d, AV.as<double>(), ldv, scale, T.as<double>(), nrows);
    hand_over(nrows, Ut);
    panel_transpose_scale(c, ncols, d, V.as<double>(), ldv, nullptr, T.as<double>(), ncols);
    hand_over(ncols, Vt);
  }
};

}  // namespace

int32_t las2_solve(svdh_ctx *ctx, svdh_op *op, int64_t dimensions, int64_t iterations, double endl, double endr,
                   double kappa, uint32_t random_seed, const double *start, int32_t precision, svdh_result *out,
                   const RsvdDeviceOut *dev, bool legacy) {
  if (precision != SVDH_PREC_F64 && precision != SVDH_PREC_F32)
    fail(SVDH_ERR_ARG, "LAS2 precision must be F64 or F32");
  if (legacy && precision != SVDH_PREC_F64) fail(SVDH_ERR_ARG, "svdLAS2 is FP64 only");
  const char *name = legacy ? "svdLAS2" : "svd_las2";
  if (op->nrows <= 0 || op->ncols <= 0) fail(SVDH_ERR_LAS2, "%s: empty matrix", name);
  if (random_seed == 0) random_seed = std::random_device{}();
  const int64_t mn = std::min(op->nrows, op->ncols);
  if (dimensions == 0 || dimensions > mn) dimensions = mn;
  if (iterations == 0 || iterations > mn)
    iterations = mn;
  else if (iterations < dimensions)
    iterations = dimensions;
  if (dimensions < 2) fail(SVDH_ERR_LAS2, "%s: insufficient dimensions: %lld", name, (long long)dimensions);

  Las2 L;
  L.c = ctx;
  L.op = op;
  L.vdt = precision == SVDH_PREC_F64 ? SVDH_F64 : SVDH_F32;
  L.eps = precision == SVDH_PREC_F64 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
  L.legacy = legacy;
  L.transposed = (double)op->ncols >= (double)op->nrows * 1.2;
  L.fwd = L.transposed ? 1 : 0;
  L.bwd = 1 - L.fwd;
  L.nrows = L.transposed ? op->ncols : op->nrows;
  L.ncols = L.transposed ? op->nrows : op->ncols;
  L.iterations = iterations;
  const int64_t n = L.ncols;
  // five working vectors and the outputs of a step and of one enqueued ahead
  L.vbufs.resize(13);
  for (int i = 0; i < 13; ++i) {
    L.vbufs[i].alloc(n * sizeof(double));
    if (i < 5) {
      L.w[i] = L.vbufs[i].as<double>();
      HIP_CHECK(hipMemsetAsync(L.w[i], 0, n * sizeof(double), ctx->stream));
    } else {
      L.vfree.push_back(L.vbufs[i].as<double>());
    }
  }
  if (const char *pe = std::getenv("SVDH_LAS2_PIPELINE")) L.pipeline = std::atoi(pe) != 0;
  HIP_CHECK(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming));
  L.tempb.alloc(L.nrows * sizeof(double));
  L.temp = L.tempb.as<double>();
  L.ldq = (n + 31) / 32 * 32;
  L.ensure_slots(std::min<int64_t>(iterations + kMaxLL + 2, 2 * dimensions + 32));
  L.alf.assign(iterations + 1, 0.0);
  L.eta.assign(iterations + 1, 0.0);
  L.oldeta.assign(iterations + 1, 0.0);
  L.bet.assign(iterations + 2, 0.0);
  L.ritz.assign(iterations + 2, 0.0);
  L.bnd.assign(iterations + 2, 1.7976931348623157e308);
  L.w5.assign(iterations + 2, 0.0);
  if (start) HIP_CHECK(hipMemcpyAsync(L.w[0], start, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));

  int64_t neig = 0, nsig = 0;
  const int64_t steps = L.lanso(dimensions, endl, endr, random_seed, &neig);
  const double kap = std::max(std::fabs(kappa), std::pow(L.eps, 0.75));
  L.ritvec(dimensions, kap, steps, neig, out, &nsig, dev);
  svdh_diagnostics &dg = out->diagnostics;
  dg.non_zero = op->nnz;
  dg.dimensions = dimensions;
  dg.iterations = iterations;
  dg.transposed = L.transposed ? 1 : 0;
  dg.lanczos_steps = steps + 1;
  dg.ritz_values_stabilized = neig;
  dg.significant_values = out->d;
  dg.singular_values = nsig;
  dg.end_interval[0] = endl;
  dg.end_interval[1] = endr;
  dg.kappa = kap;
  dg.random_seed = random_seed;
  return SVDH_OK;
}

}  // namespace svdh

extern "C" int32_t svdh_las2_start_vector(uint32_t seed, int64_t n, double *out) {
  if (!out || n < 0) return SVDH_ERR_ARG;
  const uint64_t phi = 0x9E3779B97F4A7C15ULL;
  const uint64_t key = ((uint64_t)seed * phi) ^ 0x5356444C49423031ULL;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t z = key + (uint64_t)(i + 1) * phi;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    out[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
  }
  return SVDH_OK;
}
