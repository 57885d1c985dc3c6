// Device side of the engine: upload of the host hierarchy to HBM (once), the
// multilevel cycle of mld_dmlprec_aply (mlprec/impl/mld_dmlprec_aply.f90
// :203-809) as a sequence of fused HIP kernels on one stream, and
// device-resident CG / FCG / BiCGStab.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstdio>
#include <thread>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mld_device.hpp"
#include "mld_dense.hpp"
#include "mld_devcommon.hpp"
#include "mld_kernels.hpp"

namespace mld {

using namespace mld::k;

// ---------------------------------------------------------------------------
// The device preconditioner for one value type.
// ---------------------------------------------------------------------------
template <typename V>
class DevPrecT : public DevPrec {
 public:
  DevPrecT(int device, hipStream_t stream) : device_(device), s_(stream) {}

  int upload(const Prec& p) override {
    HIPCHECK(hipSetDevice(device_));
    const int nlev = (int)p.precv.size();
    L_.resize(nlev);
    outer_sweeps_ = p.outer_sweeps;
    for (int l = 0; l < nlev; ++l) {
      const Level& H = p.precv[l];
      DLevel<V>& D = L_[l];
      D.n = H.a.nr;
      double t0 = wall_seconds();
      D.a = upload_tuned<V>(H.a, ar_, s_, p.dev_format);
      phase_log("upload A", l + 1, wall_seconds() - t0);
      if (l > 0) {
        t0 = wall_seconds();
        D.p = upload_tuned<V>(H.p, ar_, s_, p.dev_format);
        D.r = upload_tuned<V>(H.r, ar_, s_, p.dev_format);
        phase_log("upload P, R", l + 1, wall_seconds() - t0);
      }
      t0 = wall_seconds();
      D.sweeps_pre = H.parms.sweeps_pre;
      D.sweeps_post = H.parms.sweeps_post;
      D.ml_cycle = H.parms.ml_cycle;
      upload_smoother(H.smd, D.sm, p.dev_format);
      D.has_sm2 = H.sm2a.allocated;
      if (D.has_sm2) upload_smoother(H.smd2, D.sm2, p.dev_format);
      const size_t n = (size_t)D.n;
      D.x2l = ar_.alloc<V>(n); D.y2l = ar_.alloc<V>(n);
      D.tx = ar_.alloc<V>(n); D.ty = ar_.alloc<V>(n);
      D.t1 = ar_.alloc<V>(n); D.t2 = ar_.alloc<V>(n);
      if (H.smd.solver == DENSE_LU || (D.has_sm2 && H.smd2.solver == DENSE_LU)) D.wlu = ar_.alloc<double>(n);
      phase_log("upload smoothers", l + 1, wall_seconds() - t0);
    }
    // K-cycle: the inner iteration runs on levels 1 .. nlev-2 (0-based)
    for (int l = 1; l + 1 < nlev; ++l) {
      const int c = p.precv[l - 1].parms.ml_cycle;
      if (c != KCYCLE && c != KCYCLESYM) continue;
      DLevel<V>& D = L_[l];
      const size_t n = (size_t)D.n;
      D.kw = ar_.alloc<V>(n); D.kv = ar_.alloc<V>(n); D.kv1 = ar_.alloc<V>(n);
      D.kd0 = ar_.alloc<V>(n); D.kd1 = ar_.alloc<V>(n);
    }
    // Krylov matrix, summed in FP64: level 0's own copy in FP64; in the FP32
    // form, level 0's FP32 copy when every value of A is exact in FP32 (the
    // products are then bit-identical to FP64 storage), else an FP64 copy.
    akry_is32_ = false;
    if constexpr (std::is_same<V, double>::value) {
      akry_ = L_[0].a;
    } else {
      if (fp32_exact(p.precv[0].a)) {
        akry32_ = L_[0].a;
        akry_is32_ = true;
      } else {
        akry_ = upload_mat<double>(p.precv[0].a, ar_, s_, p.dev_format);
      }
    }
    n0_ = L_[0].n;
    // the residual copy of the outer sweeps, allocated here so that no
    // application allocates
    if (outer_sweeps_ > 1 && nlev > 1) xres_ = ar_.alloc<V>((size_t)n0_);
    {
      // partial sums of every reducing launch: Krylov SpMV, vector reductions,
      // fused output of the fine level's post smoother
      int nbmax = std::max(std::max(spmv_blocks(akry_), spmv_blocks(akry32_)), red_blocks(n0_));
      nbmax = std::max(nbmax, spmv_blocks(L_[0].sm.nd));
      if (L_[0].has_sm2) nbmax = std::max(nbmax, spmv_blocks(L_[0].sm2.nd));
      nbmax = std::max(nbmax, nblk(n0_));
      partial_ = ar_.alloc<double>(2 * (size_t)nbmax + 64);
    }
    scal_ = ar_.alloc<double>(S_NSCAL);
    HIPCHECK(hipMemsetAsync(scal_, 0, S_NSCAL * sizeof(double), s_));
    HIPCHECK(hipHostMalloc((void**)&hscal_, S_NSCAL * sizeof(double), hipHostMallocDefault));
    HIPCHECK(hipStreamSynchronize(s_));
    hprec_ = &p;
    const double tt = wall_seconds();
    build_tail();
    if (tail_lev_ >= 0) phase_log("dense tail", tail_lev_ + 1, wall_seconds() - tt);
    return 0;
  }

  // Dense tail: from the first level l >= 1 with at most DENSE_TAIL_MAX rows,
  // the untransposed cycle on levels l .. nlev-1 is a fixed linear operator
  // B_l of x2l (no K-cycle there), formed once here by applying that cycle to
  // the unit vectors and applied afterwards as one dense product
  // y2l = B_l x2l: about ten launches of a few microseconds each become one.
  // 'T' applications keep the level-by-level cycle.  Not taken when that
  // level is a coarsest level already solved by one dense product
  // (DENSE_LU), nor with MLD_DENSE_TAIL=0.  Same operator; the sums of the
  // product are rounded in another order.
  static constexpr int DENSE_TAIL_MAX = 1024;
  int tail_lev_ = -1;
  V* tail_b_ = nullptr;
  void build_tail() {
    tail_lev_ = -1;
    tail_b_ = nullptr;
    if (const char* e = std::getenv("MLD_DENSE_TAIL"))
      if (e[0] == '0') return;
    const int nlev = (int)L_.size();
    int l = 1;
    while (l < nlev && L_[l].n > DENSE_TAIL_MAX) ++l;
    if (l >= nlev) return;
    for (int k = l; k < nlev; ++k)
      if (L_[k].ml_cycle == KCYCLE || L_[k].ml_cycle == KCYCLESYM) return;
    DLevel<V>& D = L_[l];
    if (l == nlev - 1 && (D.sm.solver == DENSE_LU || D.sm.smoother == NOPREC)) return;
    const int n = D.n;
    V* B = ar_.alloc<V>((size_t)n * n);
    for (int j = 0; j < n; ++j) {
      unit_vec_kernel<V><<<nblk(n), BS, 0, s_>>>(n, D.x2l, j);
      int rc = inner_ml_aply(l, 'N');
      if (rc) return;  // the cycle stays as it is
      col_store_kernel<V><<<nblk(n), BS, 0, s_>>>(n, D.y2l, B, j);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s_));
    tail_b_ = B;
    tail_lev_ = l;
  }
  // y2l = B x2l at the tail level
  void tail_apply(DLevel<V>& D) {
    tail_gemv_kernel<V><<<(D.n + BS / 64 - 1) / (BS / 64), BS, 0, s_>>>(D.n, tail_b_, D.x2l, D.y2l);
    HIPCHECK(hipGetLastError());
  }

  ~DevPrecT() override {
    if (hscal_) (void)hipHostFree(hscal_);
    for (auto& e : probe_ev_) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }

  size_t device_bytes() const override { return ar_.bytes; }
  int tail_level() const override { return tail_lev_; }

  // y = op(M^-1) x (mld_dprecaply2_vect with alpha = 1, beta = 0).
  int apply(const double* x, double* y, char trans) override { return apply_impl(x, y, trans, nullptr, -1); }

  // mld_dmlprec_aply_vect's final update y = alpha vy2l(1) + beta y
  // (mld_dmlprec_aply.f90:327): the cycle's output in a work vector, then one
  // streaming pass (alpha = 1, beta = 0 writes y directly).
  int apply_ab(double alpha, const double* x, double beta, double* y, char trans) override {
    if (alpha == 1.0 && beta == 0.0) return apply_impl(x, y, trans, nullptr, -1);
    HIPCHECK(hipSetDevice(device_));
    if (!wab_) wab_ = ar_.alloc<double>(n0_);
    int rc = apply_impl(x, wab_, trans, nullptr, -1);
    if (rc) return rc;
    axpby(alpha, wab_, beta, y);
    return 0;
  }
  double* wab_ = nullptr;

  // y = op(M^-1) x; when dot_u is given, also scal_[dot_slot] = dot_u . y.
  // The fine level's last post-smoothing sweep writes y (FP64) and the dot
  // directly when it can (fused output), else a conversion / dot follows.
  int apply_impl(const double* x, double* y, char trans, const double* dot_u, int dot_slot) {
    HIPCHECK(hipSetDevice(device_));
    if (trans != 'N') ensure_transposes();
    struct Reset {  // the caller's pointers are dropped also when a HIP error throws
      DevPrecT& p;
      ~Reset() {
        p.out64_ = nullptr;
        p.dot_u_ = nullptr;
      }
    } reset{*this};
    out64_ = y;
    dot_u_ = dot_u;
    dot_slot_ = dot_slot;
    out_done_ = false;
    int rc = apply_cycle(x, y, trans);
    out64_ = nullptr;
    dot_u_ = nullptr;
    if (rc) return rc;
    if (dot_u && !out_done_) dot1(dot_u, y, dot_slot);
    return 0;
  }
  // FP32 form, untransposed multilevel cycle with one outer sweep: z = M^-1 x
  // left in the fine level's FP32 output buffer (returned), dot u.z into
  // dot_slot.  The values are those apply_impl widens to FP64.
  int apply_keep32(const double* x, const double* dot_u, int dot_slot, const V** z) {
    HIPCHECK(hipSetDevice(device_));
    if (std::is_same<V, double>::value || L_.size() < 2 || outer_sweeps_ != 1) return E_UNSUPPORTED;
    struct Reset {  // also when a HIP error throws out of the cycle
      DevPrecT& p;
      ~Reset() {
        p.keep32_ = false;
        p.dot_u_ = nullptr;
      }
    } reset{*this};
    keep32_ = true;
    out64_ = nullptr;
    dot_u_ = dot_u;
    dot_slot_ = dot_slot;
    out_done_ = false;
    int rc = apply_cycle(x, nullptr, 'N');
    keep32_ = false;
    dot_u_ = nullptr;
    if (rc) return rc;
    *z = L_[0].y2l;
    if (dot_u && !out_done_) {
      dot_kernel<double, 1, V><<<red_blocks(n0_), BS, 0, s_>>>(n0_, dot_u, L_[0].y2l, nullptr, nullptr,
                                                               red(dot_slot));
      finalize(1, dot_slot);
      HIPCHECK(hipGetLastError());
    }
    return 0;
  }

  int apply_cycle(const double* x, double* y, char trans) {
    const int nlev = (int)L_.size();
    DLevel<V>& L0 = L_[0];
    if (nlev == 1) {
      const int nsw = std::max(L0.sweeps_pre, L0.sweeps_post);
      const V* xin = in_vec(x, L0.x2l, L0.n);
      V* yout = out_vec(y, L0.y2l);
      if (L0.has_sm2) {
        // symmetrized pair (mld_dprecaply.f90:378-412): w1 = x; nsw times
        // w2 = S1 w1, w1 = S2 w2; y = w1 (S1, S2 swapped for 'T')
        DSmoother<V>& s1 = (trans == 'N') ? L0.sm : L0.sm2;
        DSmoother<V>& s2 = (trans == 'N') ? L0.sm2 : L0.sm;
        // w1 lives in the output, w2 in ty (the smoothers' own scratch is
        // tx, t1, t2)
        V* w1 = yout;
        V* w2 = L0.ty;
        add_kernel<V, V, false><<<nblk(L0.n), BS, 0, s_>>>(L0.n, xin, w1);
        for (int k = 0; k < nsw; ++k) {
          int rc = smooth(L0, s1, w1, 0, w2, 1, trans);
          if (rc) return rc;
          rc = smooth(L0, s2, w2, 0, w1, 1, trans);
          if (rc) return rc;
        }
        finish_out(y, yout, L0.n);
        return 0;
      }
      int rc = smooth(L0, L0.sm, xin, 0, yout, nsw, trans);
      if (rc) return rc;
      finish_out(y, yout, L0.n);
      return 0;
    }
    if (outer_sweeps_ > 1) {
      out64_ = nullptr;  // intermediate cycles must update y2l itself
      return apply_outer(x, y, trans);
    }
    // level-0 input/output alias the caller's vectors in FP64; the level's
    // own buffers are restored on every way out, a thrown HIP error included
    struct Restore {
      DLevel<V>& L;
      V* x;
      V* y;
      ~Restore() {
        L.x2l = x;
        L.y2l = y;
      }
    } restore{L0, L0.x2l, L0.y2l};
    L0.x2l = const_cast<V*>(in_vec(x, L0.x2l, L0.n));
    L0.y2l = out_vec(y, L0.y2l);
    int rc = inner_ml_aply(0, trans);
    if (y && !out_done_) finish_out(y, L0.y2l, L0.n);
    return rc;
  }

  // The Krylov loop enqueues the next application's first kernel, the fine
  // level's pre-smoothing of x into y (operands aliased as apply_cycle does),
  // before it waits for the residual norm on the host, so that the GPU works
  // while the host reads the norm and launches the rest; the next
  // inner_mult at level 0 then skips that launch.  When the loop stops
  // instead, the launched sweep only wrote the fine level's output buffer.
  // Returns false (nothing launched) where the cycle does not start that way.
  bool launch_pre0(const double* x, double* y, char trans) {
    if (L_.size() < 2 || outer_sweeps_ != 1) return false;
    DLevel<V>& L0 = L_[0];
    if (L0.ml_cycle != VCYCLE && L0.ml_cycle != WCYCLE && L0.ml_cycle != MULT_ML) return false;
    const bool tn = (trans == 'N');
    if (!((L0.sweeps_pre > 0 && tn) || (L0.sweeps_post > 0 && !tn))) return false;
    struct Restore {
      DLevel<V>& L;
      V* x;
      V* y;
      ~Restore() {
        L.x2l = x;
        L.y2l = y;
      }
    } restore{L0, L0.x2l, L0.y2l};
    L0.x2l = const_cast<V*>(in_vec(x, L0.x2l, L0.n));
    preconv_ = x;  // x2l now holds x (converted in the FP32 form)
    L0.y2l = out_vec(y, L0.y2l);
    DSmoother<V>& Spre = (tn || !L0.has_sm2) ? L0.sm : L0.sm2;
    if (smooth(L0, Spre, L0.x2l, 0, L0.y2l, tn ? L0.sweeps_pre : L0.sweeps_post, trans)) return false;
    pre0_launched_ = true;
    return true;
  }
  bool pre0_launched_ = false;

  // fused-output state of the current apply_impl call
  double* out64_ = nullptr;
  bool keep32_ = false;  // apply_keep32: the cycle's result stays in the FP32 y2l
  const double* dot_u_ = nullptr;
  int dot_slot_ = -1;
  bool out_done_ = false;

  int spmv_level(int lev, int what, const double* x, double* y) override {
    if (!std::is_same<V, double>::value) return E_UNSUPPORTED;
    if (lev < 0 || lev >= (int)L_.size()) return E_BAD_LEVEL;
    DLevel<V>& D = L_[lev];
    const DMat<V>* m = nullptr;
    switch (what) {
      case 0: m = &D.a; break;
      case 1: m = lev > 0 ? &D.p : nullptr; break;
      case 2: m = lev > 0 ? &D.r : nullptr; break;
      case 3: m = &D.sm.nd; break;
    }
    if (!m) return E_BAD_ARG;
    spmv_launch(*m, GatherX<V, double>{x}, EpiSet<V, double>{y}, s_);
    return 0;
  }

  int smooth_level(int lev, const double* x, int beta, double* y, int sweeps, char trans, int which) override {
    if (!std::is_same<V, double>::value) return E_UNSUPPORTED;
    if (lev < 0 || lev >= (int)L_.size()) return E_BAD_LEVEL;
    if (trans != 'N') ensure_transposes();
    DLevel<V>& D = L_[lev];
    DSmoother<V>& S = (which == 1 && D.has_sm2) ? D.sm2 : D.sm;
    return smooth(D, S, (const V*)x, beta, (V*)y, sweeps, trans);
  }

  int krylov(int method, const double* b, double* x, double eps, int itmax, int istop, int* iter,
             double* err, char trans) override;

  int sync() override {
    HIPCHECK(hipMemcpyAsync(hscal_, scal_, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIPCHECK(hipStreamSynchronize(s_));
    check_abort();
    return 0;
  }

  int op_format(int lev, int what, int* fmt, int* lanes, int* width) override {
    if (lev < 0 || lev >= (int)L_.size()) return E_BAD_LEVEL;
    DLevel<V>& D = L_[lev];
    const DMat<V>* m = nullptr;
    switch (what) {
      case 0: m = &D.a; break;
      case 1: m = &D.p; break;
      case 2: m = &D.r; break;
      case 3: m = &D.sm.nd; break;
      default: return E_BAD_ARG;
    }
    *fmt = m->fmt; *lanes = m->lanes; *width = m->w;
    return 0;
  }

  int time_step(int lev, int kind, int reps, double* avg_ms) override {
    HIPCHECK(hipSetDevice(device_));
    if (lev < 0 || lev >= (int)L_.size() || reps <= 0) return E_BAD_LEVEL;
    if ((kind == 2 || kind == 3) && lev == 0) return E_BAD_LEVEL;
    if (kind == 5 && lev != 0) return E_BAD_LEVEL;
    if (kind == 6) {
      // the residual launches of level lev inside reps whole applications of
      // the preconditioner (inputs and caches as in a solve), each bracketed
      // by events on the handle's stream
      const size_t want = (size_t)reps * 2 * (size_t)std::max(1, outer_sweeps_) * 4 + 8;
      while (probe_ev_.size() < want) {
        hipEvent_t a, b;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        probe_ev_.push_back({a, b});
      }
      double* kx = kvec(2);
      double* kq = kvec(3);
      fill_kernel<double><<<nblk(n0_), BS, 0, s_>>>(n0_, kx, 1.0);
      int rc = apply_impl(kx, kq, 'N', nullptr, -1);  // warm-up
      if (rc) return rc;
      probe_lev_ = lev;
      probe_used_ = 0;
      for (int i = 0; i < reps && !rc; ++i) rc = apply_impl(kx, kq, 'N', nullptr, -1);
      probe_lev_ = -1;
      HIPCHECK(hipStreamSynchronize(s_));
      double tot = 0.0;
      for (size_t i = 0; i < probe_used_; ++i) {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, probe_ev_[i].first, probe_ev_[i].second));
        tot += ms;
      }
      *avg_ms = probe_used_ ? tot / (double)probe_used_ : 0.0;
      return rc ? rc : (probe_used_ ? 0 : E_BAD_LEVEL);
    }
    DLevel<V>& D = L_[lev];
    // deterministic, finite operands
    fill_kernel<V><<<nblk(D.n), BS, 0, s_>>>(D.n, D.x2l, (V)1);
    fill_kernel<V><<<nblk(D.n), BS, 0, s_>>>(D.n, D.y2l, (V)0.5);
    fill_kernel<V><<<nblk(D.n), BS, 0, s_>>>(D.n, D.ty, (V)0.25);
    if (lev > 0) {
      DLevel<V>& F = L_[lev - 1];
      fill_kernel<V><<<nblk(F.n), BS, 0, s_>>>(F.n, F.ty, (V)0.25);
      fill_kernel<V><<<nblk(F.n), BS, 0, s_>>>(F.n, F.y2l, (V)0.5);
      fill_kernel<V><<<nblk(D.n), BS, 0, s_>>>(D.n, D.y2l, (V)0.5);
    }
    double* kx = nullptr;
    double* kq = nullptr;
    if (kind == 5) {
      kx = kvec(2);
      kq = kvec(3);
      fill_kernel<double><<<nblk(n0_), BS, 0, s_>>>(n0_, kx, 1.0);
    }
    HIPCHECK(hipGetLastError());
    auto run = [&]() -> int {
      switch (kind) {
        case 0: residual(D, D.x2l, D.y2l, D.ty, 'N'); return 0;
        case 1: return smooth(D, D.sm, D.x2l, 0, D.y2l, D.sweeps_pre, 'N');
        case 2: {
          DLevel<V>& F = L_[lev - 1];
          spmv_launch(D.r, GatherX<V, V>{F.ty}, EpiSet<V, V>{D.x2l}, s_);
          return 0;
        }
        case 3: {
          DLevel<V>& F = L_[lev - 1];
          spmv_launch(D.p, GatherX<V, V>{D.y2l}, EpiAdd<V>{F.y2l}, s_);
          return 0;
        }
        case 4: return smooth(D, D.has_sm2 ? D.sm2 : D.sm, D.ty, 1, D.y2l, D.sweeps_post, 'N');
        case 5:
          kspmv<1>('N', GatherX<double, double>{kx}, EpiSetDot<double>{kq, kx, nullptr}, red(S_TMP0));
          return 0;
      }
      return E_BAD_ARG;
    };
    int rc = run();  // warm-up
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, s_));
    for (int i = 0; i < reps && !rc; ++i) rc = run();
    HIPCHECK(hipEventRecord(e1, s_));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / reps;
    return rc;
  }
  void set_stream(hipStream_t s) override { s_ = s; }
  hipStream_t stream() const override { return s_; }

 private:
  int device_;
  hipStream_t s_;
  DevArena ar_;
  std::vector<DLevel<V>> L_;
  int outer_sweeps_ = 1;
  int n0_ = 0;
  DMat<double> akry_;
  DMat<float> akry32_;   // FP32-stored A (exact values), used when akry_is32_
  bool akry_is32_ = false;
  DMat<double> akry_t_;  // FP64 A^T for adjoint solves, uploaded on first use
  // Krylov product with A ('N') or A^T ('T'), FP64 arithmetic
  template <int K, class G, class E>
  void kspmv(char trans, G g, E e, Red r = Red{nullptr, nullptr}) {
    if (trans == 'T') spmv_launch<K>(akry_t_, g, e, s_, r);
    else if (akry_is32_) spmv_launch<K, double>(akry32_, g, e, s_, r);
    else spmv_launch<K>(akry_, g, e, s_, r);
  }
  bool akry_t_set_ = false;
  double* partial_ = nullptr;
  Red red(int slot) { return Red{partial_, scal_ + slot}; }
  // one-block sum of the partials of a streaming reduction over n0_ entries
  void finalize(int K, int slot, int nb = -1) {
    if (nb < 0) nb = red_blocks(n0_);
    if (K == 1) finalize_kernel<1><<<1, FIN_T, 0, s_>>>(nb, partial_, scal_ + slot);
    else finalize_kernel<2><<<1, FIN_T, 0, s_>>>(nb, partial_, scal_ + slot);
    HIPCHECK(hipGetLastError());
  }
  double* scal_ = nullptr;
  double* hscal_ = nullptr;
  const Prec* hprec_ = nullptr;
  bool transposes_ = false;
  std::vector<double*> kvec_;  // Krylov work vectors
  double* kr_ = nullptr;

  // --- vector plumbing between the FP64 interface and the V-typed cycle
  // FP64 vector whose FP32 copy the last Krylov update already left in the
  // fine level's input buffer (FP32 form); null when there is none
  const double* preconv_ = nullptr;
  const V* in_vec(const double* x, V* buf, int n) {
    if (std::is_same<V, double>::value) return (const V*)x;
    if (x == preconv_ && buf == L_[0].x2l) return buf;
    convert_kernel<V, double><<<nblk(n), BS, 0, s_>>>(n, x, buf);
    HIPCHECK(hipGetLastError());
    return buf;
  }
  V* out_vec(double* y, V* buf) {
    if (std::is_same<V, double>::value) return (V*)y;
    return buf;
  }
  void finish_out(double* y, const V* buf, int n) {
    if (std::is_same<V, double>::value) return;
    convert_kernel<double, V><<<nblk(n), BS, 0, s_>>>(n, buf, y);
    HIPCHECK(hipGetLastError());
  }

  void upload_smoother(const SmootherData& H, DSmoother<V>& S, int force) {
    S.smoother = H.smoother;
    S.solver = H.solver;
    S.nnz_nd = H.nnz_nd;
    if (H.smoother == NOPREC) return;
    if (!H.dv.empty()) {
      std::vector<V> dv(H.dv.begin(), H.dv.end());
      S.dv = ar_.upload(dv.data(), dv.size(), s_);
      HIPCHECK(hipStreamSynchronize(s_));
    }
    if (H.nnz_nd > 0) S.nd = upload_tuned<V>(H.nd, ar_, s_, force);
    // Jacobi-type solver: ND D (ND with its columns scaled by d, sharing ND's
    // indices) stored beside ND, so that the first sweep from zero, ND (D x),
    // is the plain product (ND D) x: one gather per entry, and on SELL / CSR
    // no separate pass forming D x.  The product's rounding is re-associated,
    // (a_ij d_j) x_j for a_ij (d_j x_j).
    if (H.nnz_nd > 0 && S.dv && (H.solver == DIAG_SCALE || H.solver == L1_DIAG)) {
      S.ndd = S.nd;
      S.ndd.val = ar_.alloc<V>((size_t)std::max<int64_t>(1, S.nd.stored));
      if (S.nd.fmt == FMT_DIA) {
        dia_colscale_kernel<V><<<nblk(S.nd.nr), BS, 0, s_>>>(S.nd.nr, S.nd.nc, S.nd.w, S.nd.off, S.nd.val, S.dv,
                                                             S.ndd.val);
      } else if (S.nd.fmt == FMT_SELD) {
        const int npos = (S.nd.nr + 63) / 64 * 64;
        seld_colscale_kernel<V><<<nblk(npos), BS, 0, s_>>>(npos, S.nd.nc, S.nd.sptr, S.nd.col0, S.nd.dja,
                                                             S.nd.val, S.dv, S.ndd.val);
      } else if (S.nd.stored > 0) {
        colscale_kernel<V><<<(int)((S.nd.stored + BS - 1) / BS), BS, 0, s_>>>(S.nd.stored, S.nd.ja, S.nd.val,
                                                                               S.dv, S.ndd.val);
      }
      HIPCHECK(hipGetLastError());
    }
    if (H.solver == ILU_N || H.solver == GS || H.solver == BWGS) {
      S.l = upload_mat<V>(H.l, ar_, s_, 1, 1);  // CSR (rows summed in stored order)
      S.u = upload_mat<V>(H.u, ar_, s_, 1, 1);
      {
        // lanes per row of the triangular solves, from the mean row length
        const double avg = (double)std::max(H.l.nnz(), H.u.nnz()) / std::max(1, H.l.nr);
        // Lanes per row: one up to 3.5 entries per row (the 7-point fine
        // level's factors, 3 per row: 256 rows per chunk instead of 64, a
        // quarter of the ticket draws; run-B V-cycle 47.4 -> 45.8 ms at
        // 256^3), four up to 6, else 32 (the Galerkin level's factors, about
        // 15 per row: 8 rows per chunk; 16 lanes 45.8, 32 lanes 41.4, 64 lanes
        // 48.4 ms).  MLD_TRI1_MAX and MLD_TRI_LANES override them, for
        // measurements.
        static const double one_lane_max = [] {
          const char* e = std::getenv("MLD_TRI1_MAX");
          return e ? std::atof(e) : 3.5;
        }();
        S.tri_lanes = avg <= one_lane_max ? 1 : (avg <= 6.0 ? 4 : 32);
        if (const char* e = std::getenv("MLD_TRI_LANES")) {
          const int l = std::atoi(e);
          if (avg > one_lane_max && (l == 4 || l == 16 || l == 32)) S.tri_lanes = l;
        }
      }
      S.lrows = ar_.upload(H.lsched_rows.data(), H.lsched_rows.size(), s_);
      S.urows = ar_.upload(H.usched_rows.data(), H.usched_rows.size(), s_);
      S.lptr = H.lsched_ptr;
      S.uptr = H.usched_ptr;
      S.lptr_d = ar_.upload(H.lsched_ptr.data(), H.lsched_ptr.size(), s_);
      S.uptr_d = ar_.upload(H.usched_ptr.data(), H.usched_ptr.size(), s_);
      {
        const int per = BS / S.tri_lanes;
        const std::vector<int> lc = sched_chunks(S.lptr, per), uc = sched_chunks(S.uptr, per);
        S.lch = ar_.upload(lc.data(), lc.size(), s_);
        S.uch = ar_.upload(uc.data(), uc.size(), s_);
        S.nlch = (int)lc.size() - 1;
        S.nuch = (int)uc.size() - 1;
        S.sf_ticket = ar_.alloc<int>(1);
        HIPCHECK(hipMemsetAsync(S.sf_ticket, 0, sizeof(int), s_));
      }
      if (H.l.nr > SMALL_TRI) {  // ILU(0), GS, BWGS
        std::vector<V> lg, ug;
        if (!H.dv.empty() && grid_factors(H.l, H.u, S.gnx, S.gny, S.gnz, lg, ug)) {
          skew_upload(S, lg, ug, H.dv);
          if (std::getenv("MLD_VERBOSE"))
            std::fprintf(stderr, "[mld] %d-row %s factors of a 7-point stencil on a %d x %d x %d grid: "
                                 "grid substitutions\n", H.l.nr, H.solver == ILU_N ? "ILU(0)" : "Gauss-Seidel",
                         S.gnx, S.gny, S.gnz);
        }
      }
      HIPCHECK(hipStreamSynchronize(s_));
    }
    if (H.solver == DENSE_LU) {
      const int n = (int)std::lround(std::sqrt((double)H.lu.size()));
      S.lu = ar_.upload(H.lu.data(), H.lu.size(), s_);  // factorized in place
      S.piv = ar_.alloc<int>(n);
      S.dinv = ar_.alloc<double>((size_t)n * n);
      S.ndense = n;
      dense_factor_invert(n, S.lu, S.piv, S.dinv);
    }
  }

  // Blocked LU with partial pivoting (MFMA trailing updates) and the explicit
  // inverse X = U^-1 L^-1 P of the n x n row-major matrix a (overwritten by
  // its factors).
  void dense_factor_invert(int n, double* a, int* piv, double* x) {
    const int nb = 32;
    int* info = ar_.alloc<int>(1);
    HIPCHECK(hipMemsetAsync(info, 0, sizeof(int), s_));
    auto gemm = [&](int m, int nc, int kd, const double* A, const double* B, double* C) {
      if (m <= 0 || nc <= 0 || kd <= 0) return;
      const int tiles = ((m + 15) / 16) * ((nc + 15) / 16);
      gemm_sub_mfma_kernel<<<(tiles + 3) / 4, 256, 0, s_>>>(m, nc, kd, A, n, B, n, C, n);
    };
    for (int k0 = 0; k0 < n; k0 += nb) {
      const int kend = std::min(n, k0 + nb);
      panel_factor_kernel<<<1, 1024, 0, s_>>>(n, a, k0, nb, piv, info);
      if (kend < n) {
        trsm_cols_kernel<true><<<nblk(n - kend), 256, 0, s_>>>(n, a, k0, nb, a, n, kend, n);
        gemm(n - kend, n - kend, kend - k0, a + (size_t)kend * n + k0, a + (size_t)k0 * n + kend,
             a + (size_t)kend * n + kend);
      }
    }
    perm_identity_kernel<<<nblk(n), 256, 0, s_>>>(n, piv, x);
    for (int k0 = 0; k0 < n; k0 += nb) {  // L Y = P X
      const int kend = std::min(n, k0 + nb);
      trsm_cols_kernel<true><<<nblk(n), 256, 0, s_>>>(n, a, k0, nb, x, n, 0, n);
      if (kend < n) gemm(n - kend, n, kend - k0, a + (size_t)kend * n + k0, x + (size_t)k0 * n, x + (size_t)kend * n);
    }
    for (int k0 = ((n - 1) / nb) * nb; k0 >= 0; k0 -= nb) {  // U X = Y
      const int kend = std::min(n, k0 + nb);
      trsm_cols_kernel<false><<<nblk(n), 256, 0, s_>>>(n, a, k0, nb, x, n, 0, n);
      if (k0 > 0) gemm(k0, n, kend - k0, a + k0, x + (size_t)k0 * n, x);
    }
    HIPCHECK(hipGetLastError());
    int hinfo = 0;
    HIPCHECK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s_));
    HIPCHECK(hipStreamSynchronize(s_));
    if (hinfo != 0) throw DevError("dense LU: zero pivot in column " + std::to_string(hinfo));
  }

  // Transposed operators for trans = 'T' (A^T, ND^T, U^T and L^T).
  void ensure_transposes() {
    if (transposes_) return;
    const Prec& p = *hprec_;
    for (size_t l = 0; l < L_.size(); ++l) {
      const Level& H = p.precv[l];
      DLevel<V>& D = L_[l];
      D.at = upload_mat<V>(host_transpose(H.a), ar_, s_, p.dev_format);
      auto tsm = [&](const SmootherData& hs, DSmoother<V>& ds) {
        if (hs.nnz_nd > 0) ds.ndt = upload_mat<V>(host_transpose(hs.nd), ar_, s_, p.dev_format);
        if (hs.solver == ILU_N) {
          HCsr ut = host_transpose(hs.u), lt = host_transpose(hs.l);  // U^T lower, L^T upper
          std::vector<int> p1, r1, p2, r2;
          level_schedule(ut, true, p1, r1);
          level_schedule(lt, false, p2, r2);
          ds.ut = upload_mat<V>(ut, ar_, s_, 1, 1);
          ds.lt = upload_mat<V>(lt, ar_, s_, 1, 1);
          ds.utrows = ar_.upload(r1.data(), r1.size(), s_);
          ds.ltrows = ar_.upload(r2.data(), r2.size(), s_);
          ds.utptr = p1; ds.ltptr = p2;
          {
            const int per = BS / ds.tri_lanes;
            const std::vector<int> c1 = sched_chunks(p1, per), c2 = sched_chunks(p2, per);
            ds.utch = ar_.upload(c1.data(), c1.size(), s_);
            ds.ltch = ar_.upload(c2.data(), c2.size(), s_);
            ds.nutch = (int)c1.size() - 1;
            ds.nltch = (int)c2.size() - 1;
          }
          ds.utptr_d = ar_.upload(p1.data(), p1.size(), s_);
          ds.ltptr_d = ar_.upload(p2.data(), p2.size(), s_);
          HIPCHECK(hipStreamSynchronize(s_));
        }
        ds.has_t = true;
      };
      tsm(H.smd, D.sm);
      if (D.has_sm2) tsm(H.smd2, D.sm2);
    }
    transposes_ = true;
  }

  // ---------------------------------------------------------------- solver
  // y = S^-1 x (beta 0) or y = y + S^-1 x (beta 1).
  int solver_apply(DLevel<V>& D, DSmoother<V>& S, const V* x, int beta, V* y, char trans) {
    const int n = D.n;
    switch (S.solver) {
      case F_NONE:
        if (beta) add_kernel<V, V, true><<<nblk(n), BS, 0, s_>>>(n, x, y);
        else add_kernel<V, V, false><<<nblk(n), BS, 0, s_>>>(n, x, y);
        break;
      case DIAG_SCALE:
      case L1_DIAG:
        if (beta) vmul_kernel<V, V, true><<<nblk(n), BS, 0, s_>>>(n, S.dv, x, y);
        else vmul_kernel<V, V, false><<<nblk(n), BS, 0, s_>>>(n, S.dv, x, y);
        break;
      case ILU_N:
        if (trans != 'N') ilu_apply<true>(D, S, x, beta, y);
        else ilu_apply<false>(D, S, x, beta, y);
        break;
      case DENSE_LU:  // y = A^-1 x with the inverse formed at upload (FP64)
        if (trans != 'N') return E_UNSUPPORTED;
        if (beta) dense_gemv_kernel<V, 1><<<(n + 3) / 4, 256, 0, s_>>>(n, S.dinv, x, y);
        else dense_gemv_kernel<V, 0><<<(n + 3) / 4, 256, 0, s_>>>(n, S.dinv, x, y);
        HIPCHECK(hipGetLastError());
        break;
      default:
        return E_UNSUPPORTED;
    }
    return 0;
  }

  // ILU(0) solver, y = U^-1 D L^-1 x ('N') or L^-T D U^-T x ('T'), beta 0/1
  // (mld_d_ilu_solver_apply_vect.f90:130-142).
  template <bool TR>
  void ilu_apply(DLevel<V>& D, DSmoother<V>& S, const V* x, int beta, V* y) {
    switch (S.tri_lanes) {
      case 1: ilu_apply_l<TR, 1>(D, S, x, beta, y); break;
      case 4: ilu_apply_l<TR, 4>(D, S, x, beta, y); break;
      case 16: ilu_apply_l<TR, 16>(D, S, x, beta, y); break;
      default: ilu_apply_l<TR, 32>(D, S, x, beta, y); break;
    }
  }
  template <bool TR, int LN>
  void ilu_apply_l(DLevel<V>& D, DSmoother<V>& S, const V* x, int beta, V* y) {
    const int n = D.n;
    const DMat<V>& F1 = TR ? S.ut : S.l;  // first (lower-triangular) solve
    const DMat<V>& F2 = TR ? S.lt : S.u;  // second (upper-triangular) solve
    const std::vector<int>& p1 = TR ? S.utptr : S.lptr;
    const std::vector<int>& p2 = TR ? S.ltptr : S.uptr;
    const int* r1 = TR ? S.utrows : S.lrows;
    const int* r2 = TR ? S.ltrows : S.urows;
    const int* p1d = TR ? S.utptr_d : S.lptr_d;
    const int* p2d = TR ? S.ltptr_d : S.uptr_d;
    const int nl1 = (int)p1.size() - 1, nl2 = (int)p2.size() - 1;
    V* t1 = D.tx;
    V* t2 = beta ? D.t2 : y;
    if (n <= SMALL_TRI) {
      // one workgroup; a single wavefront when the matrix is tiny
      const int nt = std::min(1024, (int)(((int64_t)n * LN + 63) / 64) * 64);
      const size_t lds = 2 * (size_t)n * sizeof(V);  // both intermediate vectors in LDS when they fit
      if (lds <= ILU_LDS_MAX) {
        if (beta)
          ilu_small_kernel<V, 1, TR, LN, true><<<1, nt, lds, s_>>>(n, F1.irp, F1.ja, F1.val, F2.irp, F2.ja, F2.val,
                                                                   S.dv, p1d, r1, nl1, p2d, r2, nl2, x, t1, t2, y);
        else
          ilu_small_kernel<V, 0, TR, LN, true><<<1, nt, lds, s_>>>(n, F1.irp, F1.ja, F1.val, F2.irp, F2.ja, F2.val,
                                                                   S.dv, p1d, r1, nl1, p2d, r2, nl2, x, t1, t2, y);
      } else if (beta) {
        ilu_small_kernel<V, 1, TR, LN><<<1, nt, 0, s_>>>(n, F1.irp, F1.ja, F1.val, F2.irp, F2.ja, F2.val, S.dv, p1d,
                                                         r1, nl1, p2d, r2, nl2, x, t1, t2, y);
      } else {
        ilu_small_kernel<V, 0, TR, LN><<<1, nt, 0, s_>>>(n, F1.irp, F1.ja, F1.val, F2.irp, F2.ja, F2.val, S.dv, p1d,
                                                         r1, nl1, p2d, r2, nl2, x, t1, t2, y);
      }
      HIPCHECK(hipGetLastError());
      return;
    }
    if (!TR && S.lgs && S.solver == ILU_N && sync_free() && grid_trsv()) {
      // x into step order (xgs), L solve into tgs, U solve into xgs, result
      // back into natural order
      grid_in(S, x);
      grid_launch<0, false>(S, S.xgs, S.tgs);
      grid_launch<1, true>(S, S.tgs, S.xgs);
      grid_out(S, S.xgs, nullptr, y, beta ? 1 : 0);
      return;
    }
    if (sync_free()) {
      const int* c1 = TR ? S.utch : S.lch;
      const int* c2 = TR ? S.ltch : S.uch;
      const int n1 = TR ? S.nutch : S.nlch, n2 = TR ? S.nltch : S.nuch;
      sf_launch<0, 2, TR, LN>(S, n, n1, nl1, c1, r1, F1, x, nullptr, t1, nullptr);
      if (beta) sf_launch<1, 1, TR, LN>(S, n, n2, nl2, c2, r2, F2, t1, S.dv, t2, y);
      else sf_launch<1, 2, TR, LN>(S, n, n2, nl2, c2, r2, F2, t1, S.dv, t2, y);
      return;
    }
    // level sets of at most RUN_ROWS rows (two passes of the workgroup):
    // consecutive ones in one single-workgroup launch; larger ones: one grid
    // launch each
    constexpr int RUN_ROWS = 2048 / LN;
    auto nbl = [](int cnt) { return (int)(((int64_t)cnt * LN + BS - 1) / BS); };
    for (int lv = 0; lv < nl1;) {
      int le = lv;
      while (le < nl1 && p1[le + 1] - p1[le] <= RUN_ROWS) ++le;
      if (le - lv >= 2) {
        trsv_run_kernel<V, 0, 2, TR, LN><<<1, 1024, 0, s_>>>(lv, le, p1d, r1, F1.irp, F1.ja, F1.val, x,
                                                                nullptr, t1, nullptr);
        lv = le;
        continue;
      }
      const int cnt = p1[lv + 1] - p1[lv];
      trsv_level_kernel<V, 0, 2, TR, LN><<<nbl(cnt), BS, 0, s_>>>(cnt, r1 + p1[lv], F1.irp, F1.ja, F1.val, x,
                                                                     nullptr, t1, nullptr);
      ++lv;
    }
    for (int lv = 0; lv < nl2;) {
      int le = lv;
      while (le < nl2 && p2[le + 1] - p2[le] <= RUN_ROWS) ++le;
      if (le - lv >= 2) {
        if (beta)
          trsv_run_kernel<V, 1, 1, TR, LN><<<1, 1024, 0, s_>>>(lv, le, p2d, r2, F2.irp, F2.ja, F2.val, t1, S.dv,
                                                                 t2, y);
        else
          trsv_run_kernel<V, 1, 2, TR, LN><<<1, 1024, 0, s_>>>(lv, le, p2d, r2, F2.irp, F2.ja, F2.val, t1, S.dv,
                                                                 t2, y);
        lv = le;
        continue;
      }
      const int cnt = p2[lv + 1] - p2[lv];
      if (beta)
        trsv_level_kernel<V, 1, 1, TR, LN><<<nbl(cnt), BS, 0, s_>>>(cnt, r2 + p2[lv], F2.irp, F2.ja, F2.val, t1,
                                                                      S.dv, t2, y);
      else
        trsv_level_kernel<V, 1, 2, TR, LN><<<nbl(cnt), BS, 0, s_>>>(cnt, r2 + p2[lv], F2.irp, F2.ja, F2.val, t1,
                                                                      S.dv, t2, y);
      ++lv;
    }
    HIPCHECK(hipGetLastError());
  }

  static constexpr int GRID_TRSV_BLOCKS = 2048;  // one wavefront each
  // One grid substitution (trsv_grid_kernel), every operand in the tiles'
  // step order: t filled with the sentinel, then the solve.
  template <int SC, bool BACK>
  void grid_launch(DSmoother<V>& S, const V* in, V* t) {
    sf_fill_kernel<V><<<(int)((S.gsz + BS - 1) / BS), BS, 0, s_>>>((int)S.gsz, t);
    const int ntiles = ((S.gny + 7) / 8) * ((S.gnz + 7) / 8);
    const int nwg = std::min(ntiles, GRID_TRSV_BLOCKS);
    trsv_grid_kernel<V, SC, BACK><<<nwg, 64, 0, s_>>>(S.gnx, S.gny, S.gnz, S.gsz, BACK ? S.ugs : S.lgs, S.dgs, in, t,
                                                      S.sf_ticket, scal_ + S_ABORT, SF_SPIN_MAX);
    HIPCHECK(hipGetLastError());
  }
  // x into the step order (S.xgs)
  void grid_in(DSmoother<V>& S, const V* x) {
    const int ntiles = ((S.gny + 7) / 8) * ((S.gnz + 7) / 8);
    grid_to_steps_kernel<V><<<ntiles * ((S.gnx + 14 + 63) / 64), BS, 0, s_>>>(S.gnx, S.gny, S.gnz, x, S.xgs);
    HIPCHECK(hipGetLastError());
  }
  // ts back into natural order: t = ts (t not null) and y by bt (0: y = ts,
  // 1: y += ts, 2: y untouched)
  void grid_out(DSmoother<V>& S, const V* ts, V* t, V* y, int bt) {
    const int nb = ((S.gny + 7) / 8) * ((S.gnz + 7) / 8) * ((S.gnx + 63) / 64);
    if (bt == 0) grid_from_steps_kernel<V, 0><<<nb, BS, 0, s_>>>(S.gnx, S.gny, S.gnz, ts, t, y);
    else if (bt == 1) grid_from_steps_kernel<V, 1><<<nb, BS, 0, s_>>>(S.gnx, S.gny, S.gnz, ts, t, y);
    else grid_from_steps_kernel<V, 2><<<nb, BS, 0, s_>>>(S.gnx, S.gny, S.gnz, ts, t, y);
    HIPCHECK(hipGetLastError());
  }
  // One sync-free substitution over a whole schedule (nch chunks) of an
  // n-row factor: t filled with the sentinel, then the solve.
  template <int SC, int BT, bool TR, int LN>
  void sf_launch(DSmoother<V>& S, int n, int nch, int nlv, const int* ch, const int* rows, const DMat<V>& F,
                 const V* b, const V* d, V* t, V* y) {
    if (nch <= 0) return;
    sf_fill_kernel<V><<<nblk(n), BS, 0, s_>>>(n, t);
    const int nb = sf_grid(nch, nlv);
    trsv_sf_kernel<V, SC, BT, TR, LN><<<nb, BS, 0, s_>>>(nch, ch, rows, F.irp, F.ja, F.val, b, d, t, y,
                                                          S.sf_ticket, scal_ + S_ABORT, SF_SPIN_MAX);
    HIPCHECK(hipGetLastError());
  }
  // Workgroups of a sync-free substitution: enough to hold the chunks of
  // SF_AHEAD level sets at once, at least SF_MIN_BLOCKS, at most SF_BLOCKS
  // (MLD_SF_AHEAD, MLD_SF_MIN_BLOCKS, MLD_SF_BLOCKS override; 0 ahead: no
  // such limit).  Every
  // workgroup beyond that only spins on values not yet written, and the
  // spinning loads crowd the L2 that the producers' stores pass through: at
  // 256^3 (run B) the V-cycle takes 89 ms with 2048 workgroups per solve,
  // 57 ms with two level sets ahead and at most 512.
  static int sf_grid(int nch, int nlv) {
    int ahead = SF_AHEAD, cap = SF_BLOCKS;
    if (const char* e = std::getenv("MLD_SF_AHEAD")) ahead = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("MLD_SF_BLOCKS")) cap = std::max(1, std::atoi(e));
    int lo = SF_MIN_BLOCKS;
    if (const char* e = std::getenv("MLD_SF_MIN_BLOCKS")) lo = std::max(1, std::atoi(e));
    int nb = std::min(nch, cap);
    if (ahead > 0 && nlv > 0) nb = std::min(nb, std::max(lo, ahead * ((nch + nlv - 1) / nlv)));
    return std::max(1, nb);
  }

  // The level-scheduled launches instead (MLD_TRSV=level).
  // The ILU(0) factors l (strict lower) and u (strict upper) of a 7-point
  // stencil on an nx x ny x nz grid in natural order: every row r = i +
  // nx (j + ny k) of l holds exactly the columns r - nx ny (k > 0), r - nx
  // (j > 0), r - 1 (i > 0), in this order, and of u r + 1, r + nx, r + nx ny
  // (i < nx - 1, j < ny - 1, k < nz - 1); nx, ny, nz >= 2, at least 16 tiles of
  // 8 x 8 grid lines.  Then the values
  // by stencil position (0 where a row has no such entry) for
  // trsv_grid_kernel (in the tiles' step order: skew_upload), and true.
  // f(r0, r1) over [0, n) in contiguous ranges on host threads
  // (OMP_NUM_THREADS, default 8); f allocates nothing and does not throw
  template <class F>
  static void host_ranges(int64_t n, F f) {
    int nt = 8;
    if (const char* e = std::getenv("OMP_NUM_THREADS")) nt = std::atoi(e);
    nt = (int)std::min<int64_t>(std::max(1, std::min(nt, 64)), std::max<int64_t>(1, n / 65536));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(f, n * t / nt, n * (t + 1) / nt);
    f((int64_t)0, n / nt);
    for (auto& th : pool) th.join();
  }
  static bool grid_factors(const HCsr& l, const HCsr& u, int& nx, int& ny, int& nz, std::vector<V>& lg,
                           std::vector<V>& ug) {
    const int64_t n = l.nr;
    if (n < 8 || u.nr != n || l.irp[n] - l.irp[n - 1] != 3) return false;
    const int64_t c0 = l.ja[l.irp[n - 1]], c1 = l.ja[l.irp[n - 1] + 1];
    const int64_t sx = (n - 1) - c1, sxy = (n - 1) - c0;  // nx and nx ny from the last row
    if (sx < 2 || sxy <= sx || sxy % sx != 0 || n % sxy != 0) return false;
    if (sxy / sx < 2 || n / sxy < 2 || sxy > INT32_MAX || n / sxy > INT32_MAX) return false;
    nx = (int)sx;
    ny = (int)(sxy / sx);
    nz = (int)(n / sxy);
    // at least 16 tiles of 8 x 8 lines (thinner grids leave the tiles' chain
    // of waits with too little work beside it: the general form then)
    if (((ny + 7) / 8) * ((nz + 7) / 8) < 16) return false;
    lg.assign(3 * (size_t)n, (V)0);
    ug.assign(3 * (size_t)n, (V)0);
    std::atomic<bool> bad{false};
    const int gy = ny, gz = nz;
    host_ranges(n, [&](int64_t r0, int64_t r1) {
      for (int64_t r = r0; r < r1 && !bad.load(std::memory_order_relaxed); ++r) {
        const int64_t i = r % sx, j = (r / sx) % gy, k = r / sxy;
        int64_t q = l.irp[r];
        const int64_t lo[3] = {r - sxy, r - sx, r - 1};
        const bool lh[3] = {k > 0, j > 0, i > 0};
        bool ok = true;
        for (int e = 0; e < 3 && ok; ++e)
          if (lh[e]) {
            if (q >= l.irp[r + 1] || l.ja[q] != lo[e]) ok = false;
            else lg[(size_t)e * n + r] = (V)l.val[q++];
          }
        if (!ok || q != l.irp[r + 1]) { bad = true; break; }
        q = u.irp[r];
        const int64_t uo[3] = {r + 1, r + sx, r + sxy};
        const bool uh[3] = {i < sx - 1, j < gy - 1, k < gz - 1};
        for (int e = 0; e < 3 && ok; ++e)
          if (uh[e]) {
            if (q >= u.irp[r + 1] || u.ja[q] != uo[e]) ok = false;
            else ug[(size_t)e * n + r] = (V)u.val[q++];
          }
        if (!ok || q != u.irp[r + 1]) { bad = true; break; }
      }
    });
    return !bad;
  }
  // Position of grid point r = i + nx (j + ny k) in the tiles' step order:
  // tile T = j / 8 + ntj (k / 8), step s = i + j % 8 + k % 8 of the forward
  // solve, lane j % 8 + 8 (k % 8): ((T ns + s) 64 + lane), ns = nx + 14.  At
  // every step of either solve the 64 lanes of a tile touch 64 consecutive
  // positions.
  static int64_t skew_pos(int64_t r, int nx, int ny, int ntj) {
    const int64_t i = r % nx, j = (r / nx) % ny, k = r / ((int64_t)nx * ny);
    const int64_t tile = j / 8 + (int64_t)ntj * (k / 8);
    const int64_t st = i + j % 8 + k % 8, lane = j % 8 + 8 * (k % 8);
    return (tile * (nx + 14) + st) * 64 + lane;
  }
  // Factors on a grid (ILU(0), or the triangular parts of Gauss-Seidel):
  // values by stencil position (lg, ug: 3 n each) and d in the tiles' step
  // order, and two work vectors of that order.
  void skew_upload(DSmoother<V>& S, const std::vector<V>& lg, const std::vector<V>& ug,
                   const std::vector<double>& dv) {
    const int nx = S.gnx, ny = S.gny, nz = S.gnz;
    const int ntj = (ny + 7) / 8, ntk = (nz + 7) / 8;
    const int64_t n = (int64_t)nx * ny * nz, sz = (int64_t)ntj * ntk * (nx + 14) * 64;
    std::vector<V> a(3 * (size_t)sz, (V)0), dd((size_t)sz, (V)0);
    host_ranges(n, [&](int64_t r0, int64_t r1) {  // skew_pos is one-to-one: the ranges write apart
      for (int64_t r = r0; r < r1; ++r) {
        const int64_t q = skew_pos(r, nx, ny, ntj);
        for (int e = 0; e < 3; ++e) a[(size_t)e * sz + q] = lg[(size_t)e * n + r];
        dd[q] = (V)dv[r];
      }
    });
    S.lgs = ar_.upload(a.data(), a.size(), s_);
    HIPCHECK(hipStreamSynchronize(s_));  // a is refilled below
    host_ranges(n, [&](int64_t r0, int64_t r1) {
      for (int64_t r = r0; r < r1; ++r) {
        const int64_t q = skew_pos(r, nx, ny, ntj);
        for (int e = 0; e < 3; ++e) a[(size_t)e * sz + q] = ug[(size_t)e * n + r];
      }
    });
    S.ugs = ar_.upload(a.data(), a.size(), s_);
    S.dgs = ar_.upload(dd.data(), dd.size(), s_);
    S.tgs = ar_.alloc<V>((size_t)sz);
    S.xgs = ar_.alloc<V>((size_t)sz);
    S.gsz = sz;
    HIPCHECK(hipStreamSynchronize(s_));
  }
  // the grid form of the ILU(0) and Gauss-Seidel substitutions
  // (MLD_TRSV_GRID=0: the general sync-free form, for measurements and tests)
  static bool grid_trsv() {
    const char* e = std::getenv("MLD_TRSV_GRID");
    return !(e && std::atoi(e) == 0);
  }

  static bool sync_free() {
    static const bool sf = [] {
      const char* e = std::getenv("MLD_TRSV");
      return !(e && std::string(e) == "level");
    }();
    return sf;
  }
  static constexpr size_t ILU_LDS_MAX = 64 * 1024;  // LDS of one workgroup for the small ILU solves
  static constexpr int SMALL_TRI = 8192;  // triangular solves up to this many rows run in one workgroup
  static constexpr int SF_BLOCKS = 512;
  static constexpr int SF_MIN_BLOCKS = 64;
  static constexpr int SF_AHEAD = 2;
  static constexpr int SF_SPIN_MAX = 1 << 22;

  // Hybrid Gauss-Seidel smoother (jacobi smoother + GS / BWGS solver, which
  // is iterative: mld_d_jac_smoother_apply_vect.f90:117-177 always runs the
  // sweep loop; ND is empty on one device): `sweeps` sweeps from zero,
  //   GS:   y <- (D + L)^-1 (x - U y),   BWGS: y <- (D + U)^-1 (x - L y)
  // (mld_d_gs_solver_apply_vect.f90:150-156, mld_d_bwgs_solver_apply_vect.f90
  // :151-155).  The first sweep's product with the zero iterate is skipped
  // (x - 0 = x exactly).  Only 'N', as in the reference solvers.  The 'Z'
  // application precedes the loop there, so sweeps = 0 also gives one sweep.
  int gs_apply(DLevel<V>& D, DSmoother<V>& S, const V* x, int beta, V* y, int sweeps, char trans) {
    if (trans != 'N') return E_UNSUPPORTED;
    sweeps = std::max(1, sweeps);
    const bool fw = (S.solver == GS);
    const DMat<V>& O = fw ? S.u : S.l;
    V* t = D.tx;
    for (int sw = 1; sw <= sweeps; ++sw) {
      const V* b = x;
      if (sw > 1) {
        spmv_launch(O, GatherX<V, V>{t}, EpiRes<V, V>{x, D.t1}, s_);  // w = x - O t
        b = D.t1;
      }
      const int bt = (sw < sweeps) ? 2 : (beta ? 1 : 0);
      switch (S.tri_lanes) {
        case 1: tri_solve<1>(D, S, fw, b, t, y, bt); break;
        case 4: tri_solve<4>(D, S, fw, b, t, y, bt); break;
        case 16: tri_solve<16>(D, S, fw, b, t, y, bt); break;
        default: tri_solve<32>(D, S, fw, b, t, y, bt); break;
      }
    }
    HIPCHECK(hipGetLastError());
    return 0;
  }
  // t = (D + T)^-1 b with T the strict lower (lower = true) or upper part:
  // level-scheduled, division by the stored diagonal; bt: 0 also y = t,
  // 1 y = y + t, 2 t only.
  template <int LN>
  void tri_solve(DLevel<V>& D, DSmoother<V>& S, bool lower, const V* b, V* t, V* y, int bt) {
    const DMat<V>& T = lower ? S.l : S.u;
    const std::vector<int>& ph = lower ? S.lptr : S.uptr;
    const int* rows = lower ? S.lrows : S.urows;
    const int* pd = lower ? S.lptr_d : S.uptr_d;
    const int nl = (int)ph.size() - 1;
    auto run = [&](int l0, int l1) {
      if (bt == 0) trsv_run_kernel<V, 2, 0, false, LN><<<1, 1024, 0, s_>>>(l0, l1, pd, rows, T.irp, T.ja, T.val, b, S.dv, t, y);
      else if (bt == 1) trsv_run_kernel<V, 2, 1, false, LN><<<1, 1024, 0, s_>>>(l0, l1, pd, rows, T.irp, T.ja, T.val, b, S.dv, t, y);
      else trsv_run_kernel<V, 2, 2, false, LN><<<1, 1024, 0, s_>>>(l0, l1, pd, rows, T.irp, T.ja, T.val, b, S.dv, t, y);
    };
    if (D.n <= SMALL_TRI) {  // whole solve in one workgroup
      run(0, nl);
      return;
    }
    if (S.lgs && sync_free() && grid_trsv()) {
      // b into step order, solve, y back (and t, which only a following
      // sweep reads)
      grid_in(S, b);
      if (lower) grid_launch<2, false>(S, S.xgs, S.tgs);
      else grid_launch<2, true>(S, S.xgs, S.tgs);
      grid_out(S, S.tgs, bt == 2 ? t : nullptr, y, bt);
      return;
    }
    if (sync_free()) {
      const int* ch = lower ? S.lch : S.uch;
      const int nch = lower ? S.nlch : S.nuch;
      if (bt == 0) sf_launch<2, 0, false, LN>(S, D.n, nch, nl, ch, rows, T, b, S.dv, t, y);
      else if (bt == 1) sf_launch<2, 1, false, LN>(S, D.n, nch, nl, ch, rows, T, b, S.dv, t, y);
      else sf_launch<2, 2, false, LN>(S, D.n, nch, nl, ch, rows, T, b, S.dv, t, y);
      return;
    }
    constexpr int RUN_ROWS = 2048 / LN;
    auto nbl = [](int cnt) { return (int)(((int64_t)cnt * LN + BS - 1) / BS); };
    for (int lv = 0; lv < nl;) {
      int le = lv;
      while (le < nl && ph[le + 1] - ph[le] <= RUN_ROWS) ++le;
      if (le - lv >= 2) {
        run(lv, le);
        lv = le;
        continue;
      }
      const int cnt = ph[lv + 1] - ph[lv];
      const int* r = rows + ph[lv];
      if (bt == 0) trsv_level_kernel<V, 2, 0, false, LN><<<nbl(cnt), BS, 0, s_>>>(cnt, r, T.irp, T.ja, T.val, b, S.dv, t, y);
      else if (bt == 1) trsv_level_kernel<V, 2, 1, false, LN><<<nbl(cnt), BS, 0, s_>>>(cnt, r, T.irp, T.ja, T.val, b, S.dv, t, y);
      else trsv_level_kernel<V, 2, 2, false, LN><<<nbl(cnt), BS, 0, s_>>>(cnt, r, T.irp, T.ja, T.val, b, S.dv, t, y);
      ++lv;
    }
  }

  // One Jacobi sweep y_out = D^-1 (b - ND y_in) (FIRST: y_in = D^-1 b).
  // FIRST ('N') is the product with ND D; for 'T', FIRST on an ELL or DIA
  // operator gathers d_j*x_j directly, on SELL/CSR, whose gathers are the
  // costly part, d.*x is formed first in one streaming pass (same products,
  // so the same result) and gathered once.
  template <bool FIRST, int K, class E>
  void jac_launch(DLevel<V>& D, DSmoother<V>& S, const V* b, const V* yin, E e, char trans, Red red) {
    const DMat<V>& nd = (trans != 'N') ? S.ndt : S.nd;
    if (FIRST && trans == 'N' && S.ndd.val) {
      spmv_launch<K>(S.ndd, GatherX<V, V>{b}, e, s_, red);
    } else if (FIRST && (nd.fmt == FMT_ELL || nd.fmt == FMT_DIA)) {
      spmv_launch<K>(nd, GatherDX<V, V>{S.dv, b}, e, s_, red);
    } else if (FIRST) {
      vmul_kernel<V, V, false><<<nblk(D.n), BS, 0, s_>>>(D.n, S.dv, b, D.t2);
      spmv_launch<K>(nd, GatherX<V, V>{D.t2}, e, s_, red);
    } else {
      spmv_launch<K>(nd, GatherX<V, V>{yin}, e, s_, red);
    }
  }
  template <bool FIRST, bool ADD>
  void jac_sweep(DLevel<V>& D, DSmoother<V>& S, const V* b, const V* yin, V* yout, char trans) {
    jac_launch<FIRST, 0>(D, S, b, yin, EpiJac<V, V, V, ADD>{S.dv, b, yout}, trans, Red{nullptr, nullptr});
  }
  // Last post-smoothing sweep of the fine level (beta = 1): writes the
  // FP64 output of the cycle (or, keeping the FP32 result, y2l itself) and
  // the requested dot in the same pass.
  template <bool FIRST>
  void jac_sweep_out(DLevel<V>& D, DSmoother<V>& S, const V* b, const V* yin, V* y2l, char trans) {
    auto run = [&](auto e) {
      if (dot_u_) jac_launch<FIRST, 1>(D, S, b, yin, e, trans, red(dot_slot_));
      else jac_launch<FIRST, 0>(D, S, b, yin, e, trans, Red{nullptr, nullptr});
    };
    if (keep32_) run(EpiJacOut<V, V, false, V>{S.dv, b, y2l, y2l, dot_u_});
    else run(EpiJacOut<V, V>{S.dv, b, y2l, out64_, dot_u_});
    out_done_ = true;
  }
  bool fuse_out(DLevel<V>& D, int beta, V* y) const {
    return (out64_ != nullptr || keep32_) && &D == &L_[0] && beta == 1 && y == D.y2l;
  }

  // mld_d_jac_smoother_apply_vect (init 'Z'), alpha = 1, beta in {0,1}.
  int smooth(DLevel<V>& D, DSmoother<V>& S, const V* x, int beta, V* y, int sweeps, char trans) {
    const int n = D.n;
    if (S.smoother == NOPREC) {
      if (beta) add_kernel<V, V, true><<<nblk(n), BS, 0, s_>>>(n, x, y);
      else add_kernel<V, V, false><<<nblk(n), BS, 0, s_>>>(n, x, y);
      return 0;
    }
    if (sweeps < 0) return E_BAD_ARG;
    if (S.solver == GS || S.solver == BWGS) return gs_apply(D, S, x, beta, y, sweeps, trans);
    const bool jac = (S.solver == DIAG_SCALE || S.solver == L1_DIAG);
    const bool fo = jac && fuse_out(D, beta, y);
    if (sweeps <= 1 || S.nnz_nd == 0) {
      if (fo) {  // y = y + d*x, written as the FP64 output (or in place)
        const int nb = nblk(n);
        if (dot_u_) {
          if (keep32_) vmul_out_kernel<V, 1, V><<<nb, BS, 0, s_>>>(n, S.dv, x, y, y, dot_u_, red(dot_slot_));
          else vmul_out_kernel<V, 1><<<nb, BS, 0, s_>>>(n, S.dv, x, y, out64_, dot_u_, red(dot_slot_));
          finalize_kernel<1><<<1, FIN_T, 0, s_>>>(nb, partial_, scal_ + dot_slot_);
        } else if (keep32_) {
          vmul_out_kernel<V, 0, V><<<nb, BS, 0, s_>>>(n, S.dv, x, y, y, nullptr, Red{nullptr, nullptr});
        } else {
          vmul_out_kernel<V, 0><<<nb, BS, 0, s_>>>(n, S.dv, x, y, out64_, nullptr, Red{nullptr, nullptr});
        }
        HIPCHECK(hipGetLastError());
        out_done_ = true;
        return 0;
      }
      return solver_apply(D, S, x, beta, y, trans);
    }
    if (!jac) return E_UNSUPPORTED;
    if (sweeps == 2) {
      if (fo) jac_sweep_out<true>(D, S, x, nullptr, y, trans);
      else if (beta) jac_sweep<true, true>(D, S, x, nullptr, y, trans);
      else jac_sweep<true, false>(D, S, x, nullptr, y, trans);
      return 0;
    }
    V* ta = D.tx;
    V* tb = D.t1;
    jac_sweep<true, false>(D, S, x, nullptr, ta, trans);
    for (int i = 2; i <= sweeps - 2; ++i) {
      jac_sweep<false, false>(D, S, x, ta, tb, trans);
      std::swap(ta, tb);
    }
    if (fo) jac_sweep_out<false>(D, S, x, ta, y, trans);
    else if (beta) jac_sweep<false, true>(D, S, x, ta, y, trans);
    else jac_sweep<false, false>(D, S, x, ta, y, trans);
    return 0;
  }

  void residual(DLevel<V>& D, const V* b, const V* yv, V* r, char trans) {
    const DMat<V>& a = (trans != 'N') ? D.at : D.a;
    const bool probe = (probe_lev_ >= 0 && &D == &L_[probe_lev_] && probe_used_ < probe_ev_.size());
    if (probe) HIPCHECK(hipEventRecord(probe_ev_[probe_used_].first, s_));
    spmv_launch(a, GatherX<V, V>{yv}, EpiRes<V, V>{b, r}, s_);
    if (probe) HIPCHECK(hipEventRecord(probe_ev_[probe_used_++].second, s_));
  }
  // in-cycle timing of one level's residual launches (time_step kind 6)
  int probe_lev_ = -1;
  size_t probe_used_ = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> probe_ev_;

  // inner_ml_aply (mld_dmlprec_aply.f90:382-464): dispatch on the level's cycle
  int inner_ml_aply(int level, char trans) {
    if (level == tail_lev_ && trans == 'N') {
      tail_apply(L_[level]);
      return 0;
    }
    switch (L_[level].ml_cycle) {
      case ADD_ML: return inner_add(level, trans);
      case MULT_ML: case VCYCLE: case WCYCLE: return inner_mult(level, trans);
      case KCYCLE: case KCYCLESYM: return inner_k_cycle(level, trans);
      default: return E_UNSUPPORTED;
    }
  }

  // mld_d_inner_add (mld_dmlprec_aply.f90:467-588): smoothing of x2l itself
  // (the sm / sm2a pair max(pre, post) times from y2l = x2l where sm2a
  // exists, in that order also for 'T'), then y2l += P (cycle on R x2l).
  int inner_add(int level, char trans) {
    const int nlev = (int)L_.size();
    DLevel<V>& D = L_[level];
    int rc;
    if (D.has_sm2) {
      add_kernel<V, V, false><<<nblk(D.n), BS, 0, s_>>>(D.n, D.x2l, D.y2l);
      const int sw = std::max(D.sweeps_pre, D.sweeps_post);
      for (int k = 0; k < sw; ++k) {  // the smoothers' own scratch is tx, t1, t2
        rc = smooth(D, D.sm, D.y2l, 0, D.ty, 1, trans);
        if (rc) return rc;
        rc = smooth(D, D.sm2, D.ty, 0, D.y2l, 1, trans);
        if (rc) return rc;
      }
    } else {
      rc = smooth(D, D.sm, D.x2l, 0, D.y2l, D.sweeps_pre, trans);
      if (rc) return rc;
    }
    if (level < nlev - 1) {
      DLevel<V>& C = L_[level + 1];
      spmv_launch(C.r, GatherX<V, V>{D.x2l}, EpiSet<V, V>{C.x2l}, s_);
      rc = inner_ml_aply(level + 1, trans);
      if (rc) return rc;
      spmv_launch(C.p, GatherX<V, V>{C.y2l}, EpiAdd<V>{D.y2l}, s_);
    }
    return 0;
  }

  // mld_d_inner_k_cycle (mld_dmlprec_aply.f90:811-1014): pre- and
  // post-smoothing without the pre/post switches; where the next level is not
  // the coarsest, its correction comes from inneritkcycle.
  int inner_k_cycle(int level, char trans) {
    const int nlev = (int)L_.size();
    DLevel<V>& D = L_[level];
    if (level == nlev - 1) return smooth(D, D.sm, D.x2l, 0, D.y2l, D.sweeps_pre, trans);
    const bool tn = (trans == 'N');
    DSmoother<V>& s2 = D.has_sm2 ? D.sm2 : D.sm;
    int rc = tn ? smooth(D, D.sm, D.x2l, 0, D.y2l, D.sweeps_pre, trans)
                : smooth(D, s2, D.x2l, 0, D.y2l, D.sweeps_post, trans);
    if (rc) return rc;
    DLevel<V>& C = L_[level + 1];
    residual(D, D.x2l, D.y2l, D.ty, trans);
    spmv_launch(C.r, GatherX<V, V>{D.ty}, EpiSet<V, V>{C.x2l}, s_);
    if (level <= nlev - 3) rc = inneritkcycle(level + 1, trans, D.ml_cycle == KCYCLESYM);
    else rc = inner_ml_aply(level + 1, trans);
    if (rc) return rc;
    spmv_launch(C.p, GatherX<V, V>{C.y2l}, EpiAdd<V>{D.y2l}, s_);
    residual(D, D.x2l, D.y2l, D.ty, trans);
    return tn ? smooth(D, s2, D.ty, 1, D.y2l, D.sweeps_post, trans)
              : smooth(D, D.sm, D.ty, 1, D.y2l, D.sweeps_pre, trans);
  }

  // mld_dinneritkcycle (mld_dmlprec_aply.f90:1017-1184): at most two steps
  // of FCG (fcg) or GCR from zero on A_l y = x2l, preconditioned by this
  // level's cycle; one step when ||w|| <= 0.25 ||x2l||.  The dots are
  // reduced on the device and read by the host (the step count and the
  // coefficients are host decisions, as in the reference); the products
  // with A_l are untransposed for 'T' too (:1095, :1137).
  int inneritkcycle(int level, char trans, bool fcg) {
    DLevel<V>& D = L_[level];
    const int n = D.n;
    const int nb = nblk(n);
    V *w = D.kw, *v = D.kv, *v1 = D.kv1, *d0 = D.kd0, *d1 = D.kd1;
    if (!w) return E_UNSUPPORTED;
    add_kernel<V, V, false><<<nb, BS, 0, s_>>>(n, D.x2l, w);  // w = x2l
    fill_kernel<V><<<nb, BS, 0, s_>>>(n, D.y2l, (V)0);
    int rc = inner_ml_aply(level, trans);
    if (rc) return rc;
    // the scalar slots are shared with the inner iterations of deeper levels,
    // so every dot of this level is formed after its cycle returns
    ldot(n, w, w, nullptr, nullptr, S_K0);  // delta0^2
    add_kernel<V, V, false><<<nb, BS, 0, s_>>>(n, D.y2l, d0);
    spmv_launch(D.a, GatherX<V, V>{d0}, EpiSet<V, V>{v}, s_);
    if (fcg) ldot(n, d0, w, d0, v, S_K1);  // delta_old, tau
    else ldot(n, v, w, v, v, S_K1);
    read_scal(S_K0);
    const double delta0 = std::sqrt(hscal_[S_K0]);
    const double delta_old = hscal_[S_K1], tau = hscal_[S_K1 + 1];
    double alpha = delta_old / tau;
    axpby_kernel<V><<<nb, BS, 0, s_>>>(n, (V)(-alpha), v, (V)1, w);  // w -= alpha v
    ldot(n, w, w, nullptr, nullptr, S_K0);
    const double l2 = std::sqrt(read_scal(S_K0));
    if (l2 <= 0.25 * delta0) {
      axpby_kernel<V><<<nb, BS, 0, s_>>>(n, (V)alpha, d0, (V)0, D.y2l);  // y = alpha d0
    } else {
      add_kernel<V, V, false><<<nb, BS, 0, s_>>>(n, w, D.x2l);  // x2l = w
      rc = inner_ml_aply(level, trans);
      if (rc) return rc;
      add_kernel<V, V, false><<<nb, BS, 0, s_>>>(n, D.y2l, d1);
      spmv_launch(D.a, GatherX<V, V>{d1}, EpiSet<V, V>{v1}, s_);
      if (fcg) {
        ldot(n, d1, v, d1, v1, S_K1);  // tau1, tau2
        ldot(n, d1, w, nullptr, nullptr, S_K0);  // tau3
      } else {
        ldot(n, v1, v, v1, v1, S_K1);
        ldot(n, v1, w, nullptr, nullptr, S_K0);
      }
      read_scal(S_K0);
      const double tau1 = hscal_[S_K1], tau2 = hscal_[S_K1 + 1], tau3 = hscal_[S_K0];
      const double tau4 = tau2 - (tau1 * tau1) / tau;
      alpha = alpha - (tau1 * tau3) / (tau * tau4);
      const double alpha1 = tau3 / tau4;
      axpby_kernel<V><<<nb, BS, 0, s_>>>(n, (V)alpha, d0, (V)0, D.y2l);   // y = alpha d0
      axpby_kernel<V><<<nb, BS, 0, s_>>>(n, (V)alpha1, d1, (V)1, D.y2l);  // y += alpha1 d1
    }
    HIPCHECK(hipGetLastError());
    return 0;
  }
  // dot products of level vectors (one or two pairs) into scalar slots
  void ldot(int n, const V* a, const V* b, const V* c, const V* d, int slot) {
    const int nb = red_blocks(n);
    if (c) dot_kernel<V, 2, V><<<nb, BS, 0, s_>>>(n, a, b, c, d, red(slot));
    else dot_kernel<V, 1, V><<<nb, BS, 0, s_>>>(n, a, b, nullptr, nullptr, red(slot));
    finalize(c ? 2 : 1, slot, nb);
  }

  // mld_d_inner_mult (mld_dmlprec_aply.f90:590-809)
  int inner_mult(int level, char trans) {
    const int nlev = (int)L_.size();
    DLevel<V>& D = L_[level];
    const bool tn = (trans == 'N');
    const bool pre = (D.sweeps_pre > 0 && tn) || (D.sweeps_post > 0 && !tn);
    const bool post = (D.sweeps_post > 0 && tn) || (D.sweeps_pre > 0 && !tn);
    DSmoother<V>& Spre = (tn || !D.has_sm2) ? D.sm : D.sm2;
    DSmoother<V>& Spost = (tn && D.has_sm2) ? D.sm2 : D.sm;
    int rc;
    if (level < nlev - 1) {
      DLevel<V>& C = L_[level + 1];
      if (pre) {
        if (level == 0 && pre0_launched_) {
          pre0_launched_ = false;  // enqueued ahead by the Krylov loop, same operands
        } else {
          rc = smooth(D, Spre, D.x2l, 0, D.y2l, tn ? D.sweeps_pre : D.sweeps_post, trans);
          if (rc) return rc;
        }
        residual(D, D.x2l, D.y2l, D.ty, trans);
        spmv_launch(C.r, GatherX<V, V>{D.ty}, EpiSet<V, V>{C.x2l}, s_);
      } else {
        fill_kernel<V><<<nblk(D.n), BS, 0, s_>>>(D.n, D.y2l, (V)0);
        spmv_launch(C.r, GatherX<V, V>{D.x2l}, EpiSet<V, V>{C.x2l}, s_);
      }
      rc = inner_ml_aply(level + 1, trans);
      if (rc) return rc;
      spmv_launch(C.p, GatherX<V, V>{C.y2l}, EpiAdd<V>{D.y2l}, s_);
      if (D.ml_cycle == WCYCLE) {
        residual(D, D.x2l, D.y2l, D.ty, trans);
        spmv_launch(C.r, GatherX<V, V>{D.ty}, EpiSet<V, V>{C.x2l}, s_);
        rc = inner_ml_aply(level + 1, trans);
        if (rc) return rc;
        spmv_launch(C.p, GatherX<V, V>{C.y2l}, EpiAdd<V>{D.y2l}, s_);
      }
      if (post) {
        residual(D, D.x2l, D.y2l, D.ty, trans);
        rc = smooth(D, Spost, D.ty, 1, D.y2l, tn ? D.sweeps_post : D.sweeps_pre, trans);
        if (rc) return rc;
      }
      return 0;
    }
    return smooth(D, D.sm, D.x2l, 0, D.y2l, D.sweeps_pre, trans);
  }

  // outer_sweeps > 1 (mld_dmlprec_aply.f90:285-327), with internal buffers.
  int apply_outer(const double* x, double* y, char trans) {
    DLevel<V>& L0 = L_[0];
    const int n = L0.n;
    V* xin = L0.x2l;  // holds x converted to V for the whole call
    if (!xres_) xres_ = ar_.alloc<V>(n);  // allocated at upload
    convert_kernel<V, double><<<nblk(n), BS, 0, s_>>>(n, x, xres_);
    V* ysum = L0.t2;
    for (int isw = 1; isw <= outer_sweeps_; ++isw) {
      if (isw == 1) {
        add_kernel<V, V, false><<<nblk(n), BS, 0, s_>>>(n, xres_, xin);
      } else {
        // x2l = x - A y
        spmv_launch(L0.a, GatherX<V, V>{ysum}, EpiRes<V, V>{xres_, xin}, s_);
      }
      int rc = inner_ml_aply(0, trans);
      if (rc) return rc;
      if (isw == 1) add_kernel<V, V, false><<<nblk(n), BS, 0, s_>>>(n, L0.y2l, ysum);
      else add_kernel<V, V, true><<<nblk(n), BS, 0, s_>>>(n, L0.y2l, ysum);
    }
    convert_kernel<double, V><<<nblk(n), BS, 0, s_>>>(n, ysum, y);
    HIPCHECK(hipGetLastError());
    return 0;
  }
  V* xres_ = nullptr;

  // --------------------------------------------------------------- Krylov
  double* kvec(int i) {
    if (ws_) {  // caller's workspace (krylov_ws): nothing is allocated
      if (i >= ws_nvec_) throw DevError("Krylov workspace holds " + std::to_string(ws_nvec_) + " vectors");
      const size_t stride = ((size_t)n0_ + 31) / 32 * 32;  // every vector 256-byte aligned, as allocations are
      return ws_ + (size_t)i * stride;
    }
    while ((int)kvec_.size() <= i) kvec_.push_back(ar_.alloc<double>(n0_));
    return kvec_[i];
  }
  double* ws_ = nullptr;
  int ws_nvec_ = 0;

 public:
  // krylov ('N') with the work vectors r, z, p, q in ws (nvec vectors of n
  // doubles): CG and FCG, one outer sweep or more; allocates nothing on the
  // device (every other buffer of the path was allocated at upload).
  int krylov_ws(int method, const double* b, double* x, double eps, int itmax, int istop, int* iter, double* err,
                double* ws, int nvec) override {
    if (method != 0 && method != 1) return E_UNSUPPORTED;
    if (!ws || nvec < 4) return E_BAD_ARG;
    struct Reset {
      DevPrecT& p;
      ~Reset() {
        p.ws_ = nullptr;
        p.ws_nvec_ = 0;
      }
    } reset{*this};
    ws_ = ws;
    ws_nvec_ = nvec;
    return krylov(method, b, x, eps, itmax, istop, iter, err, 'N');
  }

 private:
  double read_scal(int slot) {
    HIPCHECK(hipMemcpyAsync(hscal_, scal_, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIPCHECK(hipStreamSynchronize(s_));
    check_abort();
    return hscal_[slot];
  }
  // a sync-free substitution that stopped waiting (see trsv_sf_kernel)
  void check_abort() {
    if (hscal_[S_ABORT] != 0.0)
      throw DevError("sync-free triangular solve: a dependency wait exceeded its bound; results are invalid");
  }
  void dot1(const double* a, const double* b, int slot) {
    dot_kernel<double, 1, double><<<red_blocks(n0_), BS, 0, s_>>>(n0_, a, b, nullptr, nullptr, red(slot));
    finalize(1, slot);
    HIPCHECK(hipGetLastError());
  }
  void axpby(double a, const double* x, double b, double* y) {
    axpby_kernel<double><<<nblk(n0_), BS, 0, s_>>>(n0_, a, x, b, y);
    HIPCHECK(hipGetLastError());
  }
};

// ---------------------------------------------------------------------------
// Krylov solvers, all vectors device-resident in FP64.  The preconditioner
// is applied through the same apply() the C-ABI exposes (the psb_krylov ->
// prec%apply call site).  Convergence test (istop = 2): ||r||_2/||b||_2 <= eps
// after each iteration; one 8-byte device-to-host read per iteration.
// ---------------------------------------------------------------------------
template <typename V>
int DevPrecT<V>::krylov(int method, const double* b, double* x, double eps, int itmax, int istop, int* iter,
                        double* err, char trans) {
  HIPCHECK(hipSetDevice(device_));
  if (istop != 2) return E_UNSUPPORTED;
  if (trans != 'N' && trans != 'T') return E_BAD_ARG;
  const int n = n0_;
  if (trans == 'T') {
    // adjoint system A^T x = b with M^T: transposed operators of every level
    // (uploaded once, reused by later adjoint solves)
    ensure_transposes();
    if (!akry_t_set_) {
      if constexpr (std::is_same<V, double>::value) {
        akry_t_ = L_[0].at;
      } else {
        akry_t_ = upload_mat<double>(host_transpose(hprec_->precv[0].a), ar_, s_, hprec_->dev_format);
      }
      akry_t_set_ = true;
    }
  }
    double* r = kvec(0);
  double* z = kvec(1);
  double* p = kvec(2);
  double* q = kvec(3);
  // bn2 and r = b - A x
  dot1(b, b, S_TMP0);
  kspmv<0>(trans, GatherX<double, double>{x}, EpiRes<double, double>{b, r});
  dot1(r, r, S_RR);
  HIPCHECK(hipMemcpyAsync(hscal_, scal_, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHECK(hipStreamSynchronize(s_));
  check_abort();
  double bn2 = std::sqrt(hscal_[S_TMP0]);
  if (bn2 == 0.0) bn2 = 1.0;
  *err = std::sqrt(hscal_[S_RR]) / bn2;
  int it = 0;
  if (*err <= eps) { *iter = 0; return 0; }
  const int nbd = red_blocks(n);
  // FP32 form, untransposed cycle with the plain fine-level input: the
  // residual update also writes the FP32 input of the next application
  float* r32 = nullptr;
  if constexpr (!std::is_same<V, double>::value) {
    if (trans == 'N' && outer_sweeps_ == 1 && L_.size() > 1) r32 = (float*)L_[0].x2l;
  }
  preconv_ = nullptr;
  pre0_launched_ = false;
  struct PreconvReset {
    const double*& p;
    bool& pre0;
    ~PreconvReset() {
      p = nullptr;
      pre0 = false;  // a sweep launched ahead for an application that did not follow
    }
  } preconv_reset{preconv_, pre0_launched_};
  // the deferred solution update x += (num/den) p of the last iteration
  // (slots; -1: none pending), applied by the next direction update or here
  int xnum = -1, xden = -1;
  auto flush_x = [&]() {
    if (xnum >= 0) x_update<<<nblk(n), BS, 0, s_>>>(n, x, p, scal_, xnum, xden);
    xnum = xden = -1;
  };
  if (method == 0) {  // CG
    while (it < itmax) {
      ++it;
      const int rs = (it & 1) ? S_RHO1 : S_RHO0;
      if (r32) {  // z = M^-1 r kept in FP32, rho = r.z
        const V* z32 = nullptr;
        int rc = apply_keep32(r, r, rs, &z32);
        if (rc) { flush_x(); return rc; }
        cg_pupdate<<<nblk(n), BS, 0, s_>>>(n, z32, p, scal_, it, x, xnum, xden);
      } else {
        int rc = apply_impl(r, z, trans, r, rs);  // z = M^-1 r, rho = r.z
        if (rc) { flush_x(); return rc; }
        cg_pupdate<<<nblk(n), BS, 0, s_>>>(n, (const double*)z, p, scal_, it, x, xnum, xden);
      }
      kspmv<1>(trans, GatherX<double, double>{p}, EpiSetDot<double>{q, p, nullptr}, red(S_SIGMA));
      r_update<<<nbd, BS, 0, s_>>>(n, r, q, scal_, rs, S_SIGMA, red(S_RR), r32);
      xnum = rs;
      xden = S_SIGMA;
      preconv_ = r32 ? r : nullptr;
      finalize(1, S_RR);
      HIPCHECK(hipGetLastError());
      if (it < itmax) launch_pre0(r, r32 ? nullptr : z, trans);
      read_scal(S_RR);
      *err = std::sqrt(hscal_[S_RR]) / bn2;
      const double rho_new = hscal_[(it & 1) ? S_RHO1 : S_RHO0], rho_old = hscal_[(it & 1) ? S_RHO0 : S_RHO1];
      if (*err <= eps) break;
      if (hscal_[S_SIGMA] == 0.0 || (it > 1 && rho_old == 0.0) || rho_new == 0.0) break;
    }
    flush_x();
  } else if (method == 1) {  // FCG(1)
    while (it < itmax) {
      ++it;
      // z = M^-1 r and, from the second iteration, gamma = q.z in the same
      // pass (an FP32 preconditioner's z stays in FP32)
      if (r32) {
        const V* z32 = nullptr;
        int rc = apply_keep32(r, it > 1 ? q : nullptr, S_GAMMA, &z32);
        if (rc) { flush_x(); return rc; }
        fcg_pupdate<<<nblk(n), BS, 0, s_>>>(n, z32, p, scal_, it, x, xnum, xden);
      } else {
        int rc = apply_impl(r, z, trans, it > 1 ? q : nullptr, S_GAMMA);
        if (rc) { flush_x(); return rc; }
        fcg_pupdate<<<nblk(n), BS, 0, s_>>>(n, (const double*)z, p, scal_, it, x, xnum, xden);
      }
      // q = A p with p.q (S_DELTA) and p.r (S_DPR) in one pass
      kspmv<2>(trans, GatherX<double, double>{p}, EpiSetDotUW<double>{q, p, r}, red(S_DELTA));
      r_update<<<nbd, BS, 0, s_>>>(n, r, q, scal_, S_DPR, S_DELTA, red(S_RR), r32);
      xnum = S_DPR;
      xden = S_DELTA;
      preconv_ = r32 ? r : nullptr;
      finalize(1, S_RR);
      HIPCHECK(hipGetLastError());
      if (it < itmax) launch_pre0(r, r32 ? nullptr : z, trans);
      read_scal(S_RR);
      *err = std::sqrt(hscal_[S_RR]) / bn2;
      if (*err <= eps) break;
      if (hscal_[S_DELTA] == 0.0) break;
    }
    flush_x();
  } else if (method == 2) {  // BiCGStab (right preconditioned)
    double* rt = kvec(4);
    double* s = kvec(5);
    double* t = kvec(6);
    double* ph = kvec(7);
    double* sh = kvec(8);
    double* v = q;
    HIPCHECK(hipMemcpyAsync(rt, r, n * sizeof(double), hipMemcpyDeviceToDevice, s_));
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    while (it < itmax) {
      ++it;
      dot1(rt, r, S_TMP0);
      const double rho = read_scal(S_TMP0);
      if (rho == 0.0) break;
      if (it == 1) {
        HIPCHECK(hipMemcpyAsync(p, r, n * sizeof(double), hipMemcpyDeviceToDevice, s_));
      } else {
        const double beta = (rho / rho_old) * (alpha / omega);
        axpby(-omega, v, 1.0, p);
        axpby(1.0, r, beta, p);
      }
      int rc = apply(p, ph, trans);
      if (rc) return rc;
      kspmv<1>(trans, GatherX<double, double>{ph}, EpiSetDot<double>{v, rt, nullptr}, red(S_TMP1));
      const double rtv = read_scal(S_TMP1);
      if (rtv == 0.0) break;
      alpha = rho / rtv;
      HIPCHECK(hipMemcpyAsync(s, r, n * sizeof(double), hipMemcpyDeviceToDevice, s_));
      axpby(-alpha, v, 1.0, s);
      dot1(s, s, S_TMP0);
      const double sn = std::sqrt(read_scal(S_TMP0)) / bn2;
      if (sn <= eps) {
        axpby(alpha, ph, 1.0, x);
        HIPCHECK(hipMemcpyAsync(r, s, n * sizeof(double), hipMemcpyDeviceToDevice, s_));
        *err = sn;
        break;
      }
      rc = apply(s, sh, trans);
      if (rc) return rc;
      kspmv<2>(trans, GatherX<double, double>{sh}, EpiSetDot<double>{t, nullptr, s}, red(S_TMP0));  // t.t, s.t
      read_scal(S_TMP0);
      const double tt = hscal_[S_TMP0], ts = hscal_[S_TMP1];
      omega = (tt == 0.0) ? 0.0 : ts / tt;
      axpby(alpha, ph, 1.0, x);
      axpby(omega, sh, 1.0, x);
      HIPCHECK(hipMemcpyAsync(r, s, n * sizeof(double), hipMemcpyDeviceToDevice, s_));
      axpby(-omega, t, 1.0, r);
      rho_old = rho;
      dot1(r, r, S_RR);
      *err = std::sqrt(read_scal(S_RR)) / bn2;
      if (*err <= eps) break;
      if (omega == 0.0) break;
    }
  } else {
    return E_BAD_ARG;
  }
  HIPCHECK(hipStreamSynchronize(s_));
  *iter = it;
  return 0;
}

// ---------------------------------------------------------------------------
// Workspace of mldhip_krylov_noalloc: the four FP64 Krylov vectors (r, z, p,
// q) of systems of up to NOALLOC_MAX_ROWS rows, 32 MiB, in a file-scope
// device array that every GPU holds until the process ends.  Its device
// address is taken once per device (with that device current) and only that
// address is used from the host.  Calls on one GPU share the array.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[NOALLOC_VECS * NOALLOC_MAX_ROWS];

double* noalloc_workspace(int device) {
  static std::mutex mu;
  static std::vector<double*> ptr;  // per device
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0) throw DevError("no device");
  if ((int)ptr.size() <= device) ptr.resize(device + 1, nullptr);
  if (!ptr[device]) {
    int cur = 0;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != device) HIPCHECK(hipSetDevice(device));
    void* a = nullptr;
    const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (cur != device) HIPCHECK(hipSetDevice(cur));
    if (e != hipSuccess || !a)
      throw DevError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptr[device] = (double*)a;
  }
  return ptr[device];
}

DevPrec* make_dev_prec(int value_kind, int device, hipStream_t stream) {
  if (value_kind == 1) return new DevPrecT<float>(device, stream);
  return new DevPrecT<double>(device, stream);
}

int dev_guard(const std::function<int()>& f, std::string& err) {
  try {
    return f();
  } catch (const std::exception& e) {
    err = e.what();
    return E_DEVICE;
  }
}


namespace {
struct HostStage {
  static constexpr size_t CH = (size_t)32 << 20;
  char* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};
std::mutex stage_mu;
std::vector<HostStage> stages;  // per device, kept until the process ends

HostStage& stage_of_current_device() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if ((int)stages.size() <= dev) stages.resize(dev + 1);
  HostStage& st = stages[dev];
  if (!st.buf[0]) {
    for (int b = 0; b < 2; ++b) {
      HIPCHECK(hipHostMalloc((void**)&st.buf[b], HostStage::CH, hipHostMallocDefault));
      HIPCHECK(hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming));
    }
  }
  return st;
}
}  // namespace

// Both staging buffers are idle between calls: a call waits for its own
// copies before it returns, so no call waits on an event recorded on another
// call's stream (which may have been destroyed or be capturing since).
void host_to_device(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < ((size_t)4 << 20)) {
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  std::lock_guard<std::mutex> lock(stage_mu);
  HostStage& st = stage_of_current_device();
  bool used[2] = {false, false};
  int b = 0;
  for (size_t off = 0; off < bytes; off += HostStage::CH, b ^= 1) {
    const size_t len = std::min(HostStage::CH, bytes - off);
    if (used[b]) HIPCHECK(hipEventSynchronize(st.ev[b]));  // the buffer's previous copy has completed
    par_memcpy(st.buf[b], (const char*)src + off, len);
    HIPCHECK(hipMemcpyAsync((char*)dst + off, st.buf[b], len, hipMemcpyHostToDevice, s));
    HIPCHECK(hipEventRecord(st.ev[b], s));
    used[b] = true;
  }
  for (int k = 0; k < 2; ++k)
    if (used[k]) HIPCHECK(hipEventSynchronize(st.ev[k]));
}

void device_to_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < ((size_t)4 << 20)) {
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    return;
  }
  std::lock_guard<std::mutex> lock(stage_mu);
  HostStage& st = stage_of_current_device();
  const size_t nch = (bytes + HostStage::CH - 1) / HostStage::CH;
  // every chunk's event is waited for in the drain loop below before the
  // buffer is reused or the call returns
  auto issue = [&](size_t k) {
    const int b = (int)(k & 1);
    const size_t off = k * HostStage::CH, len = std::min(HostStage::CH, bytes - off);
    HIPCHECK(hipMemcpyAsync(st.buf[b], (const char*)src + off, len, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipEventRecord(st.ev[b], s));
  };
  issue(0);
  if (nch > 1) issue(1);
  for (size_t k = 0; k < nch; ++k) {
    const int b = (int)(k & 1);
    const size_t off = k * HostStage::CH, len = std::min(HostStage::CH, bytes - off);
    HIPCHECK(hipEventSynchronize(st.ev[b]));
    par_memcpy((char*)dst + off, st.buf[b], len);
    if (k + 2 < nch) issue(k + 2);
  }
}

}  // namespace mld
