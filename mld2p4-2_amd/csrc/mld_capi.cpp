// extern "C" layer: declarations in include/mld2p4_amd.h.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mld2p4_amd.h"
#include "mld_device.hpp"
#include "mld_host.hpp"
#include "mld_multi.hpp"

struct mldhip_prec {
  mld::Prec host;
  int device = -1;
  int setup = -1;  // products of the hierarchy build: 0 host, 1 device, -1 device when there is one
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::unique_ptr<mld::DevPrec> dev;
  std::unique_ptr<mld::MultiEngine> blk;  // block (32-column) engine, built on first use
  std::string err;
  // device copies of the host-vector entry points' two vectors, kept from
  // call to call (grown when the system grows, freed with the handle)
  double* hv = nullptr;
  size_t hv_n = 0;
  double* host_vectors(size_t n) {
    if (hv_n < n) {
      if (hv) (void)hipFree(hv);
      hv = nullptr;
      hv_n = 0;
      if (hipMalloc(&hv, 2 * n * sizeof(double) + 16) != hipSuccess) {
        hv = nullptr;
        return nullptr;
      }
      hv_n = n;
    }
    return hv;
  }
};

using mld::E_BAD_ARG;
using mld::E_BAD_LEVEL;
using mld::E_DEVICE;
using mld::E_NOT_BUILT;

static std::string S(const char* c) { return c ? std::string(c) : std::string(); }

static int fail(mldhip_prec* p, int code, const std::string& msg) {
  if (p) p->err = msg;
  return code;
}

static const mld::HCsr* level_csr(mldhip_prec* p, int lev, int what) {
  if (!p || lev < 0 || lev >= (int)p->host.precv.size()) return nullptr;
  const mld::Level& L = p->host.precv[lev];
  switch (what) {
    case 0: return &L.a;
    case 1: return lev > 0 ? &L.p : nullptr;
    case 2: return lev > 0 ? &L.r : nullptr;
    case 3: return &L.smd.nd;
    case 4: return &L.smd.l;
    case 5: return &L.smd.u;
    case 13: return L.sm2a.allocated ? &L.smd2.nd : nullptr;
    case 14: return L.sm2a.allocated ? &L.smd2.l : nullptr;
    case 15: return L.sm2a.allocated ? &L.smd2.u : nullptr;
  }
  return nullptr;
}

extern "C" {

int mldhip_create(mldhip_prec** pp, int device) {
  if (!pp) return E_BAD_ARG;
  auto* p = new mldhip_prec;
  p->device = device;
  if (device >= 0) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev) {
      delete p;
      *pp = nullptr;
      return E_DEVICE;
    }
  }
  *pp = p;
  return 0;
}

int mldhip_destroy(mldhip_prec* p) {
  if (!p) return 0;
  if (p->device >= 0) (void)hipSetDevice(p->device);
  p->dev.reset();
  p->blk.reset();
  if (p->hv) (void)hipFree(p->hv);
  if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return 0;
}

int mldhip_init(mldhip_prec* p, const char* ptype) {
  if (!p) return E_BAD_ARG;
  p->dev.reset();
  p->blk.reset();
  int rc = mld::prec_init(p->host, S(ptype));
  if (rc) return fail(p, rc, p->host.last_error);
  return 0;
}

int mldhip_seti(mldhip_prec* p, const char* what, int val, int ilev, int ilmax, const char* pos) {
  if (!p) return E_BAD_ARG;
  return mld::prec_set_int(p->host, S(what), val, ilev, ilmax, S(pos));
}
int mldhip_setc(mldhip_prec* p, const char* what, const char* val, int ilev, int ilmax, const char* pos) {
  if (!p) return E_BAD_ARG;
  return mld::prec_set_str(p->host, S(what), S(val), ilev, ilmax, S(pos));
}
int mldhip_setr(mldhip_prec* p, const char* what, double val, int ilev, int ilmax, const char* pos) {
  if (!p) return E_BAD_ARG;
  return mld::prec_set_real(p->host, S(what), val, ilev, ilmax, S(pos));
}
int mldhip_set_precision(mldhip_prec* p, int kind) {
  if (!p || (kind != 0 && kind != 1)) return E_BAD_ARG;
  p->host.value_kind = kind;
  return 0;
}
int mldhip_set_format(mldhip_prec* p, int fmt) {
  if (!p || fmt < 0 || fmt > 10) return E_BAD_ARG;
  p->host.dev_format = fmt;
  return 0;
}

int mldhip_set_setup(mldhip_prec* p, int where) {
  if (!p || where < -1 || where > 1) return E_BAD_ARG;
  if (where == 1 && p->device < 0) return fail(p, E_BAD_ARG, "set_setup: a host-only handle has no device");
  p->setup = where;
  return 0;
}

int mldhip_hierarchy_build(mldhip_prec* p, int n, const int64_t* irp, const int* ja, const double* val) {
  if (!p || n < 0 || !irp || (n > 0 && (!ja || !val))) return E_BAD_ARG;
  p->dev.reset();
  p->blk.reset();
  {
    const char* e = std::getenv("MLD_SETUP");
    const bool host_env = e && std::string(e) == "host";
    const bool dev = p->device >= 0 && (p->setup == 1 || (p->setup == -1 && !host_env));
    p->host.setup_device = dev ? p->device : -1;
    // automatic: the device only where its fixed costs (allocations, sorts,
    // synchronisations) are repaid: products over a level of 2^22+ rows
    // (at 256^3 the first level's products 0.30 s on the device; below a few
    // million rows the host's threads are faster)
    p->host.setup_min_rows = (p->setup == 1) ? 0 : ((int64_t)1 << 22);
  }
  int rc = mld::hierarchy_build(p->host, n, irp, ja, val);
  if (rc) return fail(p, rc, p->host.last_error.empty() ? "hierarchy_build failed" : p->host.last_error);
  return 0;
}

// The device copy of a hierarchy whose smoother data is complete (built or
// uploaded): HBM upload once, the amold/vmold conversion of
// mld_d_base_onelev_cnv.f90:56-61.
static int upload_device(mldhip_prec* p) {
  if (p->device < 0) return 0;
  return mld::dev_guard(
      [&]() -> int {
        if (hipSetDevice(p->device) != hipSuccess) return E_DEVICE;
        if (!p->stream) {
          if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) return E_DEVICE;
          p->own_stream = true;
        }
        p->dev.reset(mld::make_dev_prec(p->host.value_kind, p->device, p->stream));
        return p->dev->upload(p->host);
      },
      p->err);
}

int mldhip_smoothers_build(mldhip_prec* p) {
  if (!p) return E_BAD_ARG;
  p->dev.reset();
  p->blk.reset();
  int rc = mld::smoothers_build(p->host);
  if (rc) return fail(p, rc, "smoothers_build failed");
  return upload_device(p);
}

int mldhip_upload_level(mldhip_prec* p, int lev, int n, const int64_t* irp, const int* ja, const double* val, int nf,
                        const int64_t* pirp, const int* pja, const double* pval, const int64_t* rirp, const int* rja,
                        const double* rval) {
  if (!p) return E_BAD_ARG;
  p->dev.reset();
  p->blk.reset();
  int rc = mld::stage_level(p->host, lev, n, irp, ja, val, nf, pirp, pja, pval, rirp, rja, rval);
  if (rc) return fail(p, rc, p->host.last_error);
  return 0;
}

int mldhip_upload_smoother(mldhip_prec* p, int lev, int which, int solver, const int64_t* ndirp, const int* ndja,
                           const double* ndval, const double* dv, const int64_t* lirp, const int* lja,
                           const double* lval, const int64_t* uirp, const int* uja, const double* uval) {
  if (!p) return E_BAD_ARG;
  int rc = mld::stage_smoother(p->host, lev, which, solver, ndirp, ndja, ndval, dv, lirp, lja, lval, uirp, uja, uval);
  if (rc) return fail(p, rc, p->host.last_error);
  return 0;
}

int mldhip_finalize_hierarchy(mldhip_prec* p, int nlev) {
  if (!p) return E_BAD_ARG;
  p->dev.reset();
  p->blk.reset();
  bool done = false;
  int rc = mld::finalize_hierarchy(p->host, nlev, done);
  if (rc) return fail(p, rc, p->host.last_error);
  return done ? upload_device(p) : 0;
}

int mldhip_load_hierarchy(mldhip_prec* p, const char* prefix, int n, const int64_t* irp, const int* ja,
                          const double* val, int smoothers) {
  if (!p || n < 0 || !irp) return E_BAD_ARG;
  p->dev.reset();
  p->blk.reset();
  bool done = false;
  std::string e;
  int rc = mld::load_hierarchy(p->host, S(prefix), n, irp, ja, val, smoothers != 0, done, e);
  if (rc) return fail(p, rc, e);
  return done ? upload_device(p) : 0;
}

int mldhip_build(mldhip_prec* p, int n, const int64_t* irp, const int* ja, const double* val) {
  int rc = mldhip_hierarchy_build(p, n, irp, ja, val);
  if (rc) return rc;
  return mldhip_smoothers_build(p);
}

static int need_dev(mldhip_prec* p) {
  if (!p) return E_BAD_ARG;
  if (!p->host.smoothers_built) return fail(p, E_NOT_BUILT, "preconditioner not built");
  if (!p->dev) return fail(p, E_DEVICE, "no device copy of the preconditioner");
  return 0;
}

int mldhip_apply(mldhip_prec* p, const double* x, double* y, char trans) {
  int rc = need_dev(p);
  if (rc) return rc;
  if (trans == 'n') trans = 'N';
  if (trans == 't' || trans == 'C' || trans == 'c') trans = 'T';
  if (trans != 'N' && trans != 'T') return fail(p, E_BAD_ARG, "invalid trans");
  return mld::dev_guard([&]() { return p->dev->apply(x, y, trans); }, p->err);
}

int mldhip_mlprec_aply(mldhip_prec* p, double alpha, const double* x, double beta, double* y, char trans) {
  int rc = need_dev(p);
  if (rc) return rc;
  if (trans == 'n') trans = 'N';
  if (trans == 't' || trans == 'C' || trans == 'c') trans = 'T';
  if (trans != 'N' && trans != 'T') return fail(p, E_BAD_ARG, "invalid trans");
  return mld::dev_guard([&]() { return p->dev->apply_ab(alpha, x, beta, y, trans); }, p->err);
}

int mldhip_apply_host(mldhip_prec* p, const double* x, double* y, char trans) {
  int rc = need_dev(p);
  if (rc) return rc;
  const size_t n = (size_t)p->host.precv[0].a.nr;
  return mld::dev_guard(
      [&]() -> int {
        double* dx = p->host_vectors(n);
        if (!dx) return E_DEVICE;
        double* dy = dx + n;
        int r = 0;
        mld::host_to_device(dx, x, n * sizeof(double), p->stream);
        r = mldhip_apply(p, dx, dy, trans);
        if (!r) mld::device_to_host(y, dy, n * sizeof(double), p->stream);
        if (hipStreamSynchronize(p->stream) != hipSuccess) r = E_DEVICE;
        return r;
      },
      p->err);
}

static int method_code(const char* m) {
  std::string s = S(m);
  for (auto& c : s) c = (char)toupper((unsigned char)c);
  if (s == "CG") return 0;
  if (s == "FCG") return 1;
  if (s == "BICGSTAB") return 2;
  return -1;
}

int mldhip_krylov(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                  int istop, int* iter, double* err) {
  int rc = need_dev(p);
  if (rc) return rc;
  int mc = method_code(method);
  if (mc < 0) return fail(p, mld::E_UNSUPPORTED, "Krylov method not provided: " + S(method));
  if (!iter || !err) return E_BAD_ARG;
  return mld::dev_guard([&]() { return p->dev->krylov(mc, b, x, eps, itmax, istop, iter, err, 'N'); }, p->err);
}

int mldhip_krylov_noalloc(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                          int istop, int* iter, double* err) {
  int rc = need_dev(p);
  if (rc) return rc;
  const int mc = method_code(method);
  if (mc != 0 && mc != 1) return fail(p, mld::E_UNSUPPORTED, "krylov_noalloc: method must be CG or FCG");
  if (!iter || !err) return E_BAD_ARG;
  const size_t n = (size_t)p->host.precv[0].a.nr;
  if (n > mld::NOALLOC_MAX_ROWS)
    return fail(p, E_BAD_ARG, "krylov_noalloc: " + std::to_string(n) + " rows exceed the workspace maximum of " +
                                  std::to_string(mld::NOALLOC_MAX_ROWS));
  return mld::dev_guard(
      [&]() {
        double* ws = mld::noalloc_workspace(p->device);
        return p->dev->krylov_ws(mc, b, x, eps, itmax, istop, iter, err, ws, mld::NOALLOC_VECS);
      },
      p->err);
}

int mldhip_krylov_adjoint(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                          int istop, int* iter, double* err) {
  int rc = need_dev(p);
  if (rc) return rc;
  int mc = method_code(method);
  if (mc < 0) return fail(p, mld::E_UNSUPPORTED, "Krylov method not provided: " + S(method));
  if (!iter || !err) return E_BAD_ARG;
  return mld::dev_guard([&]() { return p->dev->krylov(mc, b, x, eps, itmax, istop, iter, err, 'T'); }, p->err);
}

int mldhip_krylov_host(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                       int istop, int* iter, double* err) {
  int rc = need_dev(p);
  if (rc) return rc;
  const size_t n = (size_t)p->host.precv[0].a.nr;
  return mld::dev_guard(
      [&]() -> int {
        double* db = p->host_vectors(n);
        if (!db) return E_DEVICE;
        double* dx = db + n;
        int r = 0;
        mld::host_to_device(db, b, n * sizeof(double), p->stream);
        mld::host_to_device(dx, x, n * sizeof(double), p->stream);
        r = mldhip_krylov(p, method, db, dx, eps, itmax, istop, iter, err);
        if (!r) mld::device_to_host(x, dx, n * sizeof(double), p->stream);
        return r;
      },
      p->err);
}

int mldhip_set_stream(mldhip_prec* p, void* s) {
  if (!p) return E_BAD_ARG;
  if (p->own_stream && p->stream && p->stream != (hipStream_t)s) {
    (void)hipStreamSynchronize(p->stream);
    (void)hipStreamDestroy(p->stream);
  }
  p->stream = (hipStream_t)s;
  p->own_stream = false;
  if (p->dev) p->dev->set_stream(p->stream);
  return 0;
}

int mldhip_synchronize(mldhip_prec* p) {
  if (!p) return E_BAD_ARG;
  if (!p->dev) return 0;
  return mld::dev_guard([&]() { return p->dev->sync(); }, p->err);
}

int mldhip_get_nlev(mldhip_prec* p, int* nlev) {
  if (!p || !nlev) return E_BAD_ARG;
  *nlev = (int)p->host.precv.size();
  return 0;
}

int mldhip_get_level_info(mldhip_prec* p, int lev, int64_t* info) {
  if (!p || !info) return E_BAD_ARG;
  if (lev < 0 || lev >= (int)p->host.precv.size()) return E_BAD_LEVEL;
  const mld::Level& L = p->host.precv[lev];
  info[0] = L.a.nr;
  info[1] = L.a.nnz();
  info[2] = L.p.nnz();
  info[3] = L.r.nnz();
  info[4] = L.smd.nnz_nd;
  info[5] = L.naggr;
  info[6] = L.parms.sweeps_pre;
  info[7] = L.parms.sweeps_post;
  info[8] = L.sm.allocated ? L.sm.solver : -1;
  info[9] = L.sm2a.allocated ? L.sm2a.solver : -1;
  return 0;
}

int mldhip_get_csr_size(mldhip_prec* p, int lev, int what, int64_t* nr, int64_t* nc, int64_t* nnz) {
  const mld::HCsr* m = level_csr(p, lev, what);
  if (!m) return E_BAD_ARG;
  if (nr) *nr = m->nr;
  if (nc) *nc = m->nc;
  if (nnz) *nnz = m->nnz();
  return 0;
}

int mldhip_get_level_csr(mldhip_prec* p, int lev, int what, int64_t* irp, int* ja, double* val) {
  const mld::HCsr* m = level_csr(p, lev, what);
  if (!m) return E_BAD_ARG;
  if (irp) {
    if (m->irp.empty()) for (int i = 0; i <= m->nr; ++i) irp[i] = 0;
    else std::memcpy(irp, m->irp.data(), m->irp.size() * sizeof(int64_t));
  }
  if (ja && !m->ja.empty()) std::memcpy(ja, m->ja.data(), m->ja.size() * sizeof(int));
  if (val && !m->val.empty()) std::memcpy(val, m->val.data(), m->val.size() * sizeof(double));
  return 0;
}

int mldhip_get_aggr(mldhip_prec* p, int lev, int* ilaggr) {
  if (!p || !ilaggr) return E_BAD_ARG;
  if (lev < 1 || lev >= (int)p->host.precv.size()) return E_BAD_LEVEL;
  const auto& v = p->host.precv[lev].ilaggr;
  std::memcpy(ilaggr, v.data(), v.size() * sizeof(int));
  return 0;
}

int mldhip_get_level_dv(mldhip_prec* p, int lev, int which, double* dv) {
  if (!p || !dv) return E_BAD_ARG;
  if (lev < 0 || lev >= (int)p->host.precv.size()) return E_BAD_LEVEL;
  const mld::Level& L = p->host.precv[lev];
  const auto& v = (which == 1 && L.sm2a.allocated) ? L.smd2.dv : L.smd.dv;
  if (!v.empty()) std::memcpy(dv, v.data(), v.size() * sizeof(double));
  return 0;
}

int mldhip_get_tri_levels(mldhip_prec* p, int lev, int which, int* nl_lower, int* nl_upper) {
  if (!p || !nl_lower || !nl_upper) return E_BAD_ARG;
  if (lev < 0 || lev >= (int)p->host.precv.size()) return E_BAD_LEVEL;
  const mld::Level& L = p->host.precv[lev];
  const mld::SmootherData& sd = (which == 1 && L.sm2a.allocated) ? L.smd2 : L.smd;
  *nl_lower = sd.lsched_ptr.empty() ? 0 : (int)sd.lsched_ptr.size() - 1;
  *nl_upper = sd.usched_ptr.empty() ? 0 : (int)sd.usched_ptr.size() - 1;
  return 0;
}

int mldhip_get_omega(mldhip_prec* p, int lev, double* omega) {
  if (!p || !omega) return E_BAD_ARG;
  if (lev < 1 || lev >= (int)p->host.precv.size()) return E_BAD_LEVEL;
  *omega = p->host.precv[lev].parms.aggr_omega_val;
  return 0;
}

int mldhip_get_complexity(mldhip_prec* p, double* c) {
  if (!p || !c) return E_BAD_ARG;
  *c = mld::operator_complexity(p->host);
  return 0;
}

int mldhip_dump(mldhip_prec* p, int istart, int iend, const char* prefix, const char* head, int flags) {
  if (!p) return E_BAD_ARG;
  if (!p->host.hier_built) return fail(p, 3112, "dump: hierarchy not built");
  std::string err;
  int rc = mld::dump_prec(p->host, istart, iend, prefix ? prefix : "", head ? head : "", flags, err);
  if (rc) return fail(p, rc, err);
  return 0;
}

int mldhip_get_dense_tail(mldhip_prec* p, int* lev) {
  if (!p || !lev) return E_BAD_ARG;
  *lev = p->dev ? p->dev->tail_level() : -1;
  return 0;
}

int mldhip_get_device_bytes(mldhip_prec* p, int64_t* bytes) {
  if (!p || !bytes) return E_BAD_ARG;
  *bytes = p->dev ? (int64_t)p->dev->device_bytes() : 0;
  return 0;
}

int mldhip_get_error(mldhip_prec* p, char* buf, int len) {
  if (!p || !buf || len <= 0) return E_BAD_ARG;
  std::strncpy(buf, p->err.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return 0;
}

int mldhip_level_spmv(mldhip_prec* p, int lev, int what, const double* x, double* y) {
  int rc = need_dev(p);
  if (rc) return rc;
  return mld::dev_guard([&]() { return p->dev->spmv_level(lev, what, x, y); }, p->err);
}

int mldhip_level_smooth(mldhip_prec* p, int lev, int which, const double* x, int beta, double* y, int sweeps,
                        char trans) {
  int rc = need_dev(p);
  if (rc) return rc;
  if (beta != 0 && beta != 1) return E_BAD_ARG;
  return mld::dev_guard([&]() { return p->dev->smooth_level(lev, x, beta, y, sweeps, trans, which); }, p->err);
}

int mldhip_time_step(mldhip_prec* p, int lev, int kind, int reps, double* avg_ms) {
  int rc = need_dev(p);
  if (rc) return rc;
  if (!avg_ms) return E_BAD_ARG;
  return mld::dev_guard([&]() { return p->dev->time_step(lev, kind, reps, avg_ms); }, p->err);
}

int mldhip_get_op_format(mldhip_prec* p, int lev, int what, int* fmt, int* lanes, int* width) {
  int rc = need_dev(p);
  if (rc) return rc;
  if (!fmt || !lanes || !width) return E_BAD_ARG;
  return p->dev->op_format(lev, what, fmt, lanes, width);
}

// ---- batched engines --------------------------------------------------------
static int ensure_block(mldhip_prec* p) {
  int rc = need_dev(p);
  if (rc) return rc;
  if (p->blk) return 0;
  std::vector<const mld::Prec*> sys{&p->host};
  std::string e;
  p->blk.reset(mld::make_multi(sys, p->host.value_kind, 32, p->device, rc, e));
  if (!p->blk) return fail(p, rc ? rc : E_DEVICE, e);
  return 0;
}

int mldhip_apply_block(mldhip_prec* p, int nrhs, const double* x, int64_t ldx, double* y, int64_t ldy) {
  int rc = ensure_block(p);
  if (rc) return rc;
  if (!x || !y) return E_BAD_ARG;
  return mld::dev_guard([&]() { return p->blk->apply_cm(nrhs, x, ldx, y, ldy, p->stream); }, p->err);
}

int mldhip_krylov_block(mldhip_prec* p, const char* method, int nrhs, const double* b, int64_t ldb, double* x,
                        int64_t ldx, double eps, int itmax, int istop, int* iters, double* errs) {
  int rc = ensure_block(p);
  if (rc) return rc;
  int mc = method_code(method);
  if (mc < 0 || mc > 1) return fail(p, mld::E_UNSUPPORTED, "block Krylov method not provided: " + S(method));
  if (istop != 2) return fail(p, mld::E_UNSUPPORTED, "istop must be 2");
  if (!b || !x || !iters || !errs) return E_BAD_ARG;
  return mld::dev_guard(
      [&]() { return p->blk->krylov_cm(mc, nrhs, b, ldb, x, ldx, eps, itmax, iters, errs, nullptr, p->stream); },
      p->err);
}

}  // extern "C"

struct mldhip_group {
  int device = 0;
  int nb = 1;
  hipStream_t stream = nullptr;
  std::vector<mldhip_prec*> sys;
  std::unique_ptr<mld::MultiEngine> eng;
  std::string err;
};

static int gfail(mldhip_group* g, int code, const std::string& msg) {
  if (g) g->err = msg;
  return code;
}

extern "C" {

int mldhip_group_create(mldhip_group** gg, int device, int nb) {
  if (!gg || (nb != 1 && nb != 32) || device < 0) return E_BAD_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev) return E_DEVICE;
  auto* g = new mldhip_group;
  g->device = device;
  g->nb = nb;
  *gg = g;
  return 0;
}

int mldhip_group_destroy(mldhip_group* g) {
  if (!g) return 0;
  (void)hipSetDevice(g->device);
  g->eng.reset();
  delete g;
  return 0;
}

int mldhip_group_add(mldhip_group* g, mldhip_prec* p) {
  if (!g || !p) return E_BAD_ARG;
  if (!p->host.smoothers_built) return gfail(g, E_NOT_BUILT, "system not built");
  g->eng.reset();
  g->sys.push_back(p);
  return 0;
}

int mldhip_group_build(mldhip_group* g, int precision) {
  if (!g || (precision != 0 && precision != 1)) return E_BAD_ARG;
  std::vector<const mld::Prec*> sys;
  for (auto* p : g->sys) sys.push_back(&p->host);
  return mld::dev_guard(
      [&]() -> int {
        if (hipSetDevice(g->device) != hipSuccess) return E_DEVICE;
        int rc = 0;
        std::string e;
        g->eng.reset(mld::make_multi(sys, precision, g->nb, g->device, rc, e));
        if (!g->eng) return gfail(g, rc ? rc : E_DEVICE, e);
        return 0;
      },
      g->err);
}

int mldhip_group_info(mldhip_group* g, int* nsys, int64_t* nrows, int64_t* offsets) {
  if (!g) return E_BAD_ARG;
  if (!g->eng) return gfail(g, E_NOT_BUILT, "group not built");
  if (nsys) *nsys = g->eng->nsys();
  if (nrows) *nrows = g->eng->rows();
  if (offsets) {
    const auto& o = g->eng->offsets();
    for (size_t i = 0; i < o.size(); ++i) offsets[i] = o[i];
  }
  return 0;
}

int mldhip_group_set_stream(mldhip_group* g, void* stream) {
  if (!g) return E_BAD_ARG;
  g->stream = (hipStream_t)stream;
  return 0;
}

int mldhip_group_apply(mldhip_group* g, int nrhs, const double* x, int64_t ldx, double* y, int64_t ldy) {
  if (!g || !x || !y) return E_BAD_ARG;
  if (!g->eng) return gfail(g, E_NOT_BUILT, "group not built");
  return mld::dev_guard([&]() { return g->eng->apply_cm(nrhs, x, ldx, y, ldy, g->stream); }, g->err);
}

int mldhip_group_krylov(mldhip_group* g, const char* method, int nrhs, const double* b, int64_t ldb, double* x,
                        int64_t ldx, double eps, int itmax, int istop, int* iters, double* errs, int* launches) {
  if (!g || !b || !x || !iters || !errs) return E_BAD_ARG;
  if (!g->eng) return gfail(g, E_NOT_BUILT, "group not built");
  int mc = method_code(method);
  if (mc < 0 || mc > 1) return gfail(g, mld::E_UNSUPPORTED, "group Krylov method not provided: " + S(method));
  if (istop != 2) return gfail(g, mld::E_UNSUPPORTED, "istop must be 2");
  return mld::dev_guard(
      [&]() { return g->eng->krylov_cm(mc, nrhs, b, ldb, x, ldx, eps, itmax, iters, errs, launches, g->stream); },
      g->err);
}

int mldhip_group_time_apply(mldhip_group* g, int reps, double* avg_ms) {
  if (!g || !avg_ms || reps <= 0) return E_BAD_ARG;
  if (!g->eng) return gfail(g, E_NOT_BUILT, "group not built");
  return mld::dev_guard([&]() { return g->eng->time_apply(reps, avg_ms); }, g->err);
}

int mldhip_group_level_info(mldhip_group* g, int lev, int64_t* info) {
  if (!g || !info) return E_BAD_ARG;
  if (!g->eng) return gfail(g, E_NOT_BUILT, "group not built");
  return g->eng->level_info(lev, info);
}

int mldhip_group_get_error(mldhip_group* g, char* buf, int len) {
  if (!g || !buf || len <= 0) return E_BAD_ARG;
  std::strncpy(buf, g->err.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return 0;
}

}  // extern "C"
