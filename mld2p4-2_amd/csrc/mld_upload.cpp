// A hierarchy built outside the engine, handed over level by level.
//
// The reference builds its hierarchy in mld_d_hierarchy_bld (aggregation by
// mld_aggrmap_bld, operators by mld_aggrmat_asb) and its smoothers in
// mld_d_smoothers_bld; after that, each precv(l) holds base_a, the map
// (restriction R and prolongation P) and the smoother objects sm / sm2a
// (mlprec/mld_d_onelev_mod.f90:122-158).  The functions here take exactly
// that data, so that the output of the reference's own setup (or of a
// prec%dump, mlprec/impl/level/mld_d_base_onelev_dump.f90:92-121) can be
// applied by the device engine:
//   stage_level       base_a, map%p (P), map%r (R) of one level
//   stage_smoother    ND and the solver data of sm or sm2a of one level
//   finalize_hierarchy  the staged levels become precv, with the configured
//                     parameters distributed as mld_d_hierarchy_bld does
//   load_hierarchy    the same from the MatrixMarket files of a dump
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mld_host.hpp"

namespace mld {
namespace {

// Copy of a caller's 0-based CSR, checked: row pointers from 0 and
// non-decreasing, columns in [0, nc) and strictly ascending in each row.
int copy_csr(HCsr& m, int nr, int nc, const int64_t* irp, const int* ja, const double* val, std::string& err,
             const char* what) {
  m.clear();
  m.nr = nr;
  m.nc = nc;
  if (nr < 0 || nc < 0 || !irp) { err = std::string(what) + ": missing or invalid CSR"; return E_BAD_ARG; }
  if (irp[0] != 0) { err = std::string(what) + ": row pointers must start at 0"; return E_BAD_ARG; }
  const int64_t nz = irp[nr];
  if (nz < 0 || nz >= ((int64_t)1 << 31) || (nz > 0 && (!ja || !val))) {
    err = std::string(what) + ": invalid number of entries";
    return E_BAD_ARG;
  }
  m.irp.assign(irp, irp + nr + 1);
  m.ja.assign(ja, ja + nz);
  m.val.assign(val, val + nz);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < nr; ++i) {
    if (irp[i + 1] < irp[i]) { bad |= 1; continue; }
    for (int64_t k = irp[i]; k < irp[i + 1]; ++k) {
      if (ja[k] < 0 || ja[k] >= nc) bad |= 2;
      if (k > irp[i] && ja[k] <= ja[k - 1]) bad |= 4;
    }
  }
  if (bad & 1) { err = std::string(what) + ": row pointers decrease"; return E_BAD_ARG; }
  if (bad & 2) { err = std::string(what) + ": column index out of range"; return E_BAD_ARG; }
  if (bad & 4) { err = std::string(what) + ": columns must be strictly ascending in each row"; return E_BAD_ARG; }
  return 0;
}

bool has_diagonal_only_outside(const HCsr& t, bool lower) {
  for (int i = 0; i < t.nr; ++i)
    for (int64_t k = t.irp[i]; k < t.irp[i + 1]; ++k)
      if (lower ? t.ja[k] >= i : t.ja[k] <= i) return false;
  return true;
}

}  // namespace

int stage_level(Prec& p, int lev, int n, const int64_t* irp, const int* ja, const double* val, int nf,
                const int64_t* pirp, const int* pja, const double* pval, const int64_t* rirp, const int* rja,
                const double* rval) {
  if (!p.initialized) return E_NOT_INIT;
  if (lev < 0 || lev >= 64) { p.last_error = "upload_level: invalid level"; return E_BAD_LEVEL; }
  if ((int)p.staged.size() <= lev) p.staged.resize(lev + 1);
  StagedLevel& S = p.staged[lev];
  S = StagedLevel();
  std::string err;
  int rc = copy_csr(S.a, n, n, irp, ja, val, err, "A");
  if (!rc && lev > 0) {
    if (nf <= 0) { p.last_error = "upload_level: P and R need the finer level's size"; return E_BAD_ARG; }
    rc = copy_csr(S.p, nf, n, pirp, pja, pval, err, "P");
    if (!rc) rc = copy_csr(S.r, n, nf, rirp, rja, rval, err, "R");
  }
  if (rc) {
    S = StagedLevel();
    p.last_error = "upload_level " + std::to_string(lev) + ": " + err;
    return rc;
  }
  S.has_a = true;
  p.hier_built = p.smoothers_built = false;
  return 0;
}

int stage_smoother(Prec& p, int lev, int which, int solver, const int64_t* ndirp, const int* ndja,
                   const double* ndval, const double* dv, const int64_t* lirp, const int* lja, const double* lval,
                   const int64_t* uirp, const int* uja, const double* uval) {
  if (!p.initialized) return E_NOT_INIT;
  if (lev < 0 || lev >= (int)p.staged.size() || !p.staged[lev].has_a) {
    p.last_error = "upload_smoother: the level's matrix must be uploaded first";
    return E_BAD_LEVEL;
  }
  if (which != 0 && which != 1) return E_BAD_ARG;
  StagedLevel& S = p.staged[lev];
  const int n = S.a.nr;
  SmootherData sd;
  sd.smoother = JAC;
  sd.solver = solver;
  std::string err;
  int rc = 0;
  sd.nd.nr = sd.nd.nc = n;
  sd.nd.irp.assign((size_t)n + 1, 0);
  auto need_dv = [&]() -> int {
    if (!dv) { err = "the solver's diagonal data is missing"; return E_BAD_ARG; }
    sd.dv.assign(dv, dv + n);
    return 0;
  };
  switch (solver) {
    case F_NONE:
      break;
    case DIAG_SCALE:
    case L1_DIAG:
      rc = need_dv();
      if (!rc && ndirp) rc = copy_csr(sd.nd, n, n, ndirp, ndja, ndval, err, "ND");
      break;
    case ILU_N:
    case GS:
    case BWGS:
      rc = need_dv();
      if (!rc && ndirp) rc = copy_csr(sd.nd, n, n, ndirp, ndja, ndval, err, "ND");
      if (!rc) rc = copy_csr(sd.l, n, n, lirp, lja, lval, err, "L");
      if (!rc) rc = copy_csr(sd.u, n, n, uirp, uja, uval, err, "U");
      if (!rc && !(has_diagonal_only_outside(sd.l, true) && has_diagonal_only_outside(sd.u, false))) {
        err = "L and U must be the strict lower and upper triangles";
        rc = E_BAD_ARG;
      }
      if (!rc) {
        level_schedule(sd.l, true, sd.lsched_ptr, sd.lsched_rows);
        level_schedule(sd.u, false, sd.usched_ptr, sd.usched_rows);
      }
      break;
    default:
      err = "solver " + std::to_string(solver) + " cannot be uploaded";
      rc = E_UNSUPPORTED;
  }
  if (rc) {
    p.last_error = "upload_smoother " + std::to_string(lev) + ": " + err;
    return rc;
  }
  sd.nnz_nd = sd.nd.nnz();
  if (sd.nnz_nd > 0 && (solver == ILU_N || solver == GS || solver == BWGS)) {
    // the block solvers' ND holds the columns beyond the local rows, which a
    // single device does not have
    p.last_error = "upload_smoother: a block solver's ND must be empty on one device";
    return E_UNSUPPORTED;
  }
  S.sd[which] = std::move(sd);
  S.has_sm[which] = true;
  p.hier_built = p.smoothers_built = false;
  return 0;
}

int finalize_hierarchy(Prec& p, int nlev, bool& smoothers_done) {
  smoothers_done = false;
  if (!p.initialized || p.precv.empty()) return E_NOT_INIT;
  if (nlev < 1 || nlev > (int)p.staged.size()) {
    p.last_error = "finalize_hierarchy: " + std::to_string(nlev) + " levels, " + std::to_string(p.staged.size()) +
                   " uploaded";
    return E_BAD_ARG;
  }
  for (int l = 0; l < nlev; ++l) {
    const StagedLevel& S = p.staged[l];
    if (!S.has_a) { p.last_error = "finalize_hierarchy: level " + std::to_string(l) + " not uploaded"; return E_BAD_ARG; }
    if (l > 0) {
      const int nf = p.staged[l - 1].a.nr;
      if (S.p.nr != nf || S.r.nc != nf) {
        p.last_error = "finalize_hierarchy: P/R of level " + std::to_string(l) + " do not match level " +
                       std::to_string(l - 1);
        return E_BAD_ARG;
      }
    }
  }
  int nsm = 0;
  for (int l = 0; l < nlev; ++l) nsm += p.staged[l].has_sm[0] ? 1 : 0;
  if (nsm != 0 && nsm != nlev) {
    p.last_error = "finalize_hierarchy: smoother data uploaded for some levels only";
    return E_BAD_ARG;
  }
  // the configured parameters and smoothers for nlev levels (a one-level
  // configuration stays one level)
  std::vector<Level> v = p.precv;
  if ((int)v.size() == 1 && nlev > 1) {
    p.last_error = "finalize_hierarchy: a one-level preconditioner takes one level";
    return E_BAD_ARG;
  }
  if (nlev == 1) v.resize(1);
  else resize_levels(v, nlev);
  for (int l = 0; l < nlev; ++l) {
    StagedLevel& S = p.staged[l];
    Level& L = v[l];
    L.a = std::move(S.a);
    L.p = std::move(S.p);
    L.r = std::move(S.r);
    L.ilaggr.clear();
    L.naggr = l > 0 ? L.a.nr : 0;
    if (nsm) {
      L.sm.allocated = true;
      L.sm.smoother = JAC;
      L.sm.solver = S.sd[0].solver;
      L.smd = std::move(S.sd[0]);
      L.sm2a.allocated = S.has_sm[1];
      if (S.has_sm[1]) {
        L.sm2a.smoother = JAC;
        L.sm2a.solver = S.sd[1].solver;
        L.smd2 = std::move(S.sd[1]);
      }
    }
  }
  p.precv = std::move(v);
  p.staged.clear();
  p.hier_built = true;
  p.smoothers_built = nsm > 0;
  smoothers_done = nsm > 0;
  return 0;
}

// ---------------------------------------------------------------------------
// MatrixMarket reading (the files mld_io.cpp writes: coordinate real general
// with 1-based indices, array real general for vectors).
// ---------------------------------------------------------------------------
namespace {

struct MMFile {
  std::string text;
  size_t pos = 0;
  bool coord = false;
  int64_t nr = 0, nc = 0, nz = 0;
};

bool mm_open(const std::string& fname, MMFile& f, std::string& err) {
  FILE* fp = std::fopen(fname.c_str(), "rb");
  if (!fp) { err = "cannot open " + fname; return false; }
  std::fseek(fp, 0, SEEK_END);
  const long len = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  f.text.resize(len > 0 ? (size_t)len : 0);
  const size_t got = len > 0 ? std::fread(&f.text[0], 1, (size_t)len, fp) : 0;
  std::fclose(fp);
  if (got != f.text.size()) { err = "cannot read " + fname; return false; }
  // header
  size_t e = f.text.find('\n');
  std::string head = f.text.substr(0, e);
  for (auto& c : head) c = (char)std::tolower((unsigned char)c);
  if (head.rfind("%%matrixmarket", 0) != 0 || head.find("real") == std::string::npos ||
      head.find("general") == std::string::npos) {
    err = fname + ": not a real general MatrixMarket file";
    return false;
  }
  f.coord = head.find("coordinate") != std::string::npos;
  f.pos = (e == std::string::npos) ? f.text.size() : e + 1;
  while (f.pos < f.text.size() && f.text[f.pos] == '%') {  // comment lines
    e = f.text.find('\n', f.pos);
    f.pos = (e == std::string::npos) ? f.text.size() : e + 1;
  }
  const char* s = f.text.c_str() + f.pos;
  char* q = nullptr;
  f.nr = std::strtoll(s, &q, 10);
  f.nc = std::strtoll(q, &q, 10);
  if (f.coord) f.nz = std::strtoll(q, &q, 10);
  f.pos = (size_t)(q - f.text.c_str());
  if (f.nr < 0 || f.nc < 0 || f.nz < 0) { err = fname + ": invalid size line"; return false; }
  return true;
}

bool read_mm_csr(const std::string& fname, HCsr& m, std::string& err) {
  MMFile f;
  if (!mm_open(fname, f, err)) return false;
  if (!f.coord) { err = fname + ": coordinate format expected"; return false; }
  std::vector<int> ri((size_t)f.nz), ci((size_t)f.nz);
  std::vector<double> v((size_t)f.nz);
  const char* s = f.text.c_str() + f.pos;
  char* q = nullptr;
  for (int64_t k = 0; k < f.nz; ++k) {
    const long long i = std::strtoll(s, &q, 10);
    const long long j = std::strtoll(q, &q, 10);
    v[k] = std::strtod(q, &q);
    s = q;
    if (i < 1 || i > f.nr || j < 1 || j > f.nc) { err = fname + ": entry out of range"; return false; }
    ri[k] = (int)(i - 1);
    ci[k] = (int)(j - 1);
  }
  m.clear();
  m.nr = (int)f.nr;
  m.nc = (int)f.nc;
  m.irp.assign((size_t)m.nr + 1, 0);
  for (int64_t k = 0; k < f.nz; ++k) ++m.irp[ri[k] + 1];
  for (int i = 0; i < m.nr; ++i) m.irp[i + 1] += m.irp[i];
  m.ja.resize((size_t)f.nz);
  m.val.resize((size_t)f.nz);
  std::vector<int64_t> pos(m.irp.begin(), m.irp.end() - 1);
  for (int64_t k = 0; k < f.nz; ++k) {
    const int64_t q2 = pos[ri[k]]++;
    m.ja[q2] = ci[k];
    m.val[q2] = v[k];
  }
  // columns ascending in each row (the file keeps the stored order; other
  // writers may not), duplicates refused
  for (int i = 0; i < m.nr; ++i) {
    const int64_t b = m.irp[i], e = m.irp[i + 1];
    bool sorted = true;
    for (int64_t k = b + 1; k < e; ++k)
      if (m.ja[k] <= m.ja[k - 1]) sorted = false;
    if (sorted) continue;
    std::vector<std::pair<int, double>> row;
    for (int64_t k = b; k < e; ++k) row.push_back({m.ja[k], m.val[k]});
    std::stable_sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    for (int64_t k = b; k < e; ++k) {
      m.ja[k] = row[k - b].first;
      m.val[k] = row[k - b].second;
      if (k > b && m.ja[k] == m.ja[k - 1]) { err = fname + ": duplicate entry"; return false; }
    }
  }
  return true;
}

bool read_mm_vec(const std::string& fname, std::vector<double>& v, std::string& err) {
  MMFile f;
  if (!mm_open(fname, f, err)) return false;
  if (f.coord || f.nc != 1) { err = fname + ": array format with one column expected"; return false; }
  v.resize((size_t)f.nr);
  const char* s = f.text.c_str() + f.pos;
  char* q = nullptr;
  for (int64_t k = 0; k < f.nr; ++k) {
    v[k] = std::strtod(s, &q);
    if (q == s) { err = fname + ": too few values"; return false; }
    s = q;
  }
  return true;
}

bool file_exists(const std::string& fname) {
  FILE* fp = std::fopen(fname.c_str(), "rb");
  if (!fp) return false;
  std::fclose(fp);
  return true;
}

std::string lev_file(const std::string& prefix, int lev, const char* what) {
  char tail[64];
  std::snprintf(tail, sizeof tail, "_p000_l%03d_%s.mtx", lev, what);
  return prefix + tail;
}

// the strict triangle of m (lower or upper) and its diagonal
void split_diag(const HCsr& m, bool lower, HCsr& t, std::vector<double>& d) {
  t.clear();
  t.nr = m.nr;
  t.nc = m.nc;
  t.irp.assign((size_t)m.nr + 1, 0);
  d.assign((size_t)m.nr, 0.0);
  for (int i = 0; i < m.nr; ++i) {
    for (int64_t k = m.irp[i]; k < m.irp[i + 1]; ++k) {
      if (m.ja[k] == i) d[i] = m.val[k];
      else if (lower ? m.ja[k] < i : m.ja[k] > i) { t.ja.push_back(m.ja[k]); t.val.push_back(m.val[k]); }
    }
    t.irp[i + 1] = (int64_t)t.ja.size();
  }
}

// Smoother files of one level (mld_d_jac_smoother_dmp + the solver's dump,
// as mld_io.cpp writes them) staged for slot `which` with the configured
// solver `solver`.
int load_smoother(Prec& p, const std::string& prefix, int lev0, int which, int solver, std::string& err) {
  const int lev = lev0 + 1;
  const int n = p.staged[lev0].a.nr;
  HCsr nd, l, u, m;
  std::vector<double> dv;
  const std::string fnd = lev_file(prefix, lev, "nd");
  bool have_nd = false;
  if (solver == DIAG_SCALE || solver == L1_DIAG) {
    if (!read_mm_csr(fnd, nd, err)) return E_BAD_ARG;
    have_nd = nd.nnz() > 0;
    if (!read_mm_vec(lev_file(prefix, lev, "diag"), dv, err)) return E_BAD_ARG;
  } else if (solver == ILU_N) {
    if (!read_mm_csr(lev_file(prefix, lev, "lower"), l, err)) return E_BAD_ARG;
    if (!read_mm_vec(lev_file(prefix, lev, "diag"), dv, err)) return E_BAD_ARG;
    if (!read_mm_csr(lev_file(prefix, lev, "upper"), u, err)) return E_BAD_ARG;
  } else if (solver == GS) {  // lower = tril(A), upper strict
    if (!read_mm_csr(lev_file(prefix, lev, "lower"), m, err)) return E_BAD_ARG;
    split_diag(m, true, l, dv);
    if (!read_mm_csr(lev_file(prefix, lev, "upper"), u, err)) return E_BAD_ARG;
  } else if (solver == BWGS) {  // lower strict, upper = triu(A)
    if (!read_mm_csr(lev_file(prefix, lev, "lower"), l, err)) return E_BAD_ARG;
    if (!read_mm_csr(lev_file(prefix, lev, "upper"), m, err)) return E_BAD_ARG;
    split_diag(m, false, u, dv);
  } else if (solver != F_NONE) {
    err = "level " + std::to_string(lev) + ": the dump holds no data of solver " + std::to_string(solver);
    return E_UNSUPPORTED;
  }
  if ((!dv.empty() && (int)dv.size() != n) || (have_nd && nd.nr != n) || (!l.irp.empty() && l.nr != n) ||
      (!u.irp.empty() && u.nr != n)) {
    err = "level " + std::to_string(lev) + ": smoother files do not match the level's size";
    return E_BAD_ARG;
  }
  const bool tri = !l.irp.empty();
  int rc = stage_smoother(p, lev0, which, solver, have_nd ? nd.irp.data() : nullptr, nd.ja.data(), nd.val.data(),
                          dv.empty() ? nullptr : dv.data(), tri ? l.irp.data() : nullptr, l.ja.data(), l.val.data(),
                          tri ? u.irp.data() : nullptr, u.ja.data(), u.val.data());
  if (rc) err = p.last_error;
  return rc;
}

}  // namespace

int load_hierarchy(Prec& p, const std::string& prefix_in, int n, const int64_t* irp, const int* ja, const double* val,
                   bool smoothers, bool& smoothers_done, std::string& err) {
  smoothers_done = false;
  if (!p.initialized || p.precv.empty()) return E_NOT_INIT;
  const std::string prefix = prefix_in.empty() ? "dump_lev_d" : prefix_in;
  p.staged.clear();
  int rc = stage_level(p, 0, n, irp, ja, val, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) { err = p.last_error; return rc; }
  int nlev = 1;
  for (int lev = 2; (int)p.precv.size() > 1 && file_exists(lev_file(prefix, lev, "ac")); ++lev) {
    HCsr a, pm, r;
    if (!read_mm_csr(lev_file(prefix, lev, "ac"), a, err) || !read_mm_csr(lev_file(prefix, lev, "p"), pm, err) ||
        !read_mm_csr(lev_file(prefix, lev, "r"), r, err)) {
      p.staged.clear();
      return E_BAD_ARG;
    }
    if (a.nr != a.nc) { err = "level " + std::to_string(lev) + ": coarse matrix is not square"; p.staged.clear(); return E_BAD_ARG; }
    rc = stage_level(p, lev - 1, a.nr, a.irp.data(), a.ja.data(), a.val.data(), pm.nr, pm.irp.data(), pm.ja.data(),
                     pm.val.data(), r.irp.data(), r.ja.data(), r.val.data());
    if (rc) { err = p.last_error; p.staged.clear(); return rc; }
    nlev = lev;
  }
  if (smoothers) {
    // the solvers of the configured smoothers for nlev levels tell which
    // files to read
    std::vector<Level> cfg = p.precv;
    if (nlev == 1) cfg.resize(1);
    else resize_levels(cfg, nlev);
    for (int l = 0; l < nlev && !rc; ++l) {
      if (!cfg[l].sm.allocated) { err = "no smoother configured at level " + std::to_string(l + 1); rc = E_NO_SOLVER; break; }
      rc = load_smoother(p, prefix + "_sm", l, 0, cfg[l].sm.solver, err);
      if (!rc && cfg[l].sm2a.allocated) rc = load_smoother(p, prefix + "_sm2a", l, 1, cfg[l].sm2a.solver, err);
    }
    if (rc) { p.staged.clear(); return rc; }
  }
  rc = finalize_hierarchy(p, nlev, smoothers_done);
  if (rc) err = p.last_error;
  return rc;
}

}  // namespace mld
