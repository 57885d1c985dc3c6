// Hierarchy dump in MatrixMarket files, with the file names of the
// reference's prec%dump (mlprec/mld_d_prec_type.f90:698-729):
//   level data   <prefix>_p000_lLLL_{ac,r,p}.mtx   (mld_d_base_onelev_dump.f90:89-119)
//   smoother ND  <prefix>_sm_p000_lLLL_nd.mtx      (mld_d_jac_smoother_dmp.f90:76-80)
//   solver data  <prefix>_sm_p000_lLLL_{diag,lower,upper}.mtx
//                (mld_d_diag_solver_dmp.f90:76-79, mld_d_ilu_solver_dmp.f90:77-85,
//                 mld_d_gs_solver_dmp.f90:77-83)
//   post smoother of the level under <prefix>_sm2a (mld_d_base_onelev_dump.f90:126-129).
// Levels are numbered from 1 as in the reference; one process, so "_p000".
// Matrices are written as coordinate real general (1-based), vectors as
// array real general; values with 17 significant digits (round trip exact).
#include <algorithm>
#include <cstdio>
#include <string>

#include "mld_host.hpp"

namespace mld {
namespace {

// n > 0: the shape of an operator held with no rows (an empty ND) is n x n
bool write_csr(const std::string& fname, const HCsr& m, const std::string& head, int n = 0) {
  FILE* f = std::fopen(fname.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
  if (!head.empty()) std::fprintf(f, "%% %s\n", head.c_str());
  const bool shaped = (m.nr == 0 && n > 0);
  std::fprintf(f, "%d %d %lld\n", shaped ? n : m.nr, shaped ? n : m.nc, (long long)m.nnz());
  for (int i = 0; i < m.nr; ++i)
    for (int64_t k = m.irp.empty() ? 0 : m.irp[i]; k < (m.irp.empty() ? 0 : m.irp[i + 1]); ++k)
      std::fprintf(f, "%d %d %.17e\n", i + 1, m.ja[k] + 1, m.val[k]);
  return std::fclose(f) == 0;
}

bool write_vec(const std::string& fname, const std::vector<double>& v, const std::string& head) {
  FILE* f = std::fopen(fname.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "%%%%MatrixMarket matrix array real general\n");
  if (!head.empty()) std::fprintf(f, "%% %s\n", head.c_str());
  std::fprintf(f, "%zu 1\n", v.size());
  for (double x : v) std::fprintf(f, "%.17e\n", x);
  return std::fclose(f) == 0;
}

// the strict triangle t with the diagonal d (not inverted) inserted in
// column order: the GS solvers' tril(A) / triu(A)
HCsr with_diag(const HCsr& t, const std::vector<double>& d, bool lower) {
  HCsr m;
  m.nr = t.nr; m.nc = t.nc;
  m.irp.assign((size_t)t.nr + 1, 0);
  for (int i = 0; i < t.nr; ++i) {
    if (!lower) { m.ja.push_back(i); m.val.push_back(d[i]); }
    for (int64_t k = t.irp[i]; k < t.irp[i + 1]; ++k) { m.ja.push_back(t.ja[k]); m.val.push_back(t.val[k]); }
    if (lower) { m.ja.push_back(i); m.val.push_back(d[i]); }
    m.irp[i + 1] = (int64_t)m.ja.size();
  }
  return m;
}

std::string lev_name(const std::string& prefix, int lev, const char* what) {
  char tail[64];
  std::snprintf(tail, sizeof tail, "_p000_l%03d_%s.mtx", lev, what);
  return prefix + tail;
}

// mld_d_jac_smoother_dmp + the solver's dump
bool dump_smoother(const SmootherData& sd, const std::string& prefix, int lev, int n, const std::string& head,
                   bool smoother, bool solver) {
  if (sd.smoother == NOPREC) return true;  // the base smoother dumps nothing
  bool ok = true;
  if (smoother) ok = ok && write_csr(lev_name(prefix, lev, "nd"), sd.nd, head, n);
  if (!solver) return ok;
  switch (sd.solver) {
    case DIAG_SCALE:
    case L1_DIAG:
      ok = ok && write_vec(lev_name(prefix, lev, "diag"), sd.dv, head);
      break;
    case ILU_N:
      ok = ok && write_csr(lev_name(prefix, lev, "lower"), sd.l, head);
      ok = ok && write_vec(lev_name(prefix, lev, "diag"), sd.dv, head);
      ok = ok && write_csr(lev_name(prefix, lev, "upper"), sd.u, head);
      break;
    case GS:  // L = tril(A, 0), U strict (mld_d_gs_solver_bld.f90:79)
      ok = ok && write_csr(lev_name(prefix, lev, "lower"), with_diag(sd.l, sd.dv, true), head);
      ok = ok && write_csr(lev_name(prefix, lev, "upper"), sd.u, head);
      break;
    case BWGS:  // L strict, U = triu(A, 0) (mld_d_bwgs_solver_bld.f90:81)
      ok = ok && write_csr(lev_name(prefix, lev, "lower"), sd.l, head);
      ok = ok && write_csr(lev_name(prefix, lev, "upper"), with_diag(sd.u, sd.dv, false), head);
      break;
    default:  // dense LU coarse solve: no file in the reference's layout
      break;
  }
  return ok;
}

}  // namespace

int dump_prec(const Prec& p, int istart, int iend, const std::string& prefix_in, const std::string& head,
              int flags, std::string& err) {
  const int nlev = (int)p.precv.size();
  if (nlev == 0) return 3112;
  const std::string prefix = prefix_in.empty() ? "dump_lev_d" : prefix_in;
  const bool ac = flags & 1, rp = flags & 2, smoother = flags & 4, solver = flags & 8;
  const int il1 = (istart > 0) ? std::max(1, istart) : std::min(2, nlev);
  const int iln = (iend > 0) ? std::min(nlev, iend) : nlev;
  for (int lev = il1; lev <= iln; ++lev) {
    const Level& L = p.precv[lev - 1];
    bool ok = true;
    if (lev >= 2) {
      if (ac) ok = ok && write_csr(lev_name(prefix, lev, "ac"), L.a, head);
      if (rp) {
        ok = ok && write_csr(lev_name(prefix, lev, "r"), L.r, head);
        ok = ok && write_csr(lev_name(prefix, lev, "p"), L.p, head);
      }
    }
    if ((smoother || solver) && p.smoothers_built) {
      ok = ok && dump_smoother(L.smd, prefix + "_sm", lev, L.a.nr, head, smoother, solver);
      if (L.sm2a.allocated)
        ok = ok && dump_smoother(L.smd2, prefix + "_sm2a", lev, L.a.nr, head, smoother, solver);
    }
    if (!ok) {
      err = "cannot write the dump files of level " + std::to_string(lev) + " under prefix " + prefix;
      return E_BAD_ARG;
    }
  }
  return 0;
}

}  // namespace mld
