// Host build phase: smoothed-aggregation hierarchy and smoother data.
//   hierarchy          : mlprec/impl/mld_d_hierarchy_bld.f90:66-413
//   aggregation        : mlprec/impl/mld_d_dec_map_bld.f90:70-321,
//                        mlprec/impl/mld_daggrmap_bld.f90:126-180
//   smoothed P, R, Ac  : mlprec/impl/mld_daggrmat_smth_asb.f90:159-401
//   smoothers          : mlprec/impl/mld_d_smoothers_bld.f90:156-169,
//                        mlprec/impl/smoother/mld_d_jac_smoother_bld.f90:73-103,
//                        mlprec/impl/solver/mld_d_diag_solver_bld.f90:74-92
//   ILU(0)             : mlprec/impl/mld_dilu0_fact.f90:284-458,551-660
// Products are computed row-parallel (OpenMP) with each output entry
// accumulated in the order of the left factor's stored row, so the result
// does not depend on the thread count.  Built with -ffp-contract=off.
#include <omp.h>

#include <algorithm>
#include <functional>
#include <queue>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>

#include "mld_host.hpp"

namespace mld {

double wall_seconds() { return omp_get_wtime(); }
void phase_log(const char* what, int level, double seconds) {
  static const bool on = std::getenv("MLD_VERBOSE") != nullptr;
  if (on) std::fprintf(stderr, "[mld] level %d %-22s %9.3f ms\n", level, what, seconds * 1e3);
}

int host_max_row(const HCsr& a) {
  int m = 0;
#pragma omp parallel for schedule(static) reduction(max : m)
  for (int i = 0; i < a.nr; ++i) m = std::max(m, (int)(a.irp[i + 1] - a.irp[i]));
  return m;
}

std::vector<int> host_dia_offsets(const HCsr& a) {
  // each thread collects the offsets of a contiguous row range (at most 9 kept:
  // a ninth means "more than 8"), then the sets are merged
  const int nt = std::max(1, std::min(omp_get_max_threads(), a.nr / 65536 + 1));
  std::vector<std::vector<int>> part(nt);
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num();
    const int r0 = (int)((int64_t)a.nr * t / nt), r1 = (int)((int64_t)a.nr * (t + 1) / nt);
    std::vector<int>& offs = part[t];
    for (int i = r0; i < r1 && offs.size() <= 8; ++i) {
      size_t p = 0;
      for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
        const int o = a.ja[k] - i;
        while (p < offs.size() && offs[p] < o) ++p;
        if (p == offs.size() || offs[p] != o) {
          offs.insert(offs.begin() + p, o);
          if (offs.size() > 8) break;
        }
        ++p;
      }
    }
  }
  std::vector<int> all;
  for (const auto& p : part) {
    std::vector<int> u;
    std::set_union(all.begin(), all.end(), p.begin(), p.end(), std::back_inserter(u));
    all.swap(u);
    if (all.size() > 8) return {};
  }
  return all;
}

bool host_sell_layout(const HCsr& a, int sigma, int c, int maxrow, std::vector<int>& perm, std::vector<int>& sptr) {
  const int nr = a.nr;
  const int nsl = (nr + c - 1) / c;
  perm.resize(nr);
  auto len = [&](int i) { return (int)(a.irp[i + 1] - a.irp[i]); };
  if (sigma <= 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; ++i) perm[i] = i;
  } else {
    // stable counting sort by decreasing length inside each window
    const int nw = (nr + sigma - 1) / sigma;
#pragma omp parallel
    {
      std::vector<int> cnt((size_t)maxrow + 2);
#pragma omp for schedule(static)
      for (int w = 0; w < nw; ++w) {
        const int w0 = w * sigma, w1 = std::min(nr, w0 + sigma);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int i = w0; i < w1; ++i) cnt[maxrow - len(i) + 1]++;
        for (size_t b = 1; b < cnt.size(); ++b) cnt[b] += cnt[b - 1];
        for (int i = w0; i < w1; ++i) perm[w0 + cnt[maxrow - len(i)]++] = i;
      }
    }
  }
  std::vector<int64_t> width(nsl);
#pragma omp parallel for schedule(static)
  for (int sl = 0; sl < nsl; ++sl) {
    int wmax = 0;
    for (int l = 0; l < c; ++l) {
      const int p = sl * c + l;
      if (p < nr) wmax = std::max(wmax, len(perm[p]));
    }
    width[sl] = (int64_t)wmax * c;
  }
  sptr.assign((size_t)nsl + 1, 0);
  int64_t acc = 0;
  for (int sl = 0; sl < nsl; ++sl) {
    acc += width[sl];
    if (acc >= ((int64_t)1 << 31)) return false;
    sptr[sl + 1] = (int)acc;
  }
  return true;
}

void par_memcpy(void* dst, const void* src, size_t n) {
  constexpr size_t PIECE = (size_t)1 << 20;
  const int64_t np = (int64_t)((n + PIECE - 1) / PIECE);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < np; ++k) {
    const size_t o = (size_t)k * PIECE;
    std::memcpy((char*)dst + o, (const char*)src + o, std::min(PIECE, n - o));
  }
}

static std::vector<double> diag_of(const HCsr& a) {
  std::vector<double> d(a.nr, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < a.nr; ++i)
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k)
      if (a.ja[k] == i) d[i] = a.val[k];
  return d;
}

// Greedy three-phase decoupled aggregation (Vanek-Mandel-Brezina) with the
// reference's strength test |a_ij| > theta*sqrt(|a_ii a_jj|), natural or
// descending-degree visiting order, phase-two snapshot and first-maximum
// tie-break.  Output ids are 0-based.
int dec_map_bld(int iorder, double theta, const HCsr& a, std::vector<int>& ilaggr, int& naggr) {
  const int nr = a.nr;
  const int64_t nz = a.nnz();
  std::vector<double> diag = diag_of(a);
  // strength flag of every stored entry (in range and strong)
  std::vector<unsigned char> strong(nz);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; ++i)
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      int j = a.ja[k];
      strong[k] = (j >= 0 && j < nr) &&
                  (std::fabs(a.val[k]) > theta * std::sqrt(std::fabs(diag[i] * diag[j])));
    }
  std::vector<int> order(nr);
  for (int i = 0; i < nr; ++i) order[i] = i;
  if (iorder == ORD_DESC_DEG) {
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
      return (a.irp[p + 1] - a.irp[p]) > (a.irp[q + 1] - a.irp[q]);
    });
  }
  const int FREE = 0;  // aggregate ids are 1-based during the build
  std::vector<int> agg(nr, FREE);
  int cnt = 0;
  // phase 1
  for (int t = 0; t < nr; ++t) {
    int i = order[t];
    if (agg[i] != FREE) continue;
    bool disjoint = true;
    for (int64_t k = a.irp[i]; k < a.irp[i + 1] && disjoint; ++k)
      if (strong[k] && agg[a.ja[k]] != FREE) disjoint = false;
    if (!disjoint) continue;
    ++cnt;
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k)
      if (strong[k]) agg[a.ja[k]] = cnt;
    agg[i] = cnt;
  }
  // phase 2 (decisions on the phase-1 snapshot; a node joins the aggregate
  // of a node aggregated in phase 1, whose id phase 2 never changes, so the
  // rows are independent)
  std::vector<int> snap(agg);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nr; ++t) {
    int i = order[t];
    if (agg[i] != FREE) continue;
    double best = 0.0;
    int64_t kbest = -1;
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      if (!strong[k] || snap[a.ja[k]] <= 0) continue;
      double c = std::fabs(a.val[k]);
      if (c > best) { best = c; kbest = k; }
    }
    if (kbest >= 0) agg[i] = agg[a.ja[kbest]];
  }
  // phase 3
  for (int t = 0; t < nr; ++t) {
    int i = order[t];
    if (agg[i] != FREE) continue;
    ++cnt;
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k)
      if (strong[k] && agg[a.ja[k]] == FREE) agg[a.ja[k]] = cnt;
    agg[i] = cnt;
  }
  ilaggr.resize(nr);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < nr; ++i) {
    if (agg[i] <= 0) bad = 1;
    ilaggr[i] = agg[i] - 1;
  }
  if (bad) return E_INTERNAL;
  naggr = cnt;
  return 0;
}

// Read-only view of a CSR matrix (lets S share A's row pointers and columns).
struct CsrView {
  int nr, nc;
  const int64_t* irp;
  const int* ja;
  const double* val;
};
static CsrView view(const HCsr& a) { return CsrView{a.nr, a.nc, a.irp.data(), a.ja.data(), a.val.data()}; }

// Row-parallel Gustavson product C = A*B; every C_ij is the sum over A's
// stored row i, in that order, of a_ik*b_kj, accumulated from 0.0 (each sum
// is the sequential loop's, whatever the thread count).  Columns of C
// ascending.  A counting pass sizes C, the second pass fills it in place.
static void spgemm(const CsrView& A, const CsrView& B, HCsr& C) {
  const int nr = A.nr, nc = B.nc;
  C.nr = nr; C.nc = nc;
  C.irp.assign((size_t)nr + 1, 0);
#pragma omp parallel
  {
    std::vector<int> mark(nc, -1);
#pragma omp for schedule(dynamic, 4096)
    for (int i = 0; i < nr; ++i) {
      int64_t c = 0;
      for (int64_t ka = A.irp[i]; ka < A.irp[i + 1]; ++ka) {
        const int k = A.ja[ka];
        for (int64_t kb = B.irp[k]; kb < B.irp[k + 1]; ++kb) {
          const int j = B.ja[kb];
          if (mark[j] != i) { mark[j] = i; ++c; }
        }
      }
      C.irp[i + 1] = c;
    }
  }
  for (int i = 0; i < nr; ++i) C.irp[i + 1] += C.irp[i];
  C.ja.resize(C.irp[nr]);
  C.val.resize(C.irp[nr]);
#pragma omp parallel
  {
    std::vector<int> mark(nc, -1);
    std::vector<double> acc(nc, 0.0);
    std::vector<int> cols;
#pragma omp for schedule(dynamic, 4096)
    for (int i = 0; i < nr; ++i) {
      cols.clear();
      for (int64_t ka = A.irp[i]; ka < A.irp[i + 1]; ++ka) {
        const int k = A.ja[ka];
        const double ak = A.val[ka];
        for (int64_t kb = B.irp[k]; kb < B.irp[k + 1]; ++kb) {
          const int j = B.ja[kb];
          if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cols.push_back(j); }
          acc[j] = acc[j] + ak * B.val[kb];
        }
      }
      std::sort(cols.begin(), cols.end());
      int64_t q = C.irp[i];
      for (int j : cols) { C.ja[q] = j; C.val[q] = acc[j]; ++q; }
    }
  }
}

// Transpose; each output row lists its columns in ascending order.  Thread t
// takes a contiguous block of rows and counts its entries per column; the
// entries of block t then go after those of blocks 0..t-1 in every output
// row, which is the sequential placement's order.
static void transpose(const HCsr& A, HCsr& T) {
  T.nr = A.nc; T.nc = A.nr;
  const int64_t nz = A.nnz();
  const int nc = A.nc;
  const int nt = std::max(1, std::min(omp_get_max_threads(), (int)(nz / 65536) + 1));
  std::vector<int> r0(nt + 1);
  for (int t = 0; t <= nt; ++t) r0[t] = (int)((int64_t)A.nr * t / nt);
  std::vector<int64_t> cnt((size_t)nt * nc, 0);  // cnt[t * nc + j]
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num();
    int64_t* c = cnt.data() + (size_t)t * nc;
    for (int i = r0[t]; i < r0[t + 1]; ++i)
      for (int64_t k = A.irp[i]; k < A.irp[i + 1]; ++k) c[A.ja[k]]++;
  }
  // start of block t in output row j: sum over rows < j, then blocks < t
  T.irp.assign((size_t)T.nr + 1, 0);
  int64_t run = 0;
  for (int j = 0; j < nc; ++j) {
    T.irp[j] = run;
    for (int t = 0; t < nt; ++t) {
      const int64_t v = cnt[(size_t)t * nc + j];
      cnt[(size_t)t * nc + j] = run;
      run += v;
    }
  }
  T.irp[nc] = run;
  T.ja.resize(nz); T.val.resize(nz);
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num();
    int64_t* pos = cnt.data() + (size_t)t * nc;
    for (int i = r0[t]; i < r0[t + 1]; ++i)
      for (int64_t k = A.irp[i]; k < A.irp[i + 1]; ++k) {
        const int64_t q = pos[A.ja[k]]++;
        T.ja[q] = i; T.val[q] = A.val[k];
      }
  }
}

int aggrmat_smth_asb(const HCsr& a, const std::vector<int>& ilaggr, int naggr, LevelParms& parms, HCsr& ac,
                     HCsr& p, HCsr& r, int device) {
  if (parms.aggr_prol != SMOOTH_PROL) return E_UNSUPPORTED;  // only SMOOTHED is provided
  const int n = a.nr;
  std::vector<double> dinv = diag_of(a);
  const bool filter = (parms.aggr_filter == FILTER_MAT);
  // S = (I - omega D^-1 A_or_Af), built in A's (or Af's) stored order
  HCsr s;
  if (filter) {
    // filtered matrix: weak entries removed and lumped on the diagonal
    s.nr = s.nc = n; s.irp.assign((size_t)n + 1, 0);
    std::vector<double> tmpv;
    for (int i = 0; i < n; ++i) {
      double lump = 0.0;
      int64_t jd = -1;
      int64_t start = (int64_t)s.ja.size();
      for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
        int j = a.ja[k];
        double v = a.val[k];
        if (j == i) jd = (int64_t)s.ja.size();
        if (std::fabs(v) < parms.aggr_thresh * std::sqrt(std::fabs(dinv[i] * dinv[j]))) {
          lump = lump + v;
          v = 0.0;
        }
        s.ja.push_back(j); s.val.push_back(v);
      }
      if (jd >= 0) s.val[jd] = s.val[jd] - lump;
      // drop exact zeros (clean_zeros)
      int64_t w = start;
      for (int64_t k = start; k < (int64_t)s.ja.size(); ++k)
        if (s.val[k] != 0.0) { s.ja[w] = s.ja[k]; s.val[w] = s.val[k]; ++w; }
      s.ja.resize(w); s.val.resize(w);
      s.irp[i + 1] = w;
    }
  }
  for (int i = 0; i < n; ++i) dinv[i] = (dinv[i] != 0.0) ? 1.0 / dinv[i] : 1.0;
  // D^-1 A row scaling (values only: the structure stays A's) and the
  // max-norm estimate ||D^-1 A||_inf
  std::vector<double> daval(a.nnz());
  double anorm = 0.0;
#pragma omp parallel for schedule(static) reduction(max : anorm)
  for (int i = 0; i < n; ++i) {
    double rs = 0.0;
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      daval[k] = a.val[k] * dinv[i];
      rs = rs + std::fabs(daval[k]);
    }
    anorm = std::max(anorm, rs);
  }
  double omega;
  if (parms.aggr_omega_alg == EIG_EST) {
    if (parms.aggr_eig != MAX_NORM) return E_UNSUPPORTED;
    omega = 4.0 / (3.0 * anorm);
    parms.aggr_omega_val = omega;
  } else if (parms.aggr_omega_alg == USER_CHOICE) {
    omega = parms.aggr_omega_val;
  } else {
    return E_BAD_ARG;
  }
  CsrView sv;
  if (filter) {
    for (int i = 0; i < n; ++i)
      for (int64_t k = s.irp[i]; k < s.irp[i + 1]; ++k) s.val[k] = s.val[k] * dinv[i];
    sv = view(s);
  } else {
    sv = CsrView{a.nr, a.nc, a.irp.data(), a.ja.data(), daval.data()};
  }
  if (device >= 0 && !filter) {
    // the products on the device (S formed there from A, D^-1 and omega)
    std::string err;
    const double t0 = wall_seconds();
    const int rc = dev_smoothed_products(device, a, nullptr, dinv, omega, ilaggr, naggr, p, r, ac, err);
    phase_log("P, AP, R, Ac on the device", -1, wall_seconds() - t0);
    if (rc == 0) return 0;
    phase_log(("device products not used: " + err).c_str(), -1, 0.0);
  }
  double* svv = const_cast<double*>(sv.val);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int64_t k = sv.irp[i]; k < sv.irp[i + 1]; ++k)
      svv[k] = (sv.ja[k] == i) ? 1.0 - omega * svv[k] : -(omega * svv[k]);
  if (device >= 0 && filter) {
    std::string err;
    const double t0 = wall_seconds();
    const int rc = dev_smoothed_products(device, a, &s, dinv, omega, ilaggr, naggr, p, r, ac, err);
    phase_log("P, AP, R, Ac on the device", -1, wall_seconds() - t0);
    if (rc == 0) return 0;
    phase_log(("device products not used: " + err).c_str(), -1, 0.0);
  }
  // tentative prolongator: row i has a single 1.0 in column ilaggr[i]
  HCsr pt;
  pt.nr = n; pt.nc = naggr;
  pt.irp.resize((size_t)n + 1);
  for (int i = 0; i <= n; ++i) pt.irp[i] = i;
  pt.ja.assign(ilaggr.begin(), ilaggr.end());
  pt.val.assign(n, 1.0);
  double t0 = wall_seconds();
  spgemm(sv, view(pt), p);  // P  = (I - omega D^-1 A) Ptilde
  double t1 = wall_seconds();
  phase_log("P = S Ptilde", -1, t1 - t0);
  HCsr ap;
  spgemm(view(a), view(p), ap);  // AP = A P
  double t2 = wall_seconds();
  phase_log("A P", -1, t2 - t1);
  transpose(p, r);          // R  = P^T
  double t3 = wall_seconds();
  phase_log("R = P^T", -1, t3 - t2);
  spgemm(view(r), view(ap), ac);  // Ac = R (A P)
  phase_log("Ac = R (A P)", -1, wall_seconds() - t3);
  return 0;
}

// IKJ ILU(0) restricted to pattern(A): L strict lower with l_ik stored,
// U strict upper scaled by 1/pivot, d = 1/pivot.
int ilu0_factor(const HCsr& a, HCsr& l, HCsr& u, std::vector<double>& d) {
  const int m = a.nr;
  l.clear(); u.clear();
  l.nr = l.nc = u.nr = u.nc = m;
  l.irp.assign((size_t)m + 1, 0); u.irp.assign((size_t)m + 1, 0);
  d.assign(m, 0.0);
  for (int i = 0; i < m; ++i) {
    double piv = 0.0;
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      int j = a.ja[k];
      if (j < i && j >= 0) { l.ja.push_back(j); l.val.push_back(a.val[k]); }
      else if (j == i) piv = a.val[k];
      else if (j > i && j < m) { u.ja.push_back(j); u.val.push_back(a.val[k]); }
    }
    l.irp[i + 1] = (int64_t)l.ja.size();
    u.irp[i + 1] = (int64_t)u.ja.size();
    const int64_t lb = l.irp[i], le = l.irp[i + 1], ub = u.irp[i], ue = u.irp[i + 1];
    for (int64_t kk = lb; kk < le; ++kk) {
      const double lik = l.val[kk];  // unscaled l_ik used for the updates
      const int k = l.ja[kk];
      l.val[kk] = lik * d[k];
      int64_t sl = kk + 1, su = ub;  // merge cursors into row i of L and U
      for (int64_t jj = u.irp[k]; jj < u.irp[k + 1]; ++jj) {
        const int j = u.ja[jj];
        if (j < i) {
          while (sl < le && l.ja[sl] < j) ++sl;
          if (sl < le && l.ja[sl] == j) { l.val[sl] = l.val[sl] - lik * u.val[jj]; ++sl; }
        } else if (j == i) {
          piv = piv - lik * u.val[jj];
        } else {
          while (su < ue && u.ja[su] < j) ++su;
          if (su < ue && u.ja[su] == j) { u.val[su] = u.val[su] - lik * u.val[jj]; ++su; }
        }
      }
    }
    if (std::fabs(piv) < 2.220446049250313e-16) return E_PIVOT;
    piv = 1.0 / piv;
    d[i] = piv;
    for (int64_t kk = ub; kk < ue; ++kk) u.val[kk] = u.val[kk] * piv;
  }
  return 0;
}

// ILU(k) and MILU(k) by rows with levels of fill (mld_diluk_fact.f90:268-960):
// row i is copied into a dense work row with level 0 on A's pattern
// (iluk_copyin, :485-566); its columns are visited in ascending order (a
// min-heap, new fill columns joining it), and every visited column k < i with
// a nonzero value and level <= fill eliminates: l_ik = row_k * d_k, row_j -=
// row_k * u_kj (U rows stored scaled by their inverse pivot), lev_j =
// min(lev_j, lev_k + lev(u_kj) + 1) (iluk_fact, :637-722).  The visited columns
// are then split (iluk_copyout, :812-960): level <= fill kept in L or U,
// others dropped (MILU: added to the pivot, in the visiting order), d_i =
// 1/pivot, row i of U scaled by d_i.  Same storage as ilu0_factor.
int iluk_factor(const HCsr& a, int fill, bool milu, HCsr& l, HCsr& u, std::vector<double>& d) {
  const int m = a.nr;
  l.clear(); u.clear();
  l.nr = l.nc = u.nr = u.nc = m;
  l.irp.assign((size_t)m + 1, 0); u.irp.assign((size_t)m + 1, 0);
  d.assign(m, 0.0);
  std::vector<int> ulev;                          // level of every stored U entry
  std::vector<double> row(m, 0.0);
  std::vector<int> lev(m, -(m + 1));              // < 0: column not in the row
  std::priority_queue<int, std::vector<int>, std::greater<int>> heap;
  std::vector<int> idx;
  for (int i = 0; i < m; ++i) {
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      const int j = a.ja[k];
      if (j < 0 || j >= m) continue;
      row[j] = a.val[k];
      lev[j] = 0;
      heap.push(j);
    }
    idx.clear();
    int last = -1;
    while (!heap.empty()) {
      const int k = heap.top();
      heap.pop();
      if (k == last) continue;
      last = k;
      idx.push_back(k);
      if (row[k] != 0.0 && lev[k] <= fill && k < i) {
        const double rk = row[k];
        row[k] = row[k] * d[k];
        const int lk = lev[k];
        for (int64_t jj = u.irp[k]; jj < u.irp[k + 1]; ++jj) {
          const int j = u.ja[jj];
          if (lev[j] < 0) { heap.push(j); lev[j] = -lev[j]; }
          row[j] = row[j] - rk * u.val[jj];
          lev[j] = std::min(lev[j], lk + ulev[jj] + 1);
        }
      }
    }
    double piv = 0.0;
    for (int j : idx) {
      if (j < i) {
        if (lev[j] <= fill) { l.ja.push_back(j); l.val.push_back(row[j]); }
        else if (milu) piv = piv + row[j];
      } else if (j == i) {
        piv = piv + row[j];
      } else {
        if (lev[j] <= fill) { u.ja.push_back(j); u.val.push_back(row[j]); ulev.push_back(lev[j]); }
        else if (milu) piv = piv + row[j];
      }
      row[j] = 0.0;
      lev[j] = -(m + 1);
    }
    l.irp[i + 1] = (int64_t)l.ja.size();
    u.irp[i + 1] = (int64_t)u.ja.size();
    if (std::fabs(piv) < 2.220446049250313e-16) return E_PIVOT;
    piv = 1.0 / piv;
    d[i] = piv;
    for (int64_t kk = u.irp[i]; kk < u.irp[i + 1]; ++kk) u.val[kk] = piv * u.val[kk];
  }
  return 0;
}

// Level sets of a triangular solve: rows in the same level depend only on
// rows of earlier levels.  Rows within a level are kept in ascending order.
void level_schedule(const HCsr& t, bool lower, std::vector<int>& ptr, std::vector<int>& rows) {
  const int n = t.nr;
  std::vector<int> lev(n, 0);
  int nlev = 0;
  if (lower) {
    for (int i = 0; i < n; ++i) {
      int lv = 0;
      for (int64_t k = t.irp[i]; k < t.irp[i + 1]; ++k) lv = std::max(lv, lev[t.ja[k]] + 1);
      lev[i] = lv;
      nlev = std::max(nlev, lv + 1);
    }
  } else {
    for (int i = n - 1; i >= 0; --i) {
      int lv = 0;
      for (int64_t k = t.irp[i]; k < t.irp[i + 1]; ++k) lv = std::max(lv, lev[t.ja[k]] + 1);
      lev[i] = lv;
      nlev = std::max(nlev, lv + 1);
    }
  }
  ptr.assign((size_t)nlev + 1, 0);
  for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
  for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
  rows.resize(n);
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < n; ++i) rows[pos[lev[i]]++] = i;
}

static int dense_lu(const HCsr& a, std::vector<double>& lu, std::vector<int>& piv) {
  const int n = a.nr;
  lu.assign((size_t)n * n, 0.0);
  piv.resize(n);
  for (int i = 0; i < n; ++i)
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) lu[(size_t)i * n + a.ja[k]] += a.val[k];
  for (int k = 0; k < n; ++k) {
    int pr = k;
    double mx = std::fabs(lu[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(lu[(size_t)i * n + k]) > mx) { mx = std::fabs(lu[(size_t)i * n + k]); pr = i; }
    piv[k] = pr;
    if (mx == 0.0) return E_PIVOT;
    if (pr != k)
      for (int j = 0; j < n; ++j) std::swap(lu[(size_t)k * n + j], lu[(size_t)pr * n + j]);
    const double inv = 1.0 / lu[(size_t)k * n + k];
#pragma omp parallel for schedule(static) if (n > 512)
    for (int i = k + 1; i < n; ++i) {
      double f = lu[(size_t)i * n + k] * inv;
      lu[(size_t)i * n + k] = f;
      for (int j = k + 1; j < n; ++j) lu[(size_t)i * n + j] = lu[(size_t)i * n + j] - f * lu[(size_t)k * n + j];
    }
  }
  return 0;
}

static int build_smoother(const HCsr& a, const SmootherSpec& spec, SmootherData& sd) {
  sd = SmootherData();
  sd.smoother = spec.smoother;
  sd.solver = spec.solver;
  const int n = a.nr;
  if (spec.smoother == NOPREC) return 0;
  // Additive Schwarz on one process: no overlap rows (mld_d_as_smoother_bld.f90:78,
  // extension only when novr > 0 and np > 1), restriction and prolongation are
  // identities and ND holds the columns beyond the local rows, so it is empty
  // (:147-151); every sweep then applies the solver to x itself
  // (mld_d_as_smoother_apply_vect.f90:108-200): the data of the block-Jacobi
  // smoother with an empty ND, whatever the solver.
  const bool as = (spec.smoother == AS);
  if (spec.smoother != JAC && !as) return E_UNSUPPORTED;
  if (spec.solver < 0) return E_NO_SOLVER;
  sd.smoother = JAC;
  sd.nd.nr = sd.nd.nc = n;
  sd.nd.irp.assign((size_t)n + 1, 0);
  switch (spec.solver) {
    case F_NONE:
      return 0;
    case DIAG_SCALE:
    case L1_DIAG: {
      const bool l1 = (spec.solver == L1_DIAG);
      std::vector<double> d = diag_of(a);
      if (as) {  // ND empty: only the (l1-)diagonal
        sd.dv.resize(n);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
          double dd = d[i];
          if (l1) {
            double s = 0.0;
            for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k)
              if (a.ja[k] != i) s = s + std::fabs(a.val[k]);
            dd = dd + s;
          }
          sd.dv[i] = (dd == 0.0) ? 1.0 : 1.0 / dd;
        }
        sd.nnz_nd = 0;
        return 0;
      }
      // ND: A without its diagonal (clip_diag); for l1-Jacobi A - D_l1,
      // i.e. the diagonal entry replaced by -sum_{j!=i}|a_ij|.
      std::vector<int64_t> cnt(n);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        int64_t c = 0;
        for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k)
          if (a.ja[k] != i) ++c;
        cnt[i] = c + (l1 ? 1 : 0);  // l1: the diagonal slot is always present
      }
      for (int i = 0; i < n; ++i) sd.nd.irp[i + 1] = sd.nd.irp[i] + cnt[i];
      sd.nd.ja.resize(sd.nd.irp[n]);
      sd.nd.val.resize(sd.nd.irp[n]);
      sd.dv.resize(n);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        int64_t q = sd.nd.irp[i];
        double dd = d[i];
        if (l1) {
          double s = 0.0;
          bool has = false;
          for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k)
            if (a.ja[k] != i) s = s + std::fabs(a.val[k]); else has = true;
          dd = dd + s;
          bool placed = false;
          for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
            int j = a.ja[k];
            if (!has && !placed && j > i) { sd.nd.ja[q] = i; sd.nd.val[q] = -s; ++q; placed = true; }
            if (j == i) { sd.nd.ja[q] = i; sd.nd.val[q] = -s; ++q; placed = true; }
            else { sd.nd.ja[q] = j; sd.nd.val[q] = a.val[k]; ++q; }
          }
          if (!placed) { sd.nd.ja[q] = i; sd.nd.val[q] = -s; ++q; }
        } else {
          for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k)
            if (a.ja[k] != i) { sd.nd.ja[q] = a.ja[k]; sd.nd.val[q] = a.val[k]; ++q; }
        }
        sd.dv[i] = (dd == 0.0) ? 1.0 : 1.0 / dd;
      }
      sd.nnz_nd = sd.nd.nnz();
      return 0;
    }
    case ILU_N:
    case MILU_N: {
      // mld_d_ilu_solver_bld.f90:126-160: ILU(0) by its own routine, ILU(k>0)
      // and MILU(k) (k = 0 included) by the level-of-fill routine
      if (spec.fillin < 0) return E_BAD_ARG;
      sd.nnz_nd = 0;                               // off-block part: empty on one device
      sd.solver = ILU_N;                           // the same factors' apply
      int rc = (spec.solver == ILU_N && spec.fillin == 0)
                   ? ilu0_factor(a, sd.l, sd.u, sd.dv)
                   : iluk_factor(a, spec.fillin, spec.solver == MILU_N, sd.l, sd.u, sd.dv);
      if (rc) return rc;
      level_schedule(sd.l, true, sd.lsched_ptr, sd.lsched_rows);
      level_schedule(sd.u, false, sd.usched_ptr, sd.usched_rows);
      return 0;
    }
    case DENSE_LU: {
      // dense row-major copy; factorized and inverted on the device
      sd.nnz_nd = 0;
      if (n > 16384) return E_UNSUPPORTED;
      sd.lu.assign((size_t)n * n, 0.0);
      for (int i = 0; i < n; ++i)
        for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) sd.lu[(size_t)i * n + a.ja[k]] += a.val[k];
      return 0;
    }
    case GS:
    case BWGS: {
      // mld_d_gs_solver_bld.f90 / mld_d_bwgs_solver_bld.f90 (a%tril / a%triu
      // with the diagonal on the solved side): kept as the strict lower part
      // l, the strict upper part u and the diagonal dv (not inverted: the
      // substitution divides by it).  ND (columns beyond the local rows) is
      // empty on one device.
      sd.nnz_nd = 0;
      sd.l.nr = sd.l.nc = sd.u.nr = sd.u.nc = n;
      sd.l.irp.assign((size_t)n + 1, 0);
      sd.u.irp.assign((size_t)n + 1, 0);
      sd.dv.assign(n, 0.0);
      for (int i = 0; i < n; ++i) {
        for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
          const int j = a.ja[k];
          if (j < i) { sd.l.ja.push_back(j); sd.l.val.push_back(a.val[k]); }
          else if (j > i) { sd.u.ja.push_back(j); sd.u.val.push_back(a.val[k]); }
          else sd.dv[i] = sd.dv[i] + a.val[k];
        }
        sd.l.irp[i + 1] = (int64_t)sd.l.ja.size();
        sd.u.irp[i + 1] = (int64_t)sd.u.ja.size();
      }
      level_schedule(sd.l, true, sd.lsched_ptr, sd.lsched_rows);
      level_schedule(sd.u, false, sd.usched_ptr, sd.usched_rows);
      return 0;
    }
    default:
      return E_UNSUPPORTED;  // MILU/ILUT and third-party solvers
  }
}

void resize_levels(std::vector<Level>& v, int nplevs) {
  // resize to the target number of levels (mld_d_hierarchy_bld.f90:199-241):
  // base parms/smoothers at 1, intermediate ones from level 2, coarse ones at
  // the end
  const int iszv0 = (int)v.size();
  if (iszv0 == nplevs) return;
  std::vector<Level> t(nplevs);
  Level base = v[0], med = v[std::min(1, iszv0 - 1)], crs = v[iszv0 - 1];
  t[0].parms = base.parms; t[0].sm = base.sm; t[0].sm2a = base.sm2a;
  for (int i = 2; i <= std::min(iszv0, nplevs) - 1; ++i) {
    t[i - 1].parms = med.parms; t[i - 1].sm = v[i - 1].sm; t[i - 1].sm2a = v[i - 1].sm2a;
  }
  for (int i = iszv0 - 1; i <= nplevs - 1; ++i) {  // as in the reference, may start at level 1
    t[i - 1].parms = med.parms; t[i - 1].sm = med.sm; t[i - 1].sm2a = med.sm2a;
  }
  t[nplevs - 1].parms = crs.parms; t[nplevs - 1].sm = crs.sm; t[nplevs - 1].sm2a = crs.sm2a;
  t[0].a = std::move(v[0].a);
  v = std::move(t);
}

int hierarchy_build(Prec& p, int n, const int64_t* irp, const int* ja, const double* val) {
  if (!p.initialized || p.precv.empty()) return E_NOT_INIT;
  p.hier_built = p.smoothers_built = false;
  std::vector<Level>& v = p.precv;
  Level& l0 = v[0];
  const double tc = wall_seconds();
  l0.a.clear();
  l0.a.nr = l0.a.nc = n;
  const int64_t nz = irp[n];
  l0.a.irp.resize((size_t)n + 1);
  l0.a.ja.resize(nz);
  l0.a.val.resize(nz);
  int bad = 0;
  // copy (the caller keeps its arrays) and check the rows, in parallel
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    l0.a.irp[i + 1] = irp[i + 1];
    for (int64_t k = irp[i]; k < irp[i + 1]; ++k) {
      l0.a.ja[k] = ja[k];
      l0.a.val[k] = val[k];
      if (k > irp[i] && ja[k] <= ja[k - 1]) bad = 1;
    }
  }
  l0.a.irp[0] = irp[0];
  phase_log("copy of A", 1, wall_seconds() - tc);
  if (bad) { p.last_error = "CSR rows must have strictly ascending columns"; return E_BAD_ARG; }
  if ((int)v.size() == 1) { p.hier_built = true; return 0; }
  // default coarse size: 40 * int(N^(1/3)) (mld_d_hierarchy_bld.f90:187-197)
  int casize = p.min_coarse_size;
  if (casize < 0) {
    casize = (int)std::pow((double)n, 1.0 / (1.0 * 3.0));
    casize = std::max(casize, 1) * 40;
    p.min_coarse_size = casize;
  }
  const int nplevs = std::max(2, p.max_levs);
  resize_levels(v, nplevs);
  const int iszv = (int)v.size();
  const LevelParms coarseparms = v[iszv - 1].parms;
  const SmootherSpec coarse_sm = v[iszv - 1].sm, coarse_sm2a = v[iszv - 1].sm2a;
  int newsz = 0;
  for (int i = 2; i <= iszv; ++i) {  // array_build_loop, 1-based level numbers
    Level& prev = v[i - 2];
    Level& cur = v[i - 1];
    std::vector<int> ilaggr;
    int naggr = 0;
    if (cur.parms.par_aggr_alg != DEC_AGGR) return E_UNSUPPORTED;
    double t0 = wall_seconds();
    int rc = dec_map_bld(cur.parms.aggr_ord, cur.parms.aggr_thresh, prev.a, ilaggr, naggr);
    if (rc) return rc;
    phase_log("aggregation", i, wall_seconds() - t0);
    double sizeratio = (double)naggr;
    sizeratio = (i == 2) ? (double)n / sizeratio : (double)prev.naggr / sizeratio;
    cur.szratio = sizeratio;
    if (naggr <= casize) newsz = i;
    if (i > 2) {
      if (sizeratio < p.min_cr_ratio) newsz = (sizeratio > 1.0) ? i : i - 1;
      if (naggr == prev.naggr) newsz = i - 1;
    }
    if (newsz > 0) {
      Level& tgt = v[newsz - 1];
      double athresh = tgt.parms.aggr_thresh, aomega = tgt.parms.aggr_omega_val;
      tgt.parms = coarseparms;
      tgt.parms.aggr_thresh = athresh;
      tgt.parms.aggr_omega_val = aomega;
      tgt.sm = coarse_sm; tgt.sm2a = coarse_sm2a;
      if (newsz < i) { ilaggr = tgt.ilaggr; naggr = tgt.naggr; }
      rc = aggrmat_smth_asb(v[newsz - 2].a, ilaggr, naggr, tgt.parms, tgt.a, tgt.p, tgt.r,
                            v[newsz - 2].a.nr >= p.setup_min_rows ? p.setup_device : -1);
      if (rc) return rc;
      tgt.ilaggr = std::move(ilaggr);
      tgt.naggr = naggr;
      break;
    }
    rc = aggrmat_smth_asb(prev.a, ilaggr, naggr, cur.parms, cur.a, cur.p, cur.r,
                          prev.a.nr >= p.setup_min_rows ? p.setup_device : -1);
    if (rc) return rc;
    cur.ilaggr = std::move(ilaggr);
    cur.naggr = naggr;
  }
  if (newsz > 0) v.resize(newsz);
  for (auto& lv : v)
    if ((int64_t)lv.a.nnz() >= (int64_t)1 << 31) { p.last_error = "level matrix exceeds 2^31 entries"; return E_UNSUPPORTED; }
  p.hier_built = true;
  return 0;
}

int smoothers_build(Prec& p) {
  if (!p.initialized) return E_NOT_INIT;
  if (!p.hier_built) return E_NOT_BUILT;
  int lev = 0;
  for (auto& lv : p.precv) {
    ++lev;
    if (!lv.sm.allocated) return E_NO_SOLVER;
    double t0 = wall_seconds();
    int rc = build_smoother(lv.a, lv.sm, lv.smd);
    phase_log("smoother build", lev, wall_seconds() - t0);
    if (rc) return rc;
    if (lv.sm2a.allocated) {
      rc = build_smoother(lv.a, lv.sm2a, lv.smd2);
      if (rc) return rc;
    }
  }
  p.smoothers_built = true;
  return 0;
}

double operator_complexity(const Prec& p) {  // mld_d_prec_type.f90:465-497
  if (p.precv.empty() || p.precv[0].a.nnz() == 0) return 0.0;
  double s = 0.0;
  for (auto& lv : p.precv) s += (double)lv.a.nnz();
  return s / (double)p.precv[0].a.nnz();
}

}  // namespace mld
