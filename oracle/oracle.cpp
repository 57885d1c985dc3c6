// ============================================================================
// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of MLD2P4 2.1's smoothed-aggregation AMG preconditioner
// (build + apply) and of the PSBLAS kernels its apply path calls; sequential
// except for the row loops of products and vector updates, which OpenMP
// threads split without changing any row's order of evaluation.  It exists to CHECK the MI355X product library; it is never linked
// into, called by, or shipped as part of that library.  Only tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
//
// Every routine names the reference file:line it restates (paths relative to
// the reference tree).  Loops keep the reference's order of evaluation so
// that results can be compared bit-for-bit where the product keeps the same
// per-row order.  Compiled with -ffp-contract=off: no fused multiply-adds.
//
// PSBLAS (the reference's BLAS layer, >= 3.5) is not part of the reference
// tree.  Its kernels are restated from their documented semantics; the
// per-row summation order "sum in stored CSR order, then combine with y" is
// an assumption and is marked "PSBLAS order assumed" where it matters.
//
// Pinning status of each part (what checks it, see tests/test_oracle_*.py):
//   gen_pde3d           : closed form of the 7-point stencil values/pattern
//   dec_map_bld         : independent pure-Python restatement on small grids,
//                         invariants (partition, contiguous ids, step-1
//                         aggregates = full strong neighbourhoods)
//   smoothed P/R/Ac     : R == P^T exactly, Ac == P^T A P (dense numpy),
//                         omega == 2/3 closed form on the 7-point Poisson
//   ILU(0)              : (L D^-1 U)_ij == a_ij on pattern(A); exact LU on
//                         tridiagonal matrices
//   jacobi / l1-jacobi  : closed forms D^-1 x and D^-1 (x - ND y)
//   V-cycle             : linearity, symmetry for symmetric A, 'T' == 'N',
//                         2-level exact-coarse identity P (P^T A P)^-1 P^T
//   CG/FCG/BiCGStab     : numpy.linalg.solve, golden mld_sol.mtx vector
//   PSBLAS last-ulp orders (spspmm, spnmi, spsm): parity unpinned.
// ============================================================================
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <set>
#include <vector>

namespace {

typedef int64_t ix;

struct Csr {
  ix nr = 0, nc = 0;
  std::vector<ix> irp;      // nr+1, 0-based
  std::vector<int> ja;      // column indices, 0-based
  std::vector<double> val;
  ix nnz() const { return irp.empty() ? 0 : irp[nr]; }
};

// ---------------------------------------------------------------------------
// Parameter codes.  Values follow mlprec/mld_base_prec_type.F90:128-287.
// Codes >= 100 are this project's additions (no reference counterpart).
// ---------------------------------------------------------------------------
enum {
  SM_NOPREC = 0,   // mld_noprec_
  SM_JAC = 1,      // mld_jac_   (jacobi smoother + diag solver)
  SM_BJAC = 2,     // mld_bjac_  (jacobi smoother + ILU(0) solver)
  SM_AS = 3,       // mld_as_    (additive Schwarz + ILU(0) solver; on one process the
                   //             block-Jacobi smoother: mld_d_as_smoother_bld.f90:78,147-151,
                   //             mld_d_as_smoother_apply_vect.f90:108-200)
  SM_FBGS = 4,     // mld_fbgs_  (pre: jacobi smoother + GS solver, post: + BWGS solver)
  SM_L1JAC = 100,  // l1-Jacobi (Baker, Falgout, Kolev, Yang, SISC 33(5) 2011)
  SM_DENSELU = 101, // dense LU of the coarsest matrix (coarse solver only)
  SM_GS = 102,     // jacobi smoother + forward Gauss-Seidel solver (mld_d_gs_solver)
  SM_BWGS = 103    // jacobi smoother + backward Gauss-Seidel solver (mld_d_bwgs_solver)
};
// mld_add_ml_, mld_mult_ml_, mld_vcycle_ml_, mld_wcycle_ml_, mld_kcycle_ml_,
// mld_kcyclesym_ml_ (mld_base_prec_type.F90)
enum { CYC_ADD = 1, CYC_MULT = 2, CYC_V = 3, CYC_W = 4, CYC_K = 5, CYC_KSYM = 6 };

}  // namespace

extern "C" {
struct orc_params {
  double theta;          // AGGR_THRESH
  int min_coarse_size;   // MIN_COARSE_SIZE (<0: library default)
  int max_levs;          // MAX_LEVS
  double min_cr_ratio;   // MIN_CR_RATIO
  int aggr_omega_alg;    // 0 = eig. estimate (max norm), 999 = user choice
  double aggr_omega_val; // used when aggr_omega_alg == 999
  int aggr_filter;       // 0 = no filter, 1 = filter
  int aggr_ord;          // 0 = natural, 1 = descending degree
  int ml_cycle;          // 1 ADD, 2 MULT (= V), 3 V, 4 W, 5 K (GCR inner), 6 K symmetric (FCG inner)
  int outer_sweeps;      // OUTER_SWEEPS
  int smoother;          // smoother at levels 1..nlev-1
  int sweeps_pre;
  int sweeps_post;
  int c_smoother;        // smoother at the coarsest level
  int c_sweeps;          // COARSE_SWEEPS
  int sub_fill;          // SUB_FILLIN (ILU levels)
  int sub_milu;          // SUB_SOLVE MILU
  int c_fill;            // COARSE_FILLIN
  int c_milu;            // COARSE_SUBSOLVE MILU
};
}

namespace {

// ---------------------------------------------------------------------------
// PSBLAS-semantics kernels (PSBLAS order assumed).
// ---------------------------------------------------------------------------

// Row loops of the products and elementwise updates are split over OpenMP
// threads (rows of at least PAR_MIN entries): each row is still summed alone,
// in stored order, so every result is the same whatever the thread count.
// Reductions (dot) and substitutions stay sequential.
constexpr ix PAR_MIN = 32768;

// psb_spmm(alpha, A, x, beta, y): y_i = beta*y_i + alpha*sum_j a_ij x_j, the
// row sum accumulated in stored order starting from zero.
void spmv(const Csr& a, double alpha, const double* x, double beta, double* y) {
#pragma omp parallel for schedule(static) if (a.nr > PAR_MIN)
  for (ix i = 0; i < a.nr; ++i) {
    double acc = 0.0;
    for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) acc = acc + a.val[k] * x[a.ja[k]];
    if (beta == 0.0) {
      y[i] = (alpha == 1.0) ? acc : (alpha == -1.0 ? -acc : alpha * acc);
    } else if (beta == 1.0) {
      y[i] = (alpha == 1.0) ? y[i] + acc : (alpha == -1.0 ? y[i] - acc : y[i] + alpha * acc);
    } else {
      y[i] = beta * y[i] + alpha * acc;
    }
  }
}

// Transposed product y = beta*y + alpha*A^T x (column scatter in row order).
void spmv_t(const Csr& a, double alpha, const double* x, double beta, double* y) {
  std::vector<double> acc(a.nc, 0.0);
  for (ix i = 0; i < a.nr; ++i)
    for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) acc[a.ja[k]] = acc[a.ja[k]] + a.val[k] * x[i];
  for (ix j = 0; j < a.nc; ++j) {
    if (beta == 0.0) y[j] = (alpha == 1.0) ? acc[j] : (alpha == -1.0 ? -acc[j] : alpha * acc[j]);
    else if (beta == 1.0) y[j] = (alpha == 1.0) ? y[j] + acc[j] : (alpha == -1.0 ? y[j] - acc[j] : y[j] + alpha * acc[j]);
    else y[j] = beta * y[j] + alpha * acc[j];
  }
}

// psb_geaxpby(alpha, x, beta, y): y = alpha*x + beta*y.
void axpby(ix n, double alpha, const double* x, double beta, double* y) {
#pragma omp parallel for schedule(static) if (n > PAR_MIN)
  for (ix i = 0; i < n; ++i) {
    if (beta == 0.0) y[i] = (alpha == 1.0) ? x[i] : alpha * x[i];
    else if (alpha == 1.0 && beta == 1.0) y[i] = x[i] + y[i];
    else y[i] = alpha * x[i] + beta * y[i];
  }
}

// psb_d_vect%mlt(alpha, d, x, beta): y = alpha*d*x + beta*y, y not read when
// beta == 0 (mld_d_diag_solver_apply_vect.f90:102).
void vmlt(ix n, double alpha, const double* d, const double* x, double beta, double* y) {
#pragma omp parallel for schedule(static) if (n > PAR_MIN)
  for (ix i = 0; i < n; ++i) {
    double t = (alpha == 1.0) ? d[i] * x[i] : alpha * d[i] * x[i];
    if (beta == 0.0) y[i] = t;
    else if (beta == 1.0) y[i] = y[i] + t;
    else y[i] = t + beta * y[i];
  }
}

double dot(ix n, const double* x, const double* y) {
  double s = 0.0;
  for (ix i = 0; i < n; ++i) s = s + x[i] * y[i];
  return s;
}
double nrm2(ix n, const double* x) { return std::sqrt(dot(n, x, x)); }

// psb_spspmm(A, B, C): Gustavson product.  For each output entry the terms
// are added in the order of A's stored row (SMMP numbmm order, assumed).
// Output columns sorted ascending.
Csr spspmm(const Csr& a, const Csr& b) {
  Csr c;
  c.nr = a.nr; c.nc = b.nc;
  c.irp.assign(a.nr + 1, 0);
  std::vector<ix> mark(b.nc, -1);
  std::vector<double> acc(b.nc, 0.0);
  std::vector<int> cols;
  for (ix i = 0; i < a.nr; ++i) {
    cols.clear();
    for (ix ka = a.irp[i]; ka < a.irp[i + 1]; ++ka) {
      int k = a.ja[ka];
      double ak = a.val[ka];
      for (ix kb = b.irp[k]; kb < b.irp[k + 1]; ++kb) {
        int j = b.ja[kb];
        if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cols.push_back(j); }
        acc[j] = acc[j] + ak * b.val[kb];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int j : cols) { c.ja.push_back(j); c.val.push_back(acc[j]); }
    c.irp[i + 1] = (ix)c.ja.size();
  }
  return c;
}

// Transpose with columns ascending in each output row (COO transp + cscnv).
Csr transpose(const Csr& a) {
  Csr t;
  t.nr = a.nc; t.nc = a.nr;
  t.irp.assign(t.nr + 1, 0);
  for (ix k = 0; k < a.nnz(); ++k) t.irp[a.ja[k] + 1]++;
  for (ix i = 0; i < t.nr; ++i) t.irp[i + 1] += t.irp[i];
  t.ja.resize(a.nnz()); t.val.resize(a.nnz());
  std::vector<ix> pos(t.irp.begin(), t.irp.end() - 1);
  for (ix i = 0; i < a.nr; ++i)
    for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      ix p = pos[a.ja[k]]++;
      t.ja[p] = (int)i; t.val[p] = a.val[k];
    }
  return t;
}

// a%get_diag: diagonal entries, 0 where not stored.
std::vector<double> get_diag(const Csr& a) {
  std::vector<double> d(a.nr, 0.0);
  for (ix i = 0; i < a.nr; ++i)
    for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k)
      if (a.ja[k] == i) d[i] = a.val[k];
  return d;
}

// ---------------------------------------------------------------------------
// Problem generator: tests/pdegen/mld_d_pde3d.f90:98-412 with a1=a2=a3=1,
// b1=b2=b3=c=0.  h = 1/(idim+2) (:165); global row (ix-1)*idim^2+(iy-1)*idim+iz
// (:258-268); entries inserted x-1, y-1, z-1, centre, z+1, y+1, x+1
// (:278-336), which is already ascending column order.
// ---------------------------------------------------------------------------
Csr gen_pde3d(int idim) {
  Csr a;
  ix m = (ix)idim * idim * idim;
  a.nr = a.nc = m;
  a.irp.assign(m + 1, 0);
  a.ja.reserve(7 * m); a.val.reserve(7 * m);
  double deltah = 1.0 / (idim + 2);
  double sqdeltah = deltah * deltah;
  double off = -1.0 / sqdeltah - 0.0 / (2.0 * deltah);   // -a1/h^2 - b1/(2h)
  double cen = 2.0 * (1.0 + 1.0 + 1.0) / sqdeltah + 0.0;  // 2(a1+a2+a3)/h^2 + c
  ix id2 = (ix)idim * idim;
  for (ix g = 0; g < m; ++g) {
    ix x = g / id2, y = (g / idim) % idim, z = g % idim;
    if (x > 0) { a.ja.push_back((int)(g - id2)); a.val.push_back(off); }
    if (y > 0) { a.ja.push_back((int)(g - idim)); a.val.push_back(off); }
    if (z > 0) { a.ja.push_back((int)(g - 1)); a.val.push_back(off); }
    a.ja.push_back((int)g); a.val.push_back(cen);
    if (z < idim - 1) { a.ja.push_back((int)(g + 1)); a.val.push_back(off); }
    if (y < idim - 1) { a.ja.push_back((int)(g + idim)); a.val.push_back(off); }
    if (x < idim - 1) { a.ja.push_back((int)(g + id2)); a.val.push_back(off); }
    a.irp[g + 1] = (ix)a.ja.size();
  }
  return a;
}

// ---------------------------------------------------------------------------
// Decoupled aggregation: mlprec/impl/mld_d_dec_map_bld.f90:120-312.
// Returns 0-based aggregate ids; naggr set.
// ---------------------------------------------------------------------------
int dec_map_bld(int iorder, double theta, const Csr& a, std::vector<int>& ilaggr, int& naggr) {
  ix nr = a.nr;
  const int UNAGG = -(int)(nr + 1);
  std::vector<double> diag = get_diag(a);
  ilaggr.assign(nr, UNAGG);
  std::vector<ix> idxs(nr);
  for (ix i = 0; i < nr; ++i) idxs[i] = i;
  if (iorder == 1) {  // mld_aggr_ord_desc_deg_: stable sort by degree, descending (:127-135)
    std::vector<ix> deg(nr);
    for (ix i = 0; i < nr; ++i) deg[i] = a.irp[i + 1] - a.irp[i];
    std::stable_sort(idxs.begin(), idxs.end(), [&](ix p, ix q) { return deg[p] > deg[q]; });
  }
  auto strong = [&](ix i, ix k) {
    int j = a.ja[k];
    return std::fabs(a.val[k]) > theta * std::sqrt(std::fabs(diag[i] * diag[j]));
  };
  naggr = 0;
  std::vector<int> icol;
  // Phase one (:148-189): a node whose strong neighbourhood is entirely
  // unaggregated starts a new aggregate made of that neighbourhood.
  for (ix ii = 0; ii < nr; ++ii) {
    ix i = idxs[ii];
    if (ilaggr[i] != UNAGG) continue;
    icol.clear();
    for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      int j = a.ja[k];
      if (j >= 0 && j < nr && strong(i, k)) icol.push_back(j);
    }
    bool disjoint = true;
    for (int j : icol) if (ilaggr[j] != UNAGG) { disjoint = false; break; }
    if (disjoint) {
      naggr++;
      for (int j : icol) ilaggr[j] = naggr;
      ilaggr[i] = naggr;
    }
  }
  // Phase two (:196-230): join the strongest already-aggregated neighbour,
  // judged on the snapshot tmpaggr; first maximum wins.
  std::vector<int> tmpaggr = ilaggr;
  for (ix ii = 0; ii < nr; ++ii) {
    ix i = idxs[ii];
    if (ilaggr[i] != UNAGG) continue;
    double cpling = 0.0;
    ix ip = -1;
    for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      int j = a.ja[k];
      if (j >= 0 && j < nr) {
        if (strong(i, k) && tmpaggr[j] > 0 && std::fabs(a.val[k]) > cpling) {
          ip = k;
          cpling = std::fabs(a.val[k]);
        }
      }
    }
    if (ip >= 0) ilaggr[i] = ilaggr[a.ja[ip]];
  }
  // Phase three (:236-278): leftovers form aggregates with their
  // unaggregated strong neighbours.
  for (ix ii = 0; ii < nr; ++ii) {
    ix i = idxs[ii];
    if (ilaggr[i] >= 0) continue;
    icol.clear();
    for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      int j = a.ja[k];
      if (j >= 0 && j < nr && strong(i, k) && ilaggr[j] < 0) icol.push_back(j);
    }
    naggr++;
    ilaggr[i] = naggr;
    for (int j : icol) ilaggr[j] = naggr;
  }
  for (ix i = 0; i < nr; ++i) {
    if (ilaggr[i] < 0) return -1;  // "Fatal error: some leftovers" (:281-285)
    ilaggr[i] -= 1;                // to 0-based
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Smoothed-aggregation level operators: mlprec/impl/mld_daggrmat_smth_asb.f90
// :159-401 (single process: no halo, all aggregates local).
// ---------------------------------------------------------------------------
struct LevelOps {
  Csr ac, p, r;
  double omega = 0.0;
};

int aggrmat_smth_asb(const Csr& a, const std::vector<int>& ilaggr, int naggr, const orc_params& prm,
                     LevelOps& out) {
  ix nrow = a.nr;
  std::vector<double> adiag = get_diag(a);  // :174
  Csr acsr3 = a;
  Csr acsrf;
  bool filter_mat = (prm.aggr_filter == 1);
  if (filter_mat) {  // :194-217 filtered matrix Af
    acsrf = a;
    for (ix i = 0; i < nrow; ++i) {
      double tmp = 0.0;
      ix jd = -1;
      for (ix j = acsrf.irp[i]; j < acsrf.irp[i + 1]; ++j) {
        if (acsrf.ja[j] == i) jd = j;
        if (std::fabs(acsrf.val[j]) < prm.theta * std::sqrt(std::fabs(adiag[i] * adiag[acsrf.ja[j]]))) {
          tmp = tmp + acsrf.val[j];
          acsrf.val[j] = 0.0;
        }
      }
      if (jd >= 0) acsrf.val[jd] = acsrf.val[jd] - tmp;
    }
    // clean_zeros
    Csr c; c.nr = acsrf.nr; c.nc = acsrf.nc; c.irp.assign(c.nr + 1, 0);
    for (ix i = 0; i < acsrf.nr; ++i) {
      for (ix j = acsrf.irp[i]; j < acsrf.irp[i + 1]; ++j)
        if (acsrf.val[j] != 0.0) { c.ja.push_back(acsrf.ja[j]); c.val.push_back(acsrf.val[j]); }
      c.irp[i + 1] = (ix)c.ja.size();
    }
    acsrf = c;
  }
  for (auto& d : adiag) d = (d != 0.0) ? 1.0 / d : 1.0;  // :220-226
  auto scal = [&](Csr& m) {                            // row scaling a_ij *= d_i
    for (ix i = 0; i < m.nr; ++i)
      for (ix j = m.irp[i]; j < m.irp[i + 1]; ++j) m.val[j] = m.val[j] * adiag[i];
  };
  if (filter_mat) scal(acsrf);
  scal(acsr3);
  double omega;
  if (prm.aggr_omega_alg == 0) {  // :233-246 omega = 4/(3*||D^-1 A||_inf)
    double anorm = 0.0;
    for (ix i = 0; i < acsr3.nr; ++i) {
      double s = 0.0;
      for (ix j = acsr3.irp[i]; j < acsr3.irp[i + 1]; ++j) s = s + std::fabs(acsr3.val[j]);
      anorm = std::max(anorm, s);
    }
    omega = 4.0 / (3.0 * anorm);
  } else {
    omega = prm.aggr_omega_val;
  }
  Csr& sm = filter_mat ? acsrf : acsr3;  // :258-316 (I - omega D^-1 A)
  for (ix i = 0; i < sm.nr; ++i)
    for (ix j = sm.irp[i]; j < sm.irp[i + 1]; ++j)
      sm.val[j] = (sm.ja[j] == i) ? 1.0 - omega * sm.val[j] : -(omega * sm.val[j]);
  // Tentative prolongator Ptilde: one unit entry per row (mld_daggrmap_bld.f90:171-180).
  Csr ptilde;
  ptilde.nr = nrow; ptilde.nc = naggr;
  ptilde.irp.resize(nrow + 1);
  ptilde.ja.resize(nrow); ptilde.val.assign(nrow, 1.0);
  for (ix i = 0; i <= nrow; ++i) ptilde.irp[i] = i;
  for (ix i = 0; i < nrow; ++i) ptilde.ja[i] = ilaggr[i];
  out.p = spspmm(sm, ptilde);        // spspmm 1 (:291,316)
  Csr am3 = spspmm(a, out.p);        // spspmm 2, A*P (:345)
  out.r = transpose(out.p);          // R = P^T on local aggregates (:355-375)
  out.ac = spspmm(out.r, am3);       // spspmm 3 (:399)
  out.omega = omega;
  return 0;
}

// ---------------------------------------------------------------------------
// ILU(0): mlprec/impl/mld_dilu0_fact.f90:284-458 (IKJ, single block).
// L unit lower (strict part stored), U unit upper (strict part, scaled by
// 1/pivot), d = 1/pivot.
// ---------------------------------------------------------------------------
int ilu0_fact(const Csr& a, Csr& l, Csr& u, std::vector<double>& d, bool milu = false) {
  ix m = a.nr;
  l.nr = l.nc = u.nr = u.nc = m;
  l.irp.assign(m + 1, 0); u.irp.assign(m + 1, 0);
  l.ja.clear(); l.val.clear(); u.ja.clear(); u.val.clear();
  d.assign(m, 0.0);
  for (ix i = 0; i < m; ++i) {
    d[i] = 0.0;
    // ilu_copyin (:584-598): split row i into L, diagonal, U; columns clipped to [0,m)
    for (ix j = a.irp[i]; j < a.irp[i + 1]; ++j) {
      ix k = a.ja[j];
      if (k < i && k >= 0) { l.ja.push_back((int)k); l.val.push_back(a.val[j]); }
      else if (k == i) d[i] = a.val[j];
      else if (k > i && k < m) { u.ja.push_back((int)k); u.val.push_back(a.val[j]); }
    }
    l.irp[i + 1] = (ix)l.ja.size();
    u.irp[i + 1] = (ix)u.ja.size();
    double dia = d[i];
    for (ix kk = l.irp[i]; kk < l.irp[i + 1]; ++kk) {
      double temp = l.val[kk];
      ix k = l.ja[kk];
      l.val[kk] = temp * d[k];  // :370-372
      ix low1 = kk + 1, low2 = u.irp[i];
      for (ix jj = u.irp[k]; jj < u.irp[k + 1]; ++jj) {  // updateloop :380-431
        ix j = u.ja[jj];
        bool matched = false;
        if (j < i) {
          for (ix ll = low1; ll < l.irp[i + 1]; ++ll) {
            ix lj = l.ja[ll];
            if (lj > j) { low1 = ll; break; }
            if (lj == j) { l.val[ll] = l.val[ll] - temp * u.val[jj]; low1 = ll + 1; matched = true; break; }
          }
        } else if (j == i) {
          dia = dia - temp * u.val[jj];
          matched = true;
        } else {
          for (ix ll = low2; ll < u.irp[i + 1]; ++ll) {
            ix lj = u.ja[ll];
            if (lj > j) { low2 = ll; break; }
            if (lj == j) { u.val[ll] = u.val[ll] - temp * u.val[jj]; low2 = ll + 1; matched = true; break; }
          }
        }
        if (!matched && milu) dia = dia - temp * u.val[jj];
      }
    }
    // pivot check (:436-444); d_epstol taken as machine epsilon
    if (std::fabs(dia) < 2.220446049250313e-16) return (int)(i + 1);
    dia = 1.0 / dia;
    d[i] = dia;
    for (ix kk = u.irp[i]; kk < u.irp[i + 1]; ++kk) u.val[kk] = u.val[kk] * dia;  // :455-457
  }
  return 0;
}

// ---------------------------------------------------------------------------
// ILU(k) and MILU(k): mlprec/impl/mld_diluk_fact.f90 (mld_diluk_factint
// :268-419).  Row i in a dense work row, A's entries at level 0 (iluk_copyin
// :485-566); columns visited once each in ascending order, fill columns
// joining the set as they appear (the heap of iluk_fact :637-722): a visited
// column k < i with a nonzero value and level <= fill gives l_ik = w_k d_k
// and w_j -= w_k u_kj, lev_j = min(lev_j, lev_k + lev(u_kj) + 1) over row k of
// U.  iluk_copyout (:812-960): level <= fill kept, else dropped (MILU: added
// to the pivot, in visiting order); d_i = 1/pivot; row i of U times d_i.
// ---------------------------------------------------------------------------
int iluk_fact(const Csr& a, int fill, bool milu, Csr& l, Csr& u, std::vector<double>& d) {
  ix m = a.nr;
  l.nr = l.nc = u.nr = u.nc = m;
  l.irp.assign(m + 1, 0); u.irp.assign(m + 1, 0);
  l.ja.clear(); l.val.clear(); u.ja.clear(); u.val.clear();
  d.assign(m, 0.0);
  std::vector<long> ulev;
  std::vector<double> w(m, 0.0);
  std::vector<long> lev(m, 0);
  std::vector<char> in(m, 0);
  for (ix i = 0; i < m; ++i) {
    std::set<ix> cols;
    for (ix q = a.irp[i]; q < a.irp[i + 1]; ++q) {
      ix k = a.ja[q];
      if (k < 0 || k >= m) continue;
      w[k] = a.val[q]; lev[k] = 0; in[k] = 1;
      cols.insert(k);
    }
    for (auto it = cols.begin(); it != cols.end(); ++it) {
      ix k = *it;
      if (!(w[k] != 0.0 && lev[k] <= fill && k < i)) continue;
      double rk = w[k];
      w[k] = w[k] * d[k];
      for (ix jj = u.irp[k]; jj < u.irp[k + 1]; ++jj) {
        ix j = u.ja[jj];
        if (!in[j]) { in[j] = 1; lev[j] = m + 1; cols.insert(j); }
        w[j] = w[j] - rk * u.val[jj];
        lev[j] = std::min(lev[j], lev[k] + ulev[jj] + 1);
      }
    }
    double dia = 0.0;
    for (ix j : cols) {
      if (j < i) {
        if (lev[j] <= fill) { l.ja.push_back((int)j); l.val.push_back(w[j]); }
        else if (milu) dia = dia + w[j];
      } else if (j == i) {
        dia = dia + w[j];
      } else {
        if (lev[j] <= fill) { u.ja.push_back((int)j); u.val.push_back(w[j]); ulev.push_back(lev[j]); }
        else if (milu) dia = dia + w[j];
      }
      w[j] = 0.0; in[j] = 0;
    }
    l.irp[i + 1] = (ix)l.ja.size();
    u.irp[i + 1] = (ix)u.ja.size();
    if (std::fabs(dia) < 2.220446049250313e-16) return (int)(i + 1);
    dia = 1.0 / dia;
    d[i] = dia;
    for (ix q = u.irp[i]; q < u.irp[i + 1]; ++q) u.val[q] = dia * u.val[q];
  }
  return 0;
}

// psb_spsm on a unit triangular CSR factor, PSBLAS order assumed:
//   lower : t_i = x_i - sum_{stored order} l_ij t_j,  i ascending
//   upper : t_i = x_i - sum_{stored order} u_ij t_j,  i descending
void trsv_unit(const Csr& t, bool lower, const double* x, double* y) {
  if (lower) {
    for (ix i = 0; i < t.nr; ++i) {
      double acc = 0.0;
      for (ix k = t.irp[i]; k < t.irp[i + 1]; ++k) acc = acc + t.val[k] * y[t.ja[k]];
      y[i] = x[i] - acc;
    }
  } else {
    for (ix i = t.nr - 1; i >= 0; --i) {
      double acc = 0.0;
      for (ix k = t.irp[i]; k < t.irp[i + 1]; ++k) acc = acc + t.val[k] * y[t.ja[k]];
      y[i] = x[i] - acc;
    }
  }
}
// Transposed unit solve with a CSR factor (column-oriented scatter).
//   lower^T is upper: i descending;  upper^T is lower: i ascending.
void trsv_unit_t(const Csr& t, bool lower, const double* x, double* y) {
  for (ix i = 0; i < t.nr; ++i) y[i] = x[i];
  if (!lower) {  // U^T: forward
    for (ix i = 0; i < t.nr; ++i)
      for (ix k = t.irp[i]; k < t.irp[i + 1]; ++k) y[t.ja[k]] = y[t.ja[k]] - t.val[k] * y[i];
  } else {       // L^T: backward
    for (ix i = t.nr - 1; i >= 0; --i)
      for (ix k = t.irp[i]; k < t.irp[i + 1]; ++k) y[t.ja[k]] = y[t.ja[k]] - t.val[k] * y[i];
  }
}

// Dense LU with partial pivoting (coarsest-level DENSE_LU solver; no
// reference counterpart).  Row-major n x n.
struct DenseLU {
  int n = 0;
  std::vector<double> lu;
  std::vector<int> piv;
  int factor(const Csr& a) {
    n = (int)a.nr;
    lu.assign((size_t)n * n, 0.0);
    piv.resize(n);
    for (ix i = 0; i < n; ++i)
      for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) lu[(size_t)i * n + a.ja[k]] += a.val[k];
    for (int k = 0; k < n; ++k) {
      int p = k;
      double mx = std::fabs(lu[(size_t)k * n + k]);
      for (int i = k + 1; i < n; ++i)
        if (std::fabs(lu[(size_t)i * n + k]) > mx) { mx = std::fabs(lu[(size_t)i * n + k]); p = i; }
      piv[k] = p;
      if (mx == 0.0) return k + 1;
      if (p != k)
        for (int j = 0; j < n; ++j) std::swap(lu[(size_t)k * n + j], lu[(size_t)p * n + j]);
      double inv = 1.0 / lu[(size_t)k * n + k];
      for (int i = k + 1; i < n; ++i) {
        double f = lu[(size_t)i * n + k] * inv;
        lu[(size_t)i * n + k] = f;
        for (int j = k + 1; j < n; ++j) lu[(size_t)i * n + j] = lu[(size_t)i * n + j] - f * lu[(size_t)k * n + j];
      }
    }
    return 0;
  }
  void solve(const double* b, double* x) const {
    std::vector<double> y(b, b + n);
    for (int k = 0; k < n; ++k) std::swap(y[k], y[piv[k]]);
    for (int i = 0; i < n; ++i) {
      double s = y[i];
      for (int j = 0; j < i; ++j) s = s - lu[(size_t)i * n + j] * y[j];
      y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int j = i + 1; j < n; ++j) s = s - lu[(size_t)i * n + j] * y[j];
      y[i] = s / lu[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) x[i] = y[i];
  }
};

// ---------------------------------------------------------------------------
// Smoother = jacobi smoother + solver (mld_d_jac_smoother.f90 + solvers).
// ---------------------------------------------------------------------------
struct Smoother {
  int kind = SM_JAC;
  int fill = 0;            // ILU: fill level (SUB_FILLIN / COARSE_FILLIN)
  bool milu = false;       // ILU: MILU
  Csr nd;                  // ND (mld_d_jac_smoother_bld.f90:73-79)
  ix nnz_nd = 0;
  std::vector<double> dv;  // diag solver: 1/diag (0 -> 1); ILU: 1/pivot
  Csr l, u;                // ILU(0) factors
  DenseLU dlu;
  // scratch vectors of the applications, kept between calls (every use
  // writes them in full before reading them)
  mutable std::vector<double> sw, st, stx, sty;
  static double* scratch(std::vector<double>& v, ix n) {
    if ((ix)v.size() < n) v.resize(n);
    return v.data();
  }
  // GS and BWGS are iterative (mld_d_gs_solver.f90:586-591): the jacobi
  // smoother then always runs its sweep loop; diag, ILU and LU are not.
  bool iterative() const { return kind == SM_GS || kind == SM_BWGS; }

  int build(const Csr& a) {
    ix n = a.nr;
    if (kind == SM_NOPREC) return 0;
    if (kind == SM_JAC || kind == SM_L1JAC) {
      // diag solver build: mld_d_diag_solver_bld.f90:74-92
      std::vector<double> d = get_diag(a);
      if (kind == SM_L1JAC) {
        // l1 diagonal: a_ii + sum_{j != i} |a_ij| (stored order); ND keeps
        // the off-diagonals of A and carries -sum_{j != i}|a_ij| on the
        // diagonal, so that ND = A - D_l1.
        nd.nr = nd.nc = n; nd.irp.assign(n + 1, 0); nd.ja.clear(); nd.val.clear();
        for (ix i = 0; i < n; ++i) {
          double s = 0.0;
          bool has = false;
          for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k)
            if (a.ja[k] != i) s = s + std::fabs(a.val[k]); else has = true;
          d[i] = d[i] + s;
          bool placed = false;
          for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) {
            if (!has && !placed && a.ja[k] > i) { nd.ja.push_back((int)i); nd.val.push_back(-s); placed = true; }
            if (a.ja[k] == i) { nd.ja.push_back((int)i); nd.val.push_back(-s); placed = true; }
            else { nd.ja.push_back(a.ja[k]); nd.val.push_back(a.val[k]); }
          }
          if (!placed) { nd.ja.push_back((int)i); nd.val.push_back(-s); }
          nd.irp[i + 1] = (ix)nd.ja.size();
        }
      } else {
        // clip_diag: ND = A without its diagonal entries
        nd.nr = nd.nc = n; nd.irp.assign(n + 1, 0); nd.ja.clear(); nd.val.clear();
        for (ix i = 0; i < n; ++i) {
          for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k)
            if (a.ja[k] != i) { nd.ja.push_back(a.ja[k]); nd.val.push_back(a.val[k]); }
          nd.irp[i + 1] = (ix)nd.ja.size();
        }
      }
      dv.resize(n);
      for (ix i = 0; i < n; ++i) dv[i] = (d[i] == 0.0) ? 1.0 : 1.0 / d[i];
      nnz_nd = nd.nnz();
      return 0;
    }
    if (kind == SM_BJAC) {
      // csclip jmin = nrow+1: empty on one process; ILU(0) by its own
      // routine, ILU(k>0) and MILU(k) by iluk (mld_d_ilu_solver_bld.f90:126-160)
      nd.nr = nd.nc = n; nd.irp.assign(n + 1, 0); nnz_nd = 0;
      if (fill == 0 && !milu) return ilu0_fact(a, l, u, dv);
      return iluk_fact(a, fill, milu, l, u, dv);
    }
    if (kind == SM_DENSELU) {
      nd.nr = nd.nc = n; nd.irp.assign(n + 1, 0); nnz_nd = 0;
      return dlu.factor(a);
    }
    if (kind == SM_GS || kind == SM_BWGS) {
      // mld_d_gs_solver_bld.f90 / mld_d_bwgs_solver_bld.f90: A split into its
      // triangle with the diagonal (solved) and the strict other triangle;
      // kept here as strict lower l, strict upper u and the diagonal dv.
      // ND (csclip jmin = nrow+1) is empty on one process.
      nd.nr = nd.nc = n; nd.irp.assign(n + 1, 0); nnz_nd = 0;
      l.nr = l.nc = u.nr = u.nc = n;
      l.irp.assign(n + 1, 0); u.irp.assign(n + 1, 0);
      l.ja.clear(); l.val.clear(); u.ja.clear(); u.val.clear();
      dv.assign(n, 0.0);
      for (ix i = 0; i < n; ++i) {
        for (ix k = a.irp[i]; k < a.irp[i + 1]; ++k) {
          if (a.ja[k] < i) { l.ja.push_back(a.ja[k]); l.val.push_back(a.val[k]); }
          else if (a.ja[k] > i) { u.ja.push_back(a.ja[k]); u.val.push_back(a.val[k]); }
          else dv[i] = dv[i] + a.val[k];
        }
        l.irp[i + 1] = (ix)l.ja.size();
        u.irp[i + 1] = (ix)u.ja.size();
      }
      return 0;
    }
    return -1;
  }

  // One sweep of the GS solver (mld_d_gs_solver_apply_vect.f90:150-156):
  // wv = b - O xit (O the strict other triangle); xit = T^-1 wv, T the
  // triangle with the diagonal, non-unit substitution with the off-diagonal
  // products summed in stored order, then divided by the diagonal.
  // GS: T lower, rows ascending; BWGS: T upper, rows descending.
  void gs_sweep(const double* b, double* xit, ix n) const {
    const bool fw = (kind == SM_GS);
    const Csr& tri = fw ? l : u;
    const Csr& oth = fw ? u : l;
    std::vector<double> w(b, b + n);
    spmv(oth, -1.0, xit, 1.0, w.data());
    for (ix q = 0; q < n; ++q) {
      const ix i = fw ? q : n - 1 - q;
      double acc = 0.0;
      for (ix k = tri.irp[i]; k < tri.irp[i + 1]; ++k) acc = acc + tri.val[k] * xit[tri.ja[k]];
      xit[i] = (w[i] - acc) / dv[i];
    }
  }

  // Solver apply: y = alpha*S^-1 x + beta*y.
  void sv_apply(double alpha, const double* x, double beta, double* y, char trans, ix n) const {
    if (kind == SM_NOPREC) { axpby(n, alpha, x, beta, y); return; }
    if (kind == SM_JAC || kind == SM_L1JAC) { vmlt(n, alpha, dv.data(), x, beta, y); return; }
    double* w = scratch(sw, n);
    double* t = scratch(st, n);
    if (kind == SM_DENSELU) {
      dlu.solve(x, t);
      axpby(n, alpha, t, beta, y);
      return;
    }
    // mld_d_ilu_solver_apply_vect.f90:130-142
    if (trans == 'N') {
      trsv_unit(l, true, x, t);
#pragma omp parallel for schedule(static) if (n > PAR_MIN)
      for (ix i = 0; i < n; ++i) w[i] = dv[i] * t[i];
      trsv_unit(u, false, w, t);
    } else {
      trsv_unit_t(u, false, x, t);
      for (ix i = 0; i < n; ++i) w[i] = dv[i] * t[i];
      trsv_unit_t(l, true, w, t);
    }
    axpby(n, alpha, t, beta, y);
  }

  // mld_d_jac_smoother_apply_vect.f90:38-215 with init='Z'
  // (mld_d_base_smoother_apply_vect.f90:60-66 for NOPREC).
  int apply(double alpha, const double* x, double beta, double* y, char trans, int sweeps, ix n) const {
    if (kind == SM_NOPREC) { axpby(n, alpha, x, beta, y); return 0; }
    if (iterative()) {
      // only 'N' (mld_d_gs_solver_apply_vect.f90:81-87); the sweep loop of
      // :133-175 with init 'Z' then 'Y' and an empty ND: every smoother sweep
      // is one GS sweep (solver sweeps = 1) on x, starting from zero; the
      // 'Z' application comes first, so sweeps = 0 also gives one sweep
      if (trans != 'N') return -3;
      if (sweeps < 0) return -2;
      std::vector<double> ty(n, 0.0);
      for (int i = 0; i < std::max(1, sweeps); ++i) gs_sweep(x, ty.data(), n);
      axpby(n, alpha, ty.data(), beta, y);
      return 0;
    }
    if (!iterative() && (sweeps == 1 || nnz_nd == 0)) {  // :107-109
      sv_apply(alpha, x, beta, y, trans, n);
      return 0;
    }
    if (sweeps < 0) return -2;
    double* tx = scratch(stx, n);
    double* ty = scratch(sty, n);
    sv_apply(1.0, x, 0.0, ty, trans, n);  // init 'Z' (:136)
    for (int i = 1; i <= sweeps - 1; ++i) {      // :161-175
      axpby(n, 1.0, x, 0.0, tx);
      if (trans == 'N') spmv(nd, -1.0, ty, 1.0, tx);
      else spmv_t(nd, -1.0, ty, 1.0, tx);
      sv_apply(1.0, tx, 0.0, ty, trans, n);
    }
    axpby(n, alpha, ty, beta, y);  // :177
    return 0;
  }
};

struct Level {
  Csr a;                 // base_a
  Csr p, r;              // map stored at this level (P: n_{l-1} x n_l, R: n_l x n_{l-1})
  std::vector<int> ilaggr;
  int naggr = 0;
  double omega = 0.0;
  int sweeps_pre = 1, sweeps_post = 1;
  Smoother sm;
  Smoother sm2;          // post smoother when it differs (FBGS: BWGS), else unused
  bool has_sm2 = false;
  // apply work (mld_dmlprec_aply.f90:249-275, allocated once here)
  std::vector<double> x2l, y2l, tx, ty;
  // K-cycle inner iteration at this level (mld_dinneritkcycle): w, v, v1, x, d(0:1)
  std::vector<double> kw, kv, kv1, kx, kd0, kd1;
};

struct Prec {
  orc_params prm;
  std::vector<Level> lv;
  int casize = 0;

  void zero_y_if_needed(int l) { std::fill(lv[l].y2l.begin(), lv[l].y2l.end(), 0.0); }

  // inner_ml_aply (mld_dmlprec_aply.f90:382-464): dispatch on the level's cycle.
  int inner_ml_aply(int level, char trans) {
    switch (prm.ml_cycle) {
      case CYC_ADD: return inner_add(level, trans);
      case CYC_MULT: case CYC_V: case CYC_W: return inner_mult(level, trans);
      case CYC_K: case CYC_KSYM: return inner_k_cycle(level, trans);
      default: return -4;
    }
  }

  // mld_d_inner_add (mld_dmlprec_aply.f90:467-588): smoother on x2l (the
  // sm/sm2a pair max(pre, post) times when sm2a exists, from y2l = x2l), then
  // y2l += P (cycle on R x2l).
  int inner_add(int level, char trans) {
    int nlev = (int)lv.size();
    Level& L = lv[level];
    ix n = L.a.nr;
    if (L.has_sm2) {
      axpby(n, 1.0, L.x2l.data(), 0.0, L.y2l.data());
      int sw = std::max(L.sweeps_pre, L.sweeps_post);
      for (int k = 0; k < sw; ++k) {
        int rc = L.sm.apply(1.0, L.y2l.data(), 0.0, L.tx.data(), trans, 1, n);
        if (rc) return rc;
        rc = L.sm2.apply(1.0, L.tx.data(), 0.0, L.y2l.data(), trans, 1, n);
        if (rc) return rc;
      }
    } else {
      int rc = L.sm.apply(1.0, L.x2l.data(), 0.0, L.y2l.data(), trans, L.sweeps_pre, n);
      if (rc) return rc;
    }
    if (level < nlev - 1) {
      Level& C = lv[level + 1];
      spmv(C.r, 1.0, L.x2l.data(), 0.0, C.x2l.data());
      int rc = inner_ml_aply(level + 1, trans);
      if (rc) return rc;
      spmv(C.p, 1.0, C.y2l.data(), 1.0, L.y2l.data());
    }
    return 0;
  }

  // mld_d_inner_k_cycle (mld_dmlprec_aply.f90:811-1014): smoothing with no
  // pre/post switches, and on levels whose next level is not the coarsest
  // the next level's correction comes from two steps of a flexible Krylov
  // method preconditioned by that level's own cycle.
  int inner_k_cycle(int level, char trans) {
    int nlev = (int)lv.size();
    Level& L = lv[level];
    ix n = L.a.nr;
    if (level == nlev - 1)
      return L.sm.apply(1.0, L.x2l.data(), 0.0, L.y2l.data(), trans, L.sweeps_pre, n);  // :867-873
    const Smoother& s2 = L.has_sm2 ? L.sm2 : L.sm;
    int rc = (trans == 'N') ? L.sm.apply(1.0, L.x2l.data(), 0.0, L.y2l.data(), trans, L.sweeps_pre, n)
                            : s2.apply(1.0, L.x2l.data(), 0.0, L.y2l.data(), trans, L.sweeps_post, n);  // :877-893
    if (rc) return rc;
    axpby(n, 1.0, L.x2l.data(), 0.0, L.ty.data());  // :903-912
    if (trans == 'N') spmv(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
    else spmv_t(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
    Level& C = lv[level + 1];
    spmv(C.r, 1.0, L.ty.data(), 0.0, C.x2l.data());  // :918-920
    if (level <= nlev - 3) rc = inneritkcycle(level + 1, trans, prm.ml_cycle == CYC_KSYM);  // :929-942
    else rc = inner_ml_aply(level + 1, trans);
    if (rc) return rc;
    spmv(C.p, 1.0, C.y2l.data(), 1.0, L.y2l.data());  // :953-955
    axpby(n, 1.0, L.x2l.data(), 0.0, L.ty.data());    // :966-972
    if (trans == 'N') spmv(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
    else spmv_t(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
    return (trans == 'N') ? s2.apply(1.0, L.ty.data(), 1.0, L.y2l.data(), trans, L.sweeps_post, n)
                          : L.sm.apply(1.0, L.ty.data(), 1.0, L.y2l.data(), trans, L.sweeps_pre, n);  // :982-993
  }

  // mld_dinneritkcycle (mld_dmlprec_aply.f90:1017-1184): at most two steps of
  // FCG (fcg) or GCR from x = 0 on A_l x = x2l, preconditioned by
  // inner_ml_aply(level); stops after one step when ||w|| <= 0.25 ||x2l||.
  // The products with A_l are untransposed for 'T' as well (:1095, :1137).
  int inneritkcycle(int level, char trans, bool fcg) {
    Level& L = lv[level];
    ix n = L.a.nr;
    if (L.kw.size() != (size_t)n) {
      L.kw.assign(n, 0.0); L.kv.assign(n, 0.0); L.kv1.assign(n, 0.0);
      L.kx.assign(n, 0.0); L.kd0.assign(n, 0.0); L.kd1.assign(n, 0.0);
    }
    double* w = L.kw.data(); double* v = L.kv.data(); double* v1 = L.kv1.data();
    double* x = L.kx.data(); double* d0 = L.kd0.data(); double* d1 = L.kd1.data();
    std::fill(L.kx.begin(), L.kx.end(), 0.0);
    axpby(n, 1.0, L.x2l.data(), 0.0, w);
    const double delta0 = nrm2(n, w);
    std::fill(L.y2l.begin(), L.y2l.end(), 0.0);
    int rc = inner_ml_aply(level, trans);
    if (rc) return rc;
    axpby(n, 1.0, L.y2l.data(), 0.0, d0);
    spmv(L.a, 1.0, d0, 0.0, v);
    double delta_old, tau;
    if (fcg) { delta_old = dot(n, d0, w); tau = dot(n, d0, v); }
    else { delta_old = dot(n, v, w); tau = dot(n, v, v); }
    double alpha = delta_old / tau;
    axpby(n, -alpha, v, 1.0, w);
    const double l2 = nrm2(n, w);
    if (l2 <= 0.25 * delta0) {
      axpby(n, alpha, d0, 1.0, x);
    } else {
      axpby(n, 1.0, w, 0.0, L.x2l.data());
      rc = inner_ml_aply(level, trans);
      if (rc) return rc;
      axpby(n, 1.0, L.y2l.data(), 0.0, d1);
      spmv(L.a, 1.0, d1, 0.0, v1);
      double tau1, tau2, tau3;
      if (fcg) { tau1 = dot(n, d1, v); tau2 = dot(n, d1, v1); tau3 = dot(n, d1, w); }
      else { tau1 = dot(n, v1, v); tau2 = dot(n, v1, v1); tau3 = dot(n, v1, w); }
      const double tau4 = tau2 - (tau1 * tau1) / tau;
      alpha = alpha - (tau1 * tau3) / (tau * tau4);
      axpby(n, alpha, d0, 1.0, x);
      alpha = tau3 / tau4;
      axpby(n, alpha, d1, 1.0, x);
    }
    axpby(n, 1.0, x, 0.0, L.y2l.data());
    return 0;
  }

  // mld_d_inner_mult (mld_dmlprec_aply.f90:590-809); level 0-based.
  int inner_mult(int level, char trans) {
    int nlev = (int)lv.size();
    Level& L = lv[level];
    ix n = L.a.nr;
    bool pre = ((L.sweeps_pre > 0) && trans == 'N') || ((L.sweeps_post > 0) && trans != 'N');
    bool post = ((L.sweeps_post > 0) && trans == 'N') || ((L.sweeps_pre > 0) && trans != 'N');
    if (level < nlev - 1) {
      Level& C = lv[level + 1];
      if (pre) {
        int sw = (trans == 'N') ? L.sweeps_pre : L.sweeps_post;
        const Smoother& spre = (trans == 'N' || !L.has_sm2) ? L.sm : L.sm2;
        int rc = spre.apply(1.0, L.x2l.data(), 0.0, L.y2l.data(), trans, sw, n);  // :646-657
        if (rc) return rc;
        axpby(n, 1.0, L.x2l.data(), 0.0, L.ty.data());                             // :669
        if (trans == 'N') spmv(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());         // :674-676
        else spmv_t(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
        spmv(C.r, 1.0, L.ty.data(), 0.0, C.x2l.data());                           // :682-684
      } else {
        // The reference leaves y2l as left by psb_geasb(scratch); it is
        // cleared here so that the prolongation adds into zeros.
        std::fill(L.y2l.begin(), L.y2l.end(), 0.0);
        spmv(C.r, 1.0, L.x2l.data(), 0.0, C.x2l.data());                          // :692-694
      }
      int rc = inner_ml_aply(level + 1, trans);                                    // :702
      if (rc) return rc;
      spmv(C.p, 1.0, C.y2l.data(), 1.0, L.y2l.data());                             // :707-709
      if (prm.ml_cycle == CYC_W) {                                                 // :716-745
        axpby(n, 1.0, L.x2l.data(), 0.0, L.ty.data());
        if (trans == 'N') spmv(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
        else spmv_t(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
        spmv(C.r, 1.0, L.ty.data(), 0.0, C.x2l.data());
        rc = inner_ml_aply(level + 1, trans);
        if (rc) return rc;
        spmv(C.p, 1.0, C.y2l.data(), 1.0, L.y2l.data());
      }
      if (post) {                                                                  // :748-785
        axpby(n, 1.0, L.x2l.data(), 0.0, L.ty.data());
        if (trans == 'N') spmv(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
        else spmv_t(L.a, -1.0, L.y2l.data(), 1.0, L.ty.data());
        int sw = (trans == 'N') ? L.sweeps_post : L.sweeps_pre;
        const Smoother& spost = (trans == 'N' && L.has_sm2) ? L.sm2 : L.sm;  // sm2a (:766-770)
        rc = spost.apply(1.0, L.ty.data(), 1.0, L.y2l.data(), trans, sw, n);
        if (rc) return rc;
      }
    } else {
      return L.sm.apply(1.0, L.x2l.data(), 0.0, L.y2l.data(), trans, L.sweeps_pre, n);  // :787-793
    }
    return 0;
  }

  // mld_dmlprec_aply_vect (mld_dmlprec_aply.f90:203-357) called from
  // mld_dprecaply2_vect (mld_dprecaply.f90:361-417): y = op(M^-1) x.
  int apply(const double* x, double* y, char trans) {
    int nlev = (int)lv.size();
    ix n = lv[0].a.nr;
    if (nlev == 1) {  // one-level preconditioner (:372-417)
      int nsw = std::max(lv[0].sweeps_pre, lv[0].sweeps_post);
      if (!lv[0].has_sm2) return lv[0].sm.apply(1.0, x, 0.0, y, trans, nsw, n);
      // symmetrized pair (:378-412): w1 = x; nsw times w2 = S1 w1, w1 = S2 w2
      const Smoother& s1 = (trans == 'N') ? lv[0].sm : lv[0].sm2;
      const Smoother& s2 = (trans == 'N') ? lv[0].sm2 : lv[0].sm;
      std::vector<double> w1(x, x + n), w2(n);
      for (int k = 0; k < nsw; ++k) {
        int rc = s1.apply(1.0, w1.data(), 0.0, w2.data(), trans, 1, n);
        if (rc) return rc;
        rc = s2.apply(1.0, w2.data(), 0.0, w1.data(), trans, 1, n);
        if (rc) return rc;
      }
      axpby(n, 1.0, w1.data(), 0.0, y);
      return 0;
    }
    double beta = 0.0;
    axpby(n, 1.0, x, 0.0, lv[0].x2l.data());  // :283
    for (int isw = 1; isw <= prm.outer_sweeps - 1; ++isw) {  // :285-308
      int rc = inner_ml_aply(0, trans);
      if (rc) return rc;
      axpby(n, 1.0, lv[0].y2l.data(), beta, y);
      beta = 1.0;
      axpby(n, 1.0, x, 0.0, lv[0].x2l.data());
      spmv(lv[0].a, -1.0, y, 1.0, lv[0].x2l.data());
    }
    int rc = inner_ml_aply(0, trans);
    if (rc) return rc;
    axpby(n, 1.0, lv[0].y2l.data(), beta, y);  // :327
    return 0;
  }

  // mld_d_hierarchy_bld.f90:66-413 + mld_d_smoothers_bld.f90:156-169.
  int build(const Csr& a) {
    lv.clear();
    ix nglob = a.nr;
    int casize_ = prm.min_coarse_size;
    if (casize_ < 0) {  // :187-197
      double c = std::pow((double)nglob, 1.0 / (1.0 * 3.0));
      casize_ = std::max((int)c, 1) * 40;
    }
    casize = casize_;
    int nplevs = std::max(2, prm.max_levs);
    lv.resize(1);
    lv[0].a = a;
    int newsz = 0;
    // MAX_LEVS 1: the one-level preconditioners of init ('JACOBI', 'FBGS', ...):
    // a one-entry level array is kept as it is (:162-172)
    if (prm.max_levs == 1) nplevs = 1;
    for (int i = 2; i <= nplevs; ++i) {  // array_build_loop (:251-380), 1-based levels
      const Csr& af = lv[i - 2].a;
      std::vector<int> ilaggr;
      int naggr = 0;
      if (dec_map_bld(prm.aggr_ord, prm.theta, af, ilaggr, naggr)) return -10;
      double sizeratio = (double)naggr;
      if (i == 2) sizeratio = (double)nglob / sizeratio;
      else sizeratio = (double)lv[i - 2].naggr / sizeratio;
      if (naggr <= casize_) newsz = i;
      if (i > 2) {
        if (sizeratio < prm.min_cr_ratio) newsz = (sizeratio > 1.0) ? i : i - 1;
        if (naggr == lv[i - 2].naggr) newsz = i - 1;
      }
      if (newsz > 0) {
        if (newsz < i) {
          // revisit the previous level with its own aggregation (:352-359)
          ilaggr = lv[newsz - 1].ilaggr;
          naggr = lv[newsz - 1].naggr;
        }
        LevelOps ops;
        aggrmat_smth_asb(lv[newsz - 2].a, ilaggr, naggr, prm, ops);
        lv.resize(newsz);
        Level& L = lv[newsz - 1];
        L.a = std::move(ops.ac); L.p = std::move(ops.p); L.r = std::move(ops.r);
        L.omega = ops.omega; L.ilaggr = ilaggr; L.naggr = naggr;
        break;
      }
      LevelOps ops;
      aggrmat_smth_asb(af, ilaggr, naggr, prm, ops);
      lv.resize(i);
      Level& L = lv[i - 1];
      L.a = std::move(ops.ac); L.p = std::move(ops.p); L.r = std::move(ops.r);
      L.omega = ops.omega; L.ilaggr = ilaggr; L.naggr = naggr;
    }
    int nlev = (int)lv.size();
    for (int l = 0; l < nlev; ++l) {
      Level& L = lv[l];
      bool coarse = (l == nlev - 1) && nlev > 1;
      int kind = coarse ? prm.c_smoother : prm.smoother;
      if (kind == SM_AS) kind = SM_BJAC;
      // FBGS (mld_d_base_onelev_seti.F90:128-132): GS before, BWGS after
      L.sm.kind = (kind == SM_FBGS) ? SM_GS : kind;
      L.sm.fill = coarse ? prm.c_fill : prm.sub_fill;
      L.sm.milu = (coarse ? prm.c_milu : prm.sub_milu) != 0;
      L.has_sm2 = (kind == SM_FBGS);
      L.sweeps_pre = coarse ? prm.c_sweeps : prm.sweeps_pre;
      L.sweeps_post = coarse ? prm.c_sweeps : prm.sweeps_post;
      int rc = L.sm.build(L.a);
      if (rc) return 1000 + rc;
      if (L.has_sm2) {
        L.sm2.kind = SM_BWGS;
        rc = L.sm2.build(L.a);
        if (rc) return 1000 + rc;
      }
      ix n = L.a.nr;
      L.x2l.assign(n, 0.0); L.y2l.assign(n, 0.0); L.tx.assign(n, 0.0); L.ty.assign(n, 0.0);
    }
    return 0;
  }
};

// ---------------------------------------------------------------------------
// Krylov methods (psb_krylov call site tests/pdegen/mld_d_pde3d.f90:753-755).
// PSBLAS's source is not in the reference tree: textbook formulations with
// istop=2 (||r||_2/||b||_2 <= eps, checked after each iteration).
// ---------------------------------------------------------------------------
// it_times (may be null): wall-clock seconds of each iteration
// (std::chrono::steady_clock), itmax entries.
int krylov(Prec& pr, const Csr& a, int method, const double* b, double* x, double eps, int itmax,
           int* iters, double* err, double* it_times = nullptr) {
  using clk = std::chrono::steady_clock;
  auto tick = clk::now();
  auto lap = [&](int k) {
    if (!it_times || k < 1 || k > itmax) return;
    const auto now = clk::now();
    it_times[k - 1] = std::chrono::duration<double>(now - tick).count();
    tick = now;
  };
  ix n = a.nr;
  std::vector<double> r(n), z(n), p(n), q(n), w(n), s(n), t(n), rt(n), ph(n), sh(n), zold(n);
  double bn2 = nrm2(n, b);
  if (bn2 == 0.0) bn2 = 1.0;
  axpby(n, 1.0, b, 0.0, r.data());
  spmv(a, -1.0, x, 1.0, r.data());
  int it = 0;
  *err = nrm2(n, r.data()) / bn2;
  if (*err <= eps) { *iters = 0; return 0; }
  if (method == 0) {  // CG (Saad 2003, Alg. 9.1)
    double rho = 0.0;
    while (it < itmax) {
      if (it == 0) tick = clk::now();
      else lap(it);
      ++it;
      if (pr.apply(r.data(), z.data(), 'N')) return -1;
      double rho_old = rho;
      rho = dot(n, r.data(), z.data());
      if (it == 1) axpby(n, 1.0, z.data(), 0.0, p.data());
      else {
        if (rho_old == 0.0) break;
        axpby(n, 1.0, z.data(), rho / rho_old, p.data());
      }
      spmv(a, 1.0, p.data(), 0.0, q.data());
      double sigma = dot(n, p.data(), q.data());
      if (sigma == 0.0) break;
      double alpha = rho / sigma;
      axpby(n, alpha, p.data(), 1.0, x);
      axpby(n, -alpha, q.data(), 1.0, r.data());
      *err = nrm2(n, r.data()) / bn2;
      if (*err <= eps) break;
    }
  } else if (method == 1) {  // FCG(1), Notay SISC 22(4) 2000: p_k = z_k - (z_k,q_{k-1})/(p_{k-1},q_{k-1}) p_{k-1}
    double delta = 0.0;
    while (it < itmax) {
      if (it == 0) tick = clk::now();
      else lap(it);
      ++it;
      if (pr.apply(r.data(), z.data(), 'N')) return -1;
      if (it == 1) axpby(n, 1.0, z.data(), 0.0, p.data());
      else {
        double gamma = dot(n, z.data(), q.data());
        axpby(n, 1.0, z.data(), -(gamma / delta), p.data());
      }
      spmv(a, 1.0, p.data(), 0.0, q.data());
      delta = dot(n, p.data(), q.data());
      if (delta == 0.0) break;
      double alpha = dot(n, p.data(), r.data()) / delta;
      axpby(n, alpha, p.data(), 1.0, x);
      axpby(n, -alpha, q.data(), 1.0, r.data());
      *err = nrm2(n, r.data()) / bn2;
      if (*err <= eps) break;
    }
  } else if (method == 2) {  // BiCGStab, right preconditioned (Barrett et al., Templates 1994, Fig. 2.10)
    axpby(n, 1.0, r.data(), 0.0, rt.data());
    double rho = 0.0, rho_old = 1.0, alpha = 1.0, omega = 1.0;
    while (it < itmax) {
      if (it == 0) tick = clk::now();
      else lap(it);
      ++it;
      rho = dot(n, rt.data(), r.data());
      if (rho == 0.0) break;
      if (it == 1) axpby(n, 1.0, r.data(), 0.0, p.data());
      else {
        double beta = (rho / rho_old) * (alpha / omega);
        // p = r + beta*(p - omega*v)
        axpby(n, -omega, q.data(), 1.0, p.data());
        axpby(n, 1.0, r.data(), beta, p.data());
      }
      if (pr.apply(p.data(), ph.data(), 'N')) return -1;
      spmv(a, 1.0, ph.data(), 0.0, q.data());  // v
      double rtv = dot(n, rt.data(), q.data());
      if (rtv == 0.0) break;
      alpha = rho / rtv;
      axpby(n, 1.0, r.data(), 0.0, s.data());
      axpby(n, -alpha, q.data(), 1.0, s.data());
      if (nrm2(n, s.data()) / bn2 <= eps) {
        axpby(n, alpha, ph.data(), 1.0, x);
        axpby(n, 1.0, s.data(), 0.0, r.data());
        *err = nrm2(n, r.data()) / bn2;
        break;
      }
      if (pr.apply(s.data(), sh.data(), 'N')) return -1;
      spmv(a, 1.0, sh.data(), 0.0, t.data());
      double tt = dot(n, t.data(), t.data());
      omega = (tt == 0.0) ? 0.0 : dot(n, t.data(), s.data()) / tt;
      axpby(n, alpha, ph.data(), 1.0, x);
      axpby(n, omega, sh.data(), 1.0, x);
      axpby(n, 1.0, s.data(), 0.0, r.data());
      axpby(n, -omega, t.data(), 1.0, r.data());
      rho_old = rho;
      *err = nrm2(n, r.data()) / bn2;
      if (*err <= eps) break;
      if (omega == 0.0) break;
    }
  } else {
    return -3;
  }
  lap(it);
  *iters = it;
  return 0;
}

Csr make_csr(ix n, ix nc, const int64_t* irp, const int* ja, const double* val) {
  Csr a;
  a.nr = n; a.nc = nc;
  a.irp.assign(irp, irp + n + 1);
  ix nz = a.irp[n];
  a.ja.assign(ja, ja + nz);
  a.val.assign(val, val + nz);
  return a;
}

}  // namespace

// ============================================================================
// C interface (loaded with ctypes by tests/ and bench.py's cpu_baseline).
// ============================================================================
extern "C" {

struct orc_csr {  // caller-visible view of a CSR held by the oracle
  int64_t nr, nc, nnz;
  const int64_t* irp;
  const int* ja;
  const double* val;
};

// Threads of the row loops (<= 0: OpenMP's default); returns the count in use.
int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

void* orc_gen_pde3d(int idim) { return new Csr(gen_pde3d(idim)); }
void* orc_csr_new(int64_t n, int64_t nc, const int64_t* irp, const int* ja, const double* val) {
  return new Csr(make_csr(n, nc, irp, ja, val));
}
void orc_csr_free(void* m) { delete (Csr*)m; }
void orc_csr_view(void* m, orc_csr* v) {
  Csr* a = (Csr*)m;
  v->nr = a->nr; v->nc = a->nc; v->nnz = a->nnz();
  v->irp = a->irp.data(); v->ja = a->ja.data(); v->val = a->val.data();
}

void orc_spmv(void* m, double alpha, const double* x, double beta, double* y) { spmv(*(Csr*)m, alpha, x, beta, y); }
void orc_spmv_t(void* m, double alpha, const double* x, double beta, double* y) { spmv_t(*(Csr*)m, alpha, x, beta, y); }

int orc_dec_map_bld(void* m, int iorder, double theta, int* ilaggr, int* naggr) {
  std::vector<int> agg;
  int rc = dec_map_bld(iorder, theta, *(Csr*)m, agg, *naggr);
  if (rc == 0) std::copy(agg.begin(), agg.end(), ilaggr);
  return rc;
}

// ILU(0) of m; returns handles to L and U and fills d (length n).
int orc_ilu0(void* m, void** l, void** u, double* d) {
  Csr* L = new Csr; Csr* U = new Csr;
  std::vector<double> dv;
  int rc = ilu0_fact(*(Csr*)m, *L, *U, dv);
  std::copy(dv.begin(), dv.end(), d);
  *l = L; *u = U;
  return rc;
}
// ILU(k) / MILU(k) of m; as orc_ilu0.
int orc_iluk(void* m, int fill, int milu, void** l, void** u, double* d) {
  Csr* L = new Csr; Csr* U = new Csr;
  std::vector<double> dv;
  int rc = iluk_fact(*(Csr*)m, fill, milu != 0, *L, *U, dv);
  std::copy(dv.begin(), dv.end(), d);
  *l = L; *u = U;
  return rc;
}
void orc_trsv(void* t, int lower, int trans, const double* x, double* y) {
  if (trans) trsv_unit_t(*(Csr*)t, lower != 0, x, y);
  else trsv_unit(*(Csr*)t, lower != 0, x, y);
}

void* orc_prec_new(const orc_params* p) { Prec* pr = new Prec; pr->prm = *p; return pr; }
void orc_prec_free(void* h) { delete (Prec*)h; }
int orc_prec_build(void* h, void* a) { return ((Prec*)h)->build(*(Csr*)a); }
int orc_prec_nlev(void* h) { return (int)((Prec*)h)->lv.size(); }
int orc_prec_casize(void* h) { return ((Prec*)h)->casize; }
double orc_prec_omega(void* h, int lev) { return ((Prec*)h)->lv[lev].omega; }
// what: 0=A, 1=P, 2=R, 3=ND, 4=L, 5=U
void* orc_prec_level_csr(void* h, int lev, int what) {
  Level& L = ((Prec*)h)->lv[lev];
  switch (what) {
    case 0: return &L.a;
    case 1: return &L.p;
    case 2: return &L.r;
    case 3: return &L.sm.nd;
    case 4: return &L.sm.l;
    case 5: return &L.sm.u;
  }
  return nullptr;
}
int orc_prec_level_naggr(void* h, int lev) { return ((Prec*)h)->lv[lev].naggr; }
void orc_prec_level_ilaggr(void* h, int lev, int* out) {
  Level& L = ((Prec*)h)->lv[lev];
  std::copy(L.ilaggr.begin(), L.ilaggr.end(), out);
}
void orc_prec_level_dv(void* h, int lev, double* out) {
  Level& L = ((Prec*)h)->lv[lev];
  std::copy(L.sm.dv.begin(), L.sm.dv.end(), out);
}
int orc_prec_apply(void* h, const double* x, double* y, char trans) { return ((Prec*)h)->apply(x, y, trans); }
// Apply only the smoother of one level (init 'Z'): y = alpha*S(x) + beta*y.
int orc_prec_smooth(void* h, int lev, double alpha, const double* x, double beta, double* y, char trans, int sweeps) {
  Level& L = ((Prec*)h)->lv[lev];
  return L.sm.apply(alpha, x, beta, y, trans, sweeps, L.a.nr);
}
// which = 1: the post smoother (sm2a) where the level has a separate one
int orc_prec_smooth_w(void* h, int lev, int which, double alpha, const double* x, double beta, double* y, char trans,
                      int sweeps) {
  Level& L = ((Prec*)h)->lv[lev];
  const Smoother& sm = (which == 1 && L.has_sm2) ? L.sm2 : L.sm;
  return sm.apply(alpha, x, beta, y, trans, sweeps, L.a.nr);
}
// The cycle from level lev down (inner_ml_aply(lev)) on x: y = B_lev x.
int orc_prec_level_cycle(void* h, int lev, const double* x, double* y, char trans) {
  Prec& P = *(Prec*)h;
  if (lev < 0 || lev >= (int)P.lv.size()) return -2;
  Level& L = P.lv[lev];
  const ix n = L.a.nr;
  axpby(n, 1.0, x, 0.0, L.x2l.data());
  std::fill(L.y2l.begin(), L.y2l.end(), 0.0);
  int rc = P.inner_ml_aply(lev, trans);
  if (rc) return rc;
  axpby(n, 1.0, L.y2l.data(), 0.0, y);
  return 0;
}
int orc_krylov(void* h, void* a, int method, const double* b, double* x, double eps, int itmax, int* iters,
               double* err) {
  return krylov(*(Prec*)h, *(Csr*)a, method, b, x, eps, itmax, iters, err);
}
// As orc_krylov, with the wall-clock time of each iteration in it_times
// (itmax entries) and of the whole call in *total (steady_clock; the CPU
// baseline's timers).
int orc_krylov_timed(void* h, void* a, int method, const double* b, double* x, double eps, int itmax, int* iters,
                     double* err, double* it_times, double* total) {
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = krylov(*(Prec*)h, *(Csr*)a, method, b, x, eps, itmax, iters, err, it_times);
  *total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}
// Hierarchy and smoother build timed with steady_clock (seconds in *secs).
int orc_prec_build_timed(void* h, void* a, double* secs) {
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = ((Prec*)h)->build(*(Csr*)a);
  *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // extern "C"
