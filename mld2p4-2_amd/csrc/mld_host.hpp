// Host-side data model of the MI355X MLD2P4 engine: configuration, the
// smoothed-aggregation hierarchy and the smoother data, all built on the host
// and uploaded to HBM once.  Mirrors mld_dprec_type / mld_d_onelev_type
// (mlprec/mld_d_prec_type.f90:84-142, mlprec/mld_d_onelev_mod.f90:122-158).
#pragma once
#include <cstdint>
#include <string>
#include <memory>
#include <utility>
#include <vector>

namespace mld {

// ---------------------------------------------------------------------------
// Integer codes.  Values equal those of mlprec/mld_base_prec_type.F90:128-287
// (NOT mld_const.h, whose numbering is stale).  Codes >= 100 are additions of
// this engine with no reference counterpart.
// ---------------------------------------------------------------------------
enum : int {
  // smoother types (SMOOTHER_TYPE)
  NOPREC = 0, JAC = 1, BJAC = 2, AS = 3, FBGS = 4,
  L1JAC = 100,
  // solver types (SUB_SOLVE): mld_slv_delta_ = 4
  F_NONE = 4, DIAG_SCALE = 5, GS = 6, ILU_N = 7, MILU_N = 8, ILU_T = 9, SLU = 10, UMF = 11,
  SLUDIST = 12, MUMPS = 13, BWGS = 14,
  L1_DIAG = 101, DENSE_LU = 102,
  // ML cycle
  NO_ML = 0, ADD_ML = 1, MULT_ML = 2, VCYCLE = 3, WCYCLE = 4, KCYCLE = 5, KCYCLESYM = 6,
  // aggregation
  VMB = 1, NO_SMOOTH = 0, SMOOTH_PROL = 1, MIN_ENERGY = 2,
  NO_FILTER = 0, FILTER_MAT = 1, DEC_AGGR = 0, SYM_DEC_AGGR = 1,
  ORD_NAT = 0, ORD_DESC_DEG = 1, EIG_EST = 0, USER_CHOICE = 999, MAX_NORM = 0,
  DISTR_MAT = 0, REPL_MAT = 1,
  // pre/post
  SMOOTH_PRE = 1, SMOOTH_POST = 2, SMOOTH_BOTH = 3,
};

// Error codes returned through the C-ABI (the reference's info codes where
// one exists, e.g. mld_d_hierarchy_bld.f90:113-118).
enum : int {
  E_OK = 0,
  E_NOT_INIT = 3111,        // init not called (mld_d_hierarchy_bld.f90:113-118)
  E_NOT_BUILT = 3112,       // apply before build (mld_dprecaply.f90:355-360)
  E_NO_SOLVER = 1121,       // smoother without solver
  E_NO_DIAG = 1124,
  E_PIVOT = 2,              // psb_err_pivot_too_small_
  E_BAD_ARG = -1,
  E_BAD_LEVEL = -2,
  E_UNSUPPORTED = -20,      // option of the reference not provided by this engine
  E_DEVICE = -30,
  E_INTERNAL = -40,
};

// Allocator whose resize() leaves new elements uninitialised: the large CSR
// arrays are then first written by the threads that fill them (no
// single-threaded zero fill of gigabytes).
template <class T>
struct NoInitAlloc : std::allocator<T> {
  using value_type = T;
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <class U>
  struct rebind { using other = NoInitAlloc<U>; };
  template <class U>
  void construct(U* p) noexcept { ::new ((void*)p) U; }
  template <class U, class... Args>
  void construct(U* p, Args&&... args) { ::new ((void*)p) U(std::forward<Args>(args)...); }
};
template <class T>
using bigvec = std::vector<T, NoInitAlloc<T>>;

struct HCsr {
  int nr = 0, nc = 0;
  bigvec<int64_t> irp;  // nr+1, 0-based
  bigvec<int> ja;
  bigvec<double> val;
  int64_t nnz() const { return irp.empty() ? 0 : irp[nr]; }
  void clear() { nr = nc = 0; irp.clear(); ja.clear(); val.clear(); }
};

struct SmootherSpec {
  bool allocated = false;
  int smoother = JAC;    // NOPREC (base smoother) or JAC (Jacobi/block-Jacobi)
  int solver = DIAG_SCALE;
  int fillin = 0;
  double iluthrs = 1e-4;
};

// mld_dml_parms (mld_base_prec_type.F90:89-120) + coarse-solve bookkeeping.
struct LevelParms {
  int sweeps_pre = 1, sweeps_post = 1;
  int ml_cycle = VCYCLE;
  int aggr_type = VMB, par_aggr_alg = DEC_AGGR, aggr_ord = ORD_NAT, aggr_prol = SMOOTH_PROL;
  int coarse_mat = DISTR_MAT, aggr_omega_alg = EIG_EST, aggr_eig = MAX_NORM, aggr_filter = NO_FILTER;
  double aggr_omega_val = 0.0, aggr_thresh = 0.01;
  int coarse_solve = -1;
};

// Host smoother data of one level, as the jacobi smoother + its solver hold
// it after build (mld_d_jac_smoother_bld.f90:73-103, solver bld routines).
struct SmootherData {
  int smoother = JAC, solver = DIAG_SCALE;
  HCsr nd;                    // ND
  int64_t nnz_nd = 0;
  std::vector<double> dv;     // 1/diag, 1/l1-diag, or ILU 1/pivot
  HCsr l, u;                  // ILU(0) strict factors
  // level schedules of the triangular solves (rows grouped by level)
  std::vector<int> lsched_ptr, lsched_rows, usched_ptr, usched_rows;
  std::vector<double> lu;     // DENSE_LU: the dense matrix (row-major), factorized on the device
  std::vector<int> piv;
  bool iterative() const { return false; }
};

struct Level {
  LevelParms parms;
  SmootherSpec sm, sm2a;      // sm2a.allocated: separate post smoother
  double szratio = 0.0;
  // hierarchy data
  HCsr a;                     // base_a (level 0: copy of the user matrix)
  HCsr p, r;                  // map stored at this level (levels >= 1)
  HCsr tprol;                 // tentative prolongator kept for revisits
  std::vector<int> ilaggr;
  int naggr = 0;
  // smoother data
  SmootherData smd, smd2;     // smd2 used when sm2a.allocated
};

// One level handed over by the caller (mldhip_upload_level,
// mldhip_upload_smoother, mldhip_load_hierarchy) and kept until
// finalize_hierarchy moves it into precv.
struct StagedLevel {
  bool has_a = false;
  HCsr a, p, r;
  bool has_sm[2] = {false, false};  // pre smoother (sm), post smoother (sm2a)
  SmootherData sd[2];
};

struct Prec {
  bool initialized = false;
  std::string ptype;
  int max_levs = 20;          // mld_d_prec_type.f90:92
  int min_coarse_size = -1;
  double min_cr_ratio = 1.5;
  int outer_sweeps = 1;
  std::vector<Level> precv;
  bool hier_built = false, smoothers_built = false;
  int value_kind = 0;         // 0 = FP64 hierarchy on device, 1 = FP32 (mixed)
  int dev_format = 0;         // 0 = autotuned, 1 = force CSR, 2 = force ELL
  std::vector<StagedLevel> staged;
  // device of the smoothed-aggregation products of hierarchy_build (-1: host),
  // used for the levels whose finer operator has at least setup_min_rows rows
  int setup_device = -1;
  int64_t setup_min_rows = 0;
  std::string last_error;
};

// Phase timing on standard error when MLD_VERBOSE is set (the reference
// prints its build phases at debug_level >= psb_debug_outer_).
double wall_seconds();
void phase_log(const char* what, int level, double seconds);
// memcpy of n bytes split over the OpenMP threads (host staging of uploads)
void par_memcpy(void* dst, const void* src, size_t n);
// Longest row of a (over the host threads).
int host_max_row(const HCsr& a);
// Distinct diagonal offsets j - i of a, ascending; empty when more than 8.
std::vector<int> host_dia_offsets(const HCsr& a);
// SELL row order and slice offsets: rows stably sorted by decreasing length
// inside windows of sigma rows (sigma <= 1: identity), slices of c rows
// padded to their longest row; false when the storage exceeds 2^31 entries.
bool host_sell_layout(const HCsr& a, int sigma, int c, int maxrow, std::vector<int>& perm, std::vector<int>& sptr);

// configuration (mld_dprecinit / mld_dprecset / mld_d_base_onelev_*set*)
int prec_init(Prec& p, const std::string& ptype);
int prec_set_int(Prec& p, const std::string& what, int val, int ilev, int ilmax, const std::string& pos);
int prec_set_str(Prec& p, const std::string& what, const std::string& val, int ilev, int ilmax,
                 const std::string& pos);
int prec_set_real(Prec& p, const std::string& what, double val, int ilev, int ilmax, const std::string& pos);
int stringval(const std::string& s);

// build phases
int hierarchy_build(Prec& p, int n, const int64_t* irp, const int* ja, const double* val);
int smoothers_build(Prec& p);
// The level array for nplevs levels with the configured parameters and
// smoothers distributed as hierarchy_build does (base, intermediate, coarse).
void resize_levels(std::vector<Level>& v, int nplevs);

// A hierarchy built outside the engine: levels staged one at a time (0-based
// lev; P is n_{lev-1} x n, R is n x n_{lev-1}; arrays copied, 0-based CSR
// with strictly ascending columns), optionally the smoother data of each
// level, then finalize_hierarchy places them in precv with the configured
// parameters.  smoothers_done: every level's smoother data was staged (the
// hierarchy is then ready for upload; else smoothers_build builds them).
int stage_level(Prec& p, int lev, int n, const int64_t* irp, const int* ja, const double* val, int nf,
                const int64_t* pirp, const int* pja, const double* pval, const int64_t* rirp, const int* rja,
                const double* rval);
int stage_smoother(Prec& p, int lev, int which, int solver, const int64_t* ndirp, const int* ndja,
                   const double* ndval, const double* dv, const int64_t* lirp, const int* lja, const double* lval,
                   const int64_t* uirp, const int* uja, const double* uval);
int finalize_hierarchy(Prec& p, int nlev, bool& smoothers_done);
// Levels 2.. of a prec%dump (<prefix>_p000_lLLL_{ac,p,r}.mtx, until a level's
// ac file is missing) over the fine matrix given here; smoothers: also the
// smoother files <prefix>_sm[2a]_p000_lLLL_{nd,diag,lower,upper}.mtx of every
// level (a dump from level 1 with smoother and solver data).
int load_hierarchy(Prec& p, const std::string& prefix, int n, const int64_t* irp, const int* ja, const double* val,
                   bool smoothers, bool& smoothers_done, std::string& err);

// building blocks, exposed for parity tests through the C-ABI
int dec_map_bld(int iorder, double theta, const HCsr& a, std::vector<int>& ilaggr, int& naggr);
int aggrmat_smth_asb(const HCsr& a, const std::vector<int>& ilaggr, int naggr, LevelParms& parms, HCsr& ac,
                     HCsr& p, HCsr& r, int device = -1);
// The products P = S Ptilde (S = I - omega D^-1 A, or the given filtered S),
// AP = A P, R = P^T, Ac = R AP on the device, bit-identical to the host
// products; E_UNSUPPORTED when a row outgrows the device accumulator (the
// caller then forms them on the host).
int dev_smoothed_products(int device, const HCsr& a, const HCsr* s_filtered, const std::vector<double>& dinv,
                          double omega, const std::vector<int>& ilaggr, int naggr, HCsr& p, HCsr& r, HCsr& ac,
                          std::string& err);
int ilu0_factor(const HCsr& a, HCsr& l, HCsr& u, std::vector<double>& d);
int iluk_factor(const HCsr& a, int fill, bool milu, HCsr& l, HCsr& u, std::vector<double>& d);
void level_schedule(const HCsr& t, bool lower, std::vector<int>& ptr, std::vector<int>& rows);
double operator_complexity(const Prec& p);
// prec%dump: flags 1 = ac, 2 = rp, 4 = smoother, 8 = solver; istart/iend <= 0: defaults
int dump_prec(const Prec& p, int istart, int iend, const std::string& prefix, const std::string& head, int flags,
              std::string& err);

}  // namespace mld
