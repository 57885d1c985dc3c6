// Configuration API: init / set with the routing rules of the reference.
//   mld_dprecinit     : mlprec/impl/mld_dprecinit.F90:86-218
//   mld_dprecseti     : mlprec/impl/mld_dprecset.F90:78-450
//   mld_dcprecset{i,c,r}: mlprec/impl/mld_dcprecset.F90:78-693
//   onelev set        : mlprec/impl/level/mld_d_base_onelev_seti.F90:108-243
//   mld_stringval     : mlprec/mld_base_prec_type.F90:378-485
#include <algorithm>
#include <cctype>

#include "mld_host.hpp"

namespace mld {

static std::string upper(const std::string& s) {
  std::string r;
  for (char c : s) {
    if (c == '\t') break;  // mld_stringval drops a trailing tab comment
    r.push_back((char)std::toupper((unsigned char)c));
  }
  while (!r.empty() && r.back() == ' ') r.pop_back();
  while (!r.empty() && r.front() == ' ') r.erase(r.begin());
  return r;
}

int stringval(const std::string& s0) {
  std::string s = upper(s0);
  struct KV { const char* k; int v; };
  static const KV tab[] = {
      {"NONE", 0}, {"HALO", 1}, {"SUM", 2}, {"AVG", 3}, {"FACT_NONE", F_NONE}, {"FBGS", FBGS},
      {"GS", GS}, {"FGS", GS}, {"FWGS", GS}, {"BGS", BWGS}, {"BWGS", BWGS}, {"ILU", ILU_N},
      {"MILU", MILU_N}, {"ILUT", ILU_T}, {"MUMPS", MUMPS}, {"UMF", UMF}, {"SLU", SLU},
      {"SLUDIST", SLUDIST}, {"DIAG", DIAG_SCALE}, {"ADD", ADD_ML}, {"MULT_DEV", 7}, {"MULT", MULT_ML},
      {"VCYCLE", VCYCLE}, {"WCYCLE", WCYCLE}, {"KCYCLE", KCYCLE}, {"KCYCLESYM", KCYCLESYM},
      {"VMB", VMB}, {"DEC", DEC_AGGR}, {"SYMDEC", SYM_DEC_AGGR}, {"BCMATCH", 3}, {"NAT", ORD_NAT},
      {"NATURAL", ORD_NAT}, {"DESC", ORD_DESC_DEG}, {"RDEGREE", ORD_DESC_DEG},
      {"DEGREE", ORD_DESC_DEG}, {"REPL", REPL_MAT}, {"DIST", DISTR_MAT}, {"UNSMOOTHED", NO_SMOOTH},
      {"NONSMOOTHED", NO_SMOOTH}, {"SMOOTHED", SMOOTH_PROL}, {"MINENERGY", MIN_ENERGY},
      {"NOPREC", NOPREC}, {"BJAC", BJAC}, {"JAC", JAC}, {"JACOBI", JAC}, {"AS", AS},
      {"A_NORMI", MAX_NORM}, {"USER_CHOICE", USER_CHOICE}, {"EIG_EST", EIG_EST},
      {"FILTER", FILTER_MAT}, {"NOFILTER", NO_FILTER}, {"OUTER_SWEEPS", 42},
      // additions of this engine
      {"L1-JACOBI", L1JAC}, {"L1JAC", L1JAC}, {"L1_JACOBI", L1JAC}, {"L1-DIAG", L1_DIAG},
      {"L1_DIAG", L1_DIAG}, {"DENSE_LU", DENSE_LU}, {"DLU", DENSE_LU},
  };
  for (const KV& kv : tab)
    if (s == kv.k) return kv.v;
  return -1;
}

static int pos_code(const std::string& pos) {
  std::string p = upper(pos);
  if (p == "PRE") return SMOOTH_PRE;
  if (p == "POST") return SMOOTH_POST;
  return SMOOTH_BOTH;
}

static void level_default(Level& lv) {  // mld_d_onelev_mod.f90:426-457
  lv.parms = LevelParms();
  if (lv.sm.allocated) { lv.sm.fillin = 0; lv.sm.iluthrs = 1e-4; }
  if (lv.sm2a.allocated) { lv.sm2a.fillin = 0; lv.sm2a.iluthrs = 1e-4; }
}

// lv%set(smoother_mold / solver_mold, pos): choose which smoother object(s)
// receive the new smoother or solver.
static void set_smoother_obj(Level& lv, int ipos, int smoother) {
  if (ipos == SMOOTH_PRE || ipos == SMOOTH_BOTH) { lv.sm.allocated = true; lv.sm.smoother = smoother; }
  if (ipos == SMOOTH_POST) {
    if (!lv.sm2a.allocated || lv.sm2a.smoother != smoother) lv.sm2a.solver = -1;  // no solver yet
    lv.sm2a.allocated = true; lv.sm2a.smoother = smoother;
  }
  if (ipos == SMOOTH_BOTH) lv.sm2a.allocated = false;
}
static void set_solver_obj(Level& lv, int ipos, int solver) {
  if (ipos == SMOOTH_PRE || ipos == SMOOTH_BOTH) lv.sm.solver = solver;
  if (ipos == SMOOTH_POST || ipos == SMOOTH_BOTH) {
    if (lv.sm2a.allocated) lv.sm2a.solver = solver;
  }
}

// mld_d_base_onelev_seti.F90:108-243
static int onelev_seti(Level& lv, int what, int val, int ipos) {
  switch (what) {
    case 1: {  // mld_smoother_type_
      switch (val) {
        case NOPREC: set_smoother_obj(lv, ipos, NOPREC); set_solver_obj(lv, ipos, F_NONE); break;
        case JAC: set_smoother_obj(lv, ipos, JAC); set_solver_obj(lv, ipos, DIAG_SCALE); break;
        case BJAC: set_smoother_obj(lv, ipos, JAC); set_solver_obj(lv, ipos, ILU_N); break;
        case L1JAC: set_smoother_obj(lv, ipos, JAC); set_solver_obj(lv, ipos, L1_DIAG); break;
        case AS: set_smoother_obj(lv, ipos, AS); set_solver_obj(lv, ipos, ILU_N); break;
        case FBGS:
          set_smoother_obj(lv, SMOOTH_PRE, JAC); set_solver_obj(lv, SMOOTH_PRE, GS);
          set_smoother_obj(lv, SMOOTH_POST, JAC); set_solver_obj(lv, SMOOTH_POST, BWGS);
          break;
        default: break;  // "Do nothing and hope for the best"
      }
      if (ipos == SMOOTH_PRE || ipos == SMOOTH_BOTH) { lv.sm.fillin = 0; lv.sm.iluthrs = 1e-4; }
      if ((ipos == SMOOTH_POST || ipos == SMOOTH_BOTH) && lv.sm2a.allocated) { lv.sm2a.fillin = 0; lv.sm2a.iluthrs = 1e-4; }
      return 0;
    }
    case 2:  // mld_sub_solve_
      switch (val) {
        case F_NONE: case DIAG_SCALE: case GS: case BWGS: case ILU_N: case MILU_N: case ILU_T:
        case L1_DIAG: case DENSE_LU:
          set_solver_obj(lv, ipos, val);
          break;
        default: break;  // SLU/UMF/MUMPS/SLUDIST not configured: do nothing
      }
      return 0;
    case 36:  // mld_smoother_sweeps_
      if (ipos == SMOOTH_PRE || ipos == SMOOTH_BOTH) lv.parms.sweeps_pre = val;
      if (ipos == SMOOTH_POST || ipos == SMOOTH_BOTH) lv.parms.sweeps_post = val;
      return 0;
    case 21: lv.parms.sweeps_pre = val; return 0;
    case 22: lv.parms.sweeps_post = val; return 0;
    case 20: lv.parms.ml_cycle = val; return 0;
    case 25: lv.parms.par_aggr_alg = val; return 0;
    case 26: lv.parms.aggr_ord = val; return 0;
    case 23: lv.parms.aggr_type = val; return 0;
    case 24: lv.parms.aggr_prol = val; return 0;
    case 30: lv.parms.coarse_mat = val; return 0;
    case 27: lv.parms.aggr_omega_alg = val; return 0;
    case 28: lv.parms.aggr_eig = val; return 0;
    case 29: lv.parms.aggr_filter = val; return 0;
    case 31: lv.parms.coarse_solve = val; return 0;
    case 7:  // mld_sub_fillin_ -> solver
      if (ipos == SMOOTH_PRE || ipos == SMOOTH_BOTH) lv.sm.fillin = val;
      if ((ipos == SMOOTH_POST || ipos == SMOOTH_BOTH) && lv.sm2a.allocated) lv.sm2a.fillin = val;
      return 0;
    default:
      return 0;  // passed to the smoother, which ignores unknown keys
  }
}

static int what_code(const std::string& w) {
  struct KV { const char* k; int v; };
  static const KV tab[] = {
      {"SMOOTHER_TYPE", 1}, {"SUB_SOLVE", 2}, {"SUB_RESTR", 3}, {"SUB_PROL", 4}, {"SUB_OVR", 6},
      {"SUB_FILLIN", 7}, {"ILU_SCALE", 8}, {"ML_CYCLE", 20}, {"SMOOTHER_SWEEPS_PRE", 21},
      {"SMOOTHER_SWEEPS_POST", 22}, {"AGGR_TYPE", 23}, {"AGGR_PROL", 24}, {"PAR_AGGR_ALG", 25},
      {"AGGR_ORD", 26}, {"AGGR_OMEGA_ALG", 27}, {"AGGR_EIG", 28}, {"AGGR_FILTER", 29},
      {"COARSE_MAT", 30}, {"COARSE_SOLVE", 31}, {"COARSE_SWEEPS", 32}, {"COARSE_FILLIN", 33},
      {"COARSE_SUBSOLVE", 34}, {"SMOOTHER_SWEEPS", 36}, {"SOLVER_SWEEPS", 37},
      {"MIN_COARSE_SIZE", 38}, {"MAX_LEVS", 40}, {"OUTER_SWEEPS", 42},
  };
  std::string u = upper(w);
  for (const KV& kv : tab)
    if (u == kv.k) return kv.v;
  return -1;
}

int prec_init(Prec& p, const std::string& ptype0) {
  std::string ptype = upper(ptype0);
  p.precv.clear();
  p.hier_built = p.smoothers_built = false;
  p.ptype = ptype;
  p.initialized = true;
  auto one_level = [&](int smoother, int solver) {
    p.precv.resize(1);
    Level& lv = p.precv[0];
    lv.sm.allocated = true;
    lv.sm.smoother = smoother;
    lv.sm.solver = solver;
    level_default(lv);
  };
  if (ptype == "NOPREC" || ptype == "NONE") { one_level(NOPREC, F_NONE); return 0; }
  if (ptype == "JAC" || ptype == "JACOBI" || ptype == "DIAG") { one_level(JAC, DIAG_SCALE); return 0; }
  if (ptype == "L1-JACOBI" || ptype == "L1JAC") { one_level(JAC, L1_DIAG); return 0; }
  if (ptype == "BJAC") { one_level(JAC, ILU_N); return 0; }
  if (ptype == "GS" || ptype == "FWGS") { one_level(JAC, GS); return 0; }
  if (ptype == "BWGS") { one_level(JAC, BWGS); return 0; }
  if (ptype == "FBGS") {
    p.precv.resize(1);
    onelev_seti(p.precv[0], 1, FBGS, SMOOTH_BOTH);
    level_default(p.precv[0]);
    return 0;
  }
  if (ptype == "AS") { one_level(AS, ILU_N); return 0; }
  if (ptype == "ML") {  // mld_dprecinit.F90:186-207
    int nlev = p.max_levs;
    p.precv.resize(nlev);
    for (auto& lv : p.precv) level_default(lv);
    prec_set_str(p, "ML_CYCLE", "VCYCLE", -1, -1, "");
    prec_set_str(p, "SMOOTHER_TYPE", "FBGS", -1, -1, "");
    prec_set_str(p, "COARSE_SOLVE", "ILU", -1, -1, "");
    return 0;
  }
  p.initialized = false;
  p.last_error = "unknown preconditioner type " + ptype0;
  return E_PIVOT;  // the reference returns psb_err_pivot_too_small_ here (mld_dprecinit.F90:210-214)
}

// COARSE_SOLVE expansion for this build (no UMF/SLU/MUMPS/SLUDIST
// configured): mld_dprecset.F90:335-425 "#else" branches.
static void coarse_solve_expand(Level& lv, int val, int ipos) {
  onelev_seti(lv, 31, val, ipos);
  switch (val) {
    case BJAC:
      onelev_seti(lv, 1, BJAC, ipos); onelev_seti(lv, 2, ILU_N, ipos); onelev_seti(lv, 30, DISTR_MAT, ipos); break;
    case SLU: case MUMPS: case UMF: case SLUDIST:
      onelev_seti(lv, 1, BJAC, ipos); onelev_seti(lv, 2, ILU_N, ipos); onelev_seti(lv, 30, DISTR_MAT, ipos); break;
    case ILU_N: case ILU_T: case MILU_N:
      onelev_seti(lv, 1, BJAC, ipos); onelev_seti(lv, 2, val, ipos); onelev_seti(lv, 30, REPL_MAT, ipos); break;
    case JAC:
      onelev_seti(lv, 1, BJAC, ipos); onelev_seti(lv, 2, DIAG_SCALE, ipos); onelev_seti(lv, 30, DISTR_MAT, ipos); break;
    case DENSE_LU:
      onelev_seti(lv, 1, BJAC, ipos); onelev_seti(lv, 2, DENSE_LU, ipos); onelev_seti(lv, 30, REPL_MAT, ipos); break;
    default: break;
  }
}

int prec_set_int(Prec& p, const std::string& what, int val, int ilev, int ilmax, const std::string& pos) {
  std::string w = upper(what);
  if (!p.initialized) return E_NOT_INIT;
  if (w == "MIN_COARSE_SIZE") { p.min_coarse_size = std::max(val, -1); return 0; }
  if (w == "MAX_LEVS") { p.max_levs = std::max(val, 1); return 0; }
  if (w == "OUTER_SWEEPS") { p.outer_sweeps = std::max(val, 1); return 0; }
  int nlev = (int)p.precv.size();
  int ilev_ = (ilev > 0) ? ilev : 1;
  int ilmax_ = (ilev > 0 && ilmax > 0) ? ilmax : ilev_;
  if (ilev_ < 1 || ilev_ > nlev) return E_BAD_ARG;
  if (ilmax_ < 1 || ilmax_ > nlev) return E_BAD_ARG;
  int code = what_code(w);
  int ipos = pos_code(pos);
  if (ilev > 0) {
    switch (code) {
      case 34:  // COARSE_SUBSOLVE
        if (ilev_ != nlev) return -2;
        return onelev_seti(p.precv[ilev_ - 1], 2, val, ipos);
      case 31:  // COARSE_SOLVE
        if (ilev_ != nlev) return -2;
        if (nlev > 1) coarse_solve_expand(p.precv[nlev - 1], val, ipos);
        return 0;
      case 32:  // COARSE_SWEEPS
        if (ilev_ != nlev) return -2;
        return onelev_seti(p.precv[nlev - 1], 36, val, ipos);
      case 33:  // COARSE_FILLIN
        if (ilev_ != nlev) return -2;
        return onelev_seti(p.precv[nlev - 1], 7, val, ipos);
      default:
        for (int il = ilev_; il <= ilmax_; ++il) onelev_seti(p.precv[il - 1], code, val, ipos);
        return 0;
    }
  }
  switch (code) {
    case 2: case 3: case 4: case 6: case 7: case 36: case 1:
      for (int il = 1; il <= std::max(1, nlev - 1); ++il) onelev_seti(p.precv[il - 1], code, val, ipos);
      return 0;
    case 20: case 25: case 26: case 24: case 23: case 27: case 28: case 29:
      for (int il = 1; il <= nlev; ++il) onelev_seti(p.precv[il - 1], code, val, ipos);
      return 0;
    case 30:
      if (nlev > 1) onelev_seti(p.precv[nlev - 1], 30, val, ipos);
      return 0;
    case 31:
      if (nlev > 1) coarse_solve_expand(p.precv[nlev - 1], val, ipos);
      return 0;
    case 34:
      if (nlev > 1) onelev_seti(p.precv[nlev - 1], 2, val, ipos);
      return 0;
    case 32:
      if (nlev > 1) onelev_seti(p.precv[nlev - 1], 36, val, ipos);
      return 0;
    case 33:
      if (nlev > 1) onelev_seti(p.precv[nlev - 1], 7, val, ipos);
      return 0;
    default:
      for (int il = 1; il <= nlev; ++il) onelev_seti(p.precv[il - 1], code, val, ipos);
      return 0;
  }
}

// mld_dcprecsetc: string values go through mld_stringval (mld_dcprecset.F90:494-556)
int prec_set_str(Prec& p, const std::string& what, const std::string& sval, int ilev, int ilmax,
                 const std::string& pos) {
  if (!p.initialized) return E_NOT_INIT;
  int v = stringval(sval);
  if (v >= 0) return prec_set_int(p, what, v, ilev, ilmax, pos);
  return 0;  // unknown strings are passed to the level, which ignores them
}

// mld_dcprecsetr (mld_dcprecset.F90:597-693) + onelev setr
int prec_set_real(Prec& p, const std::string& what, double val, int ilev, int ilmax, const std::string& pos) {
  std::string w = upper(what);
  if (w == "MIN_CR_RATIO") { p.min_cr_ratio = std::max(1.0, val); return 0; }
  if (!p.initialized) return E_NOT_INIT;
  int nlev = (int)p.precv.size();
  int ilev_ = (ilev > 0) ? ilev : 1;
  int ilmax_ = (ilev > 0 && ilmax > 0) ? ilmax : ilev_;
  if (ilev_ < 1 || ilev_ > nlev || ilmax_ < 1 || ilmax_ > nlev) return E_BAD_ARG;
  auto setr = [&](Level& lv, const std::string& key) {
    if (key == "AGGR_OMEGA_VAL") lv.parms.aggr_omega_val = val;
    else if (key == "AGGR_THRESH") lv.parms.aggr_thresh = val;
    else if (key == "SUB_ILUTHRS") { lv.sm.iluthrs = val; if (lv.sm2a.allocated) lv.sm2a.iluthrs = val; }
  };
  if (ilev > 0) {
    for (int il = ilev_; il <= ilmax_; ++il) setr(p.precv[il - 1], w);
    return 0;
  }
  if (w == "COARSE_ILUTHRS") { setr(p.precv[nlev - 1], "SUB_ILUTHRS"); return 0; }
  for (int il = 1; il <= nlev; ++il) setr(p.precv[il - 1], w);
  return 0;
}

}  // namespace mld
