/*
 * mld2p4_amd.h — C-ABI of the MI355X engine for MLD2P4's preconditioner
 * application phase.
 *
 * Plain C: opaque handle, plain pointers and sizes, int return codes
 * (0 = success).  This is the layer a Fortran 2003 host binds with
 * iso_c_binding, following the reference's own C-binding precedent
 * (mlprec/mld_d_umf_solver.F90:51,62,89-121 <-> mlprec/impl/mld_dumf_interface.c:
 * opaque c_ptr handle, integer(c_int) return, value scalars).
 *
 * Matrices are passed as 0-based CSR with int64 row pointers, int32 column
 * indices (strictly ascending within each row) and double values, in host
 * memory; they are copied, the caller keeps ownership.  Vectors of apply /
 * krylov are DEVICE pointers (double, length n) unless the name ends in
 * _host.  Error codes: 3111 not initialised, 3112 not built, 1121 smoother
 * without solver, 2 ILU pivot too small / unknown preconditioner type,
 * -1 invalid argument, -2 invalid level, -20 option not provided by this
 * engine, -30 device error, -40 internal error.  mldhip_get_error returns
 * the text of the last error of a handle.
 */
#ifndef MLD2P4_AMD_H
#define MLD2P4_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mldhip_prec mldhip_prec;

/* ---- lifetime ---------------------------------------------------------- */
/* Create a preconditioner object bound to HIP device `device` (-1: host
 * only; builds run but apply/krylov return -30).  Replaces the allocation
 * of a type(mld_dprec_type) variable (mlprec/mld_d_prec_type.f90:84-142). */
int mldhip_create(mldhip_prec** p, int device);
/* prec%free (mlprec/mld_d_prec_type.f90:137 free method). */
int mldhip_destroy(mldhip_prec* p);

/* ---- configuration ------------------------------------------------------ */
/* prec%init(ptype, info): mlprec/impl/mld_dprecinit.F90:86-218.
 * ptype: "NOPREC", "JACOBI", "L1-JACOBI", "BJAC", "ML" (also "GS", "FBGS",
 * "AS": accepted, their smoothers are not provided -> -20 at build). */
int mldhip_init(mldhip_prec* p, const char* ptype);
/* prec%set(what, val, info [,ilev, ilmax, pos]):
 * mlprec/impl/mld_dcprecset.F90:78-453 (integer), :494-556 (string),
 * :597-693 (real).  ilev/ilmax <= 0 mean "absent"; pos may be NULL. */
int mldhip_seti(mldhip_prec* p, const char* what, int val, int ilev, int ilmax, const char* pos);
int mldhip_setc(mldhip_prec* p, const char* what, const char* val, int ilev, int ilmax, const char* pos);
int mldhip_setr(mldhip_prec* p, const char* what, double val, int ilev, int ilmax, const char* pos);
/* Value type of the uploaded hierarchy: 0 = FP64 (default), 1 = FP32
 * (mixed form; vectors of apply stay FP64).  No reference counterpart:
 * the reference's FP32 twin is the separate mld_sprec_type. */
int mldhip_set_precision(mldhip_prec* p, int kind);
/* Storage of uploaded operators: 0 = automatic (default), 1 = CSR, 2 = ELL,
 * 3 = SELL (sliced ELL, 64-row slices, rows length-sorted in windows of
 * 4096), 4 = SELL with windows of 256 rows, 5 = SELL without sorting, 6 = DIA
 * (diagonal storage, for operators with at most 8 distinct diagonals; others
 * fall back to the automatic choice), 7 = SELL with windows of 128 rows,
 * 8 / 9 / 10 = SELD (SELL with windows of 256 / 4096 / 128 rows whose column indices are
 * 16-bit steps from the previous column of the row; operators with a larger
 * step fall back to SELL).  The automatic choice takes DIA for such banded
 * operators, else the layout it times fastest. */
int mldhip_set_format(mldhip_prec* p, int fmt);

/* ---- build -------------------------------------------------------------- */
/* Where hierarchy_build forms the smoothed-aggregation products P, A P, R
 * and Ac of each level (mld_daggrmat_smth_asb.f90:291,345,399; aggregation
 * always runs on the host): 0 host threads, 1 the handle's device, -1
 * (default) the device for the products over levels of at least 2^22 rows
 * when the handle has one and MLD_SETUP is not "host", the host otherwise.
 * Both give bit-identical hierarchies.  No reference counterpart. */
int mldhip_set_setup(mldhip_prec* p, int where);
/* prec%hierarchy_build(a, desc, info): mlprec/impl/mld_d_hierarchy_bld.f90:66-490. */
int mldhip_hierarchy_build(mldhip_prec* p, int n, const int64_t* irp, const int* ja, const double* val);
/* prec%smoothers_build(a, desc, info, amold, vmold): mlprec/impl/mld_d_smoothers_bld.f90
 * :66-307; on a device handle this also uploads the hierarchy to HBM once
 * (the amold/vmold conversion of mld_d_base_onelev_cnv.f90:56-61). */
int mldhip_smoothers_build(mldhip_prec* p);
/* prec%build = hierarchy_build + smoothers_build (mlprec/impl/mld_dprecbld.f90:135,143). */
int mldhip_build(mldhip_prec* p, int n, const int64_t* irp, const int* ja, const double* val);

/* ---- a hierarchy built outside the engine ---------------------------------
 * The data each precv(l) holds after the reference's hierarchy_build and
 * smoothers_build (mlprec/mld_d_onelev_mod.f90:122-158): base_a and the map
 * of mld_aggrmat_asb (mlprec/impl/mld_d_lev_aggrmat_asb.f90:246-247), and
 * the smoother objects sm / sm2a (mld_d_jac_smoother_bld.f90:73-103 and the
 * solvers' bld routines).  Levels are 0-based (precv(lev+1)); CSR as for
 * hierarchy_build (0-based, int64 row pointers from 0, strictly ascending
 * columns); every array is copied.  The configured parameters (set) are
 * distributed over the levels as hierarchy_build does. */
/* Level lev: A (n x n), and for lev >= 1 P (nf x n, nf = rows of level
 * lev-1: map%p, psb_map_Y2X) and R (n x nf: map%r, psb_map_X2Y); P and R
 * NULL for lev 0. */
int mldhip_upload_level(mldhip_prec* p, int lev, int n, const int64_t* irp, const int* ja, const double* val, int nf,
                        const int64_t* pirp, const int* pja, const double* pval, const int64_t* rirp, const int* rja,
                        const double* rval);
/* Smoother data of level lev (uploaded before): which 0 = sm (pre / both),
 * 1 = sm2a (post).  solver: 4 none, 5 diag (dv = 1/diag), 101 l1-diag
 * (dv = 1/l1-diag), 7 ILU (L unit lower and U unit upper strict factors,
 * dv = 1/pivot: mld_dilu0_fact.f90:176-182), 6 GS / 14 BWGS (strict lower L,
 * strict upper U, dv = the diagonal).  ND (NULL: empty) is the Jacobi
 * smoother's off-diagonal operator; empty for the block solvers on one
 * device. */
int mldhip_upload_smoother(mldhip_prec* p, int lev, int which, int solver, const int64_t* ndirp, const int* ndja,
                           const double* ndval, const double* dv, const int64_t* lirp, const int* lja,
                           const double* lval, const int64_t* uirp, const int* uja, const double* uval);
/* Levels 0 .. nlev-1 become the hierarchy (the state after hierarchy_build).
 * When smoother data was uploaded for every level, the smoothers are built
 * too and a device handle uploads to HBM (the state after smoothers_build);
 * otherwise call mldhip_smoothers_build. */
int mldhip_finalize_hierarchy(mldhip_prec* p, int nlev);
/* The hierarchy of a prec%dump (mldhip_dump with ac and rp; mlprec/impl/
 * level/mld_d_base_onelev_dump.f90:92-121): A of level 1 given here (the
 * dump does not write it), levels 2.. read from <prefix>_p000_lLLL_{ac,p,r}
 * .mtx until a level's ac file is missing.  smoothers != 0: also the smoother
 * files of every level (a dump from level 1 with smoother and solver data),
 * read for the configured solvers, and the device upload; else call
 * mldhip_smoothers_build. */
int mldhip_load_hierarchy(mldhip_prec* p, const char* prefix, int n, const int64_t* irp, const int* ja,
                          const double* val, int smoothers);

/* ---- apply (the hot path) ------------------------------------------------ */
/* prec%apply(x, y, desc, info, trans): y = op(M^-1) x;
 * mlprec/impl/mld_dprecaply.f90:305-442 (mld_dprecaply2_vect) ->
 * mlprec/impl/mld_dmlprec_aply.f90:203-357.  x, y: device, length n.
 * trans: 'N' or 'T' ('C' == 'T' for real data).  Enqueued on the handle's
 * stream; returns after enqueueing (call mldhip_synchronize to wait). */
int mldhip_apply(mldhip_prec* p, const double* x, double* y, char trans);
/* Same with host vectors (copies over PCIe, synchronous). */
int mldhip_apply_host(mldhip_prec* p, const double* x, double* y, char trans);
/* mld_dmlprec_aply_vect(alpha, prec, x, beta, y, desc, trans, work, info)
 * (mlprec/impl/mld_dmlprec_aply.f90:203-357): y = alpha op(M^-1) x + beta y,
 * the final update of :327; y is not read when beta = 0.  x, y: device. */
int mldhip_mlprec_aply(mldhip_prec* p, double alpha, const double* x, double beta, double* y, char trans);

/* psb_krylov(method, a, prec, b, x, eps, desc, info, itmax, iter, err,
 * istop) call site (tests/pdegen/mld_d_pde3d.f90:753-755) with the matrix
 * given to hierarchy_build.  method: "CG", "FCG", "BICGSTAB".  istop = 2
 * (||r||/||b||).  b, x: device; x holds the initial guess on entry.
 * Synchronous. */
int mldhip_krylov(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                  int istop, int* iter, double* err);
int mldhip_krylov_host(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                       int istop, int* iter, double* err);

/* mldhip_krylov for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * no counterpart in the reference.  Same arguments and results as
 * mldhip_krylov, methods "CG" and "FCG", trans 'N'.  The four Krylov work
 * vectors live in a fixed workspace of the library (32 MiB on every GPU it
 * runs on, held until the process ends) instead of the handle's lazily
 * allocated vectors: systems of at most 1048576 rows (-1 beyond, with the
 * reason in mldhip_get_error); the handle must be built (its cycle buffers
 * are allocated at build).  Calls that overlap on one GPU share the
 * workspace and race: serialize them. */
int mldhip_krylov_noalloc(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                          int istop, int* iter, double* err);

/* ---- differentiation (no reference counterpart: the reference computes no
 * gradients) --------------------------------------------------------------- */
/* Adjoint solve A^T x = b with the same hierarchy applied transposed
 * (prec%apply with trans = 'T', mld_dmlprec_aply.f90:203-357 'T' branch) and
 * the transposed fine matrix; arguments as mldhip_krylov.  Transposed
 * operators are uploaded on the first call and kept. */
int mldhip_krylov_adjoint(mldhip_prec* p, const char* method, const double* b, double* x, double eps, int itmax,
                          int istop, int* iter, double* err);
/* Edge-product gradient of the 7-point finite-volume diffusion operator
 * A(theta) on the unit cube (idim^3 cells, h = 1/idim, k = exp(theta),
 * harmonic face coefficients, zero Dirichlet faces contributing k_i):
 * grad[c] = -lambda^T (dA/dtheta_c) x.  theta, x, lam, grad: device, length
 * idim^3; work: device, idim^3 doubles of scratch for exp(theta) (one
 * exponential per cell), or NULL to form the exponentials in place (seven per
 * cell); enqueued on `stream` (a hipStream_t, NULL = default stream). */
int mldhip_fv7_grad_theta(int idim, const double* theta, const double* x, const double* lam, double* grad,
                          double* work, void* stream);

/* ---- batched forms (no reference counterpart: the reference applies its
 * preconditioner to one vector of one system at a time) ----------------------
 * Multi-vectors are column-major with a leading dimension (device pointers).
 *
 * Block V-cycle: y = M^-1 x for nrhs <= 32 columns at once, every operator of
 * the hierarchy applied as a sparse-times-dense product (one pass over each
 * matrix for all columns); the same cycle as mldhip_apply per column, with
 * the coarsest solve applied as its explicit operator.  The block engine is
 * uploaded on first use.  Point / l1-Jacobi smoothers above the coarsest
 * level; coarsest ILU(0), dense LU or Jacobi. */
int mldhip_apply_block(mldhip_prec* p, int nrhs, const double* x, int64_t ldx, double* y, int64_t ldy);
/* psb_krylov per column of b (method "CG" or "FCG", istop = 2), all columns
 * advanced together with per-column scalars; iters/errs: host arrays of
 * nrhs entries.  x holds the initial guesses. */
int mldhip_krylov_block(mldhip_prec* p, const char* method, int nrhs, const double* b, int64_t ldb, double* x,
                        int64_t ldx, double eps, int itmax, int istop, int* iters, double* errs);

/* Groups of independent systems, each built in its own handle (hierarchy
 * and smoothers; the handles must stay alive until mldhip_group_build
 * returns).  The group runs one V-cycle over the block-diagonal unions of
 * the systems' levels and captures the CG/FCG iterations of all systems
 * into one hipGraph, replayed until every system has converged.  Vectors
 * are the concatenation of the systems' vectors in the order of
 * mldhip_group_add (offsets from mldhip_group_info), nrhs columns of them
 * (nb = 1: one column; nb = 32: up to 32). */
typedef struct mldhip_group mldhip_group;
int mldhip_group_create(mldhip_group** g, int device, int nb);
int mldhip_group_destroy(mldhip_group* g);
int mldhip_group_add(mldhip_group* g, mldhip_prec* p);
/* precision: 0 FP64, 1 FP32 preconditioner (Krylov vectors stay FP64). */
int mldhip_group_build(mldhip_group* g, int precision);
/* nsys, total rows, offsets[nsys+1] (any may be NULL). */
int mldhip_group_info(mldhip_group* g, int* nsys, int64_t* nrows, int64_t* offsets);
/* Stream the group's calls are ordered after (NULL: the null stream). */
int mldhip_group_set_stream(mldhip_group* g, void* hip_stream);
int mldhip_group_apply(mldhip_group* g, int nrhs, const double* x, int64_t ldx, double* y, int64_t ldy);
/* iters/errs: host arrays of nsys*nrhs entries (index j*nsys + s);
 * launches: graph launches used (may be NULL).  Synchronous. */
int mldhip_group_krylov(mldhip_group* g, const char* method, int nrhs, const double* b, int64_t ldb, double* x,
                        int64_t ldx, double eps, int itmax, int istop, int* iters, double* errs, int* launches);
/* Average time of one group preconditioner application (HIP events). */
int mldhip_group_time_apply(mldhip_group* g, int reps, double* avg_ms);
/* Union level info[10]: n, continuing rows, nnz A, P, R, ND(+ND2), coarse
 * operator, sweeps pre, post, storage (format or columns per lane). */
int mldhip_group_level_info(mldhip_group* g, int lev, int64_t* info);
int mldhip_group_get_error(mldhip_group* g, char* buf, int len);

/* ---- streams ------------------------------------------------------------ */
int mldhip_set_stream(mldhip_prec* p, void* hip_stream);
int mldhip_synchronize(mldhip_prec* p);

/* ---- inspection (hierarchy, smoother data) ------------------------------ */
/* Number of levels (size(prec%precv)). */
int mldhip_get_nlev(mldhip_prec* p, int* nlev);
/* info[0..9] = n, nnz(A), nnz(P), nnz(R), nnz(ND pre), naggr, sweeps_pre,
 * sweeps_post, smoother solver code, post solver code (-1 if none). */
int mldhip_get_level_info(mldhip_prec* p, int lev, int64_t* info);
/* Copy out a level operator: what 0 = A, 1 = P, 2 = R, 3 = ND, 4 = L, 5 = U
 * (pre smoother), 13/14/15 = ND/L/U of a separate post smoother.  Any of
 * irp/ja/val may be NULL to skip; sizes from mldhip_get_csr_size. */
int mldhip_get_csr_size(mldhip_prec* p, int lev, int what, int64_t* nr, int64_t* nc, int64_t* nnz);
int mldhip_get_level_csr(mldhip_prec* p, int lev, int what, int64_t* irp, int* ja, double* val);
/* Aggregation map of level lev (>= 1) onto level lev-1 rows (0-based ids). */
int mldhip_get_aggr(mldhip_prec* p, int lev, int* ilaggr);
/* Smoother diagonal data (1/diag, 1/l1-diag or ILU 1/pivot), length n. */
int mldhip_get_level_dv(mldhip_prec* p, int lev, int which, double* dv);
int mldhip_get_omega(mldhip_prec* p, int lev, double* omega);
/* Number of level sets of the two triangular solves (lower, upper) of the
 * smoother `which` (0 pre, 1 post) of level lev: the dependency depth of the
 * ILU / Gauss-Seidel substitutions (0 when the smoother has none).  No
 * reference counterpart (measurement). */
int mldhip_get_tri_levels(mldhip_prec* p, int lev, int which, int* nl_lower, int* nl_upper);
/* First level of the dense tail (the untransposed cycle on levels lev ..
 * nlev-1 applied as one dense product, formed at smoothers_build), or -1 when
 * the cycle runs level by level (MLD_DENSE_TAIL=0 at build, no level of at
 * most 1024 rows below the fine level, a K-cycle there).  No reference
 * counterpart. */
int mldhip_get_dense_tail(mldhip_prec* p, int* lev);
/* Operator complexity sum nnz(A_l)/nnz(A_1) (mld_d_prec_type.f90:465-497). */
int mldhip_get_complexity(mldhip_prec* p, double* cplx);
int mldhip_get_device_bytes(mldhip_prec* p, int64_t* bytes);
int mldhip_get_error(mldhip_prec* p, char* buf, int len);
/* prec%dump (mlprec/mld_d_prec_type.f90:698-729, per level
 * mld_d_base_onelev_dump.f90:38-131): MatrixMarket files of levels
 * istart..iend (1-based; <= 0: from level 2, or 1 when there is one level,
 * to the last), named <prefix>_p000_lLLL_{ac,r,p}.mtx and, for the
 * smoothers, <prefix>_sm[2a]_p000_lLLL_{nd,diag,lower,upper}.mtx.  flags:
 * 1 = ac (A of levels >= 2), 2 = rp (R and P), 4 = smoother (ND),
 * 8 = solver (diagonal / triangular factors).  prefix NULL or "": "dump_lev_d";
 * head (may be NULL): a comment line of every file. */
int mldhip_dump(mldhip_prec* p, int istart, int iend, const char* prefix, const char* head, int flags);

/* ---- single-kernel entry points (parity tests of the device kernels) ---- */
/* y = M x with the device copy of a level operator (what 0 = A, 1 = P,
 * 2 = R, 3 = ND): the psb_spmm / psb_map_X2Y / psb_map_Y2X products of
 * mld_dmlprec_aply.f90:674-709.  FP64 handles only. */
int mldhip_level_spmv(mldhip_prec* p, int lev, int what, const double* x, double* y);
/* One smoother application sm%apply(1, x, beta, y, trans, sweeps, init='Z')
 * of level lev (mlprec/impl/smoother/mld_d_jac_smoother_apply_vect.f90:38-215);
 * which 0 = pre smoother, 1 = post smoother.  FP64 handles only. */
int mldhip_level_smooth(mldhip_prec* p, int lev, int which, const double* x, int beta, double* y, int sweeps,
                        char trans);

/* ---- measurement (no reference counterpart: the reference times phases
 * with psb_wtime in its drivers, tests/pdegen/mld_d_pde3d.f90:721-770) ---- */
/* Average duration in ms of one cycle step of level lev, timed with HIP
 * events on the handle's stream over reps back-to-back runs.  kind:
 * 0 residual r = x - A y, 1 pre smoother, 2 restriction, 3 prolongation,
 * 4 post smoother, 5 Krylov SpMV with its dot (level 0), 6 the residual
 * launches of level lev inside reps whole preconditioner applications (each
 * launch between its own pair of events: the in-cycle duration). */
int mldhip_time_step(mldhip_prec* p, int lev, int kind, int reps, double* avg_ms);
/* Device storage of a level operator (what 0 A, 1 P, 2 R, 3 ND):
 * fmt 0 = CSR with `lanes` lanes per row, 1 = ELL of width `width`, 2 = SELL,
 * 3 = DIA with `width` diagonals, 4 = SELL with 16-bit column steps. */
int mldhip_get_op_format(mldhip_prec* p, int lev, int what, int* fmt, int* lanes, int* width);

#ifdef __cplusplus
}
#endif
#endif /* MLD2P4_AMD_H */
