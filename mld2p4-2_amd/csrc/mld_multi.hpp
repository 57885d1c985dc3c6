// Batched engines (interface between the C-ABI layer and mld_multi.hip).
//
// One engine serves G independent systems (each with its own host-built
// hierarchy) times NB right-hand sides each, NB = 1 (grouped systems) or 32
// (block V-cycle).  Level l of the engine is the block-diagonal union of
// level l of every system that has one; systems whose coarsest level is l
// occupy the trailing rows of that union and are solved there while the
// others recurse.  Multi-vectors are row-interleaved in HBM: entry (row i,
// column j) at i*NB + j, so one sparse row's products with the NB columns
// read NB consecutive values.  At the C-ABI they are column-major.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mld_host.hpp"

namespace mld {

class MultiEngine {
 public:
  virtual ~MultiEngine() {}
  // y = M^-1 x for nrhs <= NB columns of length rows() (column-major with
  // leading dimensions ldx, ldy; device pointers).  Enqueued on the caller's
  // stream (ordered after prior work on it).
  virtual int apply_cm(int nrhs, const double* x, int64_t ldx, double* y, int64_t ldy, hipStream_t caller) = 0;
  // Preconditioned CG (method 0) or FCG(1) (method 1) on every column of
  // every system, with per-column scalars, to ||r||/||b|| <= eps.  iters and
  // errs: host arrays of nsys()*nrhs entries, index j*nsys() + s.
  // launches: graph launches used.  Synchronous.
  virtual int krylov_cm(int method, int nrhs, const double* b, int64_t ldb, double* x, int64_t ldx, double eps,
                        int itmax, int* iters, double* errs, int* launches, hipStream_t caller) = 0;
  virtual int64_t rows() const = 0;
  virtual int nsys() const = 0;
  virtual int nb() const = 0;
  virtual const std::vector<int64_t>& offsets() const = 0;  // level-0 row offset of each system (nsys+1)
  virtual size_t device_bytes() const = 0;
  // Average time (ms) of one preconditioner application to all columns,
  // with HIP events on the engine's stream over reps runs.
  virtual int time_apply(int reps, double* ms) = 0;
  // Number of levels of the union hierarchy and per-level info
  // (n, ni, nnz A, nnz P, nnz R, nnz ND, nnz coarse operator, sweeps pre, post).
  virtual int nlev() const = 0;
  virtual int level_info(int lev, int64_t* info) const = 0;
};

// Builds the union hierarchy of `systems` on the host and uploads it.
// value_kind: 0 FP64, 1 FP32 preconditioner.  nb: 1 or 32.
MultiEngine* make_multi(const std::vector<const Prec*>& systems, int value_kind, int nb, int device, int& rc,
                        std::string& err);

}  // namespace mld
