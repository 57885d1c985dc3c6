// Interface between the C-ABI layer and the device engine (one value type
// per instance: FP64, or FP32 for the mixed-precision form).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <string>

#include "mld_host.hpp"

namespace mld {

class DevPrec {
 public:
  virtual ~DevPrec() {}
  virtual int upload(const Prec& p) = 0;
  virtual int apply(const double* x, double* y, char trans) = 0;
  // y = alpha op(M^-1) x + beta y (y not read when beta == 0)
  virtual int apply_ab(double alpha, const double* x, double beta, double* y, char trans) = 0;
  virtual int spmv_level(int lev, int what, const double* x, double* y) = 0;
  virtual int smooth_level(int lev, const double* x, int beta, double* y, int sweeps, char trans, int which) = 0;
  // trans 'N': solve A x = b preconditioned by M; 'T': the adjoint system
  // A^T x = b preconditioned by M^T (same hierarchy, transposed operators).
  virtual int krylov(int method, const double* b, double* x, double eps, int itmax, int istop, int* iter,
                     double* err, char trans) = 0;
  // krylov ('N', CG or FCG) with the caller's work vectors: ws holds nvec
  // vectors of n doubles; allocates nothing on the device.
  virtual int krylov_ws(int method, const double* b, double* x, double eps, int itmax, int istop, int* iter,
                        double* err, double* ws, int nvec) = 0;
  virtual int sync() = 0;
  virtual void set_stream(hipStream_t s) = 0;
  virtual hipStream_t stream() const = 0;
  virtual size_t device_bytes() const = 0;
  // first level of the dense tail operator, -1 without one
  virtual int tail_level() const = 0;
  // Average duration (ms) of one launch sequence of a single cycle step at
  // level lev, timed with HIP events on the handle's stream over reps runs.
  // kind: 0 residual r = x - A y, 1 pre smoother, 2 restriction R r,
  //       3 prolongation y += P yc, 4 post smoother, 5 Krylov SpMV + dot.
  virtual int time_step(int lev, int kind, int reps, double* avg_ms) = 0;
  // Storage chosen for a level operator: fmt (0 CSR, 1 ELL), lanes, width.
  virtual int op_format(int lev, int what, int* fmt, int* lanes, int* width) = 0;
};

DevPrec* make_dev_prec(int value_kind, int device, hipStream_t stream);

// Device address of the no-allocation Krylov workspace on `device` (NOALLOC_VECS
// vectors of NOALLOC_MAX_ROWS doubles), taken once per device.
constexpr size_t NOALLOC_MAX_ROWS = (size_t)1 << 20;
constexpr int NOALLOC_VECS = 4;
double* noalloc_workspace(int device);

// Copies between pageable host memory and the device through two pinned
// staging buffers per device (host threads fill / drain one buffer while the
// DMA of the other is in flight).  A staged copy returns once its chunks
// have arrived (the staging buffers are idle between calls); copies under
// 4 MB go directly (host_to_device then returns with the copy enqueued on s,
// device_to_host after it).
void host_to_device(void* dst, const void* src, size_t bytes, hipStream_t s);
void device_to_host(void* dst, const void* src, size_t bytes, hipStream_t s);
int dev_guard(const std::function<int()>& f, std::string& err);

}  // namespace mld
