// CDNA4 (gfx950) kernels of the preconditioner-application path.
//
// Every sparse product is "one row -> one value" followed by an epilogue
// that consumes the row value in registers (residual, Jacobi update,
// prolongation add, Krylov dot).  Two storage formats:
//   ELL  : column-major (entry k of row i at k*ld+i), one lane per row; each
//          load instruction of a wavefront reads 64 consecutive values, and
//          the row sum follows the stored order exactly.
//   CSR  : LANES lanes per row (1..64) reading the row in coalesced strides,
//          reduced with cross-lane shuffles (DPP/permute); LANES == 1 keeps
//          the stored order.
// Reductions are deterministic: fixed per-block partials, then the last
// block to finish (device-scope atomic ticket) adds them in index order.
// Built with -ffp-contract=off: products and sums are rounded separately, as
// in the CPU restatement the results are checked against.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mld {
namespace k {

constexpr int BS = 256;  // threads per block (4 wavefronts)

// ---------------------------------------------------------------- gathers
template <typename V, typename XT>
struct GatherX {
  const XT* __restrict__ x;
  __device__ __forceinline__ V operator()(int j) const { return (V)x[j]; }
};
// first Jacobi sweep from a zero guess: y_j = d_j * x_j
template <typename V, typename XT>
struct GatherDX {
  const V* __restrict__ d;
  const XT* __restrict__ x;
  __device__ __forceinline__ V operator()(int j) const { return d[j] * (V)x[j]; }
};

// -------------------------------------------------------------- epilogues
template <typename V, typename YT>
struct EpiSet {  // y = A x
  YT* __restrict__ y;
  __device__ __forceinline__ void operator()(int i, V acc) const { y[i] = (YT)acc; }
};
template <typename V>
struct EpiAdd {  // y = y + A x      (prolongation, psb_map_Y2X with beta = 1)
  V* __restrict__ y;
  __device__ __forceinline__ void operator()(int i, V acc) const { y[i] = y[i] + acc; }
};
template <typename V, typename BT>
struct EpiRes {  // r = b - A x     (psb_geaxpby + psb_spmm(-1,A,x,1,r))
  const BT* __restrict__ b;
  V* __restrict__ r;
  __device__ __forceinline__ void operator()(int i, V acc) const { r[i] = (V)b[i] - acc; }
};
template <typename V, typename BT, typename YT, bool ADD>
struct EpiJac {  // t = d*(b - ND y);  y = t  or  y = y + t
  const V* __restrict__ d;
  const BT* __restrict__ b;
  YT* __restrict__ y;
  __device__ __forceinline__ void operator()(int i, V acc) const {
    V t = d[i] * ((V)b[i] - acc);
    if (ADD) y[i] = (YT)((V)y[i] + t);
    else y[i] = (YT)t;
  }
};

// ------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum of v over the block, valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) s += lds[w];
  }
  return s;
}

// K partial sums per block -> partial[K*nblocks]; the last block adds them in
// block order and writes result[0..K).  counter is returned to zero.
template <int K>
__device__ __forceinline__ void grid_sum(const double (&v)[K], double* partial, unsigned* counter,
                                         double* result) {
  __shared__ double lds[BS / 64];
  __shared__ int is_last;
  const int nb = gridDim.x;
  double bs[K];
#pragma unroll
  for (int q = 0; q < K; ++q) bs[q] = block_sum(v[q], lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) partial[(size_t)q * nb + blockIdx.x] = bs[q];
    __threadfence();
    unsigned t = atomicAdd(counter, 1u);
    is_last = (t == (unsigned)nb - 1);
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
#pragma unroll
    for (int q = 0; q < K; ++q) {
      double s = 0.0;
      for (int i = threadIdx.x; i < nb; i += BS) s += __builtin_nontemporal_load(&partial[(size_t)q * nb + i]);
      s = block_sum(s, lds);
      if (threadIdx.x == 0) result[q] = s;
    }
    if (threadIdx.x == 0) *counter = 0u;
  }
}

// ----------------------------------------------------------- SpMV kernels
template <typename V, int LANES, class G, class E>
__global__ __launch_bounds__(BS) void csr_kernel(int nr, const int* __restrict__ irp,
                                                  const int* __restrict__ ja, const V* __restrict__ val,
                                                  G g, E e) {
  const int gid = blockIdx.x * BS + threadIdx.x;
  const int row = gid / LANES;
  const int lane = gid & (LANES - 1);
  V acc = 0;
  if (row < nr) {
    const int b = irp[row], en = irp[row + 1];
    for (int k = b + lane; k < en; k += LANES) acc = acc + val[k] * g(ja[k]);
  }
  if (LANES > 1) {
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc = acc + __shfl_down(acc, off, LANES);
  }
  if (lane == 0 && row < nr) e(row, acc);
}

template <typename V, int W, class G, class E>
__global__ __launch_bounds__(BS) void ell_kernel(int nr, int w, int ld, const int* __restrict__ ja,
                                                  const V* __restrict__ val, G g, E e) {
  const int row = blockIdx.x * BS + threadIdx.x;
  if (row >= nr) return;
  V acc = 0;
  if constexpr (W > 0) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const int idx = k * ld + row;
      acc = acc + val[idx] * g(ja[idx]);
    }
  } else {
    for (int k = 0; k < w; ++k) {
      const int idx = k * ld + row;
      acc = acc + val[idx] * g(ja[idx]);
    }
  }
  e(row, acc);
}

// y = A x with the Krylov dot(s) folded in: D1 = sum_i u_i*y_i (and D2 =
// sum_i w_i*y_i when W2) over rows, deterministic grid sum.
template <typename V, int LANES, bool W2>
__global__ __launch_bounds__(BS) void csr_dot_kernel(int nr, const int* __restrict__ irp,
                                                      const int* __restrict__ ja, const V* __restrict__ val,
                                                      const V* __restrict__ x, V* __restrict__ y,
                                                      const V* __restrict__ u, const V* __restrict__ w2,
                                                      double* partial, unsigned* counter, double* result) {
  const int nrows_blk = BS / LANES;
  double d1 = 0.0, d2 = 0.0;
  for (int base = blockIdx.x * nrows_blk; base < nr; base += gridDim.x * nrows_blk) {
    const int row = base + threadIdx.x / LANES;
    const int lane = threadIdx.x & (LANES - 1);
    V acc = 0;
    if (row < nr) {
      const int b = irp[row], en = irp[row + 1];
      for (int k = b + lane; k < en; k += LANES) acc = acc + val[k] * x[ja[k]];
    }
    if (LANES > 1) {
#pragma unroll
      for (int off = LANES / 2; off > 0; off >>= 1) acc = acc + __shfl_down(acc, off, LANES);
    }
    if (lane == 0 && row < nr) {
      y[row] = acc;
      d1 += (u ? (double)u[row] : (double)acc) * (double)acc;  // u == nullptr: y.y
      if (W2) d2 += (double)w2[row] * (double)acc;
    }
  }
  double v[2] = {d1, d2};
  if (W2) grid_sum<2>(v, partial, counter, result);
  else {
    double v1[1] = {d1};
    grid_sum<1>(v1, partial, counter, result);
  }
}

template <typename V, int W, bool W2>
__global__ __launch_bounds__(BS) void ell_dot_kernel(int nr, int w, int ld, const int* __restrict__ ja,
                                                      const V* __restrict__ val, const V* __restrict__ x,
                                                      V* __restrict__ y, const V* __restrict__ u,
                                                      const V* __restrict__ w2, double* partial,
                                                      unsigned* counter, double* result) {
  double d1 = 0.0, d2 = 0.0;
  for (int row = blockIdx.x * BS + threadIdx.x; row < nr; row += gridDim.x * BS) {
    V acc = 0;
    if constexpr (W > 0) {
#pragma unroll
      for (int k = 0; k < W; ++k) {
        const int idx = k * ld + row;
        acc = acc + val[idx] * x[ja[idx]];
      }
    } else {
      for (int k = 0; k < w; ++k) {
        const int idx = k * ld + row;
        acc = acc + val[idx] * x[ja[idx]];
      }
    }
    y[row] = acc;
    d1 += (u ? (double)u[row] : (double)acc) * (double)acc;  // u == nullptr: y.y
    if (W2) d2 += (double)w2[row] * (double)acc;
  }
  if (W2) {
    double v[2] = {d1, d2};
    grid_sum<2>(v, partial, counter, result);
  } else {
    double v1[1] = {d1};
    grid_sum<1>(v1, partial, counter, result);
  }
}

// --------------------------------------------------------- vector kernels
template <typename V, typename XT, bool ADD>
__global__ __launch_bounds__(BS) void vmul_kernel(int n, const V* __restrict__ d, const XT* __restrict__ x,
                                                   V* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  V t = d[i] * (V)x[i];
  if (ADD) y[i] = y[i] + t;
  else y[i] = t;
}

template <typename TO, typename TI>
__global__ __launch_bounds__(BS) void convert_kernel(int n, const TI* __restrict__ x, TO* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = (TO)x[i];
}

// y = x + y (axpby with alpha = beta = 1) or y = x (beta = 0)
template <typename V, typename XT, bool ADD>
__global__ __launch_bounds__(BS) void add_kernel(int n, const XT* __restrict__ x, V* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (ADD) y[i] = (V)x[i] + y[i];
  else y[i] = (V)x[i];
}

template <typename V>
__global__ __launch_bounds__(BS) void fill_kernel(int n, V* __restrict__ y, V v) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = v;
}

// y = a*x + b*y, general coefficients (PSBLAS geaxpby)
template <typename V>
__global__ __launch_bounds__(BS) void axpby_kernel(int n, V a, const V* __restrict__ x, V b, V* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (b == (V)0) y[i] = a * x[i];
  else y[i] = a * x[i] + b * y[i];
}

// dot products: K pairs (x_q, y_q) in one pass
template <typename V, int K>
__global__ __launch_bounds__(BS) void dot_kernel(int n, const V* __restrict__ x0, const V* __restrict__ y0,
                                                  const V* __restrict__ x1, const V* __restrict__ y1,
                                                  double* partial, unsigned* counter, double* result) {
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    s0 += (double)x0[i] * (double)y0[i];
    if (K > 1) s1 += (double)x1[i] * (double)y1[i];
  }
  if (K > 1) {
    double v[2] = {s0, s1};
    grid_sum<2>(v, partial, counter, result);
  } else {
    double v[1] = {s0};
    grid_sum<1>(v, partial, counter, result);
  }
}

// ------------------------------------------------------------ Krylov steps
// Scalar slots in device memory.
enum { S_RHO0 = 0, S_RHO1 = 1, S_SIGMA = 2, S_RR = 3, S_PR = 4, S_GAMMA = 5, S_DELTA = 6, S_TMP0 = 8,
       S_TMP1 = 9, S_OMEGA = 10, S_ALPHA = 11, S_NSCAL = 16 };

// CG: p = z (first) or p = z + (rho_new/rho_old) p
__global__ __launch_bounds__(BS) void cg_pupdate(int n, const double* __restrict__ z, double* __restrict__ p,
                                                  const double* __restrict__ scal, int it) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (it == 1) {
    p[i] = z[i];
  } else {
    const double rn = scal[(it & 1) ? S_RHO1 : S_RHO0], ro = scal[(it & 1) ? S_RHO0 : S_RHO1];
    const double beta = rn / ro;
    p[i] = z[i] + beta * p[i];
  }
}

// x += alpha p; r -= alpha q; sum r_i^2.  alpha = num/den from device slots;
// den == 0 leaves x and r unchanged (breakdown, detected on the host).
__global__ __launch_bounds__(BS) void xr_update(int n, double* __restrict__ x, double* __restrict__ r,
                                                 const double* __restrict__ p, const double* __restrict__ q,
                                                 const double* __restrict__ scal, int num_slot, int den_slot,
                                                 double* partial, unsigned* counter, double* result) {
  const double den = scal[den_slot];
  const double alpha = (den != 0.0) ? scal[num_slot] / den : 0.0;
  double s = 0.0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    x[i] = alpha * p[i] + x[i];
    const double ri = (-alpha) * q[i] + r[i];
    r[i] = ri;
    s += ri * ri;
  }
  double v[1] = {s};
  grid_sum<1>(v, partial, counter, result);
}

// FCG(1): p = z - (gamma/delta) p, and sum_i p_i r_i for the step length.
__global__ __launch_bounds__(BS) void fcg_pupdate(int n, const double* __restrict__ z, double* __restrict__ p,
                                                   const double* __restrict__ r, const double* __restrict__ scal,
                                                   int it, double* partial, unsigned* counter, double* result) {
  const double c = (it == 1) ? 0.0 : -(scal[S_GAMMA] / scal[S_DELTA]);
  double s = 0.0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const double pi = (it == 1) ? z[i] : z[i] + c * p[i];
    p[i] = pi;
    s += pi * r[i];
  }
  double v[1] = {s};
  grid_sum<1>(v, partial, counter, result);
}

// ---------------------------------------------------- triangular solves
// Unit lower / upper solve of the rows of one level set, one lane per row,
// row sum in stored order:  t_i = b_i - sum_j l_ij t_j  (b_i = x_i, or for
// the upper solve of ILU b_i = d_i * t1_i, the D scaling of psb_spsm).
// Upper solve can also finish y = t (beta 0) or y = y + t (beta 1).
template <typename V, bool UPPER, int BETA>
__global__ __launch_bounds__(BS) void trsv_level_kernel(int cnt, const int* __restrict__ rows,
                                                         const int* __restrict__ irp, const int* __restrict__ ja,
                                                         const V* __restrict__ val, const V* __restrict__ b,
                                                         const V* __restrict__ d, V* __restrict__ t,
                                                         V* __restrict__ y) {
  const int idx = blockIdx.x * BS + threadIdx.x;
  if (idx >= cnt) return;
  const int i = rows[idx];
  V acc = 0;
  for (int k = irp[i]; k < irp[i + 1]; ++k) acc = acc + val[k] * t[ja[k]];
  V bi = UPPER ? d[i] * b[i] : b[i];
  V ti = bi - acc;
  t[i] = ti;
  if (UPPER) {
    if (BETA == 0) y[i] = ti;
    else if (BETA == 1) y[i] = y[i] + ti;
  }
}

// Whole ILU(0) apply for a small matrix in one workgroup: level sets
// processed in order with a barrier in between.
template <typename V, int BETA>
__global__ __launch_bounds__(1024) void ilu_small_kernel(
    int n, const int* __restrict__ lirp, const int* __restrict__ lja, const V* __restrict__ lval,
    const int* __restrict__ uirp, const int* __restrict__ uja, const V* __restrict__ uval,
    const V* __restrict__ d, const int* __restrict__ lptr, const int* __restrict__ lrows, int nlevl,
    const int* __restrict__ uptr, const int* __restrict__ urows, int nlevu, const V* __restrict__ x,
    V* __restrict__ t1, V* __restrict__ t2, V* __restrict__ y) {
  for (int lv = 0; lv < nlevl; ++lv) {
    for (int q = lptr[lv] + threadIdx.x; q < lptr[lv + 1]; q += blockDim.x) {
      const int i = lrows[q];
      V acc = 0;
      for (int k = lirp[i]; k < lirp[i + 1]; ++k) acc = acc + lval[k] * t1[lja[k]];
      t1[i] = x[i] - acc;
    }
    __syncthreads();
  }
  for (int lv = 0; lv < nlevu; ++lv) {
    for (int q = uptr[lv] + threadIdx.x; q < uptr[lv + 1]; q += blockDim.x) {
      const int i = urows[q];
      V acc = 0;
      for (int k = uirp[i]; k < uirp[i + 1]; ++k) acc = acc + uval[k] * t2[uja[k]];
      const V ti = d[i] * t1[i] - acc;
      t2[i] = ti;
      if (BETA == 0) y[i] = ti;
      else y[i] = y[i] + ti;
    }
    __syncthreads();
  }
}

// Dense LU solve (partial pivoting, row-major LU) in one workgroup;
// coarsest-level DENSE_LU solver for small coarse matrices.
template <typename V, int BETA>
__global__ __launch_bounds__(1024) void dense_lu_small_kernel(int n, const double* __restrict__ lu,
                                                              const int* __restrict__ piv,
                                                              const V* __restrict__ x, double* __restrict__ w,
                                                              V* __restrict__ y) {
  __shared__ double red[1024 / 64];
  for (int i = threadIdx.x; i < n; i += blockDim.x) w[i] = (double)x[i];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      if (p != k) { double tmp = w[k]; w[k] = w[p]; w[p] = tmp; }
    }
  __syncthreads();
  // forward (unit lower)
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = threadIdx.x; j < i; j += blockDim.x) s += lu[(size_t)i * n + j] * w[j];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
      w[i] = w[i] - t;
    }
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = 0.0;
    for (int j = i + 1 + threadIdx.x; j < n; j += blockDim.x) s += lu[(size_t)i * n + j] * w[j];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
      w[i] = (w[i] - t) / lu[(size_t)i * n + i];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (BETA == 0) y[i] = (V)w[i];
    else y[i] = y[i] + (V)w[i];
  }
}

}  // namespace k
}  // namespace mld
