// Dense FP64 LU of the coarsest matrix on the device, its inverse, and the
// per-cycle coarse solve as a dense matrix-vector product.
//
// Blocked right-looking LU with partial pivoting (row-major, lda = n):
//   panel_factor_kernel : unblocked factorization of columns [k0, k0+nb),
//                         pivot rows swapped across the whole matrix;
//   trsm_lower_kernel   : U12 = L11^-1 A12 (unit lower), one thread per column;
//   gemm_sub_mfma_kernel: C -= A B with v_mfma_f64_16x16x4f64, one 16x16 tile
//                         per wavefront (the O(n^3) part of both the
//                         factorization and the inverse).
// Inverse: X = U^-1 L^-1 P, by the same blocked triangular solves on the n
// columns of P.  The coarse solve of a cycle is then y = X x (FP64 accumulate).
// No reference counterpart: the reference's direct coarse solvers are
// third-party sparse LU libraries (UMFPACK, SuperLU, MUMPS).
#pragma once
#include <hip/hip_runtime.h>

namespace mld {
namespace k {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Unblocked LU of the panel columns [k0, k0+nb) for rows k0..n-1.  One
// workgroup.  piv[j] = row swapped with row j (0-based); *info = j+1 for the
// first zero pivot (the column is then left unscaled).
__global__ __launch_bounds__(1024) void panel_factor_kernel(int n, double* a, int k0, int nb, int* piv,
                                                             int* info) {
  __shared__ double sval[1024 / 64];
  __shared__ int sidx[1024 / 64];
  __shared__ int prow;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int kend = min(n, k0 + nb);
  for (int j = k0; j < kend; ++j) {
    // pivot: max |a_ij|, i >= j, smallest index on ties
    double best = -1.0;
    int bi = j;
    for (int i = j + tid; i < n; i += nt) {
      const double v = fabs(a[(size_t)i * n + j]);
      if (v > best) { best = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_down(best, off, 64);
      const int oi = __shfl_down(bi, off, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { sval[tid >> 6] = best; sidx[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      double b = sval[0];
      int p = sidx[0];
      for (int w = 1; w < nt / 64; ++w)
        if (sval[w] > b || (sval[w] == b && sidx[w] < p)) { b = sval[w]; p = sidx[w]; }
      prow = p;
      piv[j] = p;
      if (b == 0.0 && *info == 0) *info = j + 1;
    }
    __syncthreads();
    const int p = prow;
    if (p != j)
      for (int c = tid; c < n; c += nt) {
        const double t = a[(size_t)j * n + c];
        a[(size_t)j * n + c] = a[(size_t)p * n + c];
        a[(size_t)p * n + c] = t;
      }
    __syncthreads();
    const double d = a[(size_t)j * n + j];
    if (d != 0.0)
      for (int i = j + 1 + tid; i < n; i += nt) a[(size_t)i * n + j] = a[(size_t)i * n + j] / d;
    __syncthreads();
    const int w = kend - (j + 1);  // remaining panel columns
    if (w > 0)
      for (int q = tid; q < (n - j - 1) * w; q += nt) {
        const int i = j + 1 + q / w, c = j + 1 + q % w;
        a[(size_t)i * n + c] = a[(size_t)i * n + c] - a[(size_t)i * n + j] * a[(size_t)j * n + c];
      }
    __syncthreads();
  }
}

// Triangular solve of the nb x nb diagonal block T = rows/cols [r0, r0+nb) of
// t (row-major, ld n) against columns [c0, c1) of b (row-major, ld ldb), in
// place, one thread per column.  UNIT: unit lower (forward); else upper with
// its diagonal (backward).
template <bool UNIT_LOWER>
__global__ __launch_bounds__(256) void trsm_cols_kernel(int n, const double* __restrict__ t, int r0, int nb,
                                                         double* b, int ldb, int c0, int c1) {
  const int c = c0 + blockIdx.x * 256 + threadIdx.x;
  if (c >= c1) return;
  const int rend = min(n, r0 + nb);
  if (UNIT_LOWER) {
    for (int r = r0; r < rend; ++r) {
      double s = b[(size_t)r * ldb + c];
      for (int k = r0; k < r; ++k) s = s - t[(size_t)r * n + k] * b[(size_t)k * ldb + c];
      b[(size_t)r * ldb + c] = s;
    }
  } else {
    for (int r = rend - 1; r >= r0; --r) {
      double s = b[(size_t)r * ldb + c];
      for (int k = r + 1; k < rend; ++k) s = s - t[(size_t)r * n + k] * b[(size_t)k * ldb + c];
      b[(size_t)r * ldb + c] = s / t[(size_t)r * n + r];
    }
  }
}

// C[m x nc] -= A[m x kd] * B[kd x nc] (row-major, leading dims lda/ldb/ldc);
// k beyond kd is zero-padded.  One wavefront per 16x16 tile of C, four tiles per
// block; v_mfma_f64_16x16x4f64 with A[i][k] in lane i%16 + 16*k and
// B[k][j] in lane j + 16*k; result register r of a lane holds row
// lane/16 + 4*r, column lane%16 (layout measured on gfx950).
__global__ __launch_bounds__(256) void gemm_sub_mfma_kernel(int m, int nc, int kd, const double* __restrict__ A,
                                                             int lda, const double* __restrict__ B, int ldb,
                                                             double* C, int ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_n = (nc + 15) / 16;
  const int tile = blockIdx.x * 4 + wave;
  const int ti = tile / tiles_n, tj = tile % tiles_n;
  if (ti * 16 >= m) return;
  const int r0 = ti * 16, c0 = tj * 16;
  const int ar = r0 + (lane & 15), bc = c0 + (lane & 15), kq = lane >> 4;
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < kd; k += 4) {
    const bool kin = (k + kq) < kd;
    const double av = (ar < m && kin) ? A[(size_t)ar * lda + k + kq] : 0.0;
    const double bv = (bc < nc && kin) ? B[(size_t)(k + kq) * ldb + bc] : 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  const int col = c0 + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = r0 + (lane >> 4) + 4 * r;
    if (row < m && col < nc) C[(size_t)row * ldc + col] = C[(size_t)row * ldc + col] - acc[r];
  }
}

// x <- P x applied to the rows of the n x n identity: row j swapped with
// piv[j] for j = 0..n-1 in order.  One thread per column.
__global__ __launch_bounds__(256) void perm_identity_kernel(int n, const int* __restrict__ piv, double* x) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  for (int r = 0; r < n; ++r) x[(size_t)r * n + c] = (r == c) ? 1.0 : 0.0;
  for (int j = 0; j < n; ++j) {
    const int p = piv[j];
    if (p != j) {
      const double t = x[(size_t)j * n + c];
      x[(size_t)j * n + c] = x[(size_t)p * n + c];
      x[(size_t)p * n + c] = t;
    }
  }
}

// y = X x (BETA 0) or y = y + X x (BETA 1); one wavefront per row, FP64
// accumulation.
template <typename V, int BETA>
__global__ __launch_bounds__(256) void dense_gemv_kernel(int n, const double* __restrict__ X,
                                                          const V* __restrict__ x, V* __restrict__ y) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  for (int c = lane; c < n; c += 64) s += X[(size_t)row * n + c] * (double)x[c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) {
    if (BETA == 0) y[row] = (V)s;
    else y[row] = (V)((double)y[row] + s);
  }
}

}  // namespace k
}  // namespace mld
