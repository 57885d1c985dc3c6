// Gradient of an AMG-solved linear system with respect to the cell
// log-coefficients of a 7-point finite-volume diffusion operator.
//
// A(theta) = sum over interior faces f=(i,j) of c_f (e_i - e_j)(e_i - e_j)^T / h^2
//          + sum over boundary faces of cell i of k_i e_i e_i^T / h^2,
// k = exp(theta), c_f = 2 k_i k_j / (k_i + k_j) (harmonic mean), h = 1/idim.
// For L(x(theta)) with A x = b and the adjoint A^T lambda = dL/dx:
//   dL/db = lambda,  dL/dtheta_c = -lambda^T (dA/dtheta_c) x
//   = -(1/h^2) [ sum_{faces of c} dc_f/dtheta_c (lambda_c - lambda_j)(x_c - x_j)
//                + (boundary faces of c) k_c lambda_c x_c ],
// dc_f/dtheta_c = 2 k_c k_j^2 / (k_c + k_j)^2.
// One thread per cell; the six face products are "edge products" over the
// stencil.  HBM-bound: reads theta, x, lambda (neighbours from cache), writes
// the gradient.  No reference counterpart (the reference has no gradients).
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/mld2p4_amd.h"

namespace {

// KFROM = true: k[] already holds exp(theta) (one exponential per cell
// instead of seven: the kernel is bound by FP64 exp otherwise).
template <bool KFROM>
__global__ __launch_bounds__(256) void fv7_grad_theta_kernel(int idim, double inv_h2, const double* __restrict__ th,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ lam,
                                                             double* __restrict__ grad) {
  // 32-bit index arithmetic (idim^3 < 2^31 is checked by the launcher):
  // 64-bit division by a runtime value is a long software sequence
  const unsigned n = (unsigned)idim * idim * idim;
  const unsigned c = blockIdx.x * 256u + threadIdx.x;
  if (c >= n) return;
  const unsigned ud = (unsigned)idim, line = c / ud;
  const int iz = (int)(c - line * ud), ix = (int)(line / ud), iy = (int)(line - (unsigned)ix * ud);
  // all six neighbours' values are loaded up front from clamped indices
  // (branch-free), so the 18 gathers of a thread are in flight together
  const int coord[3] = {ix, iy, iz};
  const int strides[3] = {idim * idim, idim, 1};
  unsigned jj[6];
  bool in[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int ax = f >> 1, sg = (f & 1) ? 1 : -1;
    const int q = coord[ax] + sg;
    in[f] = (q >= 0 && q < idim);
    jj[f] = in[f] ? c + sg * strides[ax] : c;
  }
  double tj[6], lj[6], xj[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    tj[f] = th[jj[f]];
    lj[f] = lam[jj[f]];
    xj[f] = x[jj[f]];
  }
  const double kc = KFROM ? th[c] : exp(th[c]);
  const double xc = x[c], lc = lam[c];
  double s = 0.0;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    if (in[f]) {
      const double kj = KFROM ? tj[f] : exp(tj[f]);
      const double den = kc + kj;
      const double dcf = 2.0 * kc * kj * kj / (den * den);
      s += dcf * (lc - lj[f]) * (xc - xj[f]);
    } else {
      s += kc * lc * xc;
    }
  }
  grad[c] = -s * inv_h2;
}

__global__ __launch_bounds__(256) void exp_kernel(long n, const double* __restrict__ th, double* __restrict__ k) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) k[i] = exp(th[i]);
}

}  // namespace

extern "C" int mldhip_fv7_grad_theta(int idim, const double* theta, const double* x, const double* lam,
                                     double* grad, double* work, void* stream) {
  if (idim <= 0 || idim > 1290 || !theta || !x || !lam || !grad) return -1;
  const long n = (long)idim * idim * idim;
  const double inv_h2 = (double)idim * (double)idim;
  const int nb = (int)((n + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  if (work) {
    exp_kernel<<<nb, 256, 0, s>>>(n, theta, work);
    fv7_grad_theta_kernel<true><<<nb, 256, 0, s>>>(idim, inv_h2, work, x, lam, grad);
  } else {
    fv7_grad_theta_kernel<false><<<nb, 256, 0, s>>>(idim, inv_h2, theta, x, lam, grad);
  }
  return hipGetLastError() == hipSuccess ? 0 : -30;
}
