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

__global__ __launch_bounds__(256) void fv7_grad_theta_kernel(int idim, double inv_h2, const double* __restrict__ th,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ lam,
                                                             double* __restrict__ grad) {
  const long n = (long)idim * idim * idim;
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int iz = (int)(c % idim), iy = (int)((c / idim) % idim), ix = (int)(c / ((long)idim * idim));
  const double kc = exp(th[c]), xc = x[c], lc = lam[c];
  double s = 0.0;
  const long strides[3] = {(long)idim * idim, (long)idim, 1};
  const int coord[3] = {ix, iy, iz};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
    for (int sg = -1; sg <= 1; sg += 2) {
      const int q = coord[ax] + sg;
      if (q >= 0 && q < idim) {
        const long j = c + sg * strides[ax];
        const double kj = exp(th[j]);
        const double den = kc + kj;
        const double dcf = 2.0 * kc * kj * kj / (den * den);
        s += dcf * (lc - lam[j]) * (xc - x[j]);
      } else {
        s += kc * lc * xc;
      }
    }
  }
  grad[c] = -s * inv_h2;
}

}  // namespace

extern "C" int mldhip_fv7_grad_theta(int idim, const double* theta, const double* x, const double* lam,
                                     double* grad, void* stream) {
  if (idim <= 0 || !theta || !x || !lam || !grad) return -1;
  const long n = (long)idim * idim * idim;
  const double inv_h2 = (double)idim * (double)idim;
  const int nb = (int)((n + 255) / 256);
  fv7_grad_theta_kernel<<<nb, 256, 0, (hipStream_t)stream>>>(idim, inv_h2, theta, x, lam, grad);
  return hipGetLastError() == hipSuccess ? 0 : -30;
}
