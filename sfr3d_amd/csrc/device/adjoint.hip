// Adjoint of x = A^-1 b for a loss L(x): with lambda = A^-T dL/dx,
//   dL/db = lambda,  dL/da_ij = -lambda_i x_j,  dL/dad_i = -lambda_i x_i.
// The gradient is returned in the caller's raw storage order: a raw value
// that stands for two matrix entries (the symmetric storages keep a_ij once
// for (i,j) and (j,i)) gets the sum of both terms.  Every raw value knows its
// entries (rc: up to two (row, col) pairs), so each output element is written
// by one thread -- no atomics, deterministic.
#include "dev_common.hpp"

__global__ __launch_bounds__(SFR_BLOCK) void k_adj_grad(int64_t nraw,
    const int32_t *__restrict__ rc, const double *__restrict__ lam,
    const double *__restrict__ x, double *__restrict__ g, int64_t n,
    double *__restrict__ gad)
{
  const int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (g && k < nraw) {
    const int32_t r0 = rc[4 * k], c0 = rc[4 * k + 1];
    const int32_t r1 = rc[4 * k + 2], c1 = rc[4 * k + 3];
    double v = 0.0;
    if (r0 >= 0) v -= lam[r0] * x[c0];
    if (r1 >= 0) v -= lam[r1] * x[c1];
    g[k] = v;
  }
  if (gad && k < n) gad[k] = -lam[k] * x[k];
}

extern "C" int sfr_k_adj_grad(hipStream_t s, int64_t nraw, const int32_t *rc,
                              const double *lam, const double *x,
                              double *grad_raw, int64_t n, double *grad_ad)
{
  const int64_t m = nraw > n ? nraw : n;
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_grad, dim3((unsigned)((m + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, nraw, rc, lam, x, grad_raw, n, grad_ad);
  return hipGetLastError();
}
