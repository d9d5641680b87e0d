// Device assembly: per-cell coefficients scattered into a resident matrix's
// raw value arrays by a plan made once on the host (see host/assembly.c).
#include "dev_common.hpp"

__global__ __launch_bounds__(SFR_BLOCK) void k_scatter(int64_t n,
    const int32_t *__restrict__ map, const double *__restrict__ src,
    double *__restrict__ dst)
{
  for (int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * SFR_BLOCK) {
    const int32_t m = map[k];
    if (m > 0) dst[k] = src[m - 1];
    else if (m == 0) dst[k] = 0.0;
  }
}

extern "C" int sfr_k_scatter(hipStream_t s, int64_t n, const int32_t *map,
                             const double *src, double *dst)
{
  if (n <= 0) return hipSuccess;
  int64_t g = (n + SFR_BLOCK - 1) / SFR_BLOCK;
  if (g > 8 * SFR_GRIDCAP) g = 8 * SFR_GRIDCAP;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)g), dim3(SFR_BLOCK), 0, s, n,
                     map, src, dst);
  return hipGetLastError();
}
