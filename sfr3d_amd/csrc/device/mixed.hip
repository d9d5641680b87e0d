// Mixed-precision iterative refinement for the Jacobi-PCG path.
//
// Outer loop, FP64: x += D^-1/2 y, r = b - A x, stop when
// sqrt|r.D^-1 r| < tol * sqrt|b.D^-1 b| -- the reference PCG's own test
// (src/Cg.c:73-78,141-142), evaluated on the true FP64 residual.
// Inner loop, FP32 storage: CG on the Jacobi-scaled system
//   (D^-1/2 A D^-1/2) y = D^-1/2 r,
// which is Jacobi-PCG on A written with a unit diagonal: no diagonal and no
// preconditioner vector is read, and the correction is D^-1/2 y.  Matrix and
// vectors are FP32; dot products are accumulated in FP64.  The inner solve
// stops when its residual has dropped by inner_tol (r.r <= inner_tol^2 r0.r0)
// or after inner_maxit iterations.
// Per inner iteration (three kernels):
//   k_mx_spmv   q = Ap, p.q -> alpha       matrix + 2 words
//   k_mx_resid  r -= alpha q, r.r -> beta   3 words
//   k_mx_dir    y += alpha p, p = r + beta p  5 words
#include "dev_common.hpp"

template <class Op, int RPT, bool ZERO_X>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_resid64(Op A,
    const double *__restrict__ b, const double *__restrict__ x,
    const double *__restrict__ dinv, const double *__restrict__ dscale,
    float *__restrict__ rh, float *__restrict__ ph, float *__restrict__ y,
    int first, double inner_tol, unsigned inner_maxit, sfr_scalars *sc,
    double *part)
{
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double ax = ZERO_X ? 0.0 : A.row(i, valid, x);
      if (valid && lead) {
        const double ri = b[i] - ax;
        const float ci = (float)(ri * dscale[i]);
        rh[i] = ci;
        ph[i] = ci;
        y[i] = 0.0f;
        acc[0] += ri * (ri * dinv[i]);
        acc[1] += ri * ri;
        acc[2] += (double)ci * (double)ci;
      }
    }
  }
  if (grid_sum<3>(acc, part, &sc->counter) && threadIdx.x == 0) {
    sc->res[2] = acc[0];
    sc->res[3] = acc[1];
    if (first) {
      sc->outer = 0;
      sc->iter = 0;
    } else {
      sc->outer = sc->outer + 1;
    }
    if (sqrt(fabs(acc[0])) < sc->conv) {
      sc->done = SFR_DONE_CONV;
    } else if (!isfinite(acc[0])) {
      sc->done = SFR_DONE_BREAK;
    } else {
      sc->done = sc->iter >= sc->maxit ? SFR_DONE_MAXIT : SFR_DONE_RUN;
    }
    sc->d = acc[2];
    sc->rho0 = acc[2];
    sc->inner_tol2 = inner_tol * inner_tol;
    sc->inner_iter = 0;
    sc->inner_maxit = inner_maxit;
    sc->inner_done = sc->done != SFR_DONE_RUN ? 1u : (acc[2] == 0.0 ? 1u : 0u);
    sc->seal = 0;
  }
}

// x += D^-1/2 y
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_update(int64_t n,
    const float *__restrict__ y, const double *__restrict__ dscale,
    double *__restrict__ x)
{
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) x[i] += dscale[i] * (double)y[i];
    }
  }
}

template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_spmv(Op A,
    const float *__restrict__ p, float *__restrict__ q, sfr_scalars *sc,
    double *part)
{
  if (sc->inner_done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->seal = 1;
    return;
  }
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const float qi = A.row(i, valid, p);
      if (valid && lead) {
        q[i] = qi;
        acc[0] += (double)qi * (double)p[i];
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    const double alpha = sc->d / acc[0];
    sc->alpha = alpha;
    if (!(acc[0] > 0.0) || !isfinite(alpha)) sc->inner_done = 2;
  }
}

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_resid32(int64_t n,
    const float *__restrict__ q, float *__restrict__ r, sfr_scalars *sc,
    double *part)
{
  if (sc->inner_done) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const float alpha = (float)sc->alpha;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const float ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    const double rr = acc[0];
    sc->beta = rr / sc->d;
    sc->d = rr;
    const unsigned it = sc->inner_iter + 1;
    sc->inner_iter = it;
    sc->iter = sc->iter + 1;
    if (rr <= sc->inner_tol2 * sc->rho0 || it >= sc->inner_maxit ||
        sc->iter >= sc->maxit)
      sc->inner_done = 1;
    else if (!isfinite(rr))
      sc->inner_done = 2;
  }
}

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_dir(int64_t n,
    const float *__restrict__ r, float *__restrict__ y,
    float *__restrict__ p, sfr_scalars *sc)
{
  if (sc->seal || sc->inner_done == 2) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const float alpha = (float)sc->alpha, beta = (float)sc->beta;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const float pi = p[i];
        y[i] += alpha * pi;
        p[i] = r[i] + beta * pi;
      }
    }
  }
}

#define LAUNCH_V(K, g, s, ...)                                                 \
  do {                                                                         \
    switch ((g).rpt) {                                                         \
    case 1: hipLaunchKernelGGL(K<1>, dim3((g).grid), dim3(SFR_BLOCK), 0, s,    \
                               __VA_ARGS__); break;                            \
    case 2: hipLaunchKernelGGL(K<2>, dim3((g).grid), dim3(SFR_BLOCK), 0, s,    \
                               __VA_ARGS__); break;                            \
    default: hipLaunchKernelGGL(K<4>, dim3((g).grid), dim3(SFR_BLOCK), 0, s,   \
                                __VA_ARGS__); break;                           \
    }                                                                          \
  } while (0)

extern "C" int sfr_k_mx_outer(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, const double *b, double *x,
                              const float *y, int first, int newX, float *rh,
                              float *ph, float *yz, double inner_tol,
                              unsigned inner_maxit, sfr_scalars *sc,
                              double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  if (!first) LAUNCH_V(k_mx_update, gv, s, A->n, y, A->dscale, x);
  const bool zero_x = first && newX;
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    switch (gm.rpt) {
#define CASE(R)                                                                \
    case R:                                                                    \
      if (zero_x)                                                              \
        hipLaunchKernelGGL((k_mx_resid64<Op, R, true>), dim3(gm.grid),         \
                           dim3(SFR_BLOCK), 0, s, op, b, x, A->dinv,           \
                           A->dscale, rh, ph, yz, first, inner_tol, inner_maxit,\
                           sc, part);                                          \
      else                                                                     \
        hipLaunchKernelGGL((k_mx_resid64<Op, R, false>), dim3(gm.grid),        \
                           dim3(SFR_BLOCK), 0, s, op, b, x, A->dinv,           \
                           A->dscale, rh, ph, yz, first, inner_tol, inner_maxit,\
                           sc, part);                                          \
      break;
    CASE(1) CASE(2)
    default: CASE(4)
#undef CASE
    }
  });
  return hipGetLastError();
}

extern "C" int sfr_k_mx_inner_iter(hipStream_t s, const sfr_mat_view *A,
                                   sfr_geom g0, float *y, float *rh,
                                   float *ph, float *qh, sfr_scalars *sc,
                                   double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  with_op32(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_mx_spmv, Op, gm, s, op, ph, qh, sc, part);
  });
  LAUNCH_V(k_mx_resid32, gv, s, A->n, qh, rh, sc, part);
  LAUNCH_V(k_mx_dir, gv, s, A->n, rh, y, ph, sc);
  return hipGetLastError();
}

// the FP32 SpMV + p.q kernel of one inner iteration alone (measurement hook)
extern "C" int sfr_k_mx_spmv_only(hipStream_t s, const sfr_mat_view *A,
                                  sfr_geom g0, const float *p, float *q,
                                  sfr_scalars *sc, double *part)
{
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  with_op32(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_mx_spmv, Op, gm, s, op, p, q, sc, part);
  });
  return hipGetLastError();
}
