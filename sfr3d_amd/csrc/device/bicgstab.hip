// Right-Jacobi-preconditioned BiCGSTAB in FP64 on gfx950, following the
// reference's pbicgstab (src/Bicgstab.c:50-245):
//   setup  conv = tol*||b||, r0 = b - Ax, p = r = r0, z = D^-1 p, rr0 = r.r0
//   iter   v = Az, alpha = rr0/(v.r0)
//          x += alpha z, s = r - alpha v, z = D^-1 s   stop if ||s|| < conv
//          t = Az, w = (t.s)/(t.t)
//          x += w z, r = s - w t                      stop if ||r|| < conv
//          beta = (r.r0/rr0)(alpha/w), p = r + beta(p - w v), z = D^-1 p
// Five kernels per iteration, two of them with the SpMV.  r.r0 of the beta
// step is the next iteration's rr0 (the reference computes it twice).
#include "dev_common.hpp"

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_setup1(int64_t n,
    const double *__restrict__ b, double *__restrict__ x, int newX,
    double tol, unsigned maxit, sfr_scalars *sc, double *part)
{
  constexpr int RPC = RPT * SFR_BLOCK;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int64_t i = c * RPC + r * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double bi = b[i];
        acc[0] += bi * bi;
        if (newX) x[i] = 0.0;
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    sc->bb = acc[0];
    sc->norm_b = sqrt(acc[0]);
    sc->conv = tol * sc->norm_b;
    sc->tol = tol;
    sc->maxit = maxit;
  }
}

template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_setup2(Op A,
    const double *__restrict__ b, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ r, double *__restrict__ r0,
    double *__restrict__ p, double *__restrict__ z, int newX,
    sfr_scalars *sc, double *part)
{
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  const bool zero_b = sc->norm_b == 0.0;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double ax = (newX || zero_b) ? 0.0 : A.row(i, valid, x);
      if (valid && lead) {
        const double ri = zero_b ? 0.0 : b[i] - ax;
        if (zero_b) x[i] = 0.0;
        r0[i] = ri;
        p[i] = ri;
        r[i] = ri;
        z[i] = ri * dinv[i];
        acc[0] += ri * ri;
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    sc->d = acc[0];
    sc->iter = 0;
    sc->seal = 0;
    sc->done = zero_b ? SFR_DONE_CONV : SFR_DONE_RUN;
  }
}

// v = Az, alpha = rr0 / (v.r0)
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_v(Op A,
    const double *__restrict__ z, const double *__restrict__ r0,
    double *__restrict__ v, sfr_scalars *sc, double *part)
{
  if (sc->done) return;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double vi = A.row(i, valid, z);
      if (valid && lead) {
        v[i] = vi;
        acc[0] += vi * r0[i];
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    const double alpha = sc->d / acc[0];
    sc->alpha = alpha;
    if (!(acc[0] != 0.0) || !isfinite(alpha)) sc->done = SFR_DONE_BREAK;
  }
}

// x += alpha z, s = r - alpha v, z = D^-1 s, stop if ||s|| < conv
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_s(int64_t n,
    const double *__restrict__ v, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
    sfr_scalars *sc, double *part)
{
  if (sc->done) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const double alpha = sc->alpha;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        x[i] += alpha * z[i];
        const double si = r[i] - alpha * v[i];
        r[i] = si;
        z[i] = si * dinv[i];
        acc[0] += si * si;
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    if (sqrt(acc[0]) < sc->conv) {
      sc->iter = sc->iter + 1;
      sc->done = SFR_DONE_CONV;
    } else if (!isfinite(acc[0])) {
      sc->done = SFR_DONE_BREAK;
    }
  }
}

// t = Az, w = (t.s)/(t.t)
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_t(Op A,
    const double *__restrict__ z, const double *__restrict__ r,
    double *__restrict__ t, sfr_scalars *sc, double *part)
{
  if (sc->done) return;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[2] = {0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double ti = A.row(i, valid, z);
      if (valid && lead) {
        t[i] = ti;
        acc[0] += ti * r[i];
        acc[1] += ti * ti;
      }
    }
  }
  if (grid_sum<2>(acc, part, &sc->counter) && threadIdx.x == 0) {
    const double w = acc[0] / acc[1];
    sc->omega = w;
    if (!(acc[1] != 0.0) || !(w != 0.0) || !isfinite(w))
      sc->done = SFR_DONE_BREAK;
  }
}

// x += w z, r = s - w t, stop if ||r|| < conv, beta = (r.r0/rr0)(alpha/w)
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_r(int64_t n,
    const double *__restrict__ z, const double *__restrict__ t,
    const double *__restrict__ r0, double *__restrict__ x,
    double *__restrict__ r, sfr_scalars *sc, double *part)
{
  if (sc->done) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const double w = sc->omega;
  double acc[2] = {0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        x[i] += w * z[i];
        const double ri = r[i] - w * t[i];
        r[i] = ri;
        acc[0] += ri * ri;
        acc[1] += ri * r0[i];
      }
    }
  }
  if (grid_sum<2>(acc, part, &sc->counter) && threadIdx.x == 0) {
    const unsigned it = sc->iter + 1;
    sc->iter = it;
    if (sqrt(acc[0]) < sc->conv) {
      sc->done = SFR_DONE_CONV;
    } else if (!isfinite(acc[0]) || !(sc->d != 0.0)) {
      sc->done = SFR_DONE_BREAK;
    } else {
      sc->beta = (acc[1] / sc->d) * (sc->alpha / w);
      sc->d = acc[1];
      if (it >= sc->maxit) sc->done = SFR_DONE_MAXIT;
    }
  }
}

// p = r + beta (p - w v), z = D^-1 p
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_p(int64_t n,
    const double *__restrict__ r, const double *__restrict__ v,
    const double *__restrict__ dinv, double *__restrict__ p,
    double *__restrict__ z, sfr_scalars *sc)
{
  if (sc->done) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const double beta = sc->beta, w = sc->omega;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double pi = r[i] + beta * (p[i] - w * v[i]);
        p[i] = pi;
        z[i] = pi * dinv[i];
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

extern "C" int sfr_k_bicg_setup(hipStream_t s, const sfr_mat_view *A,
                                sfr_geom g0, const double *b, double *x,
                                double *r, double *r0, double *p, double *z,
                                int newX, double tol, unsigned maxit,
                                sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  LAUNCH_V(k_bicg_setup1, gv, s, A->n, b, x, newX, tol, maxit, sc, part);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_bicg_setup2, Op, gm, s, op, b, A->dinv, x, r, r0, p, z,
                   newX, sc, part);
  });
  return hipGetLastError();
}

extern "C" int sfr_k_bicg_iter(hipStream_t s, const sfr_mat_view *A,
                               sfr_geom g0, double *x, double *r,
                               const double *r0, double *p, double *z,
                               double *v, double *t, sfr_scalars *sc,
                               double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_bicg_v, Op, gm, s, op, z, r0, v, sc, part);
  });
  LAUNCH_V(k_bicg_s, gv, s, A->n, v, A->dinv, x, r, z, sc, part);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_bicg_t, Op, gm, s, op, z, r, t, sc, part);
  });
  LAUNCH_V(k_bicg_r, gv, s, A->n, z, t, r0, x, r, sc, part);
  LAUNCH_V(k_bicg_p, gv, s, A->n, r, v, A->dinv, p, z, sc);
  return hipGetLastError();
}
