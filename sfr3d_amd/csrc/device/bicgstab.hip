// Right-Jacobi-preconditioned BiCGSTAB in FP64 on gfx950, following the
// reference's pbicgstab (src/Bicgstab.c:50-245):
//   setup  conv = tol*||b||, r0 = b - Ax, p = r = r0, z = D^-1 p, rr0 = r.r0
//   iter   v = Az, alpha = rr0/(v.r0)
//          x += alpha z, s = r - alpha v, z = D^-1 s   stop if ||s|| < conv
//          t = Az, w = (t.s)/(t.t)
//          x += w z, r = s - w t                      stop if ||r|| < conv
//          beta = (r.r0/rr0)(alpha/w), p = r + beta(p - w v), z = D^-1 p
// Five kernels per iteration, two with the SpMV:
//   B1 k_bicg_v  v = Az, partials v.r0
//   B2 k_bicg_s  alpha; x += alpha z, s = r - alpha v, z = D^-1 s; s.s
//   B3 k_bicg_t  half-step stop test; t = Az, partials t.s, t.t
//   B4 k_bicg_r  w; x += w z, r = s - w t; r.r, r.r0
//   B5 k_bicg_p  full-step stop test; beta; p = r + beta(p - w v), z = D^-1 p
// r.r0 of the beta step is the next iteration's rr0 (the reference computes
// it twice).  Reductions and flag hand-over as in pcg.hip: B1 alone reads
// `done`, each later kernel reads the run flag its predecessor set.
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_s1(int64_t n,
    const double *__restrict__ b, double *__restrict__ x, int newX,
    double *pout)
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
  store_partials<1>(acc, pout);
}

template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_s2(Op A,
    const double *__restrict__ b, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ r, double *__restrict__ r0,
    double *__restrict__ p, double *__restrict__ z, int newX, double tol,
    unsigned maxit, const double *pin, int gin, double *pout,
    sfr_scalars *sc)
{
  double bb[1];
  reduce_partials<1>(pin, gin, bb);
  const double norm_b = sqrt(bb[0]);
  const bool zero_b = norm_b == 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->bb = bb[0];
    sc->norm_b = norm_b;
    sc->conv = tol * norm_b;
    sc->tol = tol;
    sc->maxit = maxit;
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
  store_partials<1>(acc, pout);
}

__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_start(const double *part,
    int g, sfr_scalars *sc)
{
  double t[1];
  reduce_partials<1>(part, g, t);
  if (threadIdx.x == 0) {
    sc->d = t[0];
    sc->iter = 0;
    sc->outer = 0;
    sc->ru = 0;
    sc->done = (sc->norm_b == 0.0 || t[0] == 0.0) ? SFR_DONE_CONV : SFR_DONE_RUN;
  }
}

// B1: v = Az, partials v.r0
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_v(Op A,
    const double *__restrict__ z, const double *__restrict__ r0,
    double *__restrict__ v, double *pout, sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (sc->done) {
    if (first) sc->run2 = 0;
    return;
  }
  if (first) sc->run2 = 1;
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
  store_partials<1>(acc, pout);
}

// B2: alpha = rr0/(v.r0); x += alpha z, s = r - alpha v, z = D^-1 s; s.s
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_s(int64_t n,
    const double *__restrict__ v, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
    const double *pin, int gin, double *pout, sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!sc->run2) {
    if (first) sc->run3 = 0;
    return;
  }
  double vr0[1];
  reduce_partials<1>(pin, gin, vr0);
  const double alpha = sc->d / vr0[0];
  if (!(vr0[0] != 0.0) || !isfinite(alpha)) {
    if (first) { sc->done = SFR_DONE_BREAK; sc->run3 = 0; }
    return;
  }
  if (first) { sc->alpha = alpha; sc->run3 = 1; }
  constexpr int RPC = RPT * SFR_BLOCK;
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
  store_partials<1>(acc, pout);
}

// B3: stop if ||s|| < conv; t = Az, partials t.s, t.t
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_t(Op A,
    const double *__restrict__ z, const double *__restrict__ r,
    double *__restrict__ t, const double *pin, int gin, double *pout,
    sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!sc->run3) {
    if (first) sc->run4 = 0;
    return;
  }
  double ss[1];
  reduce_partials<1>(pin, gin, ss);
  if (sqrt(ss[0]) < sc->conv || ss[0] == 0.0) {
    if (first) { sc->iter = sc->iter + 1; sc->done = SFR_DONE_CONV; sc->run4 = 0; }
    return;
  }
  if (!isfinite(ss[0])) {
    if (first) { sc->done = SFR_DONE_BREAK; sc->run4 = 0; }
    return;
  }
  if (first) sc->run4 = 1;
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
  store_partials<2>(acc, pout);
}

// B4: w = t.s / t.t; x += w z, r = s - w t; partials r.r, r.r0
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_r(int64_t n,
    const double *__restrict__ z, const double *__restrict__ t,
    const double *__restrict__ r0, double *__restrict__ x,
    double *__restrict__ r, const double *pin, int gin, double *pout,
    sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!sc->run4) {
    if (first) sc->run5 = 0;
    return;
  }
  double tt[2];
  reduce_partials<2>(pin, gin, tt);
  const double w = tt[0] / tt[1];
  if (!(tt[1] != 0.0) || !(w != 0.0) || !isfinite(w)) {
    if (first) { sc->done = SFR_DONE_BREAK; sc->run5 = 0; }
    return;
  }
  if (first) { sc->omega = w; sc->d_prev = sc->d; sc->run5 = 1; }
  constexpr int RPC = RPT * SFR_BLOCK;
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
  store_partials<2>(acc, pout);
}

// B5: stop if ||r|| < conv; beta; p = r + beta (p - w v), z = D^-1 p
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_bicg_p(int64_t n,
    const double *__restrict__ r, const double *__restrict__ v,
    const double *__restrict__ dinv, double *__restrict__ p,
    double *__restrict__ z, const double *pin, int gin, sfr_scalars *sc)
{
  if (!sc->run5) return;
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  double rr[2];
  reduce_partials<2>(pin, gin, rr);
  const double rr0 = sc->d_prev, w = sc->omega;
  const double beta = (rr[1] / rr0) * (sc->alpha / w);
  const bool conv = sqrt(rr[0]) < sc->conv || rr[0] == 0.0;
  const bool brk = !isfinite(rr[0]) || !(rr0 != 0.0) || !isfinite(beta);
  if (first) {
    const unsigned it = sc->iter + 1;
    sc->iter = it;
    sc->d = rr[1];
    if (conv) sc->done = SFR_DONE_CONV;
    else if (brk) sc->done = SFR_DONE_BREAK;
    else if (it >= sc->maxit) sc->done = SFR_DONE_MAXIT;
  }
  if (conv || brk) return;
  constexpr int RPC = RPT * SFR_BLOCK;
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

extern "C" int sfr_k_bicg_setup(hipStream_t s, const sfr_mat_view *A,
                                sfr_geom g0, const double *b, double *x,
                                double *r, double *r0, double *p, double *z,
                                int newX, double tol, unsigned maxit,
                                sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *rc = REGION(part, 2), *rb = REGION(part, 1);
  SFR_LAUNCH_V(k_bicg_s1, gv, s, A->n, b, x, newX, rc);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_bicg_s2, Op, gm, s, op, b, A->dinv, x, r, r0, p, z,
                   newX, tol, maxit, (const double *)rc, gv.grid, rb, sc);
  });
  hipLaunchKernelGGL(k_bicg_start, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)rb, gm.grid, sc);
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
  double *ra = REGION(part, 0), *rb = REGION(part, 1);
  double *rc = REGION(part, 2), *rd = REGION(part, 3);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_bicg_v, Op, gm, s, op, z, r0, v, ra, sc);
  });
  SFR_LAUNCH_V(k_bicg_s, gv, s, A->n, v, A->dinv, x, r, z, (const double *)ra,
               gm.grid, rb, sc);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_bicg_t, Op, gm, s, op, z, r, t, (const double *)rb,
                   gv.grid, rc, sc);
  });
  SFR_LAUNCH_V(k_bicg_r, gv, s, A->n, z, t, r0, x, r, (const double *)rc,
               gm.grid, rd, sc);
  SFR_LAUNCH_V(k_bicg_p, gv, s, A->n, r, v, A->dinv, p, z, (const double *)rd,
               gv.grid, sc);
  return hipGetLastError();
}
