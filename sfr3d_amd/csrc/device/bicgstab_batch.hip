// Shared-pattern batch of BiCGSTAB solves (SFR3D's u/v/w momentum systems,
// reference src/Simple.c:2980-3036: three solverC calls on the same ia/ja/al
// with their own ad, b and x; and its transported species,
// src/Combustion.c:222-240, on the same ia/ja with their own al, ad, au, b
// and x, when the view carries per-system values).  NB systems run in one
// set of launches:
//   - vectors are interleaved per row, x[i*NB + c], so the neighbour gathers
//     of all NB systems hit the same cache lines;
//   - every SpMV reads the shared columns and off-diagonal values once for
//     all NB products; each system adds its own diagonal;
//   - each system keeps its own scalars (sc[c]) and stops on its own test;
//     a finished system's vectors are left untouched while the others go on.
// The recurrences, stop tests and iteration counts per system are those of
// the single-system kernels in bicgstab.hip (reference src/Bicgstab.c:50-245).
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)
// a region holds up to 4*NB partial values of up to SFR_GRIDCAP workgroups
static_assert(4 * 4 * SFR_GRIDCAP <= 4 * SFR_MAXG, "partials region too small");

// owner thread of system c's scalars: thread c of workgroup 0
#define OWNER(c) (blockIdx.x == 0 && threadIdx.x == (unsigned)(c))

template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_s1(int64_t n,
    const double *__restrict__ b, double *__restrict__ x, int newX,
    double *pout)
{
  constexpr int RPC = RPT * SFR_BLOCK;
  double acc[NB];
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = 0.0;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < n;
       ch += gridDim.x) {
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int64_t i = ch * RPC + r * SFR_BLOCK + threadIdx.x;
      if (i < n) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          const double bi = b[i * NB + c];
          acc[c] += bi * bi;
          if (newX) x[i * NB + c] = 0.0;
        }
      }
    }
  }
  store_partials<NB>(acc, pout);
}

template <class Op, int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_s2(Op A,
    const double *__restrict__ ad, const double *__restrict__ b,
    const double *__restrict__ dinv, double *__restrict__ x,
    double *__restrict__ r, double *__restrict__ r0, double *__restrict__ p,
    double *__restrict__ z, int newX, double tol, unsigned maxit,
    const double *pin, int gin, double *pout, sfr_scalars *sc)
{
  double bb[NB];
  reduce_partials<NB>(pin, gin, bb);
#pragma unroll
  for (int c = 0; c < NB; c++)
    if (OWNER(c)) {
      sc[c].bb = bb[c];
      sc[c].norm_b = sqrt(bb[c]);
      sc[c].conv = tol * sqrt(bb[c]);
      sc[c].tol = tol;
      sc[c].maxit = maxit;
    }
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[NB];
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = 0.0;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < A.n;
       ch += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = ch * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      double ax[NB];
      if (newX) {
#pragma unroll
        for (int c = 0; c < NB; c++) ax[c] = 0.0;
      } else {
        row_multi(A, i, valid, ad, x, ax);
      }
      if (valid && lead) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          const int64_t k = i * NB + c;
          const double ri = bb[c] == 0.0 ? 0.0 : b[k] - ax[c];
          if (bb[c] == 0.0) x[k] = 0.0;
          r0[k] = ri;
          p[k] = ri;
          r[k] = ri;
          z[k] = ri * dinv[k];
          acc[c] += ri * ri;
        }
      }
    }
  }
  store_partials<NB>(acc, pout);
}

template <int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_start(const double *part,
    int g, sfr_scalars *sc)
{
  double t[NB];
  reduce_partials<NB>(part, g, t);
  if (threadIdx.x < NB) {
    const int c = threadIdx.x;
    sc[c].d = t[c];
    sc[c].iter = 0;
    sc[c].outer = 0;
    sc[c].ru = 0;
    sc[c].done = (sc[c].norm_b == 0.0 || t[c] == 0.0) ? SFR_DONE_CONV : SFR_DONE_RUN;
  }
}

// B1: v = A z; partials v.r0
template <class Op, int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_v(Op A,
    const double *__restrict__ ad, const double *__restrict__ z,
    const double *__restrict__ r0, double *__restrict__ v, double *pout,
    sfr_scalars *sc)
{
  bool any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    const bool act = sc[c].done == SFR_DONE_RUN;
    any |= act;
    if (OWNER(c)) sc[c].run2 = act;
  }
  if (!any) return;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[NB];
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = 0.0;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < A.n;
       ch += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = ch * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      double vi[NB];
      row_multi(A, i, valid, ad, z, vi);
      if (valid && lead) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          v[i * NB + c] = vi[c];
          acc[c] += vi[c] * r0[i * NB + c];
        }
      }
    }
  }
  store_partials<NB>(acc, pout);
}

// B2: alpha; x += alpha z, s = r - alpha v, z = D^-1 s; partials s.s
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_s(int64_t n,
    const double *__restrict__ v, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
    const double *pin, int gin, double *pout, sfr_scalars *sc)
{
  bool act[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) { act[c] = sc[c].run2 != 0; any |= act[c]; }
  if (!any) {
#pragma unroll
    for (int c = 0; c < NB; c++) if (OWNER(c)) sc[c].run3 = 0;
    return;
  }
  double vr0[NB], alpha[NB];
  reduce_partials<NB>(pin, gin, vr0);
#pragma unroll
  for (int c = 0; c < NB; c++) {
    alpha[c] = sc[c].d / vr0[c];
    const bool brk = act[c] && (!(vr0[c] != 0.0) || !isfinite(alpha[c]));
    if (OWNER(c)) {
      if (brk) sc[c].done = SFR_DONE_BREAK;
      sc[c].alpha = alpha[c];
      sc[c].run3 = act[c] && !brk;
    }
    act[c] = act[c] && !brk;
  }
  constexpr int RPC = RPT * SFR_BLOCK;
  double acc[NB];
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = 0.0;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < n;
       ch += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = ch * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (!act[c]) continue;
          const int64_t k = i * NB + c;
          x[k] += alpha[c] * z[k];
          const double si = r[k] - alpha[c] * v[k];
          r[k] = si;
          z[k] = si * dinv[k];
          acc[c] += si * si;
        }
      }
    }
  }
  store_partials<NB>(acc, pout);
}

// B3: half-step stop test; t = A z; partials t.s, t.t
template <class Op, int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_t(Op A,
    const double *__restrict__ ad, const double *__restrict__ z,
    const double *__restrict__ r, double *__restrict__ t, const double *pin,
    int gin, double *pout, sfr_scalars *sc)
{
  bool act[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) act[c] = sc[c].run3 != 0;
  double ss[NB];
  bool any_run = false;
#pragma unroll
  for (int c = 0; c < NB; c++) any_run |= act[c];
  if (!any_run) {
#pragma unroll
    for (int c = 0; c < NB; c++) if (OWNER(c)) sc[c].run4 = 0;
    return;
  }
  reduce_partials<NB>(pin, gin, ss);
#pragma unroll
  for (int c = 0; c < NB; c++) {
    const bool conv = act[c] && (sqrt(ss[c]) < sc[c].conv || ss[c] == 0.0);
    const bool brk = act[c] && !conv && !isfinite(ss[c]);
    if (OWNER(c)) {
      if (conv) { sc[c].iter = sc[c].iter + 1; sc[c].done = SFR_DONE_CONV; }
      if (brk) sc[c].done = SFR_DONE_BREAK;
      sc[c].run4 = act[c] && !conv && !brk;
    }
    act[c] = act[c] && !conv && !brk;
    any |= act[c];
  }
  if (!any) return;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[2 * NB];
#pragma unroll
  for (int c = 0; c < 2 * NB; c++) acc[c] = 0.0;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < A.n;
       ch += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = ch * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      double ti[NB];
      row_multi(A, i, valid, ad, z, ti);
      if (valid && lead) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (!act[c]) continue;
          t[i * NB + c] = ti[c];
          acc[c] += ti[c] * r[i * NB + c];
          acc[NB + c] += ti[c] * ti[c];
        }
      }
    }
  }
  store_partials<2 * NB>(acc, pout);
}

// B4: w = t.s/t.t; x += w z, r = s - w t; partials r.r, r.r0
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_r(int64_t n,
    const double *__restrict__ z, const double *__restrict__ t,
    const double *__restrict__ r0, double *__restrict__ x,
    double *__restrict__ r, const double *pin, int gin, double *pout,
    sfr_scalars *sc)
{
  bool act[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) { act[c] = sc[c].run4 != 0; any |= act[c]; }
  if (!any) {
#pragma unroll
    for (int c = 0; c < NB; c++) if (OWNER(c)) sc[c].run5 = 0;
    return;
  }
  double tt[2 * NB], w[NB];
  reduce_partials<2 * NB>(pin, gin, tt);
#pragma unroll
  for (int c = 0; c < NB; c++) {
    w[c] = tt[c] / tt[NB + c];
    const bool brk = act[c] && (!(tt[NB + c] != 0.0) || !(w[c] != 0.0) || !isfinite(w[c]));
    if (OWNER(c)) {
      if (brk) sc[c].done = SFR_DONE_BREAK;
      sc[c].omega = w[c];
      sc[c].d_prev = sc[c].d;
      sc[c].run5 = act[c] && !brk;
    }
    act[c] = act[c] && !brk;
  }
  constexpr int RPC = RPT * SFR_BLOCK;
  double acc[2 * NB];
#pragma unroll
  for (int c = 0; c < 2 * NB; c++) acc[c] = 0.0;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < n;
       ch += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = ch * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (!act[c]) continue;
          const int64_t k = i * NB + c;
          x[k] += w[c] * z[k];
          const double ri = r[k] - w[c] * t[k];
          r[k] = ri;
          acc[c] += ri * ri;
          acc[NB + c] += ri * r0[k];
        }
      }
    }
  }
  store_partials<2 * NB>(acc, pout);
}

// B5: stop test; beta; p = r + beta (p - w v), z = D^-1 p
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_p(int64_t n,
    const double *__restrict__ r, const double *__restrict__ v,
    const double *__restrict__ dinv, double *__restrict__ p,
    double *__restrict__ z, const double *pin, int gin, sfr_scalars *sc)
{
  bool act[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) { act[c] = sc[c].run5 != 0; any |= act[c]; }
  if (!any) return;
  double rr[2 * NB], beta[NB], w[NB];
  reduce_partials<2 * NB>(pin, gin, rr);
  any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    const double rr0 = sc[c].d_prev;
    w[c] = sc[c].omega;
    beta[c] = (rr[NB + c] / rr0) * (sc[c].alpha / w[c]);
    const bool conv = act[c] && (sqrt(rr[c]) < sc[c].conv || rr[c] == 0.0);
    const bool brk = act[c] && !conv &&
                     (!isfinite(rr[c]) || !(rr0 != 0.0) || !isfinite(beta[c]));
    if (OWNER(c) && act[c]) {
      const unsigned it = sc[c].iter + 1;
      sc[c].iter = it;
      sc[c].d = rr[NB + c];
      if (conv) sc[c].done = SFR_DONE_CONV;
      else if (brk) sc[c].done = SFR_DONE_BREAK;
      else if (it >= sc[c].maxit) sc[c].done = SFR_DONE_MAXIT;
    }
    act[c] = act[c] && !conv && !brk;
    any |= act[c];
  }
  if (!any) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < n;
       ch += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = ch * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (!act[c]) continue;
          const int64_t k = i * NB + c;
          const double pi = r[k] + beta[c] * (p[k] - w[c] * v[k]);
          p[k] = pi;
          z[k] = pi * dinv[k];
        }
      }
    }
  }
}

// per system: x.Ax, x.x, r.D^-1 r, r.r with r = b - Ax
template <class Op, int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_final(Op A,
    const double *__restrict__ ad, const double *__restrict__ b,
    const double *__restrict__ x, const double *__restrict__ dinv,
    double *pout)
{
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[4 * NB];
#pragma unroll
  for (int c = 0; c < 4 * NB; c++) acc[c] = 0.0;
  for (int64_t ch = xcd_chunk(blockIdx.x, gridDim.x); ch * RPC < A.n;
       ch += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = ch * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      double ax[NB];
      row_multi(A, i, valid, ad, x, ax);
      if (valid && lead) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          const int64_t k = i * NB + c;
          const double xi = x[k], ri = b[k] - ax[c];
          acc[c] += xi * ax[c];
          acc[NB + c] += xi * xi;
          acc[2 * NB + c] += ri * (ri * dinv[k]);
          acc[3 * NB + c] += ri * ri;
        }
      }
    }
  }
  store_partials<4 * NB>(acc, pout);
}

template <int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_bb_final_sum(const double *part,
    int g, sfr_scalars *sc)
{
  double t[4 * NB];
  reduce_partials<4 * NB>(part, g, t);
  if (threadIdx.x < NB) {
    const int c = threadIdx.x;
    for (int k = 0; k < 4; k++) sc[c].res[k] = t[k * NB + c];
  }
}

// system-major (src[c*n + i]) <-> interleaved (dst[i*NB + c]); also forms
// dinv = 1/ad when `inv` is set
__global__ __launch_bounds__(SFR_BLOCK) void k_interleave(int64_t n, int nb,
    const double *__restrict__ src, double *__restrict__ dst,
    double *__restrict__ inv)
{
  const int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (k < n * nb) {
    const int64_t i = k / nb, c = k % nb;
    const double v = src[c * n + i];
    dst[k] = v;
    if (inv) inv[k] = 1.0 / v;
  }
}

__global__ __launch_bounds__(SFR_BLOCK) void k_deinterleave(int64_t n, int nb,
    const double *__restrict__ src, double *__restrict__ dst)
{
  const int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (k < n * nb) {
    const int64_t i = k / nb, c = k % nb;
    dst[c * n + i] = src[k];
  }
}

// ---- launchers ------------------------------------------------------------

#define BB_LAUNCH_M(K, Op, NB, g, s, ...)                                      \
  do {                                                                         \
    switch ((g).rpt) {                                                         \
    case 1: hipLaunchKernelGGL((K<Op, 1, NB>), dim3((g).grid), dim3(SFR_BLOCK),\
                               0, s, __VA_ARGS__); break;                      \
    case 2: hipLaunchKernelGGL((K<Op, 2, NB>), dim3((g).grid), dim3(SFR_BLOCK),\
                               0, s, __VA_ARGS__); break;                      \
    case 4: hipLaunchKernelGGL((K<Op, 4, NB>), dim3((g).grid), dim3(SFR_BLOCK),\
                               0, s, __VA_ARGS__); break;                      \
    default: hipLaunchKernelGGL((K<Op, 8, NB>), dim3((g).grid),                \
                                dim3(SFR_BLOCK), 0, s, __VA_ARGS__); break;    \
    }                                                                          \
  } while (0)

#define BB_LAUNCH_V(K, NB, g, s, ...)                                          \
  do {                                                                         \
    switch ((g).rpt) {                                                         \
    case 1: hipLaunchKernelGGL((K<1, NB>), dim3((g).grid), dim3(SFR_BLOCK), 0, \
                               s, __VA_ARGS__); break;                         \
    case 2: hipLaunchKernelGGL((K<2, NB>), dim3((g).grid), dim3(SFR_BLOCK), 0, \
                               s, __VA_ARGS__); break;                         \
    case 4: hipLaunchKernelGGL((K<4, NB>), dim3((g).grid), dim3(SFR_BLOCK), 0, \
                               s, __VA_ARGS__); break;                         \
    default: hipLaunchKernelGGL((K<8, NB>), dim3((g).grid), dim3(SFR_BLOCK),   \
                                0, s, __VA_ARGS__); break;                     \
    }                                                                          \
  } while (0)

template <int NB>
static void bb_setup(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                     const double *ad, const double *dinv, const double *b,
                     double *x, double *r, double *r0, double *p, double *z,
                     int newX, double tol, unsigned maxit, sfr_scalars *sc,
                     double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *rc = REGION(part, 2), *rb = REGION(part, 1);
  BB_LAUNCH_V(k_bb_s1, NB, gv, s, A->n, b, x, newX, rc);
  with_op64_b(A, [&](auto op) {
    using Op = decltype(op);
    BB_LAUNCH_M(k_bb_s2, Op, NB, gm, s, op, ad, b, dinv, x, r, r0, p, z, newX,
                tol, maxit, (const double *)rc, gv.grid, rb, sc);
  });
  hipLaunchKernelGGL(k_bb_start<NB>, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)rb, gm.grid, sc);
}

template <int NB>
static void bb_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                    const double *ad, const double *dinv, double *x, double *r,
                    const double *r0, double *p, double *z, double *v,
                    double *t, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *ra = REGION(part, 0), *rb = REGION(part, 1);
  double *rc = REGION(part, 2), *rd = REGION(part, 3);
  with_op64_b(A, [&](auto op) {
    using Op = decltype(op);
    BB_LAUNCH_M(k_bb_v, Op, NB, gm, s, op, ad, (const double *)z, r0, v, ra, sc);
  });
  BB_LAUNCH_V(k_bb_s, NB, gv, s, A->n, (const double *)v, dinv, x, r, z,
              (const double *)ra, gm.grid, rb, sc);
  with_op64_b(A, [&](auto op) {
    using Op = decltype(op);
    BB_LAUNCH_M(k_bb_t, Op, NB, gm, s, op, ad, (const double *)z,
                (const double *)r, t, (const double *)rb, gv.grid, rc, sc);
  });
  BB_LAUNCH_V(k_bb_r, NB, gv, s, A->n, (const double *)z, (const double *)t,
              r0, x, r, (const double *)rc, gm.grid, rd, sc);
  BB_LAUNCH_V(k_bb_p, NB, gv, s, A->n, (const double *)r, (const double *)v,
              dinv, p, z, (const double *)rd, gv.grid, sc);
}

template <int NB>
static void bb_final(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                     const double *ad, const double *dinv, const double *b,
                     const double *x, sfr_scalars *sc, double *part)
{
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *rd = REGION(part, 3);
  with_op64_b(A, [&](auto op) {
    using Op = decltype(op);
    BB_LAUNCH_M(k_bb_final, Op, NB, gm, s, op, ad, b, x, dinv, rd);
  });
  hipLaunchKernelGGL(k_bb_final_sum<NB>, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)rd, gm.grid, sc);
}

#define NB_DISPATCH(nb, CALL)                                                  \
  switch (nb) {                                                                \
  case 2: CALL(2); break;                                                      \
  case 3: CALL(3); break;                                                      \
  case 4: CALL(4); break;                                                      \
  default: return hipErrorInvalidValue;                                        \
  }

extern "C" int sfr_k_bb_setup(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g, int nb, const double *ad,
                              const double *dinv, const double *b, double *x,
                              double *r, double *r0, double *p, double *z,
                              int newX, double tol, unsigned maxit,
                              sfr_scalars *sc, double *part)
{
#define CALL(NB) bb_setup<NB>(s, A, g, ad, dinv, b, x, r, r0, p, z, newX, tol, maxit, sc, part)
  NB_DISPATCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

extern "C" int sfr_k_bb_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                             int nb, const double *ad, const double *dinv,
                             double *x, double *r, const double *r0,
                             double *p, double *z, double *v, double *t,
                             sfr_scalars *sc, double *part)
{
#define CALL(NB) bb_iter<NB>(s, A, g, ad, dinv, x, r, r0, p, z, v, t, sc, part)
  NB_DISPATCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

extern "C" int sfr_k_bb_final(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g, int nb, const double *ad,
                              const double *dinv, const double *b,
                              const double *x, sfr_scalars *sc, double *part)
{
#define CALL(NB) bb_final<NB>(s, A, g, ad, dinv, b, x, sc, part)
  NB_DISPATCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

__global__ __launch_bounds__(SFR_BLOCK) void k_gather_nb(int64_t nnz,
    const int32_t *__restrict__ vmap, const double *__restrict__ raw,
    double *__restrict__ valnb, int nb, int c)
{
  const int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (k < nnz) {
    const int32_t m = vmap[k];
    valnb[k * nb + c] = m >= 0 ? raw[m] : 0.0;
  }
}

extern "C" int sfr_k_gather_nb(hipStream_t s, int64_t nnz, const int32_t *vmap,
                               const double *raw, double *valnb, int nb, int c)
{
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_nb, dim3((unsigned)((nnz + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, nnz, vmap, raw, valnb, nb, c);
  return hipGetLastError();
}

extern "C" int sfr_k_interleave(hipStream_t s, int64_t n, int nb,
                                const double *src, double *dst, double *inv)
{
  const int64_t tot = n * nb;
  hipLaunchKernelGGL(k_interleave, dim3((unsigned)((tot + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, n, nb, src, dst, inv);
  return hipGetLastError();
}

extern "C" int sfr_k_deinterleave(hipStream_t s, int64_t n, int nb,
                                  const double *src, double *dst)
{
  const int64_t tot = n * nb;
  hipLaunchKernelGGL(k_deinterleave, dim3((unsigned)((tot + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, n, nb, src, dst);
  return hipGetLastError();
}
