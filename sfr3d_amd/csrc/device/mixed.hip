// Mixed-precision Jacobi-PCG: FP32 CG iterations with FP64 reliable updates.
//
// The Krylov iteration runs on the Jacobi-scaled system
//     (D^-1/2 A D^-1/2) y = D^-1/2 r,
// i.e. Jacobi-PCG on A written with a unit diagonal, with the off-diagonals
// and the vectors y, r, p, q stored in FP32 (dot products accumulated in
// FP64).  The correction y is folded into the FP64 solution x every time the
// FP32 residual has dropped by a factor delta below its largest value since
// the previous fold (reliable updates: G. Sleijpen, H. van der Vorst,
// "Reliable updated residuals in hybrid Bi-CG methods", Computing 56 (1996);
// the CG form used by M. Clark et al., "Solving lattice QCD systems of
// equations using mixed precision solvers on GPUs", CPC 181 (2010)):
//   x += D^-1/2 y, y = 0               (FP64)
//   r64 = b - A x                      (FP64 matrix, FP64 arithmetic)
//   stop if sqrt|r64.D^-1 r64| < tol * sqrt|b.D^-1 b|   -- the reference
//        PCG's own test (src/Cg.c:73-78,141-142) on the true FP64 residual
//   r = D^-1/2 r64 (to FP32); p = r + beta (p - (r.p / r.r) r),
//   beta = r.r / (r.r before the fold)
// Unlike restarted refinement, the search direction survives the fold, so
// the FP32 loop keeps CG's convergence rate.
// Per FP32 iteration (two kernels; the direction lives in two buffers that
// alternate, pb[k & 1] -> pb[(k + 1) & 1], and the inner r.r in dd[k & 1]):
//   MA k_mx_a   beta = r.r / r.r_prev; fold decision; y += alpha p_old;
//               p_new = r + beta p_old (own rows, stored);
//               q = A^ p_new, formed as A^ (r + beta p_old) in the gathers;
//               partials p_new.q                    matrix + 6 words
//   MB k_mx_b   alpha = r.r / p.q; r -= alpha q; partials r.r    3 words
// The first MA after a fold (flag `fresh`) finds the direction already
// built and only forms q = A^ p.  A fold is requested by MA (or a
// breakdown by MB) through the flag `ru`, which only the host reads; it
// sees it at its next poll and enqueues the fold kernels R1-R3.
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)

// one workgroup: norms of b (partials of k_pcg_s1) -> tolerance, loop state
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_start(const double *pin,
    int gin, double tol, unsigned maxit, double delta, sfr_scalars *sc)
{
  double nb[2];
  reduce_partials<2>(pin, gin, nb);
  if (threadIdx.x == 0) {
    const double norm_b = sqrt(fabs(nb[0]));
    sc->norm_b = norm_b;
    sc->conv = tol * norm_b;
    sc->bb = nb[1];
    sc->tol = tol;
    sc->maxit = maxit;
    sc->delta2 = delta * delta;
    sc->iter = 0;
    sc->outer = 0;
    sc->ru = 0;
    sc->fresh = 0;
    sc->run5 = 0;
    sc->pbuf = 1;
    sc->done = norm_b == 0.0 ? SFR_DONE_CONV : SFR_DONE_RUN;
  }
}

// R1: x += D^-1/2 y, y = 0
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_fold(int64_t n,
    float *__restrict__ y, const double *__restrict__ dscale,
    double *__restrict__ x)
{
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        x[i] += dscale[i] * (double)y[i];
        y[i] = 0.0f;
      }
    }
  }
}

// R2: r64 = b - A x in FP64; r = D^-1/2 r64 in FP32 (and p_new = r, y = 0 on
// the first call); partials of r64.D^-1 r64, r64.r64, r.r, r.p_old
template <class Op, int RPT, bool ZERO_X>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_true(Op A,
    const double *__restrict__ b, const double *__restrict__ x,
    const double *__restrict__ dinv, const double *__restrict__ dscale,
    float *__restrict__ r, const float *po, float *pn,
    float *__restrict__ y, int init, double *pout, sfr_scalars *sc)
{
  if (blockIdx.x == 0 && threadIdx.x == 0)
    sc->run2 = sc->done ? 0u : ((init || sc->ru == 2) ? 2u : 1u);
  if (sc->done) return;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
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
        r[i] = ci;
        acc[0] += ri * (ri * dinv[i]);
        acc[1] += ri * ri;
        acc[2] += (double)ci * (double)ci;
        if (init) {
          pn[i] = ci;
          y[i] = 0.0f;
        } else {
          acc[3] += (double)ci * (double)po[i];
        }
      }
    }
  }
  store_partials<4>(acc, pout);
}

// restart after an FP32 breakdown: p = r (R2 writes p only on the first call)
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_copy(int64_t n,
    const float *__restrict__ r, float *__restrict__ p, int init,
    sfr_scalars *sc)
{
  if (init || sc->run2 != 2) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) p[i] = r[i];
    }
  }
}

// R3: FP64 stop test; p_new = r + beta (p_old - (r.p/r.r) r) after a fold
// (p_old and p_new may be the same buffer: elementwise, in place)
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_restart(int64_t n,
    const float *__restrict__ r, const float *po, float *pn, const double *pin,
    int gin, int init, sfr_scalars *sc)
{
  const unsigned mode = sc->run2;          // 0 skip, 1 fold, 2 restart (p = r)
  if (!mode) return;
  double t[4];
  reduce_partials<4>(pin, gin, t);
  const double rmr = t[0], rr = t[2];
  const bool conv = sqrt(fabs(rmr)) < sc->conv || rmr == 0.0;
  const bool brk = !isfinite(rmr) || !isfinite(rr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->res[2] = rmr;
    sc->res[3] = t[1];
    if (!init) sc->outer = sc->outer + 1;
    if (conv) sc->done = SFR_DONE_CONV;
    else if (brk) sc->done = SFR_DONE_BREAK;
    else if (sc->iter >= sc->maxit) sc->done = SFR_DONE_MAXIT;
    sc->dd[1] = rr;                        // read by MB at parity 0
    sc->rho_max = rr;
    sc->ru = 0;
    sc->fresh = 1;
    sc->run5 = 1;
  }
  if (conv || brk || mode == 2) return;
  const double beta = rr / sc->d_prev;
  const float fb = (float)beta, fg = (float)(t[3] / rr);
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const float ri = r[i];
        pn[i] = ri + fb * (po[i] - fg * ri);
      }
    }
  }
}

// MA: the fused first kernel of an FP32 iteration (see the file comment).
// Reads run5 (set by MB), never ru; writes run4, ru, iter, dd[par ^ 1].
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_a(Op A,
    const float *__restrict__ r, const float *__restrict__ po,
    float *__restrict__ pn, float *__restrict__ q, float *__restrict__ y,
    const double *pin, int gin, double *pout, int par, sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (sc->done || !sc->run5) {
    if (first) sc->run4 = 0;
    return;
  }
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[1] = {0.0};
  if (sc->fresh) {
    // first iteration after a fold: the direction is already in pn
    if (first) sc->run4 = 1;
    for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
         c += gridDim.x) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
        const bool valid = i < A.n;
        const float qi = A.row(i, valid, pn);
        if (valid && lead) {
          q[i] = qi;
          acc[0] += (double)qi * (double)pn[i];
        }
      }
    }
    store_partials<1>(acc, pout);
    return;
  }
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double rr = t[0], rr_old = sc->dd[par];
  const double conv = sc->conv;
  const bool fold = rr < sc->delta2 * sc->rho_max || rr <= conv * conv ||
                    !isfinite(rr);
  const float fa = (float)sc->alpha, fb = (float)(rr / rr_old);
  if (first) {
    const unsigned it = sc->iter + 1;
    sc->iter = it;
    sc->dd[par ^ 1] = rr;
    if (fold || it >= sc->maxit) {
      sc->ru = 1;
      sc->run4 = 0;
      sc->pbuf = fold ? (unsigned)par : (unsigned)(par ^ 1);
      sc->d_prev = rr_old;
    } else {
      sc->run4 = 1;
    }
  }
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      if (fold) {
        if (valid && lead) y[i] = fmad(fa, po[i], y[i]);
        continue;
      }
      const float qi = A.row_comb(i, valid, r, po, fb);
      if (valid && lead) {
        const float pi = po[i];
        // the same rounding as each neighbour's r_j + beta p_j in row_comb
        const float pni = fmad(fb, pi, r[i]);
        y[i] = fmad(fa, pi, y[i]);
        pn[i] = pni;
        q[i] = qi;
        acc[0] += (double)qi * (double)pni;
      }
    }
  }
  if (!fold) store_partials<1>(acc, pout);
}

// MB: alpha = r.r / p.q; r -= alpha q; partials r.r.  Reads run4 and
// dd[par ^ 1]; writes run5, alpha, rho_max, fresh (and ru on a breakdown).
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mx_b(int64_t n,
    const float *__restrict__ q, float *__restrict__ r, const double *pin,
    int gin, double *pout, int par, sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!sc->run4) {
    if (first) sc->run5 = 0;
    return;
  }
  constexpr int RPC = RPT * SFR_BLOCK;
  // the first chunk's rows are loaded before the reduction, whose latency
  // they then hide
  const int64_t c0 = xcd_chunk(blockIdx.x, gridDim.x);
  float rv[RPT], qv[RPT];
#pragma unroll
  for (int r_ = 0; r_ < RPT; r_++) {
    const int64_t i = c0 * RPC + r_ * SFR_BLOCK + threadIdx.x;
    rv[r_] = i < n ? r[i] : 0.0f;
    qv[r_] = i < n ? q[i] : 0.0f;
  }
  double pq[1];
  reduce_partials<1>(pin, gin, pq);
  const double rr = sc->dd[par ^ 1];
  const double alpha = rr / pq[0];
  if (!(pq[0] > 0.0) || !isfinite(alpha)) {
    // FP32 lost positive definiteness: restart from the true residual
    if (first) {
      sc->ru = 2;
      sc->run5 = 0;
      sc->pbuf = (unsigned)(par ^ 1);
    }
    return;
  }
  if (first) {
    sc->alpha = alpha;
    if (rr > sc->rho_max) sc->rho_max = rr;
    sc->fresh = 0;
    sc->run5 = 1;
  }
  const float fa = (float)alpha;
  double acc[1] = {0.0};
  for (int64_t c = c0; c * RPC < n; c += gridDim.x) {
    if (c != c0) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
        rv[r_] = i < n ? r[i] : 0.0f;
        qv[r_] = i < n ? q[i] : 0.0f;
      }
    }
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const float ri = rv[r_] - fa * qv[r_];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
      }
    }
  }
  store_partials<1>(acc, pout);
}

extern "C" int sfr_k_mx_start(hipStream_t s, const double *part, int nparts,
                              double tol, unsigned maxit, double delta,
                              sfr_scalars *sc)
{
  hipLaunchKernelGGL(k_mx_start, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)REGION(part, 2), nparts, tol, maxit,
                     delta, sc);
  return hipGetLastError();
}

// fold kernels R1-R3 (init: first call -- no fold, p_new = r, y = 0)
extern "C" int sfr_k_mx_fold(hipStream_t s, const sfr_mat_view *A,
                             sfr_geom g0, const double *b, double *x,
                             float *y, float *r, const float *p_old,
                             float *p_new, int init, int zero_x,
                             sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1, g0.cap)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A), g0.cap)};
  double *rc = REGION(part, 2);
  if (!init) SFR_LAUNCH_V(k_mx_fold, gv, s, A->n, y, A->dscale, x);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    switch (gm.rpt) {
#define CASE(R)                                                                \
    case R:                                                                    \
      if (zero_x && init)                                                      \
        hipLaunchKernelGGL((k_mx_true<Op, R, true>), dim3(gm.grid),            \
                           dim3(SFR_BLOCK), 0, s, op, b, x, A->dinv,           \
                           A->dscale, r, p_old, p_new, y, init, rc, sc);       \
      else                                                                     \
        hipLaunchKernelGGL((k_mx_true<Op, R, false>), dim3(gm.grid),           \
                           dim3(SFR_BLOCK), 0, s, op, b, x, A->dinv,           \
                           A->dscale, r, p_old, p_new, y, init, rc, sc);       \
      break;
    CASE(1) CASE(2) CASE(4)
    default: CASE(8)
#undef CASE
    }
  });
  SFR_LAUNCH_V(k_mx_copy, gv, s, A->n, (const float *)r, p_new, init, sc);
  SFR_LAUNCH_V(k_mx_restart, gv, s, A->n, (const float *)r, p_old, p_new,
               (const double *)rc, gm.grid, init, sc);
  return hipGetLastError();
}

// MB's launch shape.  MB streams two words a row and gathers nothing, so
// what bounds it is the loads a thread has in flight: a grid-stride pass
// issues its rows' loads only after the previous pass's arrive.  MB
// therefore takes the fewest passes the workgroup cap allows (up to 8 rows
// per thread) whatever rows per thread MA runs; its partial count is MA's
// `gin`.
static inline sfr_geom mb_geom(int64_t n, sfr_geom g0)
{
  const int64_t cap = (g0.cap > 0 && g0.cap < SFR_GRIDCAP) ? g0.cap : SFR_GRIDCAP;
  int rpt = g0.rpt;
  while (rpt < 8 && (int64_t)rpt * SFR_BLOCK * cap < n) rpt *= 2;
  return {rpt, grid_for(n, rpt, 1, g0.cap)};
}

extern "C" int sfr_k_mx_inner_iter(hipStream_t s, const sfr_mat_view *A,
                                   sfr_geom g0, float *y, float *r,
                                   float *const pb[2], float *q, int par,
                                   sfr_scalars *sc, double *part)
{
  const sfr_geom gb = mb_geom(A->n, g0);
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A), g0.cap)};
  double *ra = REGION(part, 0), *rb = REGION(part, 1);
  const float *po = pb[par];
  float *pn = pb[par ^ 1];
  with_op32(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_mx_a, Op, gm, s, op, (const float *)r, po, pn, q, y,
                   (const double *)rb, gb.grid, ra, par, sc);
  });
  SFR_LAUNCH_V(k_mx_b, gb, s, A->n, (const float *)q, r, (const double *)ra,
               gm.grid, rb, par, sc);
  return hipGetLastError();
}

// MA alone (measurement hook); the caller sets the scalars and the r.r
// partials so that every launch takes the fused path
extern "C" int sfr_k_mx_a_only(hipStream_t s, const sfr_mat_view *A,
                               sfr_geom g0, const float *r, const float *po,
                               float *pn, float *q, float *y, sfr_scalars *sc,
                               double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1, g0.cap)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A), g0.cap)};
  with_op32(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_mx_a, Op, gm, s, op, r, po, pn, q, y,
                   (const double *)REGION(part, 1), gv.grid, REGION(part, 0),
                   0, sc);
  });
  return hipGetLastError();
}
