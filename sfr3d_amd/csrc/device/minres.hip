// Jacobi-preconditioned MINRES in FP64 on gfx950, following the
// reference's pminres (src/Iters.c:309-555) and its plane rotation symOrtho
// (src/Iters.c:1171-1208), as solverC drives them (src/Solv.c:380-445):
//   setup  conv = tol ||D^-1/2 b||; v_0 = b - Ax; z_0 = D^-1 v_0;
//          beta_0 = sqrt|z_0.v_0|; eta = ||r|| = beta_0; c = 1, s = 0
//   iter j p = A z_j; alpha_j = (z_j.p)/beta_j^2
//          v_{j+1} = p/beta_j - (alpha_j/beta_j) v_j - (beta_j/beta_{j-1}) v_{j-1}
//          z_{j+1} = D^-1 v_{j+1}; beta_{j+1} = sqrt(z_{j+1}.v_{j+1})
//          QR step: delta, rho2, rho3 from the last two rotations;
//          (c, s, rho1) = symOrtho(delta, beta_{j+1})
//          w_j = (z_j/beta_j - rho2 w_{j-1} - rho3 w_{j-2}) / rho1
//          x += c eta w_j; ||r|| *= |s|, stop if < conv; eta = -s eta
// Two kernels per iteration:
//   M1 k_mr_a(j)  the QR step and the w, x update of iteration j-1 (beta_j
//                 from the partials of M2(j-1)), the stop test; then
//                 p = A z_j with partials z_j.p
//   M2 k_mr_b(j)  alpha_j; v_{j+1}, z_{j+1}; partials z_{j+1}.v_{j+1}
// The vectors rotate through rings (v: 3, w: 3, z: 2) instead of being
// copied, so one captured chunk holds a multiple of 6 iterations.  The
// recurrence scalars live in two sets alternated by iteration parity: M1(j)
// reads set[j&1] and writes set[(j+1)&1], so no kernel reads a scalar that
// it writes; M2(j) adds alpha_j to set[(j+1)&1].  A stopped M1 passes its
// stop flag on to the next set, so later kernels of the chunk return at once.
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)
#define MR_SETS(part) ((MrSet *)REGION(part, 12))

struct MrSet {
  double b1, b2;            // beta_{j-1}, beta_{j-2} as M1(j) sees them
  double c1, c2, s1, s2;    // rotations j-1 and j-2
  double eta, nr, r1;       // eta_{j-1}, recurrence ||r||, last rho1
  double alpha;             // alpha_{j-1}, written by M2(j-1)
  unsigned it, stop;        // iterations begun (j); stop flag
  double pad[4];
};

// src/Iters.c:1171-1208 as written: the third test compares |b| with a
// itself, and when no case holds (a == b > 0) c, s, r keep their values.
__device__ __forceinline__ void sym_ortho(double a, double b, double &c,
                                          double &s, double &r)
{
  const double ma = fabs(a), mb = fabs(b);
  const double sa = a > 0.0 ? 1.0 : -1.0, sb = b > 0.0 ? 1.0 : -1.0;
  if (b == 0.0) {
    s = 0.0;
    r = ma;
    c = a == 0.0 ? 1.0 : sa;
  } else if (a == 0.0) {
    c = 0.0;
    s = sb;
    r = mb;
  } else if (mb > a) {
    const double t = a / b;
    s = sb / sqrt(1.0 + t * t);
    c = s * t;
    r = b / s;
  } else if (ma > mb) {
    const double t = b / a;
    c = sa / sqrt(1.0 + t * t);
    s = c * t;
    r = a / c;
  }
}

// partials of (D^-1/2 b).(D^-1/2 b) and b.b; x = 0 when newX
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mr_s1(int64_t n,
    const double *__restrict__ b, const double *__restrict__ dinv,
    double *__restrict__ x, int newX, double *pout)
{
  constexpr int RPC = RPT * SFR_BLOCK;
  double acc[2] = {0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int64_t i = c * RPC + r * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double bi = b[i], zi = bi * sqrt(dinv[i]);
        acc[0] += zi * zi;
        acc[1] += bi * bi;
        if (newX) x[i] = 0.0;
      }
    }
  }
  store_partials<2>(acc, pout);
}

// v_0 = b - Ax, z_0 = D^-1 v_0, v_{-1} = w_{-1} = w_{-2} = 0; partials
// v_0.z_0; workgroup 0 stores the norms and opens set 0
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mr_s2(Op A,
    const double *__restrict__ b, const double *__restrict__ x,
    const double *__restrict__ dinv, double *__restrict__ v,
    double *__restrict__ z, double *__restrict__ vm1,
    double *__restrict__ wa, double *__restrict__ wb, int newX, double tol,
    unsigned maxit, const double *pin, int gin, double *pout, MrSet *sets,
    sfr_scalars *sc)
{
  double t[2];
  reduce_partials<2>(pin, gin, t);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double norm_b = sqrt(t[0]);
    sc->norm_b = norm_b;
    sc->conv = tol * norm_b;
    sc->bb = t[1];
    sc->tol = tol;
    sc->maxit = maxit;
    sc->iter = 0;
    sc->outer = 0;
    sc->ru = 0;
    sc->done = SFR_DONE_RUN;
    sets[0].it = 0;
    sets[0].stop = 0;
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
      const double ax = newX ? 0.0 : A.row(i, valid, x);
      if (valid && lead) {
        const double vi = b[i] - ax, zi = vi * dinv[i];
        v[i] = vi;
        z[i] = zi;
        vm1[i] = 0.0;
        wa[i] = 0.0;
        wb[i] = 0.0;
        acc[0] += vi * zi;
      }
    }
  }
  store_partials<1>(acc, pout);
}

// s2's partials (one per workgroup of the SpMV grid) folded into the
// vector grid's slot count that M1 reduces: the total in slot 0, zeros after
__global__ __launch_bounds__(SFR_BLOCK) void k_mr_s3(const double *pin,
    int gin, double *pout, int gout)
{
  double t[1];
  reduce_partials<1>(pin, gin, t);
  for (int k = threadIdx.x; k < gout; k += SFR_BLOCK) pout[k] = k == 0 ? t[0] : 0.0;
}

// M1(j): the end of iteration j-1 (QR step, w_{j-1}, x, stop test), then
// p = A z_j and partials z_j.p.
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mr_a(Op A,
    const double *__restrict__ zc, const double *__restrict__ zo,
    double *__restrict__ wn, const double *__restrict__ w1,
    const double *__restrict__ w2, double *__restrict__ x,
    double *__restrict__ p, const MrSet *sin, MrSet *sout, const double *pin,
    int gin, double *pout, sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (sin->stop) {
    if (first) sout->stop = 1;
    return;
  }
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const unsigned j = sin->it;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  if (j == 0) {
    const double b0 = sqrt(fabs(t[0]));
    const bool zero = b0 == 0.0, bad = !isfinite(b0);
    if (first) {
      sout->b1 = b0; sout->b2 = b0;
      sout->c1 = 1.0; sout->c2 = 1.0; sout->s1 = 0.0; sout->s2 = 0.0;
      sout->eta = b0; sout->nr = b0; sout->r1 = 1.0;
      sout->it = 1;
      sout->stop = (zero || bad) ? 1u : 0u;
      if (zero) sc->done = SFR_DONE_CONV;
      else if (bad) sc->done = SFR_DONE_BREAK;
    }
    if (zero || bad) return;
  } else {
    const double bj = sqrt(t[0]);
    const double alpha = sin->alpha, b1 = sin->b1;
    const double c1 = sin->c1, c2 = sin->c2, s1 = sin->s1, s2 = sin->s2;
    const double delta = c1 * alpha - c2 * s1 * b1;
    const double ro2 = s1 * alpha + c2 * c1 * b1;
    const double ro3 = s2 * b1;
    double c = c1, s = s1, r1 = sin->r1;
    sym_ortho(delta, bj, c, s, r1);
    const double tmp1 = 1.0 / (b1 * r1), tmp2 = -ro2 / r1, tmp3 = -ro3 / r1;
    const double tmp4 = c * sin->eta;
    const bool ok = isfinite(bj) && isfinite(tmp1) && isfinite(tmp2) &&
                    isfinite(tmp3) && isfinite(tmp4);
    const double nr = fabs(s) * sin->nr;
    const bool conv = ok && nr < sc->conv;
    const bool stop = !ok || conv || j >= sc->maxit;
    if (first) {
      sout->b1 = bj; sout->b2 = b1;
      sout->c1 = c; sout->c2 = c1; sout->s1 = s; sout->s2 = s1;
      sout->eta = -s * sin->eta; sout->nr = nr; sout->r1 = r1;
      sout->it = j + 1;
      sout->stop = stop ? 1u : 0u;
      if (stop) {
        sc->iter = j;
        sc->done = !ok ? SFR_DONE_BREAK : conv ? SFR_DONE_CONV : SFR_DONE_MAXIT;
      }
    }
    if (ok) {
      for (int64_t cc = xcd_chunk(blockIdx.x, gridDim.x); cc * RPC < A.n;
           cc += gridDim.x) {
#pragma unroll
        for (int r_ = 0; r_ < RPT; r_++) {
          const int64_t i = cc * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
          if (i < A.n && lead) {
            const double wi = tmp1 * zo[i] + tmp2 * w1[i] + tmp3 * w2[i];
            wn[i] = wi;
            x[i] += tmp4 * wi;
          }
        }
      }
    }
    if (stop) return;
  }
  double acc[1] = {0.0};
  for (int64_t cc = xcd_chunk(blockIdx.x, gridDim.x); cc * RPC < A.n;
       cc += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = cc * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double pi = A.row(i, valid, zc);
      if (valid && lead) {
        p[i] = pi;
        acc[0] += zc[i] * pi;
      }
    }
  }
  store_partials<1>(acc, pout);
}

// M2(j): alpha_j; v_{j+1} = p/beta_j - (alpha_j/beta_j) v_j
// - (beta_j/beta_{j-1}) v_{j-1} in place of p; z_{j+1} = D^-1 v_{j+1};
// partials z_{j+1}.v_{j+1}
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_mr_b(int64_t n,
    double *__restrict__ p, const double *__restrict__ v,
    const double *__restrict__ v0, const double *__restrict__ dinv,
    double *__restrict__ zn, MrSet *S, const double *pin, int gin,
    double *pout)
{
  if (S->stop) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const int64_t c0 = xcd_chunk(blockIdx.x, gridDim.x);
  double pv[RPT], vv[RPT], ov[RPT], dv[RPT];
#pragma unroll
  for (int r_ = 0; r_ < RPT; r_++) {
    const int64_t i = c0 * RPC + r_ * SFR_BLOCK + threadIdx.x;
    pv[r_] = i < n ? p[i] : 0.0;
    vv[r_] = i < n ? v[i] : 0.0;
    ov[r_] = i < n ? v0[i] : 0.0;
    dv[r_] = i < n ? dinv[i] : 0.0;
  }
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double bj = S->b1, bo = S->b2;
  const double alpha = t[0] * (1.0 / (bj * bj));
  if (blockIdx.x == 0 && threadIdx.x == 0) S->alpha = alpha;
  const double tmp1 = 1.0 / bj, tmp2 = -alpha / bj, tmp3 = -bj / bo;
  double acc[1] = {0.0};
  for (int64_t c = c0; c * RPC < n; c += gridDim.x) {
    if (c != c0) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
        pv[r_] = i < n ? p[i] : 0.0;
        vv[r_] = i < n ? v[i] : 0.0;
        ov[r_] = i < n ? v0[i] : 0.0;
        dv[r_] = i < n ? dinv[i] : 0.0;
      }
    }
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double pi = tmp1 * pv[r_] + tmp2 * vv[r_] + tmp3 * ov[r_];
        const double zi = pi * dv[r_];
        p[i] = pi;
        zn[i] = zi;
        acc[0] += zi * pi;
      }
    }
  }
  store_partials<1>(acc, pout);
}

// Ring buffers: V[k % 3] = v_k, W[k % 3] = w_k, Z[k % 2] = z_k.
extern "C" int sfr_k_mr_setup(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, const double *b, double *x,
                              double *const V[3], double *const W[3],
                              double *const Z[2], int newX, double tol,
                              unsigned maxit, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *r2 = REGION(part, 2), *r1 = REGION(part, 1), *r3 = REGION(part, 3);
  SFR_LAUNCH_V(k_mr_s1, gv, s, A->n, b, A->dinv, x, newX, r2);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_mr_s2, Op, gm, s, op, b, (const double *)x, A->dinv, V[0], Z[0],
                   V[2], W[1], W[2], newX, tol, maxit, (const double *)r2, gv.grid,
                   r3, MR_SETS(part), sc);
  });
  hipLaunchKernelGGL(k_mr_s3, dim3(1), dim3(SFR_BLOCK), 0, s, (const double *)r3,
                     gm.grid, r1, gv.grid);
  return hipGetLastError();
}

// One iteration at ring phase q = j % 6.
extern "C" int sfr_k_mr_iter(hipStream_t s, const sfr_mat_view *A,
                             sfr_geom g0, double *x, double *const V[3],
                             double *const W[3], double *const Z[2], int q,
                             sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *r0 = REGION(part, 0), *r1 = REGION(part, 1);
  MrSet *sets = MR_SETS(part);
  MrSet *sin = sets + (q & 1), *sout = sets + ((q + 1) & 1);
  double *zc = Z[q % 2], *zo = Z[(q + 1) % 2];
  double *wn = W[(q + 2) % 3], *w1 = W[(q + 1) % 3], *w2 = W[q % 3];
  double *vj = V[q % 3], *vo = V[(q + 2) % 3], *p = V[(q + 1) % 3];
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_mr_a, Op, gm, s, op, (const double *)zc, (const double *)zo, wn,
                   (const double *)w1, (const double *)w2, x, p, (const MrSet *)sin,
                   sout, (const double *)r1, gv.grid, r0, sc);
  });
  SFR_LAUNCH_V(k_mr_b, gv, s, A->n, p, (const double *)vj, (const double *)vo,
               A->dinv, zo, sout, (const double *)r0, gm.grid, r1);
  return hipGetLastError();
}
