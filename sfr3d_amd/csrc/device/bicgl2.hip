// BiCGSTAB(2) with right Jacobi preconditioning in FP64 on gfx950,
// following the reference's pbicgstabl2 (src/Bicgstab.c:1172-1378) as
// solverC drives it (src/Solv.c:217-295).  One iteration is an even and an
// odd BiCG step and a GCR(2) minimisation:
//   rr0 = -w2 rr0; rr1 = r.r0; beta = alpha rr1/rr0; rr0 = rr1
//   u = r - beta u; z = D^-1 u; v = Az; alpha = rr0/(v.r0)
//   r -= alpha v; p = D^-1 r; s = Ap; x += alpha z
//   rr1 = s.r0; beta = alpha rr1/rr0; rr0 = rr1
//   v = s - beta v; p = D^-1 v; w = Ap; alpha = rr0/(w.r0)
//   u = r - beta u; r -= alpha v; s -= alpha w; p = D^-1 s; t = Ap
//   GCR(2) from r.s, s.s, s.t, t.t, r.t -> w1, w2
//   x += w1 D^-1 r + w2 p + alpha D^-1 u; r -= w1 s + w2 t
//   stop if ||r|| < tol ||b||; u -= w1 v + w2 w
// Nine kernels per iteration, four with the SpMV:
//   L1 k_b2_u   stop test on the previous ||r||; beta; u, z   (u -= w1 v + w2 w
//               of the previous iteration rides here)
//   L2 k_b2_v   v = Az; v.r0          L3 k_b2_r   alpha; r, p = D^-1 r
//   L4 k_b2_s   s = Ap; s.r0; x += alpha z
//   L5 k_b2_vo  beta; v = s - beta v; p = D^-1 v
//   L6 k_b2_w   w = Ap; w.r0          L7 k_b2_uo  alpha; u, r, s, p = D^-1 s
//   L8 k_b2_t   t = Ap; r.s, s.s, s.t, t.t, r.t
//   L9 k_b2_x   w1, w2; x, r; r.r, r.r0
// D^-1 r and D^-1 u of the x update are recomputed in L9 from r and u (the
// same products as the reference's stored h and z), and t shares z's
// buffer: eight work vectors.  Scalars: each field of B2 is written by one
// kernel and read only by later ones; L1 reads run2 and sets run3, which
// L2-L9 read; L9 hands run2 on.
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)

struct B2 {
  double rr0b;      // rr0 after the odd step (s.r0)          L5 -> L1
  double alpha2;    // alpha of the odd step                  L7 -> L9, L1
  double om1, om2;  // GCR(2) weights                         L9 -> L1
  double rr0a;      // r.r0 of the even step                  L1 -> L3, L5
  double alpha1;    // alpha of the even step                 L3 -> L4, L5
  double beta2;     // beta of the odd step                   L5 -> L7
  unsigned it;      // completed iterations                   L9 -> L1
  unsigned first;   // no iteration completed yet             L9 -> L1
};
#define B2_OF(part) ((B2 *)REGION(part, 13))

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_s1(int64_t n,
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

// r0 = r = b - Ax; u = v = w = 0; partials r.r, r.r0
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_s2(Op A,
    const double *__restrict__ b, const double *__restrict__ x,
    double *__restrict__ r0, double *__restrict__ r, double *__restrict__ u,
    double *__restrict__ v, double *__restrict__ w, int newX, double tol,
    unsigned maxit, const double *pin, int gin, double *pout,
    sfr_scalars *sc)
{
  double bb[1];
  reduce_partials<1>(pin, gin, bb);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double norm_b = sqrt(bb[0]);
    sc->bb = bb[0];
    sc->norm_b = norm_b;
    sc->conv = tol * norm_b;
    sc->tol = tol;
    sc->maxit = maxit;
  }
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[2] = {0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double ax = newX ? 0.0 : A.row(i, valid, x);
      if (valid && lead) {
        const double ri = b[i] - ax;
        r0[i] = ri;
        r[i] = ri;
        u[i] = 0.0;
        v[i] = 0.0;
        w[i] = 0.0;
        acc[0] += ri * ri;
        acc[1] += ri * ri;
      }
    }
  }
  store_partials<2>(acc, pout);
}

// s2's two partial rows re-laid out for the vector grid L1 reduces (the
// totals in slot 0, zeros after), and the recurrence scalars' start values
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_s3(const double *pin,
    int gin, double *pout, int gout, B2 *B, sfr_scalars *sc)
{
  double t[2];
  reduce_partials<2>(pin, gin, t);
  for (int k = threadIdx.x; k < gout; k += SFR_BLOCK) {
    pout[k] = k == 0 ? t[0] : 0.0;
    pout[(size_t)gout + k] = k == 0 ? t[1] : 0.0;
  }
  if (threadIdx.x == 0) {
    B->rr0b = 1.0;
    B->alpha2 = 0.0;
    B->om1 = 0.0;
    B->om2 = 1.0;
    B->it = 0;
    B->first = 1;
    sc->iter = 0;
    sc->outer = 0;
    sc->ru = 0;
    const bool zero = t[0] == 0.0;
    sc->done = zero ? SFR_DONE_CONV : SFR_DONE_RUN;
    sc->run2 = zero ? 0u : 1u;
  }
}

// L1: stop test on ||r|| of the previous iteration; beta;
// u = r - beta (u - (w1 v + w2 w)); z = D^-1 u
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_u(int64_t n,
    const double *__restrict__ r, double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w,
    const double *__restrict__ dinv, double *__restrict__ z, const B2 *B,
    double *rr0a_out, const double *pin, int gin, sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!sc->run2) {
    if (first) sc->run3 = 0;
    return;
  }
  double t[2];
  reduce_partials<2>(pin, gin, t);
  const unsigned it = B->it;
  const bool conv = !B->first && sqrt(t[0]) < sc->conv;
  const bool bad = !isfinite(t[0]);
  const bool maxed = it >= sc->maxit;
  const double om1 = B->om1, om2 = B->om2;
  const double rr0 = -om2 * B->rr0b;
  const double beta = B->alpha2 * t[1] / rr0;
  const bool brk = !isfinite(beta);
  if (conv || bad || maxed || brk) {
    if (first) {
      sc->iter = it;
      sc->done = conv ? SFR_DONE_CONV : (bad || brk) ? SFR_DONE_BREAK : SFR_DONE_MAXIT;
      sc->run3 = 0;
    }
    return;
  }
  if (first) { *rr0a_out = t[1]; sc->run3 = 1; }
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double u1 = u[i] - (om1 * v[i] + om2 * w[i]);
        const double ui = r[i] - beta * u1;
        u[i] = ui;
        z[i] = ui * dinv[i];
      }
    }
  }
}

// L2 / L6: y = A q, partials y.r0; L4 adds x += alpha1 z (xz != NULL)
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_mv(Op A,
    const double *__restrict__ q, const double *__restrict__ r0,
    double *__restrict__ y, double *__restrict__ x,
    const double *__restrict__ zx, const B2 *B, double *pout,
    const sfr_scalars *sc)
{
  if (!sc->run3) return;
  const double a1 = x ? B->alpha1 : 0.0;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double yi = A.row(i, valid, q);
      if (valid && lead) {
        y[i] = yi;
        acc[0] += yi * r0[i];
        if (x) x[i] += a1 * zx[i];
      }
    }
  }
  store_partials<1>(acc, pout);
}

// L3: alpha1 = rr0a / (v.r0); r -= alpha1 v; p = D^-1 r
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_r(int64_t n,
    double *__restrict__ r, const double *__restrict__ v,
    const double *__restrict__ dinv, double *__restrict__ p, B2 *B,
    const double *pin, int gin, const sfr_scalars *sc)
{
  if (!sc->run3) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const int64_t c0 = xcd_chunk(blockIdx.x, gridDim.x);
  double rv[RPT], vv[RPT], dv[RPT];
#pragma unroll
  for (int r_ = 0; r_ < RPT; r_++) {
    const int64_t i = c0 * RPC + r_ * SFR_BLOCK + threadIdx.x;
    rv[r_] = i < n ? r[i] : 0.0;
    vv[r_] = i < n ? v[i] : 0.0;
    dv[r_] = i < n ? dinv[i] : 0.0;
  }
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double alpha = B->rr0a / t[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) B->alpha1 = alpha;
  for (int64_t c = c0; c * RPC < n; c += gridDim.x) {
    if (c != c0) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
        rv[r_] = i < n ? r[i] : 0.0;
        vv[r_] = i < n ? v[i] : 0.0;
        dv[r_] = i < n ? dinv[i] : 0.0;
      }
    }
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double ri = rv[r_] - alpha * vv[r_];
        r[i] = ri;
        p[i] = ri * dv[r_];
      }
    }
  }
}

// L5: rr1 = s.r0; beta2 = alpha1 rr1 / rr0a; rr0b = rr1;
// v = s - beta2 v; p = D^-1 v
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_vo(int64_t n,
    const double *__restrict__ s, double *__restrict__ v,
    const double *__restrict__ dinv, double *__restrict__ p, B2 *B,
    const double *pin, int gin, const sfr_scalars *sc)
{
  if (!sc->run3) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const int64_t c0 = xcd_chunk(blockIdx.x, gridDim.x);
  double sv[RPT], vv[RPT], dv[RPT];
#pragma unroll
  for (int r_ = 0; r_ < RPT; r_++) {
    const int64_t i = c0 * RPC + r_ * SFR_BLOCK + threadIdx.x;
    sv[r_] = i < n ? s[i] : 0.0;
    vv[r_] = i < n ? v[i] : 0.0;
    dv[r_] = i < n ? dinv[i] : 0.0;
  }
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double beta = B->alpha1 * t[0] / B->rr0a;
  if (blockIdx.x == 0 && threadIdx.x == 0) { B->beta2 = beta; B->rr0b = t[0]; }
  for (int64_t c = c0; c * RPC < n; c += gridDim.x) {
    if (c != c0) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
        sv[r_] = i < n ? s[i] : 0.0;
        vv[r_] = i < n ? v[i] : 0.0;
        dv[r_] = i < n ? dinv[i] : 0.0;
      }
    }
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double vi = sv[r_] - beta * vv[r_];
        v[i] = vi;
        p[i] = vi * dv[r_];
      }
    }
  }
}

// L7: alpha2 = rr0b / (w.r0); u = r - beta2 u; r -= alpha2 v;
// s -= alpha2 w; p = D^-1 s
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_uo(int64_t n,
    double *__restrict__ u, double *__restrict__ r,
    const double *__restrict__ v, double *__restrict__ s,
    const double *__restrict__ w, const double *__restrict__ dinv,
    double *__restrict__ p, B2 *B, const double *pin, int gin,
    const sfr_scalars *sc)
{
  if (!sc->run3) return;
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double alpha = B->rr0b / t[0], beta = B->beta2;
  if (blockIdx.x == 0 && threadIdx.x == 0) B->alpha2 = alpha;
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double ri = r[i];
        u[i] = ri - beta * u[i];
        r[i] = ri - alpha * v[i];
        const double si = s[i] - alpha * w[i];
        s[i] = si;
        p[i] = si * dinv[i];
      }
    }
  }
}

// L8: t = Ap; partials r.s, s.s, s.t, t.t, r.t
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_t(Op A,
    const double *__restrict__ p, const double *__restrict__ r,
    const double *__restrict__ s, double *__restrict__ t, double *pout,
    const sfr_scalars *sc)
{
  if (!sc->run3) return;
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double ti = A.row(i, valid, p);
      if (valid && lead) {
        const double ri = r[i], si = s[i];
        t[i] = ti;
        acc[0] += ri * si;
        acc[1] += si * si;
        acc[2] += si * ti;
        acc[3] += ti * ti;
        acc[4] += ri * ti;
      }
    }
  }
  store_partials<5>(acc, pout);
}

// L9: GCR(2) weights; x += w1 D^-1 r + w2 p + alpha2 D^-1 u;
// r -= w1 s + w2 t; partials r.r, r.r0
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_b2_x(int64_t n,
    double *__restrict__ x, double *__restrict__ r,
    const double *__restrict__ u, const double *__restrict__ p,
    const double *__restrict__ s, const double *__restrict__ t,
    const double *__restrict__ r0, const double *__restrict__ dinv, B2 *B,
    const double *pin, int gin, double *pout, sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!sc->run3) {
    if (first) sc->run2 = 0;
    return;
  }
  double g[5];
  reduce_partials<5>(pin, gin, g);
  double om1 = g[0];
  const double mi = g[1], ni = g[2];
  double tau = g[3], om2 = g[4];
  tau = tau - ni * ni / mi;
  om2 = (om2 - ni * om1 / mi) / tau;
  om1 = (om1 - ni * om2) / mi;
  const double a2 = B->alpha2;
  if (first) {
    B->om1 = om1;
    B->om2 = om2;
    B->it = B->it + 1;
    B->first = 0;
    sc->run2 = 1;
  }
  constexpr int RPC = RPT * SFR_BLOCK;
  double acc[2] = {0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double m = dinv[i], ri = r[i];
        const double hi = ri * m, zi = u[i] * m;
        x[i] += om1 * hi + om2 * p[i] + a2 * zi;
        const double rn = ri - (om1 * s[i] + om2 * t[i]);
        r[i] = rn;
        acc[0] += rn * rn;
        acc[1] += rn * r0[i];
      }
    }
  }
  store_partials<2>(acc, pout);
}

// vectors: W[0] r0, W[1] r, W[2] u, W[3] v, W[4] w, W[5] s, W[6] p,
// W[7] z and t
extern "C" int sfr_k_b2_setup(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, const double *b, double *x,
                              double *const W[8], int newX, double tol,
                              unsigned maxit, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *r2 = REGION(part, 2), *r3 = REGION(part, 3), *r1 = REGION(part, 1);
  SFR_LAUNCH_V(k_b2_s1, gv, s, A->n, b, x, newX, r2);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_b2_s2, Op, gm, s, op, b, (const double *)x, W[0], W[1], W[2],
                   W[3], W[4], newX, tol, maxit, (const double *)r2, gv.grid, r3, sc);
  });
  hipLaunchKernelGGL(k_b2_s3, dim3(1), dim3(SFR_BLOCK), 0, s, (const double *)r3,
                     gm.grid, r1, gv.grid, B2_OF(part), sc);
  return hipGetLastError();
}

extern "C" int sfr_k_b2_iter(hipStream_t s, const sfr_mat_view *A,
                             sfr_geom g0, double *x, double *const W[8],
                             sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *R0 = REGION(part, 0), *R1 = REGION(part, 1), *R2 = REGION(part, 2);
  B2 *B = B2_OF(part);
  double *r0 = W[0], *r = W[1], *u = W[2], *v = W[3], *w = W[4];
  double *sv = W[5], *p = W[6], *z = W[7];
  const double *m = A->dinv;
  const int64_t n = A->n;
  SFR_LAUNCH_V(k_b2_u, gv, s, n, (const double *)r, u, (const double *)v,
               (const double *)w, m, z, (const B2 *)B, &B->rr0a,
               (const double *)R1, gv.grid, sc);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_b2_mv, Op, gm, s, op, (const double *)z, (const double *)r0, v,
                   (double *)NULL, (const double *)NULL, (const B2 *)B, R0,
                   (const sfr_scalars *)sc);
  });
  SFR_LAUNCH_V(k_b2_r, gv, s, n, r, (const double *)v, m, p, B, (const double *)R0,
               gm.grid, (const sfr_scalars *)sc);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_b2_mv, Op, gm, s, op, (const double *)p, (const double *)r0, sv,
                   x, (const double *)z, (const B2 *)B, R1, (const sfr_scalars *)sc);
  });
  SFR_LAUNCH_V(k_b2_vo, gv, s, n, (const double *)sv, v, m, p, B, (const double *)R1,
               gm.grid, (const sfr_scalars *)sc);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_b2_mv, Op, gm, s, op, (const double *)p, (const double *)r0, w,
                   (double *)NULL, (const double *)NULL, (const B2 *)B, R0,
                   (const sfr_scalars *)sc);
  });
  SFR_LAUNCH_V(k_b2_uo, gv, s, n, u, r, (const double *)v, sv, (const double *)w, m, p,
               B, (const double *)R0, gm.grid, (const sfr_scalars *)sc);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_b2_t, Op, gm, s, op, (const double *)p, (const double *)r,
                   (const double *)sv, z, R2, (const sfr_scalars *)sc);
  });
  SFR_LAUNCH_V(k_b2_x, gv, s, n, x, r, (const double *)u, (const double *)p,
               (const double *)sv, (const double *)z, (const double *)r0, m, B,
               (const double *)R2, gm.grid, R1, sc);
  return hipGetLastError();
}
