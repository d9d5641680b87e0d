// Jacobi-preconditioned CG in FP64 on gfx950, following the recurrences of
// the reference's pcg (src/Cg.c:47-209):
//   setup  z = D^-1 b, conv = tol*sqrt|b.z|, r = b - Ax, z = D^-1 r, p = z,
//          d = r.z
//   iter   q = Ap, alpha = d/(p.q), x += alpha p, r -= alpha q, z = D^-1 r,
//          di = r.z, beta = di/d, p = z + beta p, d = di, stop when
//          sqrt|d| < conv
// One iteration is three kernels, each one streaming pass over its vectors:
//   k_pcg_spmv   q = Ap, p.q            -> alpha            (matrix + 2 words)
//   k_pcg_resid  r -= alpha q, r.D^-1 r -> beta, stop test  (4 words)
//   k_pcg_dir    x += alpha p, p = D^-1 r + beta p           (6 words)
// The x update rides with the direction update (same alpha, same old p, the
// same rounding as the reference order) so z never has to be stored.  The
// direction kernel of the iteration that converged still runs; the `seal`
// flag, raised by the first spmv kernel that finds the loop finished, stops
// it afterwards.  Reductions are deterministic (fixed order), so a solve is
// bitwise reproducible run to run.
#include "dev_common.hpp"

// ---- utility kernels --------------------------------------------------------

__global__ __launch_bounds__(SFR_BLOCK) void k_gather(int64_t nnz,
    const int32_t *__restrict__ vmap, const double *__restrict__ raw,
    double *__restrict__ val)
{
  const int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (k < nnz) {
    const int32_t m = vmap[k];
    val[k] = m >= 0 ? raw[m] : 0.0;
  }
}

__global__ __launch_bounds__(SFR_BLOCK) void k_diag(int64_t n,
    const int32_t *__restrict__ dmap, const double *__restrict__ raw,
    const double *__restrict__ raw_ad, double *__restrict__ ad,
    double *__restrict__ dinv)
{
  const int64_t i = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (i < n) {
    const double a = dmap ? raw[dmap[i]] : raw_ad[i];
    ad[i] = a;
    dinv[i] = 1.0 / a;                 // preCondDiag, src/PreCond.c:18-26
  }
}

__global__ __launch_bounds__(SFR_BLOCK) void k_fill(int64_t n, double v,
                                                    double *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (i < n) x[i] = v;
}

// Jacobi-scaled FP32 copy: a_ij / sqrt(a_ii a_jj); ELL rows are i = k mod n.
__global__ __launch_bounds__(SFR_BLOCK) void k_dscale(int64_t n,
    const double *__restrict__ ad, double *__restrict__ dscale)
{
  const int64_t i = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (i < n) dscale[i] = 1.0 / sqrt(ad[i]);
}

__global__ __launch_bounds__(SFR_BLOCK) void k_scale32_ell(int64_t n,
    int64_t nnz, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ dscale,
    float *__restrict__ val32)
{
  const int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (k < nnz) {
    const int64_t i = k % n;
    val32[k] = (float)(val[k] * dscale[i] * dscale[col[k]]);
  }
}

__global__ __launch_bounds__(SFR_BLOCK) void k_scale32_csr(int64_t n,
    const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ dscale,
    float *__restrict__ val32)
{
  const int64_t i = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (i < n)
    for (int k = ptr[i]; k < ptr[i + 1]; k++)
      val32[k] = (float)(val[k] * dscale[i] * dscale[col[k]]);
}

template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_spmv(Op A,
    const double *__restrict__ x, double *__restrict__ y)
{
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int64_t i = c * RPC + r * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double yi = A.row(i, valid, x);
      if (valid && lead) y[i] = yi;
    }
  }
}

// ---- PCG --------------------------------------------------------------------

// z = D^-1 b; b.z and b.b; x = 0 when newX
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_setup1(int64_t n,
    const double *__restrict__ b, const double *__restrict__ dinv,
    double *__restrict__ x, int newX, double tol, unsigned maxit,
    sfr_scalars *sc, double *part)
{
  constexpr int RPC = RPT * SFR_BLOCK;
  double acc[2] = {0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int64_t i = c * RPC + r * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double bi = b[i];
        acc[0] += bi * (bi * dinv[i]);
        acc[1] += bi * bi;
        if (newX) x[i] = 0.0;
      }
    }
  }
  if (grid_sum<2>(acc, part, &sc->counter) && threadIdx.x == 0) {
    sc->norm_b = sqrt(fabs(acc[0]));
    sc->conv = tol * sc->norm_b;
    sc->bb = acc[1];
    sc->tol = tol;
    sc->maxit = maxit;
  }
}

// r = b - Ax, p = D^-1 r, d = r.p   (b == 0: x = r = p = 0, done at once)
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_setup2(Op A,
    const double *__restrict__ b, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ r, double *__restrict__ p,
    int newX, sfr_scalars *sc, double *part)
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
        if (zero_b) {
          x[i] = 0.0; r[i] = 0.0; p[i] = 0.0;
        } else {
          const double ri = b[i] - ax;
          const double zi = ri * dinv[i];
          r[i] = ri;
          p[i] = zi;
          acc[0] += ri * zi;
        }
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

template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_spmv(Op A,
    const double *__restrict__ p, double *__restrict__ q, sfr_scalars *sc,
    double *part)
{
  if (sc->done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->seal = 1;
    return;
  }
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int64_t i = c * RPC + r * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double qi = A.row(i, valid, p);
      if (valid && lead) {
        q[i] = qi;
        acc[0] += qi * p[i];
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    const double alpha = sc->d / acc[0];
    sc->alpha = alpha;
    if (!(acc[0] != 0.0) || !isfinite(alpha)) sc->done = SFR_DONE_BREAK;
  }
}

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_resid(int64_t n,
    const double *__restrict__ q, const double *__restrict__ dinv,
    double *__restrict__ r, sfr_scalars *sc, double *part)
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
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc[0] += ri * (ri * dinv[i]);
      }
    }
  }
  if (grid_sum<1>(acc, part, &sc->counter) && threadIdx.x == 0) {
    const double di = acc[0];
    sc->beta = di / sc->d;
    sc->d = di;
    const unsigned it = sc->iter + 1;
    sc->iter = it;
    if (sqrt(fabs(di)) < sc->conv) sc->done = SFR_DONE_CONV;
    else if (!isfinite(di)) sc->done = SFR_DONE_BREAK;
    else if (it >= sc->maxit) sc->done = SFR_DONE_MAXIT;
  }
}

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_dir(int64_t n,
    const double *__restrict__ r, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ p, sfr_scalars *sc)
{
  if (sc->seal || sc->done == SFR_DONE_BREAK) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  const double alpha = sc->alpha, beta = sc->beta;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double pi = p[i];
        x[i] += alpha * pi;
        p[i] = r[i] * dinv[i] + beta * pi;
      }
    }
  }
}

// x.Ax, x.x, r.D^-1 r and r.r with r = b - Ax  (src/Cg.c:155-176)
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_final(Op A,
    const double *__restrict__ b, const double *__restrict__ x,
    const double *__restrict__ dinv, sfr_scalars *sc, double *part)
{
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double ax = A.row(i, valid, x);
      if (valid && lead) {
        const double xi = x[i], ri = b[i] - ax;
        acc[0] += xi * ax;
        acc[1] += xi * xi;
        acc[2] += ri * (ri * dinv[i]);
        acc[3] += ri * ri;
      }
    }
  }
  if (grid_sum<4>(acc, part, &sc->counter) && threadIdx.x == 0) {
    for (int k = 0; k < 4; k++) sc->res[k] = acc[k];
  }
}

// ---- launchers ----------------------------------------------------------------

extern "C" int sfr_k_gather(hipStream_t s, int64_t nnz, const int32_t *vmap,
                            const double *raw, double *val)
{
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((nnz + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, nnz, vmap, raw, val);
  return hipGetLastError();
}

extern "C" int sfr_k_diag(hipStream_t s, int64_t n, const int32_t *dmap,
                          const double *raw, const double *raw_ad, double *ad,
                          double *dinv)
{
  hipLaunchKernelGGL(k_diag, dim3((unsigned)((n + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, n, dmap, raw, raw_ad, ad, dinv);
  return hipGetLastError();
}

extern "C" int sfr_k_fill(hipStream_t s, int64_t n, double v, double *x)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, n, v, x);
  return hipGetLastError();
}

extern "C" int sfr_k_scale32(hipStream_t s, const sfr_mat_view *A,
                             double *dscale, float *val32)
{
  const int64_t n = A->n;
  const unsigned gn = (unsigned)((n + SFR_BLOCK - 1) / SFR_BLOCK);
  hipLaunchKernelGGL(k_dscale, dim3(gn), dim3(SFR_BLOCK), 0, s, n, A->ad,
                     dscale);
  if (A->fmt == SFR_FMT_ELL) {
    const int64_t nnz = n * A->slots;
    if (nnz > 0)
      hipLaunchKernelGGL(k_scale32_ell,
                         dim3((unsigned)((nnz + SFR_BLOCK - 1) / SFR_BLOCK)),
                         dim3(SFR_BLOCK), 0, s, n, nnz, A->col, A->val,
                         (const double *)dscale, val32);
  } else {
    hipLaunchKernelGGL(k_scale32_csr, dim3(gn), dim3(SFR_BLOCK), 0, s, n,
                       A->rowptr, A->col, A->val, (const double *)dscale,
                       val32);
  }
  return hipGetLastError();
}

extern "C" int sfr_k_spmv(hipStream_t s, const sfr_mat_view *A,
                          const double *x, double *y)
{
  const sfr_geom g = {2, grid_for(A->n, 2, op_lanes(A))};
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_spmv, Op, g, s, op, x, y);
  });
  return hipGetLastError();
}

extern "C" int sfr_k_pcg_setup(hipStream_t s, const sfr_mat_view *A,
                               sfr_geom g0, const double *b, double *x,
                               double *r, double *p, int newX, double tol,
                               unsigned maxit, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  switch (gv.rpt) {
  case 1: hipLaunchKernelGGL(k_pcg_setup1<1>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s,
                             A->n, b, A->dinv, x, newX, tol, maxit, sc, part); break;
  case 2: hipLaunchKernelGGL(k_pcg_setup1<2>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s,
                             A->n, b, A->dinv, x, newX, tol, maxit, sc, part); break;
  default: hipLaunchKernelGGL(k_pcg_setup1<4>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s,
                              A->n, b, A->dinv, x, newX, tol, maxit, sc, part); break;
  }
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_pcg_setup2, Op, gm, s, op, b, A->dinv, x, r, p, newX, sc,
                   part);
  });
  return hipGetLastError();
}

extern "C" int sfr_k_pcg_iter(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, double *x, double *r, double *p,
                              double *q, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_pcg_spmv, Op, gm, s, op, p, q, sc, part);
  });
  switch (gv.rpt) {
  case 1:
    hipLaunchKernelGGL(k_pcg_resid<1>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s, A->n, q, A->dinv, r, sc, part);
    hipLaunchKernelGGL(k_pcg_dir<1>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s, A->n, r, A->dinv, x, p, sc);
    break;
  case 2:
    hipLaunchKernelGGL(k_pcg_resid<2>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s, A->n, q, A->dinv, r, sc, part);
    hipLaunchKernelGGL(k_pcg_dir<2>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s, A->n, r, A->dinv, x, p, sc);
    break;
  default:
    hipLaunchKernelGGL(k_pcg_resid<4>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s, A->n, q, A->dinv, r, sc, part);
    hipLaunchKernelGGL(k_pcg_dir<4>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s, A->n, r, A->dinv, x, p, sc);
    break;
  }
  return hipGetLastError();
}

extern "C" int sfr_k_final(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                           const double *b, const double *x, sfr_scalars *sc,
                           double *part)
{
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_final, Op, gm, s, op, b, x, A->dinv, sc, part);
  });
  return hipGetLastError();
}

// b.D^-1 b, b.b, tolerance and x = 0 when newX (first half of the PCG setup)
extern "C" int sfr_k_pcg_norm(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, const double *b, double *x,
                              int newX, double tol, unsigned maxit,
                              sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  switch (gv.rpt) {
  case 1: hipLaunchKernelGGL(k_pcg_setup1<1>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s,
                             A->n, b, A->dinv, x, newX, tol, maxit, sc, part); break;
  case 2: hipLaunchKernelGGL(k_pcg_setup1<2>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s,
                             A->n, b, A->dinv, x, newX, tol, maxit, sc, part); break;
  default: hipLaunchKernelGGL(k_pcg_setup1<4>, dim3(gv.grid), dim3(SFR_BLOCK), 0, s,
                              A->n, b, A->dinv, x, newX, tol, maxit, sc, part); break;
  }
  return hipGetLastError();
}

__global__ __launch_bounds__(SFR_BLOCK) void k_fill32(int64_t n, float v,
                                                      float *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (i < n) x[i] = v;
}

extern "C" int sfr_k_fill32(hipStream_t s, int64_t n, float v, float *x)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill32, dim3((unsigned)((n + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, n, v, x);
  return hipGetLastError();
}

// the SpMV + p.q kernel of one PCG iteration alone (measurement hook)
extern "C" int sfr_k_pcg_spmv_only(hipStream_t s, const sfr_mat_view *A,
                                   sfr_geom g0, const double *p, double *q,
                                   sfr_scalars *sc, double *part)
{
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_pcg_spmv, Op, gm, s, op, p, q, sc, part);
  });
  return hipGetLastError();
}
