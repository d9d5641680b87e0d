// Restarted GMRES(k) with left Jacobi preconditioning, FP64 on gfx950,
// following the reference's gmres (src/Iters.c:587-790) as solverC drives
// it (k = 15, cycles = maxIt; src/Solv.c:302-372):
//   cycle   g_0 = D^-1 (b - A x), e_0 = ||g_0||, g_0 /= e_0
//   step i  w = D^-1 A g_i; modified Gram-Schmidt against g_0..g_i
//           (h_ji = w.g_j, w -= h_ji g_j, one j after the other);
//           h_{i+1,i} = ||w||, g_{i+1} = w / h_{i+1,i};
//           the earlier Givens rotations applied to column i, a new one
//           zeroes h_{i+1,i}, e_{i+1} = -s_i e_i, e_i = c_i e_i
//   end     back substitution for y with the reference's truncated sums
//           (inner j < k-1, Iters.c:714-719), x += sum_{j<k-1} y_j g_j
//           (Iters.c:723-728); stop when |e_k| < tol ||D^-1 b||.
// Every cycle runs all k steps: the reference's in-cycle test is an empty
// statement (Iters.c:702).
//
// Kernels of one cycle (137 for k = 15), captured in a graph:
//   G0 k_gm_res        raw residual w = D^-1 (b - Ax), partials w.w
//   G1 k_gm_arn(i)     h_{i,i-1} = ||w|| (the Givens step of column i-1 in
//                      workgroup 0); g_i = w/||w||; w' = D^-1 A g_i computed
//                      as (A w)/||w||; partials w'.g_0
//   G2 k_gm_mgs(i, j)  h_ji from the partials; w' -= h_ji g_j; partials
//                      w'.g_{j+1}, or w'.w' after the last j
//   G3 k_gm_end        Givens step of the last column and the back
//                      substitution (every workgroup, in LDS), x update,
//                      stop test
// The raw vectors w alternate between two buffers, since step i reads w at
// neighbouring rows while it writes the next one.  Reductions as in
// pcg.hip: block partials, reduced redundantly by every workgroup of the
// next kernel.  Workgroup 0 alone writes the small Hessenberg state and no
// kernel reads what it writes there.  Flags: G0 alone reads `done` and sets
// run2 for the cycle; run4 records an exact zero residual (2) or a
// breakdown (1) for G3, which alone decides `done`.
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)
#define GK SFR_GM_K
#define GH(i, j) ((i) * GK + (j))
#define GM_C ((GK + 1) * GK)
#define GM_S (GM_C + GK)
#define GM_E (GM_S + GK)

// partials of (D^-1 b).(D^-1 b) and b.b; x = 0 when newX
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_gm_s1(int64_t n,
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
        const double bi = b[i], gi = bi * dinv[i];
        acc[0] += gi * gi;
        acc[1] += bi * bi;
        if (newX) x[i] = 0.0;
      }
    }
  }
  store_partials<2>(acc, pout);
}

__global__ __launch_bounds__(SFR_BLOCK) void k_gm_start(const double *pin,
    int gin, double tol, unsigned maxit, sfr_scalars *sc)
{
  double t[2];
  reduce_partials<2>(pin, gin, t);
  if (threadIdx.x == 0) {
    const double norm_b = sqrt(t[0]);
    sc->norm_b = norm_b;
    sc->conv = tol * norm_b;
    sc->bb = t[1];
    sc->tol = tol;
    sc->maxit = maxit;
    sc->iter = 0;
    sc->outer = 0;
    sc->ru = 0;
    sc->run2 = 0;
    sc->run4 = 0;
    sc->res[4] = 0.0;
    sc->done = maxit == 0 ? SFR_DONE_MAXIT : SFR_DONE_RUN;
  }
}

// G0: w = D^-1 (b - Ax); partials w.w
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_gm_res(Op A,
    const double *__restrict__ b, const double *__restrict__ x,
    const double *__restrict__ dinv, double *__restrict__ w, double *pout,
    sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (sc->done) {
    if (first) sc->run2 = 0;
    return;
  }
  if (first) { sc->run2 = 1; sc->run4 = 0; }
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[1] = {0.0};
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
      const bool valid = i < A.n;
      const double ax = A.row(i, valid, x);
      if (valid && lead) {
        const double wi = (b[i] - ax) * dinv[i];
        w[i] = wi;
        acc[0] += wi * wi;
      }
    }
  }
  store_partials<1>(acc, pout);
}

// The Givens step of column `col` (src/Iters.c:683-701) on the state in
// global memory, by workgroup 0: h_{col+1,col} = nrm is set first.  The
// column, c, s and e_col are staged in LDS by many lanes, the recurrence
// runs in lane 0, the results go back with plain stores.
__device__ __forceinline__ void gm_givens_global(double *st, int col,
                                                 double nrm)
{
  __shared__ double hc[GK + 1], cs[GK], ss[GK], ec;
  const int t = threadIdx.x;
  if (t <= col) hc[t] = st[GH(t, col)];
  if (t < col) { cs[t] = st[GM_C + t]; ss[t] = st[GM_S + t]; }
  if (t == 0) { hc[col + 1] = nrm; ec = st[GM_E + col]; }
  __syncthreads();
  if (t == 0) {
    double hj = hc[0];
    for (int j = 0; j < col; j++) {
      const double h1 = hj, h2 = hc[j + 1];
      hc[j] = cs[j] * h1 + ss[j] * h2;
      hj = -ss[j] * h1 + cs[j] * h2;
    }
    const double h1 = hj, h2 = hc[col + 1];
    const double r = sqrt(h1 * h1 + h2 * h2), tmp = 1.0 / r;
    const double c = h1 * tmp, s = h2 * tmp;
    for (int j = 0; j < col; j++) st[GH(j, col)] = hc[j];
    st[GH(col, col)] = r;
    st[GH(col + 1, col)] = 0.0;
    st[GM_C + col] = c;
    st[GM_S + col] = s;
    st[GM_E + col + 1] = -s * ec;
    st[GM_E + col] = c * ec;
  }
  __syncthreads();
}

// G1 of step ni: the raw w_in becomes g_ni; w_out = D^-1 A g_ni; partials
// w_out.g_0.  Workgroup 0 stores e_0 (ni = 0) or h_{ni,ni-1} and the Givens
// step of column ni-1.
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_gm_arn(Op A,
    const double *__restrict__ win, const double *__restrict__ dinv,
    double *__restrict__ gk, const double *__restrict__ g0,
    double *__restrict__ wout, int ni, double *st, const double *pin,
    int gin, double *pout, sfr_scalars *sc)
{
  if (!sc->run2) return;
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double nrm = sqrt(t[0]);
  const double tmp = 1.0 / nrm;
  const bool ok = nrm > 0.0 && isfinite(tmp);
  if (blockIdx.x == 0) {
    if (ni == 0) {
      if (threadIdx.x == 0) {
        st[GM_E] = nrm;
        if (!ok) sc->run4 = nrm == 0.0 ? 2u : 1u;
      }
    } else {
      gm_givens_global(st, ni - 1, nrm);
      if (threadIdx.x == 0 && !ok && sc->run4 == 0) sc->run4 = 1;
    }
  }
  constexpr int L = Op::L, RPC = RPT * (SFR_BLOCK / L);
  const bool lead = (threadIdx.x % L) == 0;
  double acc[1] = {0.0};
  if (ok) {
    for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < A.n;
         c += gridDim.x) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * (SFR_BLOCK / L) + threadIdx.x / L;
        const bool valid = i < A.n;
        const double ax = A.row(i, valid, win);
        if (valid && lead) {
          const double gi = win[i] * tmp;
          gk[i] = gi;
          const double wi = (ax * tmp) * dinv[i];
          wout[i] = wi;
          acc[0] += wi * (ni == 0 ? gi : g0[i]);
        }
      }
    }
  }
  store_partials<1>(acc, pout);
}

// G2: h_{j,ni} = w.g_j (reduced); w -= h_{j,ni} g_j; partials w.g_{j+1}
// (gnext) or w.w (gnext == NULL, after the last j)
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_gm_mgs(int64_t n,
    double *__restrict__ w, const double *__restrict__ gj,
    const double *__restrict__ gnext, int hidx, double *st,
    const double *pin, int gin, double *pout, sfr_scalars *sc)
{
  if (!sc->run2) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  // the first chunk's rows are loaded before the reduction (hides it)
  const int64_t c0 = xcd_chunk(blockIdx.x, gridDim.x);
  double wv[RPT], gv[RPT], nv[RPT];
#pragma unroll
  for (int r_ = 0; r_ < RPT; r_++) {
    const int64_t i = c0 * RPC + r_ * SFR_BLOCK + threadIdx.x;
    wv[r_] = i < n ? w[i] : 0.0;
    gv[r_] = i < n ? gj[i] : 0.0;
    nv[r_] = (i < n && gnext) ? gnext[i] : 0.0;
  }
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double beta = t[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) st[hidx] = beta;
  double acc[1] = {0.0};
  for (int64_t c = c0; c * RPC < n; c += gridDim.x) {
    if (c != c0) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
        wv[r_] = i < n ? w[i] : 0.0;
        gv[r_] = i < n ? gj[i] : 0.0;
        nv[r_] = (i < n && gnext) ? gnext[i] : 0.0;
      }
    }
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double wi = wv[r_] - beta * gv[r_];
        w[i] = wi;
        acc[0] += wi * (gnext ? nv[r_] : wi);
      }
    }
  }
  store_partials<1>(acc, pout);
}

// G3: h_{k,k-1} = ||w||; the last Givens step and the back substitution in
// LDS (every workgroup, bitwise the same); x += sum_{j<k-1} y_j g_j; the
// stop test and the cycle count in workgroup 0.
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_gm_end(int64_t n,
    double *__restrict__ x, const double *__restrict__ basis, int64_t ld,
    const double *st, const double *pin, int gin, sfr_scalars *sc)
{
  if (!sc->run2) return;
  __shared__ double H[(GK + 1) * GK], cs[GK], ss[GK], e[GK + 1], y[GK];
  __shared__ int flag;
  double t[1];
  reduce_partials<1>(pin, gin, t);
  const double nrm = sqrt(t[0]);
  for (int k = threadIdx.x; k < (GK + 1) * GK; k += SFR_BLOCK) H[k] = st[k];
  if (threadIdx.x < GK) {
    cs[threadIdx.x] = st[GM_C + threadIdx.x];
    ss[threadIdx.x] = st[GM_S + threadIdx.x];
  }
  if (threadIdx.x <= GK) e[threadIdx.x] = st[GM_E + threadIdx.x];
  const unsigned code = sc->run4;
  __syncthreads();
  constexpr int last = GK - 1;
  if (threadIdx.x == 0) {
    // Givens step of column k-1 (its result feeds only e_k and y_{k-1})
    H[GH(last + 1, last)] = nrm;
    for (int j = 0; j < last; j++) {
      const double h1 = H[GH(j, last)], h2 = H[GH(j + 1, last)];
      H[GH(j, last)] = cs[j] * h1 + ss[j] * h2;
      H[GH(j + 1, last)] = -ss[j] * h1 + cs[j] * h2;
    }
    const double h1 = H[GH(last, last)], h2 = H[GH(last + 1, last)];
    const double r = sqrt(h1 * h1 + h2 * h2), tmp = 1.0 / r;
    const double c = h1 * tmp, s = h2 * tmp;
    H[GH(last, last)] = r;
    e[last + 1] = -s * e[last];
    e[last] = c * e[last];
    // back substitution, inner sums over j < k-1 as the reference's
    y[last] = e[last] / H[GH(last, last)];
    bool fin = true;
    for (int i = last - 1; i >= 0; i--) {
      double yi = 0.0;
      for (int j = i + 1; j < last; j++) yi -= H[GH(i, j)] * y[j];
      yi = (yi + e[i]) / H[GH(i, i)];
      y[i] = yi;
      fin = fin && isfinite(yi);
    }
    flag = (code == 0 && fin && isfinite(e[last + 1])) ? 1 : 0;
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double ek = fabs(e[last + 1]);
    if (code == 2) {
      sc->res[4] = 0.0;
      sc->done = SFR_DONE_CONV;
    } else if (!flag) {
      sc->done = SFR_DONE_BREAK;
    } else {
      const unsigned outer = sc->outer + 1;
      sc->outer = outer;
      sc->iter = sc->iter + GK;
      sc->res[4] = ek;
      if (ek < sc->conv) sc->done = SFR_DONE_CONV;
      else if (outer >= sc->maxit) sc->done = SFR_DONE_MAXIT;
    }
  }
  if (!flag) return;
  double yr[last];
#pragma unroll
  for (int j = 0; j < last; j++) yr[j] = y[j];
  constexpr int RPC = RPT * SFR_BLOCK;
  for (int64_t c = xcd_chunk(blockIdx.x, gridDim.x); c * RPC < n;
       c += gridDim.x) {
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        double xi = x[i];
#pragma unroll
        for (int j = 0; j < last; j++) xi += basis[j * ld + i] * yr[j];
        x[i] = xi;
      }
    }
  }
}

extern "C" int sfr_k_gm_setup(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, const double *b, double *x,
                              int newX, double tol, unsigned maxit,
                              sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  double *r2 = REGION(part, 2);
  SFR_LAUNCH_V(k_gm_s1, gv, s, A->n, b, A->dinv, x, newX, r2);
  hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)r2, gv.grid, tol, maxit, sc);
  return hipGetLastError();
}

// One restart cycle: basis holds g_0..g_{k-1} at stride ld, w[0] and w[1]
// are the raw vectors, st the Hessenberg state (SFR_GM_STATE doubles).
extern "C" int sfr_k_gm_cycle(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, const double *b, double *x,
                              double *basis, int64_t ld, double *const w[2],
                              double *st, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *reg[2] = {REGION(part, 0), REGION(part, 1)};
  int par = 0, gprev = gm.grid;
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_gm_res, Op, gm, s, op, b, (const double *)x, A->dinv,
                   w[0], reg[0], sc);
  });
  for (int ni = 0; ni < GK; ni++) {
    double *win = w[ni & 1], *wout = w[(ni + 1) & 1];
    with_op64(A, [&](auto op) {
      using Op = decltype(op);
      SFR_LAUNCH_RPT(k_gm_arn, Op, gm, s, op, (const double *)win, A->dinv,
                     basis + ni * ld, (const double *)basis, wout, ni, st,
                     (const double *)reg[par], gprev, reg[par ^ 1], sc);
    });
    par ^= 1;
    gprev = gm.grid;
    for (int j = 0; j <= ni; j++) {
      const double *gnext = j < ni ? basis + (j + 1) * ld : NULL;
      SFR_LAUNCH_V(k_gm_mgs, gv, s, A->n, wout, (const double *)(basis + j * ld),
                   gnext, GH(j, ni), st, (const double *)reg[par], gprev,
                   reg[par ^ 1], sc);
      par ^= 1;
      gprev = gv.grid;
    }
  }
  SFR_LAUNCH_V(k_gm_end, gv, s, A->n, x, (const double *)basis, ld,
               (const double *)st, (const double *)reg[par], gprev, sc);
  return hipGetLastError();
}
