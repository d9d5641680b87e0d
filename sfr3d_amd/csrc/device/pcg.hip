// Jacobi-preconditioned CG in FP64 on gfx950, following the recurrences of
// the reference's pcg (src/Cg.c:47-209):
//   setup  z = D^-1 b, conv = tol*sqrt|b.z|, r = b - Ax, z = D^-1 r, p = z,
//          d = r.z
//   iter   q = Ap, alpha = d/(p.q), x += alpha p, r -= alpha q, z = D^-1 r,
//          di = r.z, beta = di/d, p = z + beta p, d = di, stop when
//          sqrt|d| < conv
// One iteration is three streaming kernels:
//   K1 k_pcg_spmv   q = Ap; partials of p.q                 (matrix + 2 words)
//   K2 k_pcg_resid  alpha = d/(p.q); r -= alpha q; partials of r.D^-1 r
//                                                           (4 words)
//   K3 k_pcg_dir    beta = r.D^-1 r / d; x += alpha p; p = D^-1 r + beta p;
//                   stop test                               (6 words)
// The x update rides with the direction update (same alpha, same old p,
// same rounding as the reference's order), so z is never stored.
//
// Reductions: each workgroup stores its partial sum with a plain store and
// every workgroup of the next kernel adds all partials in the same fixed
// order, so all of them derive bitwise the same alpha / beta / stop decision
// without atomics, fences or grid barriers.  A scalar is never written by the
// kernel that reads it: K1 alone reads `done`; K2 reads `run2` (set by K1)
// and sets `run3`; K3 reads `run3`; only workgroup 0 writes scalars.
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)

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

__global__ __launch_bounds__(SFR_BLOCK) void k_fill32(int64_t n, float v,
                                                      float *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (i < n) x[i] = v;
}

// Streaming copy of n2 16-byte elements, four per thread per pass with all
// loads issued before the stores, non-temporal both ways: the HBM rate a
// pure stream attains (the measurement hook sfr_copy_time).  An odd
// trailing double is copied by the first thread.
typedef double sfr_v2d __attribute__((vector_size(16)));

__global__ __launch_bounds__(SFR_BLOCK) void k_stream_copy(int64_t n2,
    const sfr_v2d *__restrict__ src, sfr_v2d *__restrict__ dst,
    const double *__restrict__ tsrc, double *__restrict__ tdst)
{
  constexpr int U = 4;
  const int64_t stride = (int64_t)gridDim.x * SFR_BLOCK * U;
  for (int64_t base = (int64_t)blockIdx.x * SFR_BLOCK * U + threadIdx.x;
       base < n2; base += stride) {
    sfr_v2d v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = base + (int64_t)u * SFR_BLOCK;
      if (i < n2) v[u] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = base + (int64_t)u * SFR_BLOCK;
      if (i < n2) __builtin_nontemporal_store(v[u], dst + i);
    }
  }
  if (tsrc && blockIdx.x == 0 && threadIdx.x == 0) *tdst = *tsrc;
}

// Jacobi-scaled FP32 copy for the mixed solve: a_ij / sqrt(a_ii a_jj).
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
    const int64_t i = k % n;           // slot-major: row of entry k
    // a_ij (s_i s_j): the same bits for a_ji (s_j s_i) when a_ij == a_ji, so
    // a symmetric A gives a symmetric copy
    val32[k] = (float)(val[k] * (dscale[i] * dscale[col[k]]));
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
      val32[k] = (float)(val[k] * (dscale[i] * dscale[col[k]]));
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

// one workgroup: r.D^-1 r of the setup -> d, and the loop state
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_start(const double *part,
    int g, sfr_scalars *sc)
{
  double t[1];
  reduce_partials<1>(part, g, t);
  if (threadIdx.x == 0) {
    sc->d = t[0];
    sc->iter = 0;
    sc->outer = 0;
    sc->ru = 0;
    // b = 0, or an exact initial guess: nothing to iterate
    sc->done = (sc->norm_b == 0.0 || t[0] == 0.0) ? SFR_DONE_CONV : SFR_DONE_RUN;
  }
}

// one workgroup: x.Ax, x.x, r.D^-1 r, r.r of the final kernel -> res
__global__ __launch_bounds__(SFR_BLOCK) void k_final_sum(const double *part,
    int g, sfr_scalars *sc)
{
  double t[4];
  reduce_partials<4>(part, g, t);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) sc->res[k] = t[k];
}

// ---- PCG ----------------------------------------------------------------------

// partials of b.D^-1 b and b.b; x = 0 when newX
template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_s1(int64_t n,
    const double *__restrict__ b, const double *__restrict__ dinv,
    double *__restrict__ x, int newX, double *part)
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
  store_partials<2>(acc, part);
}

// norms of b; r = b - Ax, p = D^-1 r; partials of r.p
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_s2(Op A,
    const double *__restrict__ b, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ r, double *__restrict__ p,
    int newX, double tol, unsigned maxit, const double *pin, int gin,
    double *pout, sfr_scalars *sc)
{
  double nb[2];
  reduce_partials<2>(pin, gin, nb);
  const double norm_b = sqrt(fabs(nb[0]));
  const bool zero_b = norm_b == 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->norm_b = norm_b;
    sc->conv = tol * norm_b;
    sc->bb = nb[1];
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
  store_partials<1>(acc, pout);
}

template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_spmv(Op A,
    const double *__restrict__ p, double *__restrict__ q, double *pout,
    sfr_scalars *sc)
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
  store_partials<1>(acc, pout);
}

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_resid(int64_t n,
    const double *__restrict__ q, const double *__restrict__ dinv,
    double *__restrict__ r, const double *pin, int gin, double *pout,
    sfr_scalars *sc)
{
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!sc->run2) {
    if (first) sc->run3 = 0;
    return;
  }
  constexpr int RPC = RPT * SFR_BLOCK;
  // the first chunk's rows are loaded before the reduction (hides it)
  const int64_t c0 = xcd_chunk(blockIdx.x, gridDim.x);
  double rv[RPT], qv[RPT], dv[RPT];
#pragma unroll
  for (int r_ = 0; r_ < RPT; r_++) {
    const int64_t i = c0 * RPC + r_ * SFR_BLOCK + threadIdx.x;
    rv[r_] = i < n ? r[i] : 0.0;
    qv[r_] = i < n ? q[i] : 0.0;
    dv[r_] = i < n ? dinv[i] : 0.0;
  }
  double pq[1];
  reduce_partials<1>(pin, gin, pq);
  const double d = sc->d;
  const double alpha = d / pq[0];
  if (!(pq[0] != 0.0) || !isfinite(alpha)) {
    if (first) { sc->done = SFR_DONE_BREAK; sc->run3 = 0; }
    return;
  }
  if (first) { sc->alpha = alpha; sc->d_prev = d; sc->run3 = 1; }
  double acc[1] = {0.0};
  for (int64_t c = c0; c * RPC < n; c += gridDim.x) {
    if (c != c0) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
        rv[r_] = i < n ? r[i] : 0.0;
        qv[r_] = i < n ? q[i] : 0.0;
        dv[r_] = i < n ? dinv[i] : 0.0;
      }
    }
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        const double ri = rv[r_] - alpha * qv[r_];
        r[i] = ri;
        acc[0] += ri * (ri * dv[r_]);
      }
    }
  }
  store_partials<1>(acc, pout);
}

template <int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_pcg_dir(int64_t n,
    const double *__restrict__ r, const double *__restrict__ dinv,
    double *__restrict__ x, double *__restrict__ p, const double *pin,
    int gin, sfr_scalars *sc)
{
  if (!sc->run3) return;
  constexpr int RPC = RPT * SFR_BLOCK;
  // the first chunk's rows are loaded before the reduction (hides it).
  // x is touched by no other kernel of the loop, so it streams past the
  // caches (non-temporal), leaving them to p, r, q and D^-1: measured
  // +1.4 % on the 200^3 FP64 solve, no loss at 128^3.
  const int64_t c0 = xcd_chunk(blockIdx.x, gridDim.x);
  double pv[RPT], xv[RPT], zv[RPT];
#pragma unroll
  for (int r_ = 0; r_ < RPT; r_++) {
    const int64_t i = c0 * RPC + r_ * SFR_BLOCK + threadIdx.x;
    pv[r_] = i < n ? p[i] : 0.0;
    xv[r_] = i < n ? __builtin_nontemporal_load(x + i) : 0.0;
    zv[r_] = i < n ? r[i] * dinv[i] : 0.0;
  }
  double rz[1];
  reduce_partials<1>(pin, gin, rz);
  const double di = rz[0];
  const double alpha = sc->alpha, beta = di / sc->d_prev;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const unsigned it = sc->iter + 1;
    sc->iter = it;
    sc->d = di;
    if (sqrt(fabs(di)) < sc->conv || di == 0.0) sc->done = SFR_DONE_CONV;
    else if (!isfinite(di)) sc->done = SFR_DONE_BREAK;
    else if (it >= sc->maxit) sc->done = SFR_DONE_MAXIT;
  }
  for (int64_t c = c0; c * RPC < n; c += gridDim.x) {
    if (c != c0) {
#pragma unroll
      for (int r_ = 0; r_ < RPT; r_++) {
        const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
        pv[r_] = i < n ? p[i] : 0.0;
        xv[r_] = i < n ? __builtin_nontemporal_load(x + i) : 0.0;
        zv[r_] = i < n ? r[i] * dinv[i] : 0.0;
      }
    }
#pragma unroll
    for (int r_ = 0; r_ < RPT; r_++) {
      const int64_t i = c * RPC + r_ * SFR_BLOCK + threadIdx.x;
      if (i < n) {
        __builtin_nontemporal_store(xv[r_] + alpha * pv[r_], x + i);
        p[i] = zv[r_] + beta * pv[r_];
      }
    }
  }
}

// partials of x.Ax, x.x, r.D^-1 r and r.r with r = b - Ax (src/Cg.c:155-176)
template <class Op, int RPT>
__global__ __launch_bounds__(SFR_BLOCK) void k_final(Op A,
    const double *__restrict__ b, const double *__restrict__ x,
    const double *__restrict__ dinv, double *pout)
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
  store_partials<4>(acc, pout);
}

// ---- launchers ----------------------------------------------------------------

static inline unsigned blocks1(int64_t n)
{
  return (unsigned)((n + SFR_BLOCK - 1) / SFR_BLOCK);
}

extern "C" int sfr_k_gather(hipStream_t s, int64_t nnz, const int32_t *vmap,
                            const double *raw, double *val)
{
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(blocks1(nnz)), dim3(SFR_BLOCK), 0, s, nnz,
                     vmap, raw, val);
  return hipGetLastError();
}

extern "C" int sfr_k_diag(hipStream_t s, int64_t n, const int32_t *dmap,
                          const double *raw, const double *raw_ad, double *ad,
                          double *dinv)
{
  hipLaunchKernelGGL(k_diag, dim3(blocks1(n)), dim3(SFR_BLOCK), 0, s, n, dmap,
                     raw, raw_ad, ad, dinv);
  return hipGetLastError();
}

extern "C" int sfr_k_fill(hipStream_t s, int64_t n, double v, double *x)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks1(n)), dim3(SFR_BLOCK), 0, s, n, v, x);
  return hipGetLastError();
}

extern "C" int sfr_k_fill32(hipStream_t s, int64_t n, float v, float *x)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill32, dim3(blocks1(n)), dim3(SFR_BLOCK), 0, s, n, v, x);
  return hipGetLastError();
}

// grid: 16,384 workgroups of 4 x 256 elements per pass (the fastest shape
// of a sweep over 1,024-65,536 workgroups on 2 GiB buffers)
extern "C" int sfr_k_stream_copy(hipStream_t s, int64_t n, const double *src,
                                 double *dst)
{
  if (n <= 0) return hipSuccess;
  const int64_t n2 = n / 2;
  const int64_t per = (int64_t)SFR_BLOCK * 4;
  const int64_t want = (n2 + per - 1) / per;
  const unsigned grid = (unsigned)(want < 16384 ? (want > 0 ? want : 1) : 16384);
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(SFR_BLOCK), 0, s, n2,
                     (const sfr_v2d *)src, (sfr_v2d *)dst,
                     (n & 1) ? src + n - 1 : (const double *)nullptr,
                     (n & 1) ? dst + n - 1 : (double *)nullptr);
  return hipGetLastError();
}

__device__ __forceinline__ uint64_t f64_bits(double v)
{
  uint64_t u;
  __builtin_memcpy(&u, &v, sizeof u);
  return u;
}

// Rows whose values do not mirror across the face pairs (see PairOp):
// per row, the second slot g of each pair with offset d either faces no
// neighbour (d == 0) and holds +0, or the neighbour j = i + d faces back
// across g's partner f with the same value bits.  One count per workgroup.
template <int S>
__global__ __launch_bounds__(SFR_BLOCK) void k_pair_check(int64_t n,
    const double *__restrict__ val, const int16_t *__restrict__ dcol,
    double *pout)
{
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * SFR_BLOCK) {
    bool bad = false;
#pragma unroll
    for (int g = 0; g < S; g++) {
      const int f = pair_src<S>(g);
      if (f == g) continue;
      const int d = dcol[g * n + i];
      const uint64_t vg = f64_bits(val[g * n + i]);
      if (d == 0) {
        bad |= vg != 0u;
        continue;
      }
      const int64_t j = i + d;
      if (j < 0 || j >= n) { bad = true; continue; }
      bad |= dcol[f * n + j] != -d;
      bad |= vg != f64_bits(val[f * n + j]);
    }
    acc[0] += bad ? 1.0 : 0.0;
  }
  store_partials<1>(acc, pout);
}

extern "C" int sfr_k_pair_check(hipStream_t s, const sfr_mat_view *A, int kind,
                                double *part, int *nparts)
{
  const int64_t n = A->n;
  const int64_t want = (n + SFR_BLOCK - 1) / SFR_BLOCK;
  const int grid = want < 1 ? 1 : (want > 1024 ? 1024 : (int)want);
  *nparts = grid;
  if (kind == SFR_PAIR_HEX)
    hipLaunchKernelGGL(k_pair_check<6>, dim3(grid), dim3(SFR_BLOCK), 0, s, n,
                       A->val, A->col16, part);
  else
    hipLaunchKernelGGL(k_pair_check<4>, dim3(grid), dim3(SFR_BLOCK), 0, s, n,
                       A->val, A->col16, part);
  return hipGetLastError();
}

extern "C" int sfr_k_scale32(hipStream_t s, const sfr_mat_view *A,
                             double *dscale, float *val32)
{
  const int64_t n = A->n;
  hipLaunchKernelGGL(k_dscale, dim3(blocks1(n)), dim3(SFR_BLOCK), 0, s, n,
                     A->ad, dscale);
  if (A->fmt == SFR_FMT_ELL) {
    const int64_t nnz = n * A->slots;
    if (nnz > 0)
      hipLaunchKernelGGL(k_scale32_ell, dim3(blocks1(nnz)), dim3(SFR_BLOCK), 0,
                         s, n, nnz, A->col, A->val, (const double *)dscale,
                         val32);
  } else {
    hipLaunchKernelGGL(k_scale32_csr, dim3(blocks1(n)), dim3(SFR_BLOCK), 0, s,
                       n, A->rowptr, A->col, A->val, (const double *)dscale,
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

// partials of b.D^-1 b, b.b into region 2; x = 0 when newX
extern "C" int sfr_k_pcg_norm(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, const double *b, double *x,
                              int newX, double *part, int *nparts)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  SFR_LAUNCH_V(k_pcg_s1, gv, s, A->n, b, A->dinv, x, newX, REGION(part, 2));
  *nparts = gv.grid;
  return hipGetLastError();
}

extern "C" int sfr_k_pcg_setup(hipStream_t s, const sfr_mat_view *A,
                               sfr_geom g0, const double *b, double *x,
                               double *r, double *p, int newX, double tol,
                               unsigned maxit, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *rc = REGION(part, 2), *rb = REGION(part, 1);
  SFR_LAUNCH_V(k_pcg_s1, gv, s, A->n, b, A->dinv, x, newX, rc);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_pcg_s2, Op, gm, s, op, b, A->dinv, x, r, p, newX, tol,
                   maxit, (const double *)rc, gv.grid, rb, sc);
  });
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)rb, gm.grid, sc);
  return hipGetLastError();
}

// The two streaming kernels of an iteration take the fewest passes a grid
// of at most 1,024 workgroups allows (up to 8 rows per thread), whatever
// rows per thread the SpMV runs: a grid-stride pass issues its loads only
// after the previous pass's arrive, so one pass keeps the most loads in
// flight.  The r-update's partial count is the direction kernel's `gin`.
// Workgroup cap of the FP64 PCG SpMV: from 2^21 rows, 1,024 workgroups
// making several grid-stride passes.  The residual kernel then reduces
// 1,024 partials instead of 2,048, and the rows one pass has in flight span
// a smaller part of the matrix, so fewer neighbour gathers miss.  Measured
// A/B on one box, twice (ms per FP64 PCG solve, 1 / 1 / 2 rows per
// thread): 128^3 63.8 -> 62.9, 160^3 193.0 -> 186.8, 200^3 454 -> 423.
static inline int pcg_spmv_cap(int64_t n)
{
  return n >= ((int64_t)1 << 21) ? 1024 : 0;
}

extern "C" int sfr_k_pcg_iter(hipStream_t s, const sfr_mat_view *A,
                              sfr_geom g0, double *x, double *r, double *p,
                              double *q, sfr_scalars *sc, double *part)
{
  int rv = g0.rpt;
  while (rv < 8 && (int64_t)rv * SFR_BLOCK * 1024 < A->n) rv *= 2;
  const sfr_geom gv = {rv, grid_for(A->n, rv, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A), pcg_spmv_cap(A->n))};
  double *ra = REGION(part, 0), *rb = REGION(part, 1);
  with_op64_pair(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_pcg_spmv, Op, gm, s, op, p, q, ra, sc);
  });
  SFR_LAUNCH_V(k_pcg_resid, gv, s, A->n, q, A->dinv, r, (const double *)ra,
               gm.grid, rb, sc);
  SFR_LAUNCH_V(k_pcg_dir, gv, s, A->n, r, A->dinv, x, p, (const double *)rb,
               gv.grid, sc);
  return hipGetLastError();
}

extern "C" int sfr_k_final(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                           const double *b, const double *x, sfr_scalars *sc,
                           double *part)
{
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *rd = REGION(part, 3);
  with_op64(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_final, Op, gm, s, op, b, x, A->dinv, rd);
  });
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)rd, gm.grid, sc);
  return hipGetLastError();
}

// the SpMV + p.q kernel of one PCG iteration alone (measurement hook)
extern "C" int sfr_k_pcg_spmv_only(hipStream_t s, const sfr_mat_view *A,
                                   sfr_geom g0, const double *p, double *q,
                                   sfr_scalars *sc, double *part)
{
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A), pcg_spmv_cap(A->n))};
  with_op64_pair(A, [&](auto op) {
    using Op = decltype(op);
    SFR_LAUNCH_RPT(k_pcg_spmv, Op, gm, s, op, p, q, REGION(part, 0), sc);
  });
  return hipGetLastError();
}
