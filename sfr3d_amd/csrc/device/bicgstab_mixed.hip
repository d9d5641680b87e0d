// Mixed-precision Jacobi-BiCGSTAB: FP32 iterations with FP64 reliable
// updates, for one system (NB = 1) or a batch of NB = 2..4 systems on one
// pattern (SFR3D's u/v/w momentum solves with their own diagonals, or its
// species with their own values).
//
// The inner loop is the reference's right-preconditioned BiCGSTAB
// (src/Bicgstab.c:50-245; z = D^-1 p, x += alpha z + omega z') with the
// vectors, the matrix values and the diagonal in FP32 and every dot product
// accumulated in FP64.  Its iterate is a correction y (in x's units, FP32)
// that is folded into the FP64 solution x:
//   - when the recurrence residual has dropped by a factor delta below its
//     largest norm since the previous fold (reliable update: G. Sleijpen,
//     H. van der Vorst, "Reliable updated residuals in hybrid Bi-CG
//     methods", Computing 56 (1996)), or
//   - when it meets the reference's test ||r|| < tol ||b||
//     (src/Bicgstab.c:79-81, 129-131, 153-154), after either half step.
// A fold computes x += y, the true residual r = b - A x in FP64 and the
// reference's test on it; if it fails, the FP32 residual is replaced by the
// true one (residual replacement: H. van der Vorst, Q. Ye, "Residual
// replacement strategies for Krylov subspace iterative methods for the
// convergence of true residuals", SIAM J. Sci. Comput. 22 (2000)) and the
// recurrence restarts from it with a new shadow, r0 = r, p = r.  Going on
// with the old shadow and direction (beta from the replaced residual) was
// measured on the CPU and dropped: in FP32 the replaced residual and the
// old Krylov quantities disagree enough that BiCGSTAB's irregular
// convergence turns into stagnation (80^3 Poisson, delta 1e-3: 1,498
// iterations against 645 with restarts and 554 in FP64; on the 128^3
// system the continued recurrence ended in a breakdown).  A fold requested
// at the half step (omega unknown, the residual met the test) keeps the
// shadow with p = r.  alpha, omega and beta must be finite in FP32 as well:
// a value beyond FLT_MAX is a breakdown, not an update.  The fold threshold
// defaults to delta = 0.1 (option ru_delta_bicg): each restarted FP32 cycle
// must reduce its residual by delta, and on harder systems the attainable
// FP32 reduction (about kappa * eps32) is short of 1e-3 (128^3 Poisson:
// 1,397 FP32 iterations with delta 0.1, stagnation with 1e-3).  When the
// FP64 true residual stops improving (SFR_MB_STALL folds without halving)
// the loop ends as a breakdown and the host finishes in FP64 from x.
//
// Vectors are interleaved per row, x[i*NB + c].  Each system keeps its own
// scalars (sc[c]) and its own fold requests: a system waiting for its fold
// is skipped by the iteration kernels while the others go on, and the fold
// kernels touch only the systems whose request flag is set.
//   M1 v = A z; partials v.r0                               matrix + 3 words
//   M2 alpha; y += alpha z; s = r - alpha v; z = D^-1 s; s.s  8 words
//   M3 half-step test; t = A z; partials t.s, t.t           matrix + 3 words
//   M4 omega; y += omega z; r = s - omega t; r.r, r.r0      7 words
//   M5 fold decision; beta; p = r + beta (p - omega v); z = D^-1 p   5 words
// (FP32 words of 4 bytes per row and system.)
#include <type_traits>
#include "dev_common.hpp"

#define REGION(part, k) ((part) + (size_t)(k) * 4 * SFR_MAXG)
static_assert(4 * 4 * SFR_GRIDCAP <= 4 * SFR_MAXG, "partials region too small");

#define OWNER(c) (blockIdx.x == 0 && threadIdx.x == (unsigned)(c))

// folds in a row without progress that end the FP32 loop (see F2)
#define SFR_MB_STALL 8.0

// a finite double that stays finite in FP32
__device__ __forceinline__ bool fits32(double v)
{
  return fabs(v) <= 3.4028234663852886e38;
}

// Scalars of system c (no kernel reads a value its own first workgroup
// writes): run2..run5 hand the run flag M1->M2->M3->M4->M5; ru = fold
// requested (1 at the end of an iteration, 2 at the half step, 3 after a
// breakdown), read by the fold kernels, cleared by F2; pbuf = system is
// iterating (set by F2, cleared with a request); d = rho (r.r0), copied to
// d_prev by M4 for M5; the largest ||r||^2 since the last fold is staged in
// res[5] by M5 / F2 and copied to rho_max by M1; fresh = the fold mode,
// staged by F1 for F2.

// The FP32 state of a batch
struct MbState {
  float *y, *r, *r0, *p, *z, *v, *t;
  const float *ad, *dinv;       // FP32 diagonal and inverse, interleaved
};

// M1: v = A z; partials v.r0
template <class Op, int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_v(Op A, MbState st,
    double *pout, sfr_scalars *sc)
{
  bool act[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    act[c] = sc[c].done == SFR_DONE_RUN && sc[c].pbuf != 0;
    any |= act[c];
    if (OWNER(c)) {
      sc[c].run2 = act[c];
      sc[c].rho_max = sc[c].res[5];       // staged by M5 / F2
    }
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
      float vi[NB];
      row_multi(A, i, valid, st.ad, st.z, vi);
      if (valid && lead) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (!act[c]) continue;
          st.v[i * NB + c] = vi[c];
          acc[c] += (double)vi[c] * (double)st.r0[i * NB + c];
        }
      }
    }
  }
  store_partials<NB>(acc, pout);
}

// M2: alpha = rho / v.r0; y += alpha z, s = r - alpha v, z = D^-1 s; s.s
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_s(int64_t n, MbState st,
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
    const bool brk = act[c] && (!(vr0[c] != 0.0) || !fits32(alpha[c]));
    if (OWNER(c)) {
      if (brk) { sc[c].ru = 3; sc[c].pbuf = 0; }
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
          const float fa = (float)alpha[c];
          st.y[k] += fa * st.z[k];
          const float si = st.r[k] - fa * st.v[k];
          st.r[k] = si;
          st.z[k] = si * st.dinv[k];
          acc[c] += (double)si * (double)si;
        }
      }
    }
  }
  store_partials<NB>(acc, pout);
}

// M3: half-step test (a fold request when ||s|| meets the reference's test:
// y already holds alpha z, the correction of the half step); t = A z;
// partials t.s, t.t
template <class Op, int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_t(Op A, MbState st,
    const double *pin, int gin, double *pout, sfr_scalars *sc)
{
  bool act[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) { act[c] = sc[c].run3 != 0; any |= act[c]; }
  if (!any) {
#pragma unroll
    for (int c = 0; c < NB; c++) if (OWNER(c)) sc[c].run4 = 0;
    return;
  }
  double ss[NB];
  reduce_partials<NB>(pin, gin, ss);
  any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    const bool hit = act[c] && (sqrt(ss[c]) < sc[c].conv || ss[c] == 0.0 || !isfinite(ss[c]));
    if (OWNER(c)) {
      if (act[c]) sc[c].iter = sc[c].iter + (hit ? 1u : 0u);
      if (hit) { sc[c].ru = 2; sc[c].pbuf = 0; }
      sc[c].run4 = act[c] && !hit;
    }
    act[c] = act[c] && !hit;
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
      float ti[NB];
      row_multi(A, i, valid, st.ad, st.z, ti);
      if (valid && lead) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (!act[c]) continue;
          const int64_t k = i * NB + c;
          st.t[k] = ti[c];
          acc[c] += (double)ti[c] * (double)st.r[k];
          acc[NB + c] += (double)ti[c] * (double)ti[c];
        }
      }
    }
  }
  store_partials<2 * NB>(acc, pout);
}

// M4: omega = t.s / t.t; y += omega z, r = s - omega t; partials r.r, r.r0
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_r(int64_t n, MbState st,
    const double *pin, int gin, double *pout, sfr_scalars *sc)
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
    const bool brk = act[c] && (!(tt[NB + c] != 0.0) || !(w[c] != 0.0) || !fits32(w[c]));
    if (OWNER(c)) {
      if (brk) { sc[c].ru = 3; sc[c].pbuf = 0; }
      sc[c].omega = w[c];
      sc[c].d_prev = sc[c].d;              // rho of this iteration, for M5
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
          const float fw = (float)w[c];
          st.y[k] += fw * st.z[k];
          const float ri = st.r[k] - fw * st.t[k];
          st.r[k] = ri;
          acc[c] += (double)ri * (double)ri;
          acc[NB + c] += (double)ri * (double)st.r0[k];
        }
      }
    }
  }
  store_partials<2 * NB>(acc, pout);
}

// M5: the reference's test on the recurrence residual, or a drop by delta
// since the last fold: a fold request; else beta = (r.r0 / rho) (alpha /
// omega), p = r + beta (p - omega v), z = D^-1 p
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_p(int64_t n, MbState st,
    const double *pin, int gin, sfr_scalars *sc)
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
    const double rho = sc[c].d_prev;
    w[c] = sc[c].omega;
    beta[c] = (rr[NB + c] / rho) * (sc[c].alpha / w[c]);
    const bool fold = act[c] && (sqrt(rr[c]) < sc[c].conv || rr[c] == 0.0 ||
                                 rr[c] < sc[c].delta2 * sc[c].rho_max || !isfinite(rr[c]));
    const bool brk = act[c] && !fold && (!(rho != 0.0) || !fits32(beta[c]));
    if (OWNER(c) && act[c]) {
      const unsigned it = sc[c].iter + 1;
      sc[c].iter = it;
      if (fold || brk || it >= sc[c].maxit) {
        sc[c].ru = brk ? 3u : 1u;
        sc[c].pbuf = 0;
      } else {
        sc[c].d = rr[NB + c];
        sc[c].res[5] = rr[c] > sc[c].rho_max ? rr[c] : sc[c].rho_max;
      }
    }
    act[c] = act[c] && !fold && !brk;
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
          const float pi = st.r[k] + (float)beta[c] * (st.p[k] - (float)w[c] * st.v[k]);
          st.p[k] = pi;
          st.z[k] = pi * st.dinv[k];
        }
      }
    }
  }
}

// ---- fold (FP64) -------------------------------------------------------------
// The systems whose request flag ru is set: F1 x += y, y = 0, r64 = b - A x
// (FP64 matrix and diagonal), partials ||r64||^2 and r64.r0 (the replaced
// FP32 residual r = r64 goes to r); F2 the reference's test on ||r64||, then
// the new direction and the loop state.  init = the first call: every system,
// r0 = r (the reference's shadow, src/Solv.c:187), p = r.
template <class Op, int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_f1(Op A,
    const double *__restrict__ ad, const double *__restrict__ b,
    double *__restrict__ x, MbState st, int init, int zero_x, double *pout,
    sfr_scalars *sc)
{
  bool act[NB], zero[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    act[c] = sc[c].done == SFR_DONE_RUN && (init || sc[c].ru != 0);
    zero[c] = init && sc[c].bb == 0.0;   // b = 0: x = 0, done already
    any |= act[c] || zero[c];
    // F2 reads the mode from here, not from ru, which it clears
    if (OWNER(c)) sc[c].fresh = act[c] ? (init ? 2u : sc[c].ru) : 0u;
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
      double ax[NB];
      if (zero_x) {
#pragma unroll
        for (int c = 0; c < NB; c++) ax[c] = 0.0;
      } else {
        row_multi(A, i, valid, ad, x, ax);
      }
      if (valid && lead) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          const int64_t k = i * NB + c;
          if (zero[c]) x[k] = 0.0;
          if (!act[c]) continue;
          const double ri = b[k] - ax[c];
          const float rf = (float)ri;
          st.r[k] = rf;
          if (init) {
            st.r0[k] = rf;
            st.y[k] = 0.0f;
          }
          acc[c] += ri * ri;
          acc[NB + c] += (double)rf * (double)st.r0[k];
        }
      }
    }
  }
  store_partials<2 * NB>(acc, pout);
}

// x += y, y = 0 for the systems with a fold request
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_addy(int64_t n,
    double *__restrict__ x, MbState st, sfr_scalars *sc)
{
  bool act[NB], any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    act[c] = sc[c].done == SFR_DONE_RUN && sc[c].ru != 0;
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
          x[k] += (double)st.y[k];
          st.y[k] = 0.0f;
        }
      }
    }
  }
}

// F2: stop test on the true residual; the new direction p = r (with a new
// shadow r0 = r after an end-of-iteration fold or a breakdown)
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_f2(int64_t n, MbState st,
    int init, const double *pin, int gin, sfr_scalars *sc)
{
  bool act[NB], any = false;
  unsigned mode[NB];
#pragma unroll
  for (int c = 0; c < NB; c++) {
    mode[c] = sc[c].fresh;               // 1 continue, 2 p = r, 3 r0 = r, p = r
    act[c] = mode[c] != 0;
    any |= act[c];
  }
  if (!any) return;
  double t[2 * NB];
  reduce_partials<2 * NB>(pin, gin, t);
  any = false;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    const double rr = t[c];
    double rho = t[NB + c];
    const bool conv = act[c] && (sqrt(rr) < sc[c].conv || rr == 0.0);
    const bool bad = act[c] && !conv && !isfinite(rr);
    const bool maxed = act[c] && !conv && !bad && sc[c].iter >= sc[c].maxit;
    // stagnation: SFR_MB_STALL folds in a row without the true residual
    // norm halving below the best one seen (beta holds that best ||r||^2,
    // res[4] the count) ends the FP32 loop as a breakdown; the host then
    // finishes the solve in FP64 from this x
    bool stall = false;
    if (act[c] && !conv && !bad && !maxed && !init)
      stall = !(rr < 0.25 * sc[c].beta) && sc[c].res[4] + 1.0 >= SFR_MB_STALL;
    // an end-of-iteration fold (1) or a breakdown (3): a new shadow, r
    // itself; a half-step fold (2): the same shadow
    const unsigned md = mode[c] == 2 ? 2u : 3u;
    if (md == 3) rho = rr;
    mode[c] = md;
    if (OWNER(c) && act[c]) {
      sc[c].res[3] = rr;
      if (!init) sc[c].outer = sc[c].outer + 1;
      if (init || rr < 0.25 * sc[c].beta) {
        sc[c].beta = rr;
        sc[c].res[4] = 0.0;
      } else {
        sc[c].res[4] = sc[c].res[4] + 1.0;
      }
      if (conv) sc[c].done = SFR_DONE_CONV;
      else if (bad || stall) sc[c].done = SFR_DONE_BREAK;
      else if (maxed) sc[c].done = SFR_DONE_MAXIT;
      sc[c].d = rho;
      sc[c].res[5] = rr;                   // largest ||r||^2 since this fold
      sc[c].ru = 0;
      sc[c].pbuf = 1;
    }
    act[c] = act[c] && !conv && !bad && !maxed && !stall;
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
          const float ri = st.r[k];
          if (mode[c] == 3) st.r0[k] = ri;
          st.p[k] = ri;
          st.z[k] = ri * st.dinv[k];
        }
      }
    }
  }
}

// loop state of every system: tolerance from ||b||, counters
template <int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_start(const double *pin,
    int gin, double tol, unsigned maxit, double delta, sfr_scalars *sc)
{
  double bb[NB];
  reduce_partials<NB>(pin, gin, bb);
  if (threadIdx.x < NB) {
    const int c = threadIdx.x;
    sc[c].bb = bb[c];
    sc[c].norm_b = sqrt(bb[c]);
    sc[c].conv = tol * sqrt(bb[c]);
    sc[c].tol = tol;
    sc[c].maxit = maxit;
    sc[c].delta2 = delta * delta;
    sc[c].iter = 0;
    sc[c].outer = 0;
    sc[c].ru = 0;
    sc[c].pbuf = 0;
    sc[c].omega = 1.0;
    sc[c].alpha = 1.0;
    sc[c].run2 = sc[c].run3 = sc[c].run4 = sc[c].run5 = 0;
    sc[c].done = bb[c] == 0.0 ? SFR_DONE_CONV : SFR_DONE_RUN;
  }
}

// partials of b.b per system; x = 0 when newX; the FP32 diagonal and its
// inverse from the FP64 ones
template <int RPT, int NB>
__global__ __launch_bounds__(SFR_BLOCK) void k_mb_s1(int64_t n,
    const double *__restrict__ b, const double *__restrict__ ad,
    double *__restrict__ x, float *__restrict__ ad32,
    float *__restrict__ dinv32, int newX, double *pout)
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
          const int64_t k = i * NB + c;
          const double bi = b[k];
          acc[c] += bi * bi;
          if (newX) x[k] = 0.0;
          ad32[k] = (float)ad[k];
          dinv32[k] = (float)(1.0 / ad[k]);
        }
      }
    }
  }
  store_partials<NB>(acc, pout);
}

// ---- launchers ---------------------------------------------------------------

#define MB_LAUNCH_M(K, Op, NB, g, s, ...)                                      \
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

#define MB_LAUNCH_V(K, NB, g, s, ...)                                          \
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

// FP32 operator of the batch: the FP32 copy of the shared values, or, with
// per-system values (valnb32), the same pattern over those; the diagonal is
// the batch's own (st.ad), so the operator's is not used
template <typename F>
static void with_op32raw_b(const sfr_mat_view *A, const float *val32,
                           const float *valnb32, F &&f)
{
  auto run = [&](const float *vals, auto tag) {
    using Tag = decltype(tag);
    if (A->fmt == SFR_FMT_ELL) {
      auto ell = [&](auto op) {
        if constexpr (Tag::value) f(PerSys<decltype(op)>{op});
        else f(op);
      };
      if (A->col16 && A->slots == 6) {
        ell(EllOp<float, 6, false, true>{vals, A->col, nullptr, A->n, 6, A->col16, A->exc_k, A->exc_col, A->nexc});
        return;
      }
      if (A->col16 && A->slots == 4) {
        ell(EllOp<float, 4, false, true>{vals, A->col, nullptr, A->n, 4, A->col16, A->exc_k, A->exc_col, A->nexc});
        return;
      }
      switch (A->slots) {
      case 6: ell(EllOp<float, 6, false>{vals, A->col, nullptr, A->n, 6}); return;
      case 4: ell(EllOp<float, 4, false>{vals, A->col, nullptr, A->n, 4}); return;
      case 3: ell(EllOp<float, 3, false>{vals, A->col, nullptr, A->n, 3}); return;
      default: ell(EllOp<float, 0, false>{vals, A->col, nullptr, A->n, A->slots}); return;
      }
    }
    auto csr = [&](auto op) {
      if constexpr (Tag::value) f(PerSys<decltype(op)>{op});
      else f(op);
    };
    switch (A->lanes) {
    case 1: csr(CsrOp<float, 1, false>{vals, A->col, A->rowptr, nullptr, A->n, 0}); return;
    case 2: csr(CsrOp<float, 2, false>{vals, A->col, A->rowptr, nullptr, A->n, 0}); return;
    case 4: csr(CsrOp<float, 4, false>{vals, A->col, A->rowptr, nullptr, A->n, 0}); return;
    default: csr(CsrOp<float, 8, false>{vals, A->col, A->rowptr, nullptr, A->n, 0}); return;
    }
  };
  if (valnb32) run(valnb32, std::true_type{});
  else run(val32, std::false_type{});
}

template <int NB>
static void mb_start(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                     const double *ad, const double *b, double *x,
                     float *ad32, float *dinv32, int newX, double tol,
                     unsigned maxit, double delta, sfr_scalars *sc,
                     double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  double *rc = REGION(part, 4);
  MB_LAUNCH_V(k_mb_s1, NB, gv, s, A->n, b, ad, x, ad32, dinv32, newX, rc);
  hipLaunchKernelGGL(k_mb_start<NB>, dim3(1), dim3(SFR_BLOCK), 0, s,
                     (const double *)rc, gv.grid, tol, maxit, delta, sc);
}

template <int NB>
static void mb_fold(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                    const double *ad, const double *b, double *x, MbState st,
                    int init, int zero_x, sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *rf = REGION(part, 5);
  if (!init) MB_LAUNCH_V(k_mb_addy, NB, gv, s, A->n, x, st, sc);
  with_op64_b(A, [&](auto op) {
    using Op = decltype(op);
    MB_LAUNCH_M(k_mb_f1, Op, NB, gm, s, op, ad, b, x, st, init, zero_x, rf, sc);
  });
  MB_LAUNCH_V(k_mb_f2, NB, gv, s, A->n, st, init, (const double *)rf, gm.grid, sc);
}

template <int NB>
static void mb_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g0,
                    const float *val32, const float *valnb32, MbState st,
                    sfr_scalars *sc, double *part)
{
  const sfr_geom gv = {g0.rpt, grid_for(A->n, g0.rpt, 1)};
  const sfr_geom gm = {g0.rpt, grid_for(A->n, g0.rpt, op_lanes(A))};
  double *ra = REGION(part, 0), *rb = REGION(part, 1);
  double *rc = REGION(part, 2), *rd = REGION(part, 3);
  with_op32raw_b(A, val32, valnb32, [&](auto op) {
    using Op = decltype(op);
    MB_LAUNCH_M(k_mb_v, Op, NB, gm, s, op, st, ra, sc);
  });
  MB_LAUNCH_V(k_mb_s, NB, gv, s, A->n, st, (const double *)ra, gm.grid, rb, sc);
  with_op32raw_b(A, val32, valnb32, [&](auto op) {
    using Op = decltype(op);
    MB_LAUNCH_M(k_mb_t, Op, NB, gm, s, op, st, (const double *)rb, gv.grid, rc, sc);
  });
  MB_LAUNCH_V(k_mb_r, NB, gv, s, A->n, st, (const double *)rc, gm.grid, rd, sc);
  MB_LAUNCH_V(k_mb_p, NB, gv, s, A->n, st, (const double *)rd, gv.grid, sc);
}

#define MB_DISPATCH(nb, CALL)                                                  \
  switch (nb) {                                                                \
  case 1: CALL(1); break;                                                      \
  case 2: CALL(2); break;                                                      \
  case 3: CALL(3); break;                                                      \
  case 4: CALL(4); break;                                                      \
  default: return hipErrorInvalidValue;                                        \
  }

static MbState mb_state(float *const w[7], const float *ad32, const float *dinv32)
{
  MbState st;
  st.y = w[0]; st.r = w[1]; st.r0 = w[2]; st.p = w[3]; st.z = w[4];
  st.v = w[5]; st.t = w[6]; st.ad = ad32; st.dinv = dinv32;
  return st;
}

extern "C" int sfr_k_mb_start(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                              int nb, const double *ad, const double *b,
                              double *x, float *ad32, float *dinv32, int newX,
                              double tol, unsigned maxit, double delta,
                              sfr_scalars *sc, double *part)
{
#define CALL(NB) mb_start<NB>(s, A, g, ad, b, x, ad32, dinv32, newX, tol, maxit, delta, sc, part)
  MB_DISPATCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

extern "C" int sfr_k_mb_fold(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                             int nb, const double *ad, const double *b,
                             double *x, float *const w[7], const float *ad32,
                             const float *dinv32, int init, int zero_x,
                             sfr_scalars *sc, double *part)
{
  const MbState st = mb_state(w, ad32, dinv32);
#define CALL(NB) mb_fold<NB>(s, A, g, ad, b, x, st, init, zero_x, sc, part)
  MB_DISPATCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

extern "C" int sfr_k_mb_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                             int nb, const float *val32, const float *valnb32,
                             float *const w[7], const float *ad32,
                             const float *dinv32, sfr_scalars *sc,
                             double *part)
{
  const MbState st = mb_state(w, ad32, dinv32);
#define CALL(NB) mb_iter<NB>(s, A, g, val32, valnb32, st, sc, part)
  MB_DISPATCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

// FP32 copies of values: dst[k] = (float) src[k]
__global__ __launch_bounds__(SFR_BLOCK) void k_to_f32(int64_t n,
    const double *__restrict__ src, float *__restrict__ dst)
{
  const int64_t k = (int64_t)blockIdx.x * SFR_BLOCK + threadIdx.x;
  if (k < n) dst[k] = (float)src[k];
}

extern "C" int sfr_k_to_f32(hipStream_t s, int64_t n, const double *src,
                            float *dst)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + SFR_BLOCK - 1) / SFR_BLOCK)),
                     dim3(SFR_BLOCK), 0, s, n, src, dst);
  return hipGetLastError();
}
