// Device helpers shared by the gfx950 kernels: XCD-aware workgroup mapping,
// wave64/workgroup/grid reductions with a fixed summation order, and the
// sparse row operators for the two device formats.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../sfr_internal.h"

#define SFR_WAVES (SFR_BLOCK / 64)

// Workgroup b of a grid of g (g % 8 == 0) works on chunk xcd_chunk(b, g).
// Workgroups are handed to the 8 XCDs round-robin, so XCD x receives the
// contiguous chunk range [x*g/8, (x+1)*g/8): rows that are neighbours in the
// ordering share that XCD's L2.
__device__ __forceinline__ int xcd_chunk(int b, int g)
{
  return (b & 7) * (g >> 3) + (b >> 3);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum NV values over the workgroup; totals are valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
  __shared__ double sh[NV][SFR_WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = sh[k][0];
#pragma unroll
      for (int j = 1; j < SFR_WAVES; j++) s += sh[k][j];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Block partial sums: the workgroup's sum of each of the NV values goes to
// part[k*gridDim.x + blockIdx.x] with a plain store.  No atomics and no
// fences: the partials are read by a later kernel on the same stream, and
// the kernel boundary makes them visible on every XCD.
template <int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], double *part)
{
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) part[(size_t)k * gridDim.x + blockIdx.x] = v[k];
  }
}

// Totals of the g partials a previous kernel stored (slot-major), summed in
// a fixed order by the whole workgroup and handed to every thread.  Every
// workgroup that calls it gets bitwise the same totals, so scalars derived
// from them (alpha, beta, stop tests) agree across the grid without any
// grid-wide synchronisation.  Kernels issue their own first loads before the
// call, so that the partials' latency and theirs overlap.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double *__restrict__ part,
                                                int g, double (&out)[NV])
{
  __shared__ double bc[NV];
  double t[NV];
  // Up to 8 partials a thread (2,048 a workgroup) have their loads issued
  // together before any is added: the partials were written by workgroups
  // on every XCD, so most loads leave this XCD's L2, and one round trip per
  // partial in turn was most of a short kernel's fixed cost.  They are
  // added in the same ascending order as one at a time.  Up to SFR_BLOCK
  // partials (one a thread; the per-system sets of the uneven batch) take
  // a single load.
  constexpr int U = 8;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double *pk = part + (size_t)k * g;
    double s = 0.0;
    if (g <= SFR_BLOCK) {
      if ((int)threadIdx.x < g) s += pk[threadIdx.x];
      t[k] = s;
      continue;
    }
    for (int base = 0; base < g; base += U * SFR_BLOCK) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int j = base + u * SFR_BLOCK + (int)threadIdx.x;
        v[u] = j < g ? pk[j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (base + u * SFR_BLOCK + (int)threadIdx.x < g) s += v[u];
    }
    t[k] = s;
  }
  block_sum<NV>(t);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) bc[k] = t[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) out[k] = bc[k];
  __syncthreads();
}

// The same sum with one load at a time: for kernels that reduce many small
// partial sets in turn (the uneven batch reduces its system's set for every
// chunk it takes), where the unrolled loads of reduce_partials measured
// slower (260 -> 246 batch solves/s).  The same order and bits.
template <int NV>
__device__ __forceinline__ void reduce_partials_seq(const double *__restrict__ part,
                                                    int g, double (&out)[NV])
{
  __shared__ double bs[NV];
  double t[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (int j = threadIdx.x; j < g; j += SFR_BLOCK) s += part[(size_t)k * g + j];
    t[k] = s;
  }
  block_sum<NV>(t);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) bs[k] = t[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) out[k] = bs[k];
  __syncthreads();
}

// a*b + c rounded once, written out rather than left to contraction: where
// the compiler fuses a product into an add depends on the basic blocks the
// two land in, so two kernels with the same source expression could round
// differently; the row products and the direction updates are fused here
// by the source, the same in every kernel that forms them
__device__ __forceinline__ float fmad(float a, float b, float c)
{
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fmad(double a, double b, double c)
{
  return __builtin_fma(a, b, c);
}

// Loads of a matrix stream that a product reads once (ELL values, column
// offsets, diagonals).  NT: non-temporal, so the stream does not displace
// the loop's vectors from the caches.  That pays once the loop's working
// set exceeds the 256 MB last-level cache (measured on MI355X: the 200^3
// FP32 loop kernel takes 87.8 against 101.4 us); below it the plain load
// keeps the matrix cached across iterations and is faster (128^3: 22.0
// against 26.3 us).  The host chooses per loop (sfr_loop_nt).
template <bool NT, class T>
__device__ __forceinline__ T mld(const T *p)
{
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// ---- sparse row operators ---------------------------------------------------
// row(i, valid, x) returns (A x)_i for the L lanes that share row i; lanes of
// an invalid row return 0 but still take part in the cross-lane shuffles.

// Slot-major ELL: val[f*n + i], col[f*n + i]; one lane per row.  S > 0 fixes
// the slot count at compile time, S == 0 reads it at run time.  UNIT = the
// diagonal is 1 (Jacobi-scaled copy) and ad is not read.  D16 = the columns
// are read as 16-bit offsets from the row, dcol[f*n + i] = col - i (half the
// index traffic).  The few entries whose offset does not fit hold
// SFR_D16_ESC: the gather loop reads the row's own x there (a select, no
// branch), and a row that has one adds a_k (x[col] - x[i]) afterwards, with
// the column from a short exception list (k = f*n + i, col).
__device__ __forceinline__ int64_t d16_col(const int16_t *__restrict__ dcol,
    const int64_t *xk, const int32_t *xc, int64_t nx, int64_t k, int64_t i)
{
  const int d = dcol[k];
  if (d != SFR_D16_ESC) return i + d;
  for (int64_t e = 0; e < nx; e++)
    if (xk[e] == k) return xc[e];
  return i;                    // not reached: every escape is listed
}

template <typename T, int S, bool UNIT, bool D16 = false, bool NT = false>
struct EllOp {
  static constexpr int L = 1;
  const T *val;
  const int32_t *col;
  const T *ad;
  int64_t n;
  int s;
  const int16_t *dcol;
  const int64_t *xk;
  const int32_t *xc;
  int64_t nx;
  // column of slot f of row i (exact, escapes resolved; set-up paths)
  __device__ __forceinline__ int64_t cj(int f, int64_t i) const
  {
    if constexpr (D16) return d16_col(dcol, xk, xc, nx, f * n + i, i);
    else return col[f * n + i];
  }
  // column the gather loop reads: the row itself for an escaped entry
  __device__ __forceinline__ int64_t cg(int f, int64_t i, bool &esc) const
  {
    if constexpr (D16) {
      const int d = mld<NT>(dcol + f * n + i);
      const bool e = d == SFR_D16_ESC;
      esc |= e;
      return e ? i : i + d;
    } else {
      return col[f * n + i];
    }
  }
  // sum over the escaped entries of row i of a_k (v(col) - v(i))
  template <class V>
  __device__ T esc_fix(int64_t i, V v) const
  {
    T acc = T(0);
    const T vi = v(i);
    for (int64_t e = 0; e < nx; e++) {
      const int64_t k = xk[e];
      if (k % n == i) acc += val[k] * (v((int64_t)xc[e]) - vi);
    }
    return acc;
  }
  __device__ __forceinline__ T row(int64_t i, bool valid, const T *__restrict__ x) const
  {
    if (!valid) return T(0);
    T acc = UNIT ? x[i] : mld<NT>(ad + i) * x[i];
    bool esc = false;
    if constexpr (S > 0) {
#pragma unroll
      for (int f = 0; f < S; f++) acc = fmad(mld<NT>(val + f * n + i), x[cg(f, i, esc)], acc);
    } else {
      for (int f = 0; f < s; f++) acc = fmad(mld<NT>(val + f * n + i), x[cg(f, i, esc)], acc);
    }
    if constexpr (D16)
      if (esc) acc += esc_fix(i, [&](int64_t j) { return x[j]; });
    return acc;
  }
  // (A (r + beta p))_i with a unit diagonal: the combination of each
  // neighbour is formed exactly as that neighbour's own row forms it, so the
  // result is the product with the updated vector, bit for bit (rows with an
  // escaped entry: up to the rounding of the fix-up)
  __device__ __forceinline__ T row_comb(int64_t i, bool valid,
      const T *__restrict__ r, const T *__restrict__ p, T beta) const
  {
    if (!valid) return T(0);
    T acc = fmad(beta, p[i], r[i]);
    bool esc = false;
    if constexpr (S > 0) {
#pragma unroll
      for (int f = 0; f < S; f++) {
        const int64_t j = cg(f, i, esc);
        acc = fmad(mld<NT>(val + f * n + i), fmad(beta, p[j], r[j]), acc);
      }
    } else {
      for (int f = 0; f < s; f++) {
        const int64_t j = cg(f, i, esc);
        acc = fmad(mld<NT>(val + f * n + i), fmad(beta, p[j], r[j]), acc);
      }
    }
    if constexpr (D16)
      if (esc) acc += esc_fix(i, [&](int64_t j) { return fmad(beta, p[j], r[j]); });
    return acc;
  }
};

// Symmetric slot-major ELL read through face pairs.  SFR3D's face order
// pairs the slots of a hex row as (0,1) (2,4) (3,5) and of a quad row as
// (0,2) (1,3): the neighbour j across slot f of row i faces back to i
// across the other slot of the pair.  When the values mirror too (a_ij in
// row i's slot f equals, bit for bit, a_ji in j's paired slot; a slot
// facing no neighbour, offset 0, holds +0 in the second slot of a pair),
// a row reads the first slot of each pair from its own plane and the
// second from the neighbour's first-slot plane, at the offset it reads for
// the gather anyway: half the value stream, the other half from cache.
// The values and their order are EllOp's, so the products are bit for bit
// EllOp's (16-bit offsets without escapes only).
template <int S>
__device__ __forceinline__ constexpr int pair_src(int f)
{
  if constexpr (S == 6) return f == 1 ? 0 : (f == 4 ? 2 : (f == 5 ? 3 : f));
  else return f >= 2 ? f - 2 : f;
}

template <typename T, int S, bool UNIT, bool NT = false>
struct PairOp {
  static_assert(S == 6 || S == 4, "face pairs of hex or quad rows");
  static constexpr int L = 1;
  const T *val;
  const T *ad;
  int64_t n;
  const int16_t *dcol;
  __device__ __forceinline__ T value(int f, int64_t i, int d) const
  {
    if (pair_src<S>(f) == f) return val[f * n + i];
    return d == 0 ? T(0) : val[pair_src<S>(f) * n + i + d];
  }
  __device__ __forceinline__ T row(int64_t i, bool valid, const T *__restrict__ x) const
  {
    if (!valid) return T(0);
    T acc = UNIT ? x[i] : mld<NT>(ad + i) * x[i];
#pragma unroll
    for (int f = 0; f < S; f++) {
      const int d = mld<NT>(dcol + f * n + i);
      acc = fmad(value(f, i, d), x[i + d], acc);
    }
    return acc;
  }
  __device__ __forceinline__ T row_comb(int64_t i, bool valid,
      const T *__restrict__ r, const T *__restrict__ p, T beta) const
  {
    if (!valid) return T(0);
    T acc = fmad(beta, p[i], r[i]);
#pragma unroll
    for (int f = 0; f < S; f++) {
      const int d = mld<NT>(dcol + f * n + i);
      const int64_t j = i + d;
      acc = fmad(value(f, i, d), fmad(beta, p[j], r[j]), acc);
    }
    return acc;
  }
};

// NB systems sharing the pattern and the off-diagonal values, each with its
// own diagonal: x and ad interleaved per row (x[i*NB + c]).  One pass over
// the row's columns and values serves all NB products.
template <typename T, int S, bool UNIT, bool D16, int NB>
__device__ __forceinline__ void ell_row_multi(const EllOp<T, S, UNIT, D16> &A,
    int64_t i, bool valid, const T *__restrict__ ad, const T *__restrict__ x,
    T (&acc)[NB])
{
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = T(0);
  if (!valid) return;
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = ad[i * NB + c] * x[i * NB + c];
  bool esc = false;
  if constexpr (S > 0) {
#pragma unroll
    for (int f = 0; f < S; f++) {
      const T v = A.val[f * A.n + i];
      const int64_t j = A.cg(f, i, esc);
#pragma unroll
      for (int c = 0; c < NB; c++) acc[c] = fmad(v, x[j * NB + c], acc[c]);
    }
  } else {
    for (int f = 0; f < A.s; f++) {
      const T v = A.val[f * A.n + i];
      const int64_t j = A.cg(f, i, esc);
#pragma unroll
      for (int c = 0; c < NB; c++) acc[c] = fmad(v, x[j * NB + c], acc[c]);
    }
  }
  if constexpr (D16)
    if (esc) {
#pragma unroll
      for (int c = 0; c < NB; c++)
        acc[c] += A.esc_fix(i, [&](int64_t j) { return x[j * NB + c]; });
    }
}

// Full-row CSR (off-diagonals only, diagonal apart); L consecutive lanes per
// row stride through the row and are combined with xor shuffles.
template <typename T, int L_, bool UNIT>
struct CsrOp {
  static constexpr int L = L_;
  const T *val;
  const int32_t *col;
  const int32_t *ptr;
  const T *ad;
  int64_t n;
  int s;
  __device__ __forceinline__ T row(int64_t i, bool valid, const T *__restrict__ x) const
  {
    T acc = T(0);
    if (valid) {
      const int sub = threadIdx.x & (L - 1);
      const int e = ptr[i + 1];
      for (int k = ptr[i] + sub; k < e; k += L) acc = fmad(val[k], x[col[k]], acc);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
    if (valid) acc = UNIT ? acc + x[i] : fmad(ad[i], x[i], acc);
    return acc;
  }
  // (A (r + beta p))_i with a unit diagonal (see EllOp::row_comb)
  __device__ __forceinline__ T row_comb(int64_t i, bool valid,
      const T *__restrict__ r, const T *__restrict__ p, T beta) const
  {
    T acc = T(0);
    if (valid) {
      const int sub = threadIdx.x & (L - 1);
      const int e = ptr[i + 1];
      for (int k = ptr[i] + sub; k < e; k += L) {
        const int64_t j = col[k];
        acc = fmad(val[k], fmad(beta, p[j], r[j]), acc);
      }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
    if (valid) acc += fmad(beta, p[i], r[i]);
    return acc;
  }
};

template <typename T, int L, bool UNIT, int NB>
__device__ __forceinline__ void csr_row_multi(const CsrOp<T, L, UNIT> &A,
    int64_t i, bool valid, const T *__restrict__ ad, const T *__restrict__ x,
    T (&acc)[NB])
{
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = T(0);
  if (valid) {
    const int sub = threadIdx.x & (L - 1);
    const int e = A.ptr[i + 1];
    for (int k = A.ptr[i] + sub; k < e; k += L) {
      const T v = A.val[k];
      const int64_t j = A.col[k];
#pragma unroll
      for (int c = 0; c < NB; c++) acc[c] = fmad(v, x[j * NB + c], acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < NB; c++)
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, L);
  if (valid) {
#pragma unroll
    for (int c = 0; c < NB; c++) acc[c] = fmad(ad[i * NB + c], x[i * NB + c], acc[c]);
  }
}

template <typename T, int S, bool UNIT, bool D16, int NB>
__device__ __forceinline__ void row_multi(const EllOp<T, S, UNIT, D16> &A,
    int64_t i, bool valid, const T *__restrict__ ad, const T *__restrict__ x,
    T (&acc)[NB])
{
  ell_row_multi<T, S, UNIT, D16, NB>(A, i, valid, ad, x, acc);
}

template <typename T, int L, bool UNIT, int NB>
__device__ __forceinline__ void row_multi(const CsrOp<T, L, UNIT> &A, int64_t i,
    bool valid, const T *__restrict__ ad, const T *__restrict__ x, T (&acc)[NB])
{
  csr_row_multi<T, L, UNIT, NB>(A, i, valid, ad, x, acc);
}

// NB systems sharing the pattern, each with its own values: the operator's
// val holds them interleaved per entry, val[k*NB + c] for device entry k
// (SFR3D's transported species, one matrix per species on one pattern).
// The wave's loads of slot f stay one contiguous segment (64*NB values).
template <class Op>
struct PerSys : Op {};

template <typename T, int S, bool UNIT, bool D16, int NB>
__device__ __forceinline__ void row_multi(const PerSys<EllOp<T, S, UNIT, D16>> &A,
    int64_t i, bool valid, const T *__restrict__ ad, const T *__restrict__ x,
    T (&acc)[NB])
{
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = T(0);
  if (!valid) return;
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = ad[i * NB + c] * x[i * NB + c];
  bool esc = false;
  auto slot = [&](int f) {
    const int64_t k = ((int64_t)f * A.n + i) * NB;
    const int64_t j = A.cg(f, i, esc);
#pragma unroll
    for (int c = 0; c < NB; c++) acc[c] = fmad(A.val[k + c], x[j * NB + c], acc[c]);
  };
  if constexpr (S > 0) {
#pragma unroll
    for (int f = 0; f < S; f++) slot(f);
  } else {
    for (int f = 0; f < A.s; f++) slot(f);
  }
  if constexpr (D16)
    if (esc) {
      for (int64_t e = 0; e < A.nx; e++) {
        const int64_t k = A.xk[e];
        if (k % A.n != i) continue;
        const int64_t j = A.xc[e];
#pragma unroll
        for (int c = 0; c < NB; c++)
          acc[c] += A.val[k * NB + c] * (x[j * NB + c] - x[i * NB + c]);
      }
    }
}

template <typename T, int L, bool UNIT, int NB>
__device__ __forceinline__ void row_multi(const PerSys<CsrOp<T, L, UNIT>> &A,
    int64_t i, bool valid, const T *__restrict__ ad, const T *__restrict__ x,
    T (&acc)[NB])
{
#pragma unroll
  for (int c = 0; c < NB; c++) acc[c] = T(0);
  if (valid) {
    const int sub = threadIdx.x & (L - 1);
    const int e = A.ptr[i + 1];
    for (int k = A.ptr[i] + sub; k < e; k += L) {
      const int64_t j = A.col[k];
#pragma unroll
      for (int c = 0; c < NB; c++) acc[c] = fmad(A.val[(int64_t)k * NB + c], x[j * NB + c], acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < NB; c++)
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, L);
  if (valid) {
#pragma unroll
    for (int c = 0; c < NB; c++) acc[c] = fmad(ad[i * NB + c], x[i * NB + c], acc[c]);
  }
}

// Call f(op) with the FP64 operator matching the view.
template <typename F>
__host__ inline void with_op64(const sfr_mat_view *A, F &&f)
{
  if (A->fmt == SFR_FMT_ELL) {
    if (A->col16 && A->slots == 6) {
      f(EllOp<double, 6, false, true>{A->val, A->col, A->ad, A->n, 6, A->col16, A->exc_k, A->exc_col, A->nexc});
      return;
    }
    if (A->col16 && A->slots == 4) {
      f(EllOp<double, 4, false, true>{A->val, A->col, A->ad, A->n, 4, A->col16, A->exc_k, A->exc_col, A->nexc});
      return;
    }
    switch (A->slots) {
    case 6: f(EllOp<double, 6, false>{A->val, A->col, A->ad, A->n, 6}); return;
    case 4: f(EllOp<double, 4, false>{A->val, A->col, A->ad, A->n, 4}); return;
    case 3: f(EllOp<double, 3, false>{A->val, A->col, A->ad, A->n, 3}); return;
    default: f(EllOp<double, 0, false>{A->val, A->col, A->ad, A->n, A->slots}); return;
    }
  }
  switch (A->lanes) {
  case 1: f(CsrOp<double, 1, false>{A->val, A->col, A->rowptr, A->ad, A->n, 0}); return;
  case 2: f(CsrOp<double, 2, false>{A->val, A->col, A->rowptr, A->ad, A->n, 0}); return;
  case 4: f(CsrOp<double, 4, false>{A->val, A->col, A->rowptr, A->ad, A->n, 0}); return;
  default: f(CsrOp<double, 8, false>{A->val, A->col, A->rowptr, A->ad, A->n, 0}); return;
  }
}

// The batched kernels' operator: the view's own values, or, when the view
// carries per-system values (valnb), the same pattern over those.
template <typename F>
__host__ inline void with_op64_b(const sfr_mat_view *A, F &&f)
{
  if (!A->valnb) {
    with_op64(A, f);
    return;
  }
  sfr_mat_view v = *A;
  v.val = A->valnb;
  with_op64(&v, [&](auto op) { f(PerSys<decltype(op)>{op}); });
}

// Call f(op) with the FP32 Jacobi-scaled (unit diagonal) operator; a view
// with nt set streams a 16-bit-offset ELL with non-temporal loads.
template <typename F>
__host__ inline void with_op32(const sfr_mat_view *A, F &&f)
{
  if (A->fmt == SFR_FMT_ELL) {
    if (A->col16 && A->slots == 6 && A->nt) {
      f(EllOp<float, 6, true, true, true>{A->val32, A->col, nullptr, A->n, 6, A->col16, A->exc_k, A->exc_col, A->nexc});
      return;
    }
    if (A->col16 && A->slots == 4 && A->nt) {
      f(EllOp<float, 4, true, true, true>{A->val32, A->col, nullptr, A->n, 4, A->col16, A->exc_k, A->exc_col, A->nexc});
      return;
    }
    if (A->col16 && A->slots == 6) {
      f(EllOp<float, 6, true, true>{A->val32, A->col, nullptr, A->n, 6, A->col16, A->exc_k, A->exc_col, A->nexc});
      return;
    }
    if (A->col16 && A->slots == 4) {
      f(EllOp<float, 4, true, true>{A->val32, A->col, nullptr, A->n, 4, A->col16, A->exc_k, A->exc_col, A->nexc});
      return;
    }
    switch (A->slots) {
    case 6: f(EllOp<float, 6, true>{A->val32, A->col, nullptr, A->n, 6}); return;
    case 4: f(EllOp<float, 4, true>{A->val32, A->col, nullptr, A->n, 4}); return;
    case 3: f(EllOp<float, 3, true>{A->val32, A->col, nullptr, A->n, 3}); return;
    default: f(EllOp<float, 0, true>{A->val32, A->col, nullptr, A->n, A->slots}); return;
    }
  }
  switch (A->lanes) {
  case 1: f(CsrOp<float, 1, true>{A->val32, A->col, A->rowptr, nullptr, A->n, 0}); return;
  case 2: f(CsrOp<float, 2, true>{A->val32, A->col, A->rowptr, nullptr, A->n, 0}); return;
  case 4: f(CsrOp<float, 4, true>{A->val32, A->col, A->rowptr, nullptr, A->n, 0}); return;
  default: f(CsrOp<float, 8, true>{A->val32, A->col, A->rowptr, nullptr, A->n, 0}); return;
  }
}

// The FP64 PCG SpMV's operator: the face-pair read of the ELL when the view
// allows it (the values checked to mirror; nt: column offsets and
// diagonals by non-temporal loads), else with_op64.  (The FP32
// iteration kernel keeps the full read: it saves 12 of its 60 B/row but
// adds a third gather stream, and measured no faster at 128^3-160^3 and
// 3 % slower at 200^3.)
template <typename F>
__host__ inline void with_op64_pair(const sfr_mat_view *A, F &&f)
{
  if (A->fmt == SFR_FMT_ELL && A->col16 && A->nexc == 0) {
    if (A->pair == SFR_PAIR_HEX && A->slots == 6) {
      if (A->nt) f(PairOp<double, 6, false, true>{A->val, A->ad, A->n, A->col16});
      else f(PairOp<double, 6, false>{A->val, A->ad, A->n, A->col16});
      return;
    }
    if (A->pair == SFR_PAIR_QUAD && A->slots == 4) {
      if (A->nt) f(PairOp<double, 4, false, true>{A->val, A->ad, A->n, A->col16});
      else f(PairOp<double, 4, false>{A->val, A->ad, A->n, A->col16});
      return;
    }
  }
  with_op64(A, f);
}

// Workgroups for n rows at rpt rows per thread and `lanes` lanes per row,
// a multiple of 8 (XCD remap) and at most SFR_GRIDCAP, or `cap` when that
// is smaller (kernels grid-stride beyond it), so a consumer never reduces
// more than SFR_GRIDCAP partials.
static inline int grid_for(int64_t n, int rpt, int lanes, int cap = 0)
{
  const int64_t rows = (int64_t)rpt * (SFR_BLOCK / lanes);
  const int64_t c = (cap > 0 && cap < SFR_GRIDCAP) ? ((cap + 7) / 8) * 8 : SFR_GRIDCAP;
  int64_t g = (n + rows - 1) / rows;
  g = ((g + 7) / 8) * 8;
  if (g < 8) g = 8;
  if (g > c) g = c;
  return (int)g;
}

static inline int op_lanes(const sfr_mat_view *A)
{
  return A->fmt == SFR_FMT_ELL ? 1 : A->lanes;
}

// Launch kernel template K<Op, RPT> / K<RPT> with the rows-per-thread of g.
#define SFR_LAUNCH_RPT(K, Op, g, s, ...)                                      \
  do {                                                                        \
    switch ((g).rpt) {                                                        \
    case 1: hipLaunchKernelGGL((K<Op, 1>), dim3((g).grid), dim3(SFR_BLOCK),   \
                               0, s, __VA_ARGS__); break;                     \
    case 2: hipLaunchKernelGGL((K<Op, 2>), dim3((g).grid), dim3(SFR_BLOCK),   \
                               0, s, __VA_ARGS__); break;                     \
    case 4: hipLaunchKernelGGL((K<Op, 4>), dim3((g).grid), dim3(SFR_BLOCK),   \
                               0, s, __VA_ARGS__); break;                     \
    default: hipLaunchKernelGGL((K<Op, 8>), dim3((g).grid), dim3(SFR_BLOCK),  \
                                0, s, __VA_ARGS__); break;                    \
    }                                                                         \
  } while (0)

#define SFR_LAUNCH_V(K, g, s, ...)                                            \
  do {                                                                        \
    switch ((g).rpt) {                                                        \
    case 1: hipLaunchKernelGGL(K<1>, dim3((g).grid), dim3(SFR_BLOCK), 0, s,   \
                               __VA_ARGS__); break;                           \
    case 2: hipLaunchKernelGGL(K<2>, dim3((g).grid), dim3(SFR_BLOCK), 0, s,   \
                               __VA_ARGS__); break;                           \
    case 4: hipLaunchKernelGGL(K<4>, dim3((g).grid), dim3(SFR_BLOCK), 0, s,   \
                               __VA_ARGS__); break;                           \
    default: hipLaunchKernelGGL(K<8>, dim3((g).grid), dim3(SFR_BLOCK), 0, s,  \
                                __VA_ARGS__); break;                          \
    }                                                                         \
  } while (0)
