// Uneven batch: many independent SPD systems of different sizes and patterns,
// each solved by mixed-precision Jacobi-PCG (FP32 CG on D^-1/2 A D^-1/2 with
// FP64 reliable updates, the method of mixed.hip), all in one stream of
// launches replayed from one captured graph.
//
// Work items.  System s is cut into g_s entries of rpt_s * 256 rows (rpt_s,
// and the device format of its matrix, are per-system choices).  A compact
// list of the systems still running (act: ids and the prefix sums of their
// g) is rebuilt on the device by k_ub_compact, one workgroup, once per
// graph chunk.  Every other kernel runs a fixed grid and walks the active
// entries grid-stride; entry e maps to (system, local chunk) by a binary
// search of the prefix in LDS.  Finished systems therefore cost nothing once
// compacted away, and the captured graph never changes.
//
// Each system moves through its own states (sc[s].ru):
//   INIT / RESTART / FOLD  a reliable update in the three slot kernels
//       s1  r64 = b - A (x + D^-1/2 y) in FP64, r = D^-1/2 r64 (FP32);
//           partials r64.D^-1 r64, r64.r64, r.r, r.p
//       s2  the reference PCG's stop test on r64; x += D^-1/2 y; y = 0
//       s3  p = r + beta (p - (r.p / r.r) r), or p = r (INIT, RESTART)
//   CG  one FP32 iteration: s1 q = A^ p, p.q | s2 alpha, r -= alpha q, r.r |
//       s3 beta, y += alpha p, p = r + beta p, fold decision
//   DONE  nothing (completion stamped from the device wall clock).
// Per-system partials live at fixed offsets (4 g_s doubles per region), so
// the reductions are the deterministic two-stage ones of dev_common.hpp.
// Scalar ownership as in pcg.hip: within one kernel no workgroup reads a
// scalar field that a workgroup of the same kernel writes.
#include "dev_common.hpp"

enum { UB_CG = 1, UB_FOLD = 2, UB_RESTART = 3, UB_INIT = 4, UB_DONE = 5 };

typedef sfr_ub_sys UbSys;

// Walk the active entries of the launch: body(system, local chunk) for each
// entry this workgroup owns.  The calls are uniform over the workgroup.
template <class F>
__device__ __forceinline__ void walk(const sfr_ub_view &v, F &&body)
{
  __shared__ int s_pre[SFR_UB_MAXSYS + 1];
  __shared__ int s_sys[SFR_UB_MAXSYS];
  const int *act = v.act;
  const int nact = act[0], total = act[1];
  const int *asys = act + 2, *apre = act + 2 + v.nsys;
  for (int k = threadIdx.x; k < nact; k += SFR_BLOCK) {
    s_sys[k] = asys[k];
    s_pre[k] = apre[k];
  }
  if (threadIdx.x == 0) s_pre[nact] = total;
  __syncthreads();
  const int G = gridDim.x;
  for (int e = xcd_chunk(blockIdx.x, G); e < total; e += G) {
    int lo = 0, hi = nact - 1;            // last k with s_pre[k] <= e
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    // the same for every lane: keep it in scalar registers, so the system's
    // descriptor and scalars are read with scalar loads
    const int s = __builtin_amdgcn_readfirstlane(s_sys[lo]);
    const int lb = __builtin_amdgcn_readfirstlane(e - s_pre[lo]);
    body(s, lb);
  }
}

// workgroup sum of NV values -> part[k*g + lb] (the system's partial slots)
template <int NV>
__device__ __forceinline__ void store_at(double (&v)[NV], double *part, int g,
                                         int lb)
{
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) part[(size_t)k * g + lb] = v[k];
  }
}

__device__ __forceinline__ double *slots_of(double *part, const UbSys &S)
{
  return part + 4 * (size_t)S.gfirst;
}

// column of slot f of row i in an ELL system (16-bit offsets when present)
__device__ __forceinline__ int64_t ell_col(const UbSys &S, int f, int64_t i)
{
  const int64_t k = f * (int64_t)S.n + i;
  return S.col16 ? d16_col(S.col16, S.exc_k, S.exc_col, S.nexc, k, i)
                 : (int64_t)S.col[k];
}

// (A^ x)_i of the unit-diagonal FP32 operator of system S
__device__ __forceinline__ float row32(const UbSys &S, int64_t i,
                                       const float *__restrict__ x)
{
  float acc = x[i];
  if (S.fmt == SFR_FMT_ELL) {
    const int64_t n = S.n;
    if (S.slots == 6) {
#pragma unroll
      for (int f = 0; f < 6; f++) acc += S.val32[f * n + i] * x[ell_col(S, f, i)];
    } else {
      for (int f = 0; f < S.slots; f++) acc += S.val32[f * n + i] * x[ell_col(S, f, i)];
    }
  } else {
    const int32_t e = S.rowptr[i + 1];
    for (int32_t k = S.rowptr[i]; k < e; k++) acc += S.val32[k] * x[S.col[k]];
  }
  return acc;
}

// Rows a thread keeps in flight: the slot kernels handle UB_U rows of a
// work item at once, all loads first, then the arithmetic and the stores,
// so a thread has UB_U rows' memory traffic outstanding instead of one.
#define UB_U 4

// rows r0 .. r0+UB_U-1 of this thread in the work item starting at `base`;
// rows outside the item or the system are clamped to the last row (loads
// stay in bounds) and flagged !ok (no store, no contribution)
struct Rows {
  int64_t i[UB_U];
  bool ok[UB_U];
};

__device__ __forceinline__ Rows rows_at(const UbSys &S, int64_t base, int r0)
{
  Rows R;
#pragma unroll
  for (int u = 0; u < UB_U; u++) {
    const int64_t i = base + (int64_t)(r0 + u) * SFR_BLOCK + threadIdx.x;
    R.ok[u] = r0 + u < S.rpt && i < S.n;
    R.i[u] = R.ok[u] ? i : S.n - 1;
  }
  return R;
}

// (A^ x)_i for the UB_U rows of R
__device__ __forceinline__ void rows32(const UbSys &S, const Rows &R,
                                       const float *__restrict__ x,
                                       float (&out)[UB_U])
{
  if (S.fmt == SFR_FMT_ELL && S.slots == 6) {
    const int64_t n = S.n;
    const float *__restrict__ val = S.val32;
    int64_t cj[UB_U][6];
    float a[UB_U][6];
    bool esc[UB_U] = {};
    if (S.col16) {
      // escaped offsets read the row itself here and are fixed up below
      const int16_t *__restrict__ d16 = S.col16;
#pragma unroll
      for (int u = 0; u < UB_U; u++)
#pragma unroll
        for (int f = 0; f < 6; f++) {
          const int d = d16[f * n + R.i[u]];
          const bool e = d == SFR_D16_ESC;
          esc[u] |= e;
          cj[u][f] = e ? R.i[u] : R.i[u] + d;
          a[u][f] = val[f * n + R.i[u]];
        }
    } else {
      const int32_t *__restrict__ col = S.col;
#pragma unroll
      for (int u = 0; u < UB_U; u++)
#pragma unroll
        for (int f = 0; f < 6; f++) {
          cj[u][f] = col[f * n + R.i[u]];
          a[u][f] = val[f * n + R.i[u]];
        }
    }
#pragma unroll
    for (int u = 0; u < UB_U; u++) {
      float acc = x[R.i[u]];
#pragma unroll
      for (int f = 0; f < 6; f++) acc += a[u][f] * x[cj[u][f]];
      if (esc[u]) {
        const float xi = x[R.i[u]];
        for (int64_t e = 0; e < S.nexc; e++) {
          const int64_t k = S.exc_k[e];
          if (k % n == R.i[u]) acc += val[k] * (x[S.exc_col[e]] - xi);
        }
      }
      out[u] = acc;
    }
  } else {
#pragma unroll
    for (int u = 0; u < UB_U; u++) out[u] = row32(S, R.i[u], x);
  }
}

// value of the not yet folded solution x + D^-1/2 y at row j
__device__ __forceinline__ double xval(const UbSys &S, int64_t j,
                                       const double *__restrict__ x,
                                       const float *__restrict__ y)
{
  return x[j] + S.dscale[j] * (double)y[j];
}

// (A (x + D^-1/2 y))_i in FP64
__device__ __forceinline__ double row64_fold(const UbSys &S, int64_t i,
                                             const double *__restrict__ x,
                                             const float *__restrict__ y)
{
  double acc = S.ad[i] * xval(S, i, x, y);
  if (S.fmt == SFR_FMT_ELL) {
    const int64_t n = S.n;
    for (int f = 0; f < S.slots; f++)
      acc += S.val[f * n + i] * xval(S, ell_col(S, f, i), x, y);
  } else {
    const int32_t e = S.rowptr[i + 1];
    for (int32_t k = S.rowptr[i]; k < e; k++) acc += S.val[k] * xval(S, S.col[k], x, y);
  }
  return acc;
}

// (A x)_i in FP64
__device__ __forceinline__ double row64(const UbSys &S, int64_t i,
                                        const double *__restrict__ x)
{
  double acc = S.ad[i] * x[i];
  if (S.fmt == SFR_FMT_ELL) {
    const int64_t n = S.n;
    for (int f = 0; f < S.slots; f++) acc += S.val[f * n + i] * x[ell_col(S, f, i)];
  } else {
    const int32_t e = S.rowptr[i + 1];
    for (int32_t k = S.rowptr[i]; k < e; k++) acc += S.val[k] * x[S.col[k]];
  }
  return acc;
}

// ---- active list -----------------------------------------------------------

// One workgroup.  all = 1: every system (set-up and final passes; also the
// start stamp); all = 0: the systems not DONE.  Systems are listed in index
// order, so the list (and every entry's mapping) is deterministic.
__global__ __launch_bounds__(SFR_BLOCK) void k_ub_compact(sfr_ub_view v,
                                                          int all)
{
  __shared__ int sc_n[SFR_BLOCK], sc_e[SFR_BLOCK];
  const int nsys = v.nsys;
  const int per = (nsys + SFR_BLOCK - 1) / SFR_BLOCK;
  const int lo = threadIdx.x * per < nsys ? threadIdx.x * per : nsys;
  const int hi = lo + per < nsys ? lo + per : nsys;
  int cn = 0, ce = 0;
  for (int s = lo; s < hi; s++)
    if (all || v.sc[s].ru != UB_DONE) { cn++; ce += v.sys[s].g; }
  sc_n[threadIdx.x] = cn;
  sc_e[threadIdx.x] = ce;
  __syncthreads();
  // inclusive scan (Hillis-Steele) of the two counts
  for (int off = 1; off < SFR_BLOCK; off <<= 1) {
    const int an = threadIdx.x >= off ? sc_n[threadIdx.x - off] : 0;
    const int ae = threadIdx.x >= off ? sc_e[threadIdx.x - off] : 0;
    __syncthreads();
    sc_n[threadIdx.x] += an;
    sc_e[threadIdx.x] += ae;
    __syncthreads();
  }
  int k = sc_n[threadIdx.x] - cn, e = sc_e[threadIdx.x] - ce;
  int *asys = v.act + 2, *apre = v.act + 2 + nsys;
  for (int s = lo; s < hi; s++)
    if (all || v.sc[s].ru != UB_DONE) {
      asys[k] = s;
      apre[k] = e;
      k++;
      e += v.sys[s].g;
    }
  if (threadIdx.x == SFR_BLOCK - 1) {
    v.act[0] = sc_n[SFR_BLOCK - 1];
    v.act[1] = sc_e[SFR_BLOCK - 1];
    if (all == 1) v.ts[nsys] = wall_clock64();
  }
}

// ---- set-up ----------------------------------------------------------------

// partials of b.D^-1 b and b.b; x = 0 when newX; y = p = 0
__global__ __launch_bounds__(SFR_BLOCK) void k_ub_init(sfr_ub_view v, int newX)
{
  walk(v, [&](int s, int lb) {
    const UbSys S = v.sys[s];
    const double *b = v.b + S.row0;
    double acc[2] = {0.0, 0.0};
    const int64_t base = (int64_t)lb * S.rpt * SFR_BLOCK;
    for (int r = 0; r < S.rpt; r++) {
      const int64_t i = base + r * SFR_BLOCK + threadIdx.x;
      if (i < S.n) {
        const double bi = b[i];
        acc[0] += bi * (bi * S.dinv[i]);
        acc[1] += bi * bi;
        const int64_t k = S.row0 + i;
        if (newX) v.x[k] = 0.0;
        v.y[k] = 0.0f;
        v.p[k] = 0.0f;
      }
    }
    store_at<2>(acc, slots_of(v.partB, S), S.g, lb);
  });
}

// tolerance and loop state (the workgroup of chunk 0 of each system)
__global__ __launch_bounds__(SFR_BLOCK) void k_ub_init2(sfr_ub_view v,
    double tol, unsigned maxit, double delta)
{
  walk(v, [&](int s, int lb) {
    if (lb != 0) return;
    const UbSys S = v.sys[s];
    double nb[2];
    reduce_partials_seq<2>(slots_of(v.partB, S), S.g, nb);
    if (threadIdx.x == 0) {
      sfr_scalars *c = &v.sc[s];
      c->norm_b = sqrt(fabs(nb[0]));
      c->conv = tol * c->norm_b;
      c->bb = nb[1];
      c->tol = tol;
      c->maxit = maxit;
      c->delta2 = delta * delta;
      c->d = 0.0;
      c->d_prev = 1.0;
      c->rho_max = 0.0;
      c->alpha = 0.0;
      c->iter = 0;
      c->outer = 0;
      c->done = SFR_DONE_RUN;
      c->ru = UB_INIT;
      c->run2 = c->run3 = 0;
    }
  });
}

// ---- the three slot kernels of an iteration ---------------------------------

__global__ __launch_bounds__(SFR_BLOCK) void k_ub_s1(sfr_ub_view v)
{
  walk(v, [&](int s, int lb) {
    const UbSys S = v.sys[s];
    sfr_scalars *c = &v.sc[s];
    const unsigned mode = c->ru;
    const bool owner = lb == 0 && threadIdx.x == 0;
    if (mode == UB_DONE) {
      if (owner) c->run2 = 0;
      return;
    }
    if (owner) c->run2 = mode;
    const int64_t base = (int64_t)lb * S.rpt * SFR_BLOCK;
    const float *p = v.p + S.row0;
    if (mode == UB_CG) {                       // q = A^ p; p.q
      const float *__restrict__ pp = p;
      float *__restrict__ q = v.q + S.row0;
      double acc[1] = {0.0};
      for (int r0 = 0; r0 < S.rpt; r0 += UB_U) {
        const Rows R = rows_at(S, base, r0);
        float qv[UB_U], pv[UB_U];
        rows32(S, R, pp, qv);
#pragma unroll
        for (int u = 0; u < UB_U; u++) pv[u] = pp[R.i[u]];
#pragma unroll
        for (int u = 0; u < UB_U; u++)
          if (R.ok[u]) {
            q[R.i[u]] = qv[u];
            acc[0] += (double)qv[u] * (double)pv[u];
          }
      }
      store_at<1>(acc, slots_of(v.partA, S), S.g, lb);
      return;
    }
    // fold: true residual of x + D^-1/2 y in FP64
    const double *x = v.x + S.row0, *b = v.b + S.row0;
    const float *y = v.y + S.row0;
    float *rr = v.r + S.row0;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int r = 0; r < S.rpt; r++) {
      const int64_t i = base + r * SFR_BLOCK + threadIdx.x;
      if (i < S.n) {
        const double ri = b[i] - row64_fold(S, i, x, y);
        const float ci = (float)(ri * S.dscale[i]);
        rr[i] = ci;
        acc[0] += ri * (ri * S.dinv[i]);
        acc[1] += ri * ri;
        acc[2] += (double)ci * (double)ci;
        acc[3] += (double)ci * (double)p[i];
      }
    }
    store_at<4>(acc, slots_of(v.partA, S), S.g, lb);
  });
}

__global__ __launch_bounds__(SFR_BLOCK) void k_ub_s2(sfr_ub_view v)
{
  walk(v, [&](int s, int lb) {
    const UbSys S = v.sys[s];
    sfr_scalars *c = &v.sc[s];
    const unsigned run = c->run2;
    const bool owner = lb == 0 && threadIdx.x == 0;
    if (!run) {
      if (owner) c->run3 = 0;
      return;
    }
    const int64_t base = (int64_t)lb * S.rpt * SFR_BLOCK;
    if (run == UB_CG) {
      double pq[1];
      reduce_partials_seq<1>(slots_of(v.partA, S), S.g, pq);
      const double d = c->d, alpha = d / pq[0];
      if (!(pq[0] > 0.0) || !isfinite(alpha)) {
        // FP32 lost positive definiteness: fold, then restart with p = r
        if (owner) c->run3 = UB_RESTART;
        return;
      }
      if (owner) {
        c->alpha = alpha;
        c->d_prev = d;
        if (d > c->rho_max) c->rho_max = d;
        c->run3 = UB_CG;
      }
      const float fa = (float)alpha;
      const float *__restrict__ q = v.q + S.row0;
      float *__restrict__ r = v.r + S.row0;
      double acc[1] = {0.0};
      for (int r0 = 0; r0 < S.rpt; r0 += UB_U) {
        const Rows R = rows_at(S, base, r0);
        float rv[UB_U], qv[UB_U];
#pragma unroll
        for (int u = 0; u < UB_U; u++) {
          rv[u] = r[R.i[u]];
          qv[u] = q[R.i[u]];
        }
#pragma unroll
        for (int u = 0; u < UB_U; u++)
          if (R.ok[u]) {
            const float ri = rv[u] - fa * qv[u];
            r[R.i[u]] = ri;
            acc[0] += (double)ri * (double)ri;
          }
      }
      store_at<1>(acc, slots_of(v.partB, S), S.g, lb);
      return;
    }
    // fold: x += D^-1/2 y, y = 0; the stop test on the FP64 residual
    double *x = v.x + S.row0;
    float *y = v.y + S.row0;
    for (int k = 0; k < S.rpt; k++) {
      const int64_t i = base + k * SFR_BLOCK + threadIdx.x;
      if (i < S.n) {
        x[i] += S.dscale[i] * (double)y[i];
        y[i] = 0.0f;
      }
    }
    if (lb != 0) return;
    double t[4];
    reduce_partials_seq<4>(slots_of(v.partA, S), S.g, t);
    if (threadIdx.x == 0) {
      const double rmr = t[0];
      const bool conv = sqrt(fabs(rmr)) < c->conv || rmr == 0.0;
      const bool brk = !isfinite(rmr) || !isfinite(t[2]) ||
                       (!conv && t[2] == 0.0);
      const unsigned outer = c->outer + (run == UB_INIT ? 0u : 1u);
      c->res[2] = rmr;
      c->res[3] = t[1];
      c->outer = outer;
      if (conv || brk || c->iter >= c->maxit || outer >= c->maxit) {
        c->done = conv ? SFR_DONE_CONV : (brk ? SFR_DONE_BREAK : SFR_DONE_MAXIT);
        c->ru = UB_DONE;
        c->run3 = 0;
        v.ts[s] = wall_clock64();
      } else {
        c->run3 = run;
      }
    }
  });
}

__global__ __launch_bounds__(SFR_BLOCK) void k_ub_s3(sfr_ub_view v)
{
  walk(v, [&](int s, int lb) {
    const UbSys S = v.sys[s];
    sfr_scalars *c = &v.sc[s];
    const unsigned run = c->run3;
    if (!run) return;
    const bool owner = lb == 0 && threadIdx.x == 0;
    const int64_t base = (int64_t)lb * S.rpt * SFR_BLOCK;
    float *__restrict__ p = v.p + S.row0;
    const float *__restrict__ r = v.r + S.row0;
    if (run == UB_CG) {
      double t[1];
      reduce_partials_seq<1>(slots_of(v.partB, S), S.g, t);
      const double rr = t[0];
      const double conv = c->conv;
      const bool fold = rr < c->delta2 * c->rho_max || rr <= conv * conv ||
                        !isfinite(rr);
      const float fa = (float)c->alpha, fb = (float)(rr / c->d_prev);
      float *__restrict__ y = v.y + S.row0;
      for (int r0 = 0; r0 < S.rpt; r0 += UB_U) {
        const Rows R = rows_at(S, base, r0);
        float pv[UB_U], rv[UB_U], yv[UB_U];
#pragma unroll
        for (int u = 0; u < UB_U; u++) {
          pv[u] = p[R.i[u]];
          rv[u] = r[R.i[u]];
          yv[u] = y[R.i[u]];
        }
#pragma unroll
        for (int u = 0; u < UB_U; u++)
          if (R.ok[u]) {
            y[R.i[u]] = yv[u] + fa * pv[u];
            if (!fold) p[R.i[u]] = rv[u] + fb * pv[u];
          }
      }
      if (owner) {
        const unsigned it = c->iter + 1;
        c->iter = it;
        c->d = rr;
        if (fold || it >= c->maxit) c->ru = UB_FOLD;
      }
      return;
    }
    if (run == UB_RESTART && c->run2 == UB_CG) {
      // breakdown seen in s2 of this CG iteration: fold next
      if (owner) c->ru = UB_RESTART;
      return;
    }
    // after a fold: new direction from the refreshed residual
    double t[4];
    reduce_partials_seq<4>(slots_of(v.partA, S), S.g, t);
    const double c2 = t[2];
    const bool fresh = run != UB_FOLD;          // INIT or RESTART: p = r
    const float fb = (float)(c2 / c->d_prev), fg = (float)(t[3] / c2);
    for (int k = 0; k < S.rpt; k++) {
      const int64_t i = base + k * SFR_BLOCK + threadIdx.x;
      if (i < S.n) {
        const float ri = r[i];
        p[i] = fresh ? ri : ri + fb * (p[i] - fg * ri);
      }
    }
    if (owner) {
      c->d = c2;
      c->rho_max = c2;
      c->ru = UB_CG;
    }
  });
}

// ---- final quantities: x.Ax, x.x, r.D^-1 r, r.r in FP64 ---------------------

__global__ __launch_bounds__(SFR_BLOCK) void k_ub_final(sfr_ub_view v)
{
  walk(v, [&](int s, int lb) {
    const UbSys S = v.sys[s];
    const double *x = v.x + S.row0, *b = v.b + S.row0;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t base = (int64_t)lb * S.rpt * SFR_BLOCK;
    for (int k = 0; k < S.rpt; k++) {
      const int64_t i = base + k * SFR_BLOCK + threadIdx.x;
      if (i < S.n) {
        const double ax = row64(S, i, x);
        const double xi = x[i], ri = b[i] - ax;
        acc[0] += xi * ax;
        acc[1] += xi * xi;
        acc[2] += ri * (ri * S.dinv[i]);
        acc[3] += ri * ri;
      }
    }
    store_at<4>(acc, slots_of(v.partB, S), S.g, lb);
  });
}

__global__ __launch_bounds__(SFR_BLOCK) void k_ub_final2(sfr_ub_view v)
{
  walk(v, [&](int s, int lb) {
    if (lb != 0) return;
    const UbSys S = v.sys[s];
    double t[4];
    reduce_partials_seq<4>(slots_of(v.partB, S), S.g, t);
    if (threadIdx.x == 0)
      for (int k = 0; k < 4; k++) v.sc[s].res[k] = t[k];
  });
}

// ---- launchers -------------------------------------------------------------

extern "C" int sfr_k_ub_compact(hipStream_t s, const sfr_ub_view *v, int all)
{
  hipLaunchKernelGGL(k_ub_compact, dim3(1), dim3(SFR_BLOCK), 0, s, *v, all);
  return hipGetLastError();
}

extern "C" int sfr_k_ub_init(hipStream_t s, const sfr_ub_view *v, int newX,
                             double tol, unsigned maxit, double delta)
{
  hipLaunchKernelGGL(k_ub_init, dim3(v->grid), dim3(SFR_BLOCK), 0, s, *v, newX);
  hipLaunchKernelGGL(k_ub_init2, dim3(v->grid), dim3(SFR_BLOCK), 0, s, *v,
                     tol, maxit, delta);
  return hipGetLastError();
}

extern "C" int sfr_k_ub_iter(hipStream_t s, const sfr_ub_view *v)
{
  hipLaunchKernelGGL(k_ub_s1, dim3(v->grid), dim3(SFR_BLOCK), 0, s, *v);
  hipLaunchKernelGGL(k_ub_s2, dim3(v->grid), dim3(SFR_BLOCK), 0, s, *v);
  hipLaunchKernelGGL(k_ub_s3, dim3(v->grid), dim3(SFR_BLOCK), 0, s, *v);
  return hipGetLastError();
}

extern "C" int sfr_k_ub_final(hipStream_t s, const sfr_ub_view *v)
{
  hipLaunchKernelGGL(k_ub_final, dim3(v->grid), dim3(SFR_BLOCK), 0, s, *v);
  hipLaunchKernelGGL(k_ub_final2, dim3(v->grid), dim3(SFR_BLOCK), 0, s, *v);
  return hipGetLastError();
}
