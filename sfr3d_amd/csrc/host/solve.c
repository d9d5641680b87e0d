/*
 * Krylov drivers (C host side of the device solve) and the public solves.
 *
 * A solve is: setup kernels, then the iteration body captured once as a HIP
 * graph of `chunk` iterations and replayed; the device decides convergence
 * itself (scalars never leave the GPU inside the loop) and every kernel of an
 * iteration that starts after convergence returns at once.  The host keeps
 * two chunks in flight and polls a 32-byte status copy after each, so the GPU
 * is never idle waiting for the host.  The final kernel computes x.Ax, ||x||,
 * the FP64 residual and its M-norm, as the reference's pcg does after its
 * loop (src/Cg.c:155-176).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../sfr_internal.h"

#define TRY(call)                                                            \
  do {                                                                       \
    hipError_t e_ = (hipError_t)(call);                                      \
    if (e_ != hipSuccess)                                                    \
      return sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s (%s:%d)", #call,      \
                           hipGetErrorString(e_), __FILE__, __LINE__);       \
  } while (0)

#define CHECK(call) do { int rc_ = (call); if (rc_ != SFR_OK) return rc_; } while (0)

int sfr_workspace(sfr_context *ctx, int64_t n);
int sfr_matrix_prepare_mixed(sfr_context *ctx, sfr_matrix *A);
int sfr_matrix_prepare_adjoint(sfr_context *ctx, sfr_matrix *A);

enum { G_PCG = 1, G_BICG = 2, G_MX = 3, G_BB = 4, G_GM = 5, G_MR = 6, G_B2 = 7,
       G_BBV = 8, G_MB = 9 };

/* mixed BiCGSTAB state of nb systems in ctx->mb32: w[0..6] = y, r, r0, p,
 * z, v, t, then the FP32 diagonal and inverse (nb*n floats each) */
static void mb_split(sfr_context *ctx, int64_t nbn, float *w[7], float **ad32,
                     float **dinv32)
{
  for (int k = 0; k < 7; k++) w[k] = ctx->mb32 + (size_t)k * (size_t)nbn;
  *ad32 = ctx->mb32 + (size_t)7 * (size_t)nbn;
  *dinv32 = ctx->mb32 + (size_t)8 * (size_t)nbn;
}

/* BiCGSTAB(2) iterations (two BiCG steps and a GCR(2) step, nine kernels)
 * per captured chunk */
static int b2_chunk(const sfr_context *ctx)
{
  const int c = ctx->chunk / 2;
  return c < 1 ? 1 : c;
}

/* MINRES ring buffers in the workspace (w64[8], w64[9] stay free for the
 * host-pointer path's b and x) */
#define MR_RINGS(ctx)                                                          \
  double *const V_[3] = {(ctx)->w64[0], (ctx)->w64[1], (ctx)->w64[2]};         \
  double *const W_[3] = {(ctx)->w64[3], (ctx)->w64[4], (ctx)->w64[5]};         \
  double *const Z_[2] = {(ctx)->w64[6], (ctx)->w64[7]}

/* MINRES iterations per captured chunk: a multiple of the rings' period 6 */
static int mr_chunk(const sfr_context *ctx)
{
  const int c = (ctx->chunk / 6) * 6;
  return c < 6 ? 6 : c;
}

/* GMRES restart cycles per captured chunk (137 kernels each) */
static int gm_cycles_per_chunk(const sfr_context *ctx)
{
  const int c = ctx->chunk / 16;
  return c < 1 ? 1 : c;
}

/* Iterations per captured chunk of a loop kind: the full chunk, or the
 * short one a solve starts with (see run_loop). */
static int chunk_of(const sfr_context *ctx, int kind, int small)
{
  /* the mixed loop alternates its direction buffers: an even chunk ends
   * where it started; MINRES's rings repeat every 6 iterations */
  switch (kind) {
  case G_MX: return small ? 2 : (ctx->chunk + 1) & ~1;
  case G_GM: return gm_cycles_per_chunk(ctx);
  case G_MR: return small ? 6 : mr_chunk(ctx);
  case G_B2: return small ? 1 : b2_chunk(ctx);
  default:   return small ? (ctx->chunk < 2 ? ctx->chunk : 2) : ctx->chunk;
  }
}

/* Capture chunk_of(kind, small) iterations of one loop body, or reuse a
 * captured one. */
static int get_graph(sfr_context *ctx, int kind, const void *k0,
                     const void *k1, const void *k2, const sfr_mat_view *v,
                     sfr_geom g, int small, hipGraphExec_t *exec)
{
  const int chunk = chunk_of(ctx, kind, small);
  for (int i = 0; i < ctx->ngraph; i++) {
    sfr_graph *e = &ctx->graphs[i];
    if (e->kind == kind && e->chunk == chunk && e->key[0] == k0 &&
        e->key[1] == k1 && e->key[2] == k2 &&
        e->key[3] == (const void *)ctx->w64[0] && e->key[4] == (const void *)(intptr_t)g.rpt &&
        e->key[5] == (const void *)ctx->gm_buf && e->key[6] == (const void *)(intptr_t)g.cap &&
        e->key[7] == (const void *)(intptr_t)(v->lanes | v->pair << 8 | v->nt << 12)) {
      /* least recently used first: a hit moves to the end, so the graph a
       * caller has just been handed is never the one a second lookup for
       * the same solve evicts */
      const sfr_graph hit = *e;
      memmove(e, e + 1, sizeof(sfr_graph) * (size_t)(ctx->ngraph - 1 - i));
      ctx->graphs[ctx->ngraph - 1] = hit;
      *exec = hit.exec;
      return SFR_OK;
    }
  }
  if (ctx->ngraph == (int)(sizeof ctx->graphs / sizeof ctx->graphs[0])) {
    /* a graph is destroyed only with no launch of it in flight */
    TRY(hipStreamSynchronize(ctx->stream));
    hipGraphExecDestroy(ctx->graphs[0].exec);
    memmove(&ctx->graphs[0], &ctx->graphs[1], sizeof(sfr_graph) * (size_t)(ctx->ngraph - 1));
    ctx->ngraph--;
  }
  hipStream_t s = ctx->stream;
  double *const *w = ctx->w64;
  float *const *w32 = ctx->w32;
  TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int lrc = hipSuccess;
  for (int it = 0; it < chunk && lrc == hipSuccess; it++) {
    if (kind == G_PCG)
      lrc = sfr_k_pcg_iter(s, v, g, (double *)k1, w[0], w[1], w[2], ctx->d_sc, ctx->part);
    else if (kind == G_BICG)
      lrc = sfr_k_bicg_iter(s, v, g, (double *)k1, w[0], w[1], w[2], w[3], w[4], w[5],
                            ctx->d_sc, ctx->part);
    else if (kind == G_MX) {
      float *const pb[2] = {w32[2], w32[4]};
      lrc = sfr_k_mx_inner_iter(s, v, g, w32[0], w32[1], pb, w32[3], it & 1, ctx->d_sc,
                                ctx->part);
    }
    else if (kind == G_GM) {  /* k2 is b */
      double *basis = ctx->gm_buf;
      const int64_t ld = ctx->gm_ld;
      double *const wr[2] = {basis + SFR_GM_K * ld, basis + (SFR_GM_K + 1) * ld};
      lrc = sfr_k_gm_cycle(s, v, g, (const double *)k2, (double *)k1, basis, ld, wr,
                           basis + (SFR_GM_K + 2) * ld, ctx->d_sc, ctx->part);
    }
    else if (kind == G_B2)
      lrc = sfr_k_b2_iter(s, v, g, (double *)k1, w, ctx->d_sc, ctx->part);
    else if (kind == G_MR) {
      MR_RINGS(ctx);
      lrc = sfr_k_mr_iter(s, v, g, (double *)k1, V_, W_, Z_, it % 6, ctx->d_sc, ctx->part);
    }
    else if (kind == G_MB) {  /* k1: the FP32 values, k2: the batch size */
      const int nb = (int)(intptr_t)k2;
      float *w[7], *ad32, *dinv32;
      mb_split(ctx, (int64_t)nb * v->n, w, &ad32, &dinv32);
      lrc = sfr_k_mb_iter(s, v, g, nb, v->valnb ? NULL : (const float *)k1,
                          v->valnb ? (const float *)k1 : NULL, w, ad32, dinv32, ctx->d_sc,
                          ctx->part);
    }
    else                    /* G_BB, G_BBV: k2 carries the batch size */
      lrc = sfr_k_bb_iter(s, v, g, (int)(intptr_t)k2, w[8], w[9], w[0], w[1], w[2], w[3],
                          w[4], w[5], w[6], ctx->d_sc, ctx->part);
  }
  hipGraph_t graph = NULL;
  const hipError_t erc = hipStreamEndCapture(s, &graph);
  if (lrc != hipSuccess || erc != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    return sfr_set_error(ctx, SFR_HIP_ERROR, "graph capture: %s",
                         hipGetErrorString(lrc != hipSuccess ? (hipError_t)lrc : erc));
  }
  hipGraphExec_t ex = NULL;
  const hipError_t irc = hipGraphInstantiate(&ex, graph, NULL, NULL, 0);
  hipGraphDestroy(graph);
  if (irc != hipSuccess)
    return sfr_set_error(ctx, SFR_HIP_ERROR, "graph instantiate: %s", hipGetErrorString(irc));
  sfr_graph *e = &ctx->graphs[ctx->ngraph++];
  memset(e, 0, sizeof *e);
  e->exec = ex; e->kind = kind; e->chunk = chunk;
  e->key[0] = k0; e->key[1] = k1; e->key[2] = k2;
  e->key[3] = (const void *)ctx->w64[0];
  e->key[4] = (const void *)(intptr_t)g.rpt;
  e->key[5] = (const void *)ctx->gm_buf;
  e->key[6] = (const void *)(intptr_t)g.cap;
  e->key[7] = (const void *)(intptr_t)(v->lanes | v->pair << 8 | v->nt << 12);
  *exec = ex;
  return SFR_OK;
}

/* Replay the loop graph until the device reports the loop finished for
 * every one of the nsys systems: which = 0 waits for sc->done, which = 1 for
 * sc->done or sc->ru (a mixed-precision fold request).  `inflight` chunks
 * (1 or 2) are kept in flight; after each one the scalar blocks are copied
 * to pinned memory. */
static int run_loop_n(sfr_context *ctx, hipGraphExec_t ex, int which,
                      unsigned max_chunks, int nsys, unsigned inflight)
{
  hipStream_t s = ctx->stream;
  unsigned launched = 0, seen = 0;
  for (;;) {
    while (launched - seen < inflight && launched < max_chunks) {
      const int slot = (int)(launched & 1);
      TRY(hipGraphLaunch(ex, s));
      TRY(hipMemcpyAsync(ctx->h_poll + SFR_MAXSYS * slot, ctx->d_sc,
                         sizeof(sfr_scalars) * (size_t)nsys, hipMemcpyDeviceToHost, s));
      TRY(hipEventRecord(ctx->ev_poll[slot], s));
      launched++;
    }
    if (seen == launched) break;                 /* chunk budget exhausted */
    const int slot = (int)(seen & 1);
    TRY(hipEventSynchronize(ctx->ev_poll[slot]));
    seen++;
    const sfr_scalars *p = ctx->h_poll + SFR_MAXSYS * slot;
    int finished = 1;
    for (int c = 0; c < nsys; c++)
      if (!(p[c].done || (which && p[c].ru))) finished = 0;
    if (finished) break;
  }
  return SFR_OK;
}

/* The loop of a solve: short chunks (exs, cs iterations each) for its first
 * `ramp` iterations, then full ones (exb, cb), two in flight, at most
 * max_iters iterations issued.  Most solves of a SIMPLE step converge in a
 * handful of iterations: starting short bounds the early-exit launches
 * issued past convergence to two short chunks, while a long solve soon
 * runs on full chunks.  `last` (may be NULL) receives the first system's
 * scalars as the last chunk seen left them. */
/* The loop body launched without graphs (the allocation-free solve): the
 * iteration kernels enqueued directly, `iters` iterations per chunk, with
 * the scalars, partials and work vectors wherever the caller keeps them. */
typedef struct direct_loop {
  int           solver;         /* SFR_PCG or SFR_PBICGSTAB                */
  sfr_mat_view  v;
  sfr_geom      g;
  double       *x, *w[6];
  sfr_scalars  *sc;
  double       *part;
} direct_loop;

static int direct_chunk(sfr_context *ctx, const direct_loop *d, unsigned iters)
{
  for (unsigned it = 0; it < iters; it++) {
    if (d->solver == SFR_PCG)
      TRY(sfr_k_pcg_iter(ctx->stream, &d->v, d->g, d->x, d->w[0], d->w[1], d->w[2], d->sc,
                         d->part));
    else
      TRY(sfr_k_bicg_iter(ctx->stream, &d->v, d->g, d->x, d->w[0], d->w[1], d->w[2], d->w[3],
                          d->w[4], d->w[5], d->sc, d->part));
  }
  return SFR_OK;
}

static int run_loop_ramp(sfr_context *ctx, hipGraphExec_t exs, unsigned cs,
                         hipGraphExec_t exb, unsigned cb, unsigned ramp,
                         int which, unsigned max_iters, int nsys,
                         sfr_scalars *last, const direct_loop *dl)
{
  hipStream_t s = ctx->stream;
  const sfr_scalars *dsc = dl ? dl->sc : ctx->d_sc;
  unsigned launched = 0, seen = 0, iters = 0;
  if (last) {
    memset(last, 0, sizeof *last);
    last->done = SFR_DONE_RUN;
  }
  for (;;) {
    while (launched - seen < 2 && iters < max_iters) {
      const int slot = (int)(launched & 1);
      const int small = iters < ramp;
      if (dl) CHECK(direct_chunk(ctx, dl, small ? cs : cb));
      else TRY(hipGraphLaunch(small ? exs : exb, s));
      TRY(hipMemcpyAsync(ctx->h_poll + SFR_MAXSYS * slot, dsc,
                         sizeof(sfr_scalars) * (size_t)nsys, hipMemcpyDeviceToHost, s));
      TRY(hipEventRecord(ctx->ev_poll[slot], s));
      iters += small ? cs : cb;
      launched++;
    }
    if (seen == launched) break;                 /* iteration budget spent */
    const int slot = (int)(seen & 1);
    TRY(hipEventSynchronize(ctx->ev_poll[slot]));
    seen++;
    const sfr_scalars *p = ctx->h_poll + SFR_MAXSYS * slot;
    if (last) memcpy(last, p, sizeof(sfr_scalars) * (size_t)nsys);
    /* which = 0: every system done; 1: every system done or asking for a
     * fold; 2: every system done, or any running one asking for a fold */
    int finished = 1, asks = 0;
    for (int c = 0; c < nsys; c++) {
      if (!(p[c].done || (which && p[c].ru))) finished = 0;
      if (!p[c].done && p[c].ru) asks = 1;
    }
    if (finished || (which == 2 && asks)) break;
  }
  return SFR_OK;
}

/* The short and the full chunk of a loop kind (captured before a solve
 * enqueues anything) and the loop over them for up to maxit iterations. */
typedef struct { hipGraphExec_t s, b; unsigned cs, cb; } chunks;

static int get_chunks(sfr_context *ctx, int kind, const void *k0, const void *k1,
                      const void *k2, const sfr_mat_view *v, sfr_geom g, chunks *c)
{
  CHECK(get_graph(ctx, kind, k0, k1, k2, v, g, 1, &c->s));
  CHECK(get_graph(ctx, kind, k0, k1, k2, v, g, 0, &c->b));
  c->cs = (unsigned)chunk_of(ctx, kind, 1);
  c->cb = (unsigned)chunk_of(ctx, kind, 0);
  return SFR_OK;
}

static int run_chunks(sfr_context *ctx, const chunks *c, int which, unsigned maxit,
                      int nsys)
{
  return run_loop_ramp(ctx, c->s, c->cs, c->b, c->cb, c->cb, which, maxit + 2 * c->cb, nsys,
                       NULL, NULL);
}

void sfr_fill_stats(const sfr_scalars *sc, sfr_stats *st, float ms)
{
  if (!st) return;
  memset(st, 0, sizeof *st);
  st->iters = sc->iter;
  st->outer = sc->outer;
  st->tol = sc->tol;
  st->norm_b = sc->norm_b;
  st->conv = sc->conv;
  st->xKx = sc->res[0];
  st->norm_x = sqrt(sc->res[1]);
  st->norm_r_m = sqrt(fabs(sc->res[2]));
  st->norm_r = sqrt(sc->res[3]);
  st->relres = sc->bb > 0.0 ? st->norm_r / sqrt(sc->bb) : 0.0;
  st->time_s = ms * 1e-3;
  switch (sc->done) {
  case SFR_DONE_CONV: st->status = SFR_OK; break;
  case SFR_DONE_MAXIT: st->status = SFR_MAXIT; break;
  case SFR_DONE_BREAK: st->status = SFR_BREAKDOWN; break;
  default: st->status = SFR_MAXIT; break;
  }
}

/* The final norms of a solve whose scalars and partials are at sc / part,
 * then its statistics. */
static int finish_at(sfr_context *ctx, const sfr_mat_view *v, sfr_geom g,
                     const double *d_b, const double *d_x, sfr_scalars *sc,
                     double *part, sfr_stats *st)
{
  TRY(sfr_k_final(ctx->stream, v, g, d_b, d_x, sc, part));
  TRY(hipEventRecord(ctx->ev_stop, ctx->stream));
  TRY(hipMemcpyAsync(ctx->h_sc, sc, sizeof(sfr_scalars), hipMemcpyDeviceToHost,
                     ctx->stream));
  TRY(hipStreamSynchronize(ctx->stream));
  float ms = 0.0f;
  TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  sfr_stats tmp;
  sfr_fill_stats(ctx->h_sc, &tmp, ms);
  if (st) *st = tmp;
  return tmp.status;
}

static int finish(sfr_context *ctx, const sfr_mat_view *v, sfr_geom g,
                  const double *d_b, const double *d_x, sfr_stats *st)
{
  return finish_at(ctx, v, g, d_b, d_x, ctx->d_sc, ctx->part, st);
}


static int solve_pcg64(sfr_context *ctx, sfr_matrix *A, double tol,
                       unsigned maxit, const double *d_b, double *d_x, int newX,
                       sfr_stats *st)
{
  CHECK(sfr_workspace(ctx, A->n));
  CHECK(sfr_matrix_pair_check(ctx, A));
  sfr_mat_view v = sfr_view(A);
  v.nt = sfr_loop_nt(ctx, A, 0);
  sfr_geom g = {sfr_pcg_rpt(A), 0, 0};
  chunks ch;
  CHECK(get_chunks(ctx, G_PCG, A, d_x, NULL, &v, g, &ch));
  TRY(hipEventRecord(ctx->ev_start, ctx->stream));
  TRY(sfr_k_pcg_setup(ctx->stream, &v, g, d_b, d_x, ctx->w64[0], ctx->w64[1], newX, tol,
                      maxit, ctx->d_sc, ctx->part));
  CHECK(run_chunks(ctx, &ch, 0, maxit, 1));
  return finish(ctx, &v, g, d_b, d_x, st);
}

/* Rows per thread of the FP64 PCG loop by measurement: 64 iterations of
 * PCG (tol 0, so the iteration count is fixed) on a fixed pseudo-random
 * right-hand side at each of 1, 2, 4 and 8 rows per thread, untimed once
 * (graph capture) and timed twice; the fastest time per iteration is kept.
 * Which one wins depends on the matrix size and on the caches (the 128^3
 * Poisson matrix prefers 2, the 96^3 one the default 4). */
int sfr_tune_pcg(sfr_context *ctx, sfr_matrix *A)
{
  const int64_t n = A->n;
  double *h = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  if (!h) return sfr_set_error(ctx, SFR_BAD_ARG, "tune: no host memory");
  uint64_t s = 0x9e3779b97f4a7c15ull;
  for (int64_t i = 0; i < n; i++) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    h[i] = (double)(s >> 11) * 0x1.0p-53 - 0.5;
  }
  double *d_b = NULL, *d_x = NULL;
  int rc = SFR_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc((void **)&d_b, sizeof(double) * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void **)&d_x, sizeof(double) * (size_t)n);
  if (e == hipSuccess)
    e = hipMemcpy(d_b, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
  free(h);
  if (e != hipSuccess)
    rc = sfr_set_error(ctx, SFR_HIP_ERROR, "tune: %s", hipGetErrorString(e));
  const int keep = A->rpt_pcg;
  int best = 0;
  double best_t = 0.0;
  for (int rpt = 1; rpt <= 8 && rc == SFR_OK; rpt *= 2) {
    A->rpt_pcg = rpt;
    double t = 0.0;
    for (int rep = 0; rep < 3 && rc == SFR_OK; rep++) {
      sfr_stats st;
      const int r = solve_pcg64(ctx, A, 0.0, 64, d_b, d_x, 1, &st);
      if (r != SFR_OK && r != SFR_MAXIT && r != SFR_BREAKDOWN) rc = r;
      else if (rep > 0 && st.iters > 0) {
        const double ti = st.time_s / st.iters;
        if (t == 0.0 || ti < t) t = ti;
      }
    }
    if (rc == SFR_OK && t > 0.0 && (best == 0 || t < best_t)) {
      best = rpt;
      best_t = t;
    }
  }
  A->rpt_pcg = (rc == SFR_OK && best) ? best : keep;
  /* the trial graphs hold the temporary x: drop them before it goes, so
   * they neither occupy the caller's cache slots nor outlive their buffer */
  if (d_x) sfr_drop_graphs_keyed(ctx, 1, d_x);
  if (d_b) hipFree(d_b);
  if (d_x) hipFree(d_x);
  return rc;
}

static int solve_bicg64(sfr_context *ctx, sfr_matrix *A, double tol,
                        unsigned maxit, const double *d_b, double *d_x, int newX,
                        sfr_stats *st)
{
  CHECK(sfr_workspace(ctx, A->n));
  sfr_mat_view v = sfr_view(A);
  sfr_geom g = {A->rpt, 0, 0};
  chunks ch;
  CHECK(get_chunks(ctx, G_BICG, A, d_x, NULL, &v, g, &ch));
  double *const *w = ctx->w64;
  TRY(hipEventRecord(ctx->ev_start, ctx->stream));
  TRY(sfr_k_bicg_setup(ctx->stream, &v, g, d_b, d_x, w[0], w[1], w[2], w[3], newX, tol,
                       maxit, ctx->d_sc, ctx->part));
  CHECK(run_chunks(ctx, &ch, 0, maxit, 1));
  return finish(ctx, &v, g, d_b, d_x, st);
}

/* GMRES(SFR_GM_K), reference src/Iters.c:587-790 as src/Solv.c:302-372 runs
 * it: maxit counts restart cycles.  Statistics as the reference reports
 * them: iters = Arnoldi steps + 1 (0 when the first residual is exactly
 * zero), outer = cycles, norm_b = ||D^-1 b||, norm_r_m = |e_k| of the last
 * cycle (the quantity of the stop test). */
static int solve_gmres64(sfr_context *ctx, sfr_matrix *A, double tol,
                         unsigned maxit, const double *d_b, double *d_x,
                         int newX, sfr_stats *st)
{
  CHECK(sfr_workspace(ctx, A->n));
  CHECK(sfr_gm_workspace(ctx, A->n));
  sfr_mat_view v = sfr_view(A);
  sfr_geom g = {A->rpt, 0, 0};
  hipGraphExec_t ex;
  CHECK(get_graph(ctx, G_GM, A, d_x, d_b, &v, g, 0, &ex));
  TRY(hipEventRecord(ctx->ev_start, ctx->stream));
  TRY(sfr_k_gm_setup(ctx->stream, &v, g, d_b, d_x, newX, tol, maxit, ctx->d_sc,
                     ctx->part));
  /* A restart cycle of a large system outlasts the host's poll turnaround
   * many times over: one chunk in flight then costs less than a second
   * chunk's early-exit launches after convergence. */
  const unsigned cyc = (unsigned)gm_cycles_per_chunk(ctx);
  CHECK(run_loop_n(ctx, ex, 0, maxit / cyc + 2, 1, A->n >= (1 << 17) ? 1u : 2u));
  sfr_stats tmp;
  const int rc = finish(ctx, &v, g, d_b, d_x, &tmp);
  if (rc == SFR_HIP_ERROR) return rc;
  const sfr_scalars *sc = ctx->h_sc;
  tmp.iters = sc->iter ? sc->iter + 1 : 0;
  tmp.outer = sc->outer;
  tmp.norm_r_m = sc->res[4];
  if (st) *st = tmp;
  return rc;
}

/* BiCGSTAB(2), reference src/Bicgstab.c:1172-1378 as src/Solv.c:217-295
 * runs it; iters = iterations of two BiCG steps each (the reference's j+1). */
static int solve_bicgl2_64(sfr_context *ctx, sfr_matrix *A, double tol,
                           unsigned maxit, const double *d_b, double *d_x,
                           int newX, sfr_stats *st)
{
  CHECK(sfr_workspace(ctx, A->n));
  sfr_mat_view v = sfr_view(A);
  sfr_geom g = {A->rpt, 0, 0};
  chunks ch;
  CHECK(get_chunks(ctx, G_B2, A, d_x, NULL, &v, g, &ch));
  TRY(hipEventRecord(ctx->ev_start, ctx->stream));
  TRY(sfr_k_b2_setup(ctx->stream, &v, g, d_b, d_x, ctx->w64, newX, tol, maxit,
                     ctx->d_sc, ctx->part));
  CHECK(run_chunks(ctx, &ch, 0, maxit, 1));
  return finish(ctx, &v, g, d_b, d_x, st);
}

/* MINRES, reference src/Iters.c:309-555 as src/Solv.c:380-445 runs it.
 * iters = iterations performed (the reference's j+1 on convergence). */
static int solve_minres64(sfr_context *ctx, sfr_matrix *A, double tol,
                          unsigned maxit, const double *d_b, double *d_x,
                          int newX, sfr_stats *st)
{
  CHECK(sfr_workspace(ctx, A->n));
  sfr_mat_view v = sfr_view(A);
  sfr_geom g = {A->rpt, 0, 0};
  chunks ch;
  CHECK(get_chunks(ctx, G_MR, A, d_x, NULL, &v, g, &ch));
  MR_RINGS(ctx);
  TRY(hipEventRecord(ctx->ev_start, ctx->stream));
  TRY(sfr_k_mr_setup(ctx->stream, &v, g, d_b, d_x, V_, W_, Z_, newX, tol, maxit,
                     ctx->d_sc, ctx->part));
  CHECK(run_chunks(ctx, &ch, 0, maxit, 1));
  return finish(ctx, &v, g, d_b, d_x, st);
}

/* Mixed-precision BiCGSTAB of nb systems on A's pattern (device kernels in
 * bicgstab_mixed.hip): ad, b, x FP64 and interleaved per row (nb = 1:
 * plain), valnb = per-system FP64 values interleaved per entry or NULL for
 * A's own.  The loop runs captured chunks of the five FP32 kernels; after
 * each chunk the host looks at the systems' fold requests and enqueues the
 * fold kernels, which serve every system that asked, while the others go
 * on in the next chunk.  mb_plan captures the chunks before the solve
 * enqueues anything; mb_loop leaves the systems' scalars in d_sc. */
typedef struct mb_plan {
  sfr_mat_view v;
  sfr_geom     g;
  const float *vals32;
  chunks       ch;
} mb_plan;

static int mb_plan_make(sfr_context *ctx, sfr_matrix *A, int nb, const double *valnb,
                        mb_plan *pl)
{
  CHECK(sfr_mb_buffers(ctx, A->n * nb));
  pl->v = sfr_view(A);
  pl->v.valnb = valnb;
  pl->g = (sfr_geom){A->rpt, 0, 0};
  if (!valnb && !A->val32r)
    TRY(hipMalloc((void **)&A->val32r, sizeof(float) * (size_t)(A->nnz + 1)));
  pl->vals32 = valnb ? ctx->nb_val32 : A->val32r;
  return get_chunks(ctx, G_MB, A, pl->vals32, (const void *)(intptr_t)nb, &pl->v, pl->g,
                    &pl->ch);
}

static int mb_loop(sfr_context *ctx, sfr_matrix *A, const mb_plan *pl, int nb,
                   const double *ad, const double *b, double *x, double tol,
                   unsigned maxit, int newX)
{
  hipStream_t s = ctx->stream;
  if (pl->v.valnb)
    TRY(sfr_k_to_f32(s, A->nnz * nb, pl->v.valnb, ctx->nb_val32));
  else
    CHECK(sfr_matrix_prepare_f32(ctx, A));
  const chunks *ch = &pl->ch;
  float *w[7], *ad32, *dinv32;
  mb_split(ctx, A->n * nb, w, &ad32, &dinv32);
  TRY(sfr_k_mb_start(s, &pl->v, pl->g, nb, ad, b, x, ad32, dinv32, newX, tol, maxit,
                     ctx->ru_delta_b, ctx->d_sc, ctx->part));
  TRY(sfr_k_mb_fold(s, &pl->v, pl->g, nb, ad, b, x, w, ad32, dinv32, 1, newX, ctx->d_sc,
                    ctx->part));
  /* every fold is triggered by a residual drop of delta or a stop test:
   * bound the number of stretches generously */
  for (int folds = 0; folds < 100000; folds++) {
    sfr_scalars last[SFR_MAXSYS];
    CHECK(run_loop_ramp(ctx, ch->s, ch->cs, ch->b, ch->cb, folds ? 0 : ch->cb, 2,
                        maxit + 2 * ch->cb, nb, last, NULL));
    int asks = 0, running = 0;
    for (int c = 0; c < nb; c++) {
      if (last[c].done == SFR_DONE_RUN) running = 1;
      if (last[c].done == SFR_DONE_RUN && last[c].ru) asks = 1;
    }
    if (!running || !asks) break;
    /* the fold kernels queue behind the chunk still in flight and serve
     * every system whose request is set when they run */
    TRY(sfr_k_mb_fold(s, &pl->v, pl->g, nb, ad, b, x, w, ad32, dinv32, 0, 0, ctx->d_sc,
                      ctx->part));
  }
  return SFR_OK;
}

static int solve_bicg_mixed(sfr_context *ctx, sfr_matrix *A, double tol,
                            unsigned maxit, const double *d_b, double *d_x,
                            int newX, sfr_stats *st)
{
  CHECK(sfr_workspace(ctx, A->n));
  mb_plan pl;
  CHECK(mb_plan_make(ctx, A, 1, NULL, &pl));
  TRY(hipEventRecord(ctx->ev_start, ctx->stream));
  CHECK(mb_loop(ctx, A, &pl, 1, A->ad, d_b, d_x, tol, maxit, newX));
  sfr_stats s32;
  const int rc = finish(ctx, &pl.v, pl.g, d_b, d_x, &s32);
  if (rc != SFR_BREAKDOWN) {
    if (st) *st = s32;
    return rc;
  }
  /* the FP32 loop stagnated or broke down: FP64 from its last x (from zero
   * when that x is not finite) */
  sfr_stats s64;
  const int rc64 = solve_bicg64(ctx, A, tol, maxit, d_b, d_x, isfinite(s32.norm_x) ? 0 : 1,
                                &s64);
  s64.iters += s32.iters;
  s64.outer = s32.outer;
  s64.time_s += s32.time_s;
  if (st) *st = s64;
  return rc64;
}

/* Launch geometry of the mixed loop.  From 2^21 rows up it runs 2 rows per
 * thread with the grid capped at 1,024 workgroups, so that each workgroup
 * makes several passes.  The rows in flight then span a fraction of the
 * matrix, and the gathers find their neighbours' values in cache more often:
 * on the 128^3 Poisson matrix the fused kernel takes 22.9 us instead of
 * 24.5 us, and the solve 38.5 ms instead of 39.5 ms.  A rows-per-thread
 * choice made with sfr_matrix_tune is kept as given. */
static sfr_geom mixed_geom(const sfr_matrix *A)
{
  sfr_geom g = {A->rpt, 0, 0};
  if (!A->rpt_set && A->n >= ((int64_t)1 << 21)) {
    g.rpt = 2;
    g.cap = 1024;
  }
  return g;
}

static int solve_pcg_mixed(sfr_context *ctx, sfr_matrix *A, double tol,
                           unsigned maxit, const double *d_b, double *d_x,
                           int newX, sfr_stats *st)
{
  CHECK(sfr_workspace(ctx, A->n));
  CHECK(sfr_matrix_prepare_mixed(ctx, A));
  sfr_mat_view v = sfr_view(A);
  v.nt = sfr_loop_nt(ctx, A, 1);
  const sfr_geom g = mixed_geom(A);
  chunks ch;
  CHECK(get_chunks(ctx, G_MX, A, NULL, NULL, &v, g, &ch));
  float *const *w32 = ctx->w32;
  hipStream_t s = ctx->stream;
  TRY(hipEventRecord(ctx->ev_start, s));
  int nparts = 0;
  TRY(sfr_k_pcg_norm(s, &v, g, d_b, d_x, newX, ctx->part, &nparts));
  TRY(sfr_k_mx_start(s, ctx->part, nparts, tol, maxit, ctx->ru_delta, ctx->d_sc));
  /* the direction for the first iteration of a chunk goes to w32[4] */
  TRY(sfr_k_mx_fold(s, &v, g, d_b, d_x, w32[0], w32[1], w32[4], w32[4], 1, newX,
                    ctx->d_sc, ctx->part));
  /* every fold is triggered by a residual drop of delta (or the stop test):
   * bound the number of folds generously */
  for (int folds = 0; folds < 100000; folds++) {
    /* only the first stretch starts on short chunks: one that follows a
     * fold runs about log(delta) / log(rate) iterations */
    sfr_scalars last;
    CHECK(run_loop_ramp(ctx, ch.s, ch.cs, ch.b, ch.cb, folds ? 0 : ch.cb, 1,
                        maxit + 2 * ch.cb, 1, &last, NULL));
    if (last.done != SFR_DONE_RUN) break;
    if (!last.ru) break;                  /* loop stopped with no request */
    /* Once MA or MB requests a fold, every later MA / MB returns at once
     * and writes nothing but run4 / run5, so the chunk that saw the request
     * already holds the final pbuf.  The fold kernels queue behind the
     * chunk still in flight without the host waiting for it. */
    const float *p_old = last.pbuf ? w32[4] : w32[2];
    TRY(sfr_k_mx_fold(s, &v, g, d_b, d_x, w32[0], w32[1], p_old, w32[4], 0, 0,
                      ctx->d_sc, ctx->part));
  }
  return finish(ctx, &v, g, d_b, d_x, st);
}

/* The device matrix of the host-pointer paths for sys's pattern (the
 * identity of ia/ja plus the sizes and storage), from the context's small
 * LRU set.  A hit uploads the values al/ad/au (upload = 1) or leaves them
 * (upload = 0, the caller brings its own); a miss creates the matrix from
 * them, evicting the least recently used entry. */
static int cached_matrix(sfr_context *ctx, const sfr_system *sys, const double *al,
                         const double *ad, const double *au, int upload,
                         sfr_matrix **out)
{
  const uint64_t tick = ++ctx->mtick;
  for (int i = 0; i < SFR_MCACHE; i++) {
    sfr_matrix *A = ctx->mcache[i].A;
    if (A && A->h_ia == sys->ia && A->h_ja == sys->ja && A->n == sys->neq &&
        A->nad == sys->nad && A->storage == sys->storage && A->unsym == sys->unsym) {
      ctx->mcache[i].used = tick;
      ctx->mhits++;
      if (upload) CHECK(sfr_matrix_set_values(ctx, A, al, ad, au));
      *out = A;
      return SFR_OK;
    }
  }
  int slot = 0;
  for (int i = 0; i < SFR_MCACHE; i++) {
    if (!ctx->mcache[i].A) { slot = i; break; }
    if (ctx->mcache[i].used < ctx->mcache[slot].used) slot = i;
  }
  if (ctx->mcache[slot].A) sfr_matrix_destroy(ctx, ctx->mcache[slot].A);
  sfr_system s0 = *sys;
  s0.al = al;
  s0.ad = ad;
  s0.au = au;
  sfr_matrix *A = NULL;
  CHECK(sfr_matrix_create(ctx, &s0, SFR_FMT_AUTO, &A));
  ctx->mcache[slot].A = A;
  ctx->mcache[slot].used = tick;
  ctx->mcreated++;
  *out = A;
  return SFR_OK;
}

void sfr_mcache_clear(sfr_context *ctx)
{
  for (int i = 0; i < SFR_MCACHE; i++)
    if (ctx->mcache[i].A) sfr_matrix_destroy(ctx, ctx->mcache[i].A);
}

void sfr_matrix_cache_clear(sfr_context *ctx)
{
  if (ctx) sfr_mcache_clear(ctx);
}

int sfr_cache_info(const sfr_context *ctx, int *entries, int64_t *created,
                   int64_t *hits)
{
  if (!ctx) return SFR_BAD_ARG;
  int e = 0;
  for (int i = 0; i < SFR_MCACHE; i++) e += ctx->mcache[i].A != NULL;
  if (entries) *entries = e;
  if (created) *created = ctx->mcreated;
  if (hits) *hits = ctx->mhits;
  return SFR_OK;
}

/* The arithmetic the context's "precision" option asks for, where the
 * solver has a mixed-precision variant (PCG, PBICGSTAB); FP64 otherwise. */
static int ctx_precision(const sfr_context *ctx, int solver)
{
  return (ctx->precision == SFR_PREC_MIXED && (solver == SFR_PCG || solver == SFR_PBICGSTAB))
             ? SFR_PREC_MIXED : SFR_PREC_FP64;
}

/* Shared-pattern batch: interleave the system-major inputs, run the batched
 * BiCGSTAB, de-interleave x.  Workspace (interleaved, nsys*n each):
 * w0 x, w1 r, w2 r0, w3 p, w4 z, w5 v, w6 t, w7 b, w8 ad, w9 dinv. */
/* The batched BiCGSTAB of nsys (2..4) systems on A's pattern: d_ad, d_b, d_x
 * system-major (system c at c*n); valnb = the systems' own values
 * interleaved per device entry (species), or NULL for A's shared values
 * (u/v/w momentum). */
static int shared_run(sfr_context *ctx, sfr_matrix *A, int nsys,
                      const double *d_ad, const double *d_b, double *d_x,
                      const double *valnb, double tol, unsigned maxit,
                      int newX, int precision, sfr_stats *stats);

int sfr_solve_shared_device(sfr_context *ctx, sfr_matrix *A, int nsys,
                            const double *d_ad, const double *d_b, double *d_x,
                            int solver, double tol, unsigned maxit, int newX,
                            sfr_stats *stats)
{
  if (!ctx || !A || !d_b || !d_x || nsys < 1 || nsys > 4) return SFR_BAD_ARG;
  if (nsys == 1) {
    /* one system: the single-system path, with its own diagonal */
    if (d_ad) return sfr_set_error(ctx, SFR_NOT_SUPPORTED, "nsys = 1 with a separate diagonal");
    return sfr_solve_device(ctx, A, solver, ctx_precision(ctx, solver), tol, maxit, d_b, d_x,
                            newX, stats);
  }
  if (solver != SFR_PBICGSTAB)
    return sfr_set_error(ctx, SFR_NOT_SUPPORTED, "shared-pattern batch: solver %d", solver);
  if (maxit == 0) return sfr_set_error(ctx, SFR_BAD_ARG, "maxit = 0");
  TRY(hipSetDevice(ctx->device));
  return shared_run(ctx, A, nsys, d_ad, d_b, d_x, NULL, tol, maxit, newX,
                    ctx_precision(ctx, solver), stats);
}

static int shared_run(sfr_context *ctx, sfr_matrix *A, int nsys,
                      const double *d_ad, const double *d_b, double *d_x,
                      const double *valnb, double tol, unsigned maxit,
                      int newX, int precision, sfr_stats *stats)
{
  const int64_t n = A->n, tot = n * nsys;
  CHECK(sfr_workspace(ctx, tot));
  sfr_mat_view v = sfr_view(A);
  v.valnb = valnb;
  sfr_geom g = {A->rpt, 0, 0};
  double *const *w = ctx->w64;
  hipStream_t s = ctx->stream;
  const int mixed = precision == SFR_PREC_MIXED;
  chunks ch;
  mb_plan pl;
  if (mixed)
    CHECK(mb_plan_make(ctx, A, nsys, valnb, &pl));
  else
    CHECK(get_chunks(ctx, valnb ? G_BBV : G_BB, A, w[0], (const void *)(intptr_t)nsys, &v, g,
                     &ch));
  TRY(hipEventRecord(ctx->ev_start, s));
  if (d_ad) TRY(sfr_k_interleave(s, n, nsys, d_ad, w[8], w[9]));
  else for (int c = 0; c < nsys; c++) {
    /* the matrix's own diagonal for every system */
    TRY(hipMemcpyAsync(ctx->w64[1] + c * n, A->ad, sizeof(double) * (size_t)n,
                       hipMemcpyDeviceToDevice, s));
  }
  if (!d_ad) TRY(sfr_k_interleave(s, n, nsys, w[1], w[8], w[9]));
  TRY(sfr_k_interleave(s, n, nsys, d_b, w[7], NULL));
  if (!newX) TRY(sfr_k_interleave(s, n, nsys, d_x, w[0], NULL));
  if (mixed) {
    CHECK(mb_loop(ctx, A, &pl, nsys, w[8], w[7], w[0], tol, maxit, newX));
  } else {
    TRY(sfr_k_bb_setup(s, &v, g, nsys, w[8], w[9], w[7], w[0], w[1], w[2], w[3], w[4], newX,
                       tol, maxit, ctx->d_sc, ctx->part));
    CHECK(run_chunks(ctx, &ch, 0, maxit, nsys));
  }
  TRY(sfr_k_bb_final(s, &v, g, nsys, w[8], w[9], w[7], w[0], ctx->d_sc, ctx->part));
  TRY(sfr_k_deinterleave(s, n, nsys, w[0], d_x));
  TRY(hipEventRecord(ctx->ev_stop, s));
  TRY(hipMemcpyAsync(ctx->h_sc, ctx->d_sc, sizeof(sfr_scalars) * (size_t)nsys,
                     hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  float ms = 0.0f;
  TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  int worst = SFR_OK, broke = 0;
  sfr_stats all[SFR_BATCH_MAX];
  for (int c = 0; c < nsys; c++) {
    sfr_fill_stats(&ctx->h_sc[c], &all[c], ms);
    if (all[c].status != SFR_OK && worst == SFR_OK) worst = all[c].status;
    broke |= all[c].status == SFR_BREAKDOWN;
  }
  if (mixed && broke) {
    /* a system's FP32 loop stagnated or broke down: the batch again in
     * FP64 from its solutions (a converged system stops at once; a
     * non-finite x starts from zero) */
    for (int c = 0; c < nsys; c++)
      if (!isfinite(all[c].norm_x))
        TRY(sfr_k_fill(s, n, 0.0, d_x + (size_t)c * (size_t)n));
    sfr_stats s64[SFR_BATCH_MAX];
    worst = shared_run(ctx, A, nsys, d_ad, d_b, d_x, valnb, tol, maxit, 0,
                       SFR_PREC_FP64, s64);
    for (int c = 0; c < nsys; c++) {
      s64[c].iters += all[c].iters;
      s64[c].outer = all[c].outer;
      s64[c].time_s += all[c].time_s;
      all[c] = s64[c];
    }
  }
  if (stats)
    for (int c = 0; c < nsys; c++) stats[c] = all[c];
  return worst;
}

/* Species batch, host arrays: the reference's per-species solverC loop
 * (src/Combustion.c:76-92, 222-240) on one pattern.  BiCGSTAB systems run
 * in groups of 2-4 through the batched kernels with each system's values
 * interleaved per device entry; another solver id, storage CSR (diagonal
 * inside al) or a single system are solved one after another. */
int sfr_solve_batched(sfr_context *ctx, const sfr_system *sys, int nsys,
                      const double *const *al, const double *const *ad,
                      const double *const *au, const double *const *b,
                      double *const *x, int solver, double tol, unsigned maxit,
                      int newX, sfr_stats *stats)
{
  if (!ctx || !sys || !al || !ad || !b || !x || nsys < 1 || nsys > SFR_BATCH_MAX)
    return SFR_BAD_ARG;
  if (maxit == 0) return sfr_set_error(ctx, SFR_BAD_ARG, "maxit = 0");
  TRY(hipSetDevice(ctx->device));
  if (solver != SFR_PBICGSTAB || sys->storage == SFR_CSR || nsys == 1) {
    int worst = SFR_OK;
    for (int c = 0; c < nsys; c++) {
      sfr_system sc = *sys;
      sc.al = al[c];
      sc.ad = ad[c];
      sc.au = au ? au[c] : NULL;
      sfr_stats st;
      const int rc = sfr_solve(ctx, &sc, solver, tol, maxit, b[c], x[c], newX, &st);
      if (rc == SFR_HIP_ERROR || rc == SFR_BAD_ARG || rc == SFR_NOT_SUPPORTED) return rc;
      if (stats) stats[c] = st;
      if (rc != SFR_OK && worst == SFR_OK) worst = rc;
    }
    return worst;
  }
  /* each group uploads its own values below: a cached pattern is taken as
   * it is */
  sfr_matrix *A = NULL;
  CHECK(cached_matrix(ctx, sys, al[0], ad[0], au ? au[0] : NULL, 0, &A));
  CHECK(sfr_nb_buffers(ctx, A->nnz, A->nraw));
  const int64_t n = A->n, tot = 4 * n;
  if (ctx->stage_n < 3 * tot) {
    TRY(hipStreamSynchronize(ctx->stream));
    hipFree(ctx->stage);
    ctx->stage = NULL;
    ctx->stage_n = 0;
    TRY(hipMalloc((void **)&ctx->stage, sizeof(double) * (size_t)(3 * tot)));
    ctx->stage_n = 3 * tot;
  }
  double *d_ad = ctx->stage, *d_b = ctx->stage + tot, *d_x = ctx->stage + 2 * tot;
  const size_t bytes = sizeof(double) * (size_t)n;
  const int csrc_unsym = (A->storage == SFR_CSRC || A->storage == SFR_CSRCCOO) && A->unsym;
  hipStream_t s = ctx->stream;
  /* groups of 2..4 systems, as even as possible */
  const int ng = (nsys + 3) / 4, base = nsys / ng, rem = nsys % ng;
  int worst = SFR_OK;
  for (int gi = 0, c0 = 0; gi < ng; gi++) {
    const int nb = base + (gi < rem ? 1 : 0);
    for (int j = 0; j < nb; j++) {
      const int c = c0 + j;
      if (csrc_unsym) {
        TRY(hipMemcpyAsync(ctx->nb_raw, al[c], sizeof(double) * (size_t)A->nad,
                           hipMemcpyHostToDevice, s));
        TRY(hipMemcpyAsync(ctx->nb_raw + A->nad, (au && au[c]) ? au[c] : al[c] + A->nad,
                           sizeof(double) * (size_t)A->nad, hipMemcpyHostToDevice, s));
      } else if (A->nraw > 0) {
        TRY(hipMemcpyAsync(ctx->nb_raw, al[c], sizeof(double) * (size_t)A->nraw,
                           hipMemcpyHostToDevice, s));
      }
      TRY(sfr_k_gather_nb(s, A->nnz, A->vmap, ctx->nb_raw, ctx->nb_val, nb, j));
      TRY(hipMemcpyAsync(d_ad + j * n, ad[c], bytes, hipMemcpyHostToDevice, s));
      TRY(hipMemcpyAsync(d_b + j * n, b[c], bytes, hipMemcpyHostToDevice, s));
      if (!newX) TRY(hipMemcpyAsync(d_x + j * n, x[c], bytes, hipMemcpyHostToDevice, s));
    }
    const int rc = shared_run(ctx, A, nb, d_ad, d_b, d_x, ctx->nb_val, tol, maxit, newX,
                              ctx_precision(ctx, solver), stats ? stats + c0 : NULL);
    if (rc == SFR_HIP_ERROR || rc == SFR_BAD_ARG || rc == SFR_NOT_SUPPORTED) return rc;
    if (rc != SFR_OK && worst == SFR_OK) worst = rc;
    for (int j = 0; j < nb; j++)
      TRY(hipMemcpyAsync(x[c0 + j], d_x + j * n, bytes, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    c0 += nb;
  }
  return worst;
}

/* Host arrays: the reference's u/v/w call pattern (src/Simple.c:2980-3036) */
int sfr_solve_shared(sfr_context *ctx, const sfr_system *sys, int nsys,
                     const double *const *ad, const double *const *b,
                     double *const *x, int solver, double tol, unsigned maxit,
                     int newX, sfr_stats *stats)
{
  if (!ctx || !sys || !b || !x || nsys < 1 || nsys > 4) return SFR_BAD_ARG;
  TRY(hipSetDevice(ctx->device));
  sfr_matrix *A = NULL;
  CHECK(cached_matrix(ctx, sys, sys->al, ad ? ad[0] : sys->ad, sys->au, 1, &A));
  const int64_t n = A->n, tot = n * nsys;
  if (ctx->stage_n < 3 * tot) {
    TRY(hipStreamSynchronize(ctx->stream));
    hipFree(ctx->stage);
    ctx->stage = NULL;
    ctx->stage_n = 0;
    TRY(hipMalloc((void **)&ctx->stage, sizeof(double) * (size_t)(3 * tot)));
    ctx->stage_n = 3 * tot;
  }
  double *d_ad = ctx->stage, *d_b = ctx->stage + tot, *d_x = ctx->stage + 2 * tot;
  const size_t bytes = sizeof(double) * (size_t)n;
  for (int c = 0; c < nsys; c++) {
    if (ad) TRY(hipMemcpyAsync(d_ad + c * n, ad[c], bytes, hipMemcpyHostToDevice, ctx->stream));
    TRY(hipMemcpyAsync(d_b + c * n, b[c], bytes, hipMemcpyHostToDevice, ctx->stream));
    if (!newX) TRY(hipMemcpyAsync(d_x + c * n, x[c], bytes, hipMemcpyHostToDevice, ctx->stream));
  }
  int rc;
  if (nsys == 1)
    rc = sfr_solve_device(ctx, A, solver, ctx_precision(ctx, solver), tol, maxit, d_b, d_x,
                          newX, stats);
  else
    rc = sfr_solve_shared_device(ctx, A, nsys, ad ? d_ad : NULL, d_b, d_x, solver, tol, maxit,
                                 newX, stats);
  if (rc == SFR_HIP_ERROR || rc == SFR_BAD_ARG || rc == SFR_NOT_SUPPORTED) return rc;
  for (int c = 0; c < nsys; c++)
    TRY(hipMemcpyAsync(x[c], d_x + c * n, bytes, hipMemcpyDeviceToHost, ctx->stream));
  TRY(hipStreamSynchronize(ctx->stream));
  return rc;
}

int sfr_solve_device(sfr_context *ctx, sfr_matrix *A, int solver,
                     int precision, double tol, unsigned maxit,
                     const double *d_b, double *d_x, int newX, sfr_stats *st)
{
  if (!ctx || !A || !d_b || !d_x) return SFR_BAD_ARG;
  if (maxit == 0) return sfr_set_error(ctx, SFR_BAD_ARG, "maxit = 0");
  TRY(hipSetDevice(ctx->device));
  if (solver == SFR_PCG)
    return precision == SFR_PREC_MIXED
               ? solve_pcg_mixed(ctx, A, tol, maxit, d_b, d_x, newX, st)
               : solve_pcg64(ctx, A, tol, maxit, d_b, d_x, newX, st);
  if (solver == SFR_PBICGSTAB)
    return precision == SFR_PREC_MIXED
               ? solve_bicg_mixed(ctx, A, tol, maxit, d_b, d_x, newX, st)
               : solve_bicg64(ctx, A, tol, maxit, d_b, d_x, newX, st);
  if (solver == SFR_PBICGSTABL2) {
    if (precision == SFR_PREC_MIXED)
      return sfr_set_error(ctx, SFR_NOT_SUPPORTED, "mixed BiCGSTAB(2)");
    return solve_bicgl2_64(ctx, A, tol, maxit, d_b, d_x, newX, st);
  }
  if (solver == SFR_GMRES) {
    if (precision == SFR_PREC_MIXED)
      return sfr_set_error(ctx, SFR_NOT_SUPPORTED, "mixed GMRES");
    return solve_gmres64(ctx, A, tol, maxit, d_b, d_x, newX, st);
  }
  if (solver == SFR_MINRES) {
    if (precision == SFR_PREC_MIXED)
      return sfr_set_error(ctx, SFR_NOT_SUPPORTED, "mixed MINRES");
    return solve_minres64(ctx, A, tol, maxit, d_b, d_x, newX, st);
  }
  return sfr_set_error(ctx, SFR_BAD_ARG, "solver id %d", solver);
}

/* Work vectors of the allocation-free solve: PCG r, p, q; BiCGSTAB r, r0,
 * p, z, v, t; each padded to 256 bytes in the static scratch. */
static int noalloc_vectors(int solver)
{
  return solver == SFR_PCG ? 3 : 6;
}

int sfr_solve_device_noalloc(sfr_context *ctx, sfr_matrix *A, int solver,
                             double tol, unsigned maxit, const double *d_b,
                             double *d_x, int newX, sfr_stats *st)
{
  if (!ctx || !A || !d_b || !d_x) return SFR_BAD_ARG;
  if (solver != SFR_PCG && solver != SFR_PBICGSTAB)
    return sfr_set_error(ctx, SFR_NOT_SUPPORTED,
                         "allocation-free solve: solver %d (PCG and PBICGSTAB only)", solver);
  if (maxit == 0) return sfr_set_error(ctx, SFR_BAD_ARG, "maxit = 0");
  if (!ctx->scratch) return sfr_set_error(ctx, SFR_HIP_ERROR, "static scratch unavailable");
  const int nvec = noalloc_vectors(solver);
  const int64_t ld = ((int64_t)A->n + 31) & ~(int64_t)31;
  const int64_t room = (int64_t)((SFR_SCRATCH_BYTES - SFR_SCRATCH_VEC) / sizeof(double));
  if (nvec * ld > room)
    return sfr_set_error(ctx, SFR_BAD_ARG,
                         "allocation-free solve: %d rows exceed the %zu MiB static scratch "
                         "(at most %lld rows for solver %d)", (int)A->n,
                         SFR_SCRATCH_BYTES >> 20, (long long)((room / nvec) & ~(int64_t)31),
                         solver);
  TRY(hipSetDevice(ctx->device));
  unsigned char *base = (unsigned char *)ctx->scratch;
  direct_loop d;
  memset(&d, 0, sizeof d);
  d.solver = solver;
  d.v = sfr_view(A);
  if (solver == SFR_PCG) d.v.nt = sfr_loop_nt(ctx, A, 0);
  d.g = (sfr_geom){solver == SFR_PCG ? sfr_pcg_rpt(A) : A->rpt, 0, 0};
  d.x = d_x;
  d.sc = (sfr_scalars *)base;
  d.part = (double *)(base + SFR_SCRATCH_PART);
  for (int k = 0; k < nvec; k++) d.w[k] = (double *)(base + SFR_SCRATCH_VEC) + k * ld;
  hipStream_t s = ctx->stream;
  /* the scratch keeps the previous call's scalars: start from zero */
  TRY(sfr_k_fill(s, (int64_t)((sizeof(sfr_scalars) + 7) / 8), 0.0, (double *)d.sc));
  TRY(hipEventRecord(ctx->ev_start, s));
  if (solver == SFR_PCG)
    TRY(sfr_k_pcg_setup(s, &d.v, d.g, d_b, d_x, d.w[0], d.w[1], newX, tol, maxit, d.sc,
                        d.part));
  else
    TRY(sfr_k_bicg_setup(s, &d.v, d.g, d_b, d_x, d.w[0], d.w[1], d.w[2], d.w[3], newX, tol,
                         maxit, d.sc, d.part));
  const int kind = solver == SFR_PCG ? G_PCG : G_BICG;
  const unsigned cs = (unsigned)chunk_of(ctx, kind, 1), cb = (unsigned)chunk_of(ctx, kind, 0);
  CHECK(run_loop_ramp(ctx, NULL, cs, NULL, cb, cb, 0, maxit + 2 * cb, 1, NULL, &d));
  return finish_at(ctx, &d.v, d.g, d_b, d_x, d.sc, d.part, st);
}

/* Average duration of one launch of a loop kernel, timed with HIP events on
 * the context stream: which = 0 FP64 PCG SpMV (+p.q), 1 FP32 inner SpMV. */
int sfr_kernel_time(sfr_context *ctx, sfr_matrix *A, int which, int reps,
                    double *ms_per_launch)
{
  if (!ctx || !A || reps <= 0 || !ms_per_launch) return SFR_BAD_ARG;
  TRY(hipSetDevice(ctx->device));
  CHECK(sfr_workspace(ctx, A->n));
  if (which == 1) CHECK(sfr_matrix_prepare_mixed(ctx, A));
  if (which == 0) CHECK(sfr_matrix_pair_check(ctx, A));
  sfr_mat_view v = sfr_view(A);
  v.nt = sfr_loop_nt(ctx, A, which == 1);
  sfr_geom g = {which == 0 ? sfr_pcg_rpt(A) : A->rpt, 0, 0};
  if (which == 1) g = mixed_geom(A);
  hipStream_t s = ctx->stream;
  float *const *w32 = ctx->w32;
  if (which == 0) {
    TRY(hipMemsetAsync(ctx->d_sc, 0, sizeof(sfr_scalars), s));
    TRY(sfr_k_fill(s, A->n, 1.0, ctx->w64[1]));
  } else {
    /* the fused FP32 kernel in its steady state: running, not fresh, no
     * fold due (delta 0, tol 0), alpha 0 (y unchanged), r.r partials 1 */
    sfr_scalars st;
    memset(&st, 0, sizeof st);
    st.run5 = 1;
    st.dd[0] = st.dd[1] = 1.0;
    st.maxit = 0x7fffffffu;
    TRY(hipMemcpyAsync(ctx->d_sc, &st, sizeof st, hipMemcpyHostToDevice, s));
    TRY(sfr_k_fill(s, 4 * SFR_MAXG, 1.0, ctx->part + 4 * SFR_MAXG));
    TRY(sfr_k_fill32(s, A->n, 1.0f, w32[1]));
    TRY(sfr_k_fill32(s, A->n, 1.0f, w32[2]));
    TRY(sfr_k_fill32(s, A->n, 0.0f, w32[0]));
  }
#define LAUNCH()                                                               \
  do {                                                                         \
    if (which == 0)                                                            \
      TRY(sfr_k_pcg_spmv_only(s, &v, g, ctx->w64[1], ctx->w64[2], ctx->d_sc,   \
                              ctx->part));                                     \
    else                                                                       \
      TRY(sfr_k_mx_a_only(s, &v, g, w32[1], w32[2], w32[4], w32[3], w32[0],    \
                          ctx->d_sc, ctx->part));                              \
  } while (0)
  for (int warm = 0; warm < 2; warm++) LAUNCH();
  TRY(hipEventRecord(ctx->ev_start, s));
  for (int r = 0; r < reps; r++) LAUNCH();
#undef LAUNCH
  TRY(hipEventRecord(ctx->ev_stop, s));
  TRY(hipEventSynchronize(ctx->ev_stop));
  float ms = 0.0f;
  TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  *ms_per_launch = ms / reps;
  return SFR_OK;
}

int sfr_copy_time(sfr_context *ctx, const double *d_src, double *d_dst,
                  int64_t n, int reps, double *ms_per_copy)
{
  if (!ctx || !d_src || !d_dst || n <= 0 || reps <= 0 || !ms_per_copy)
    return SFR_BAD_ARG;
  if ((((uintptr_t)d_src | (uintptr_t)d_dst) & 15) != 0)
    return sfr_set_error(ctx, SFR_BAD_ARG, "sfr_copy_time: buffers not 16-byte aligned");
  const uintptr_t a = (uintptr_t)d_src, b = (uintptr_t)d_dst,
                  len = (uintptr_t)n * sizeof(double);
  if (a < b + len && b < a + len)
    return sfr_set_error(ctx, SFR_BAD_ARG, "sfr_copy_time: buffers overlap");
  TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  TRY(sfr_k_stream_copy(s, n, d_src, d_dst));
  TRY(hipEventRecord(ctx->ev_start, s));
  for (int r = 0; r < reps; r++) TRY(sfr_k_stream_copy(s, n, d_src, d_dst));
  TRY(hipEventRecord(ctx->ev_stop, s));
  TRY(hipEventSynchronize(ctx->ev_stop));
  float ms = 0.0f;
  TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  *ms_per_copy = ms / reps;
  return SFR_OK;
}

/* Adjoint of x = A^-1 b: A^T lambda = g, then the gradients in the caller's
 * raw storage order.  solver PCG: A is symmetric (A^T = A); PBICGSTAB:
 * BiCGSTAB on the transposed operator. */
int sfr_solve_adjoint(sfr_context *ctx, sfr_matrix *A, int solver, double tol,
                      unsigned maxit, const double *d_x, const double *d_g,
                      double *d_lambda, double *d_grad_al, double *d_grad_ad,
                      sfr_stats *st)
{
  if (!ctx || !A || !d_x || !d_g || !d_lambda) return SFR_BAD_ARG;
  if (solver != SFR_PCG && solver != SFR_PBICGSTAB)
    return sfr_set_error(ctx, SFR_BAD_ARG, "solver id %d", solver);
  TRY(hipSetDevice(ctx->device));
  CHECK(sfr_matrix_prepare_adjoint(ctx, A));
  const int rc = solver == SFR_PCG
      ? sfr_solve_device(ctx, A, SFR_PCG, SFR_PREC_FP64, tol, maxit, d_g, d_lambda, 1, st)
      : sfr_solve_device(ctx, A->T, SFR_PBICGSTAB, SFR_PREC_FP64, tol, maxit, d_g, d_lambda, 1, st);
  if (rc == SFR_HIP_ERROR || rc == SFR_BAD_ARG || rc == SFR_NOT_SUPPORTED) return rc;
  TRY(sfr_k_adj_grad(ctx->stream, d_grad_al ? A->nraw : 0, A->grad_rc, d_lambda, d_x,
                     d_grad_al, d_grad_ad ? A->n : 0, d_grad_ad));
  TRY(hipStreamSynchronize(ctx->stream));
  return rc;
}

/* Host-pointer solve: each pattern is uploaded once and recognised again by
 * the identity of ia/ja (SFR3D re-assembles values on fixed patterns and
 * alternates several of them per SIMPLE step). */
int sfr_solve(sfr_context *ctx, const sfr_system *sys, int solver, double tol,
              unsigned maxit, const double *b, double *x, int newX,
              sfr_stats *st)
{
  if (!ctx || !sys || !b || !x) return SFR_BAD_ARG;
  TRY(hipSetDevice(ctx->device));
  sfr_matrix *A = NULL;
  CHECK(cached_matrix(ctx, sys, sys->al, sys->ad, sys->au, 1, &A));
  CHECK(sfr_workspace(ctx, A->n));
  double *d_b = ctx->w64[8], *d_x = ctx->w64[9];
  const size_t bytes = sizeof(double) * (size_t)A->n;
  TRY(hipMemcpyAsync(d_b, b, bytes, hipMemcpyHostToDevice, ctx->stream));
  if (!newX) TRY(hipMemcpyAsync(d_x, x, bytes, hipMemcpyHostToDevice, ctx->stream));
  const int rc = sfr_solve_device(ctx, A, solver, ctx_precision(ctx, solver), tol, maxit,
                                  d_b, d_x, newX, st);
  if (rc == SFR_HIP_ERROR || rc == SFR_BAD_ARG || rc == SFR_NOT_SUPPORTED) return rc;
  TRY(hipMemcpyAsync(x, d_x, bytes, hipMemcpyDeviceToHost, ctx->stream));
  TRY(hipStreamSynchronize(ctx->stream));
  return rc;
}

/* ---- solverC drop-in -------------------------------------------------------- */

static sfr_context *g_ctx;

int sfr_solverC_context(sfr_context **ctx)
{
  if (!ctx) return SFR_BAD_ARG;
  if (!g_ctx) {
    const char *dev = getenv("SFR_DEVICE");
    const int rc = sfr_init(&g_ctx, dev ? atoi(dev) : 0);
    if (rc != SFR_OK) {
      *ctx = NULL;
      return rc;
    }
  }
  *ctx = g_ctx;
  return SFR_OK;
}

void sfr_solverC(void *m, int nEq, int nEqNov, int nAd, int nAdR, int *ia,
                 int *ja, double *al, double *ad, double *au, double *b,
                 double *x, void *iNeq, void *bOmp, double tol,
                 unsigned int maxIt, short storage, short solver,
                 FILE *fSolvLog, bool fLog, bool newX, bool unSym,
                 const char *name)
{
  (void)m; (void)iNeq; (void)bOmp;
  if (solver != SFR_PCG && solver != SFR_PBICGSTAB && solver != SFR_PBICGSTABL2 &&
      solver != SFR_GMRES && solver != SFR_MINRES) {
    fprintf(stderr, "Opcao %d e invalida!!\n", solver);
    fprintf(stderr, "Arquivo:%s\nFonte:  %s\n", __FILE__, __func__);
    exit(EXIT_FAILURE);
  }
  if (storage < SFR_CSR || storage > SFR_CSRCCOO) {
    fprintf(stderr, "Opcao %d e invalida!!\n", storage);
    fprintf(stderr, "Arquivo:%s\nFonte:  %s\n", __FILE__, __func__);
    exit(EXIT_FAILURE);
  }
  if (nAdR != 0 || nEqNov != nEq) {
    fprintf(stderr, "sfr_solverC: distributed systems (nAdR=%d, nEqNov=%d != nEq=%d)"
            " are not supported\n", nAdR, nEqNov, nEq);
    exit(EXIT_FAILURE);
  }
  if (storage == SFR_CSR) {
    /* The reference's storage CSR does not solve the assembled system:
     * dataStruct aliases ad to al (src/Datastruct.c:85), so Jacobi is built
     * from the first neq values of al, and matVecCsrSym zeroes y[i] after
     * earlier rows have scattered into it (src/HccaBlas.c:3315-3325), which
     * drops the lower triangle.  This build computes the operator the
     * storage represents (sfr_solve accepts it), so its answer would differ
     * from solverC's: the drop-in refuses rather than return it silently. */
    fprintf(stderr, "sfr_solverC: storage CSR (%d) is refused: the reference's solverC "
            "builds Jacobi from al (ad aliases al) and its symmetric product drops the "
            "lower triangle; use CSRD, CSRC or ELLPACK, or sfr_solve for the exact "
            "operator\n", storage);
    exit(EXIT_FAILURE);
  }
  {
    sfr_context *c = NULL;
    const int rc = sfr_solverC_context(&c);
    if (rc != SFR_OK) {
      fprintf(stderr, "sfr_solverC: %s\n", sfr_status_string(rc));
      exit(EXIT_FAILURE);
    }
  }
  const sfr_system sys = {nEq, nAd, ia, ja, al, ad, au, storage, (short)unSym};
  sfr_stats st;
  const int rc = sfr_solve(g_ctx, &sys, solver, tol, maxIt, b, x, newX, &st);
  if (rc == SFR_HIP_ERROR || rc == SFR_BAD_ARG || rc == SFR_NOT_SUPPORTED ||
      rc == SFR_NO_DEVICE) {
    fprintf(stderr, "sfr_solverC: %s: %s\n", sfr_status_string(rc), sfr_last_error(g_ctx));
    exit(EXIT_FAILURE);
  }
  if (rc != SFR_OK && solver == SFR_MINRES) { /* src/Iters.c:540-544: no exit */
    printf("PMINRES: *** WARNING: no convergende reached !\n");
    printf("MAXIT = %d \n", maxIt);
  }
  if (rc != SFR_OK && solver == SFR_PBICGSTABL2) {  /* src/Bicgstab.c:1362-1366 */
    printf(" PBICGSTAB(2) *** WARNING: no convergende reached !\n");
    printf("MAXIT = %d \n", maxIt);
    exit(EXIT_FAILURE);
  }
  if (rc != SFR_OK && solver == SFR_GMRES) {  /* src/Iters.c:770-774 */
    printf(" GMRES *** WARNING: no convergende reached !\n");
    printf("MAXIT = %d \n", maxIt);
    exit(EXIT_FAILURE);
  }
  if (rc != SFR_OK && solver != SFR_MINRES) { /* j == maxIt (and breakdowns) */
    printf("name : %s\n", name ? name : "");
    if (solver == SFR_PCG) {
      printf("PCG: *** WARNING: no convergende reached !\n");
      printf("MAXIT = %d \n", maxIt);
      exit(EXIT_FAILURE);
    }
    printf(" PBICGSTAB *** WARNING: no convergende reached !\n");
    printf("MAXIT = %d \n", maxIt);
    exit(-2);                          /* EXIT_SOLVER */
  }
  if (fLog && fSolvLog) {
    if (solver == SFR_PCG)
      fprintf(fSolvLog, "PCG: tol %20.2e iteration %d xKx %20.12e norma(x*x) %20.12e "
              "time %20.5lf\n", tol, (int)st.iters, st.xKx, st.norm_x, st.time_s);
    else if (solver == SFR_PBICGSTABL2)    /* src/Bicgstab.c:1368-1376 */
      fprintf(fSolvLog, "PBICGSTAB(2): tol %20.2e iteration %d xKx %20.12e "
              "norma(x*x) %20.12e time %20.5lf\n", tol, (int)st.iters, st.xKx,
              st.norm_x, st.time_s);
    else if (solver == SFR_MINRES)         /* src/Iters.c:545-552; j+1 at maxIt */
      fprintf(fSolvLog, "PMINRES: tol %20.2e iteration %d xKx %20.12e norma(x*x) %20.12e "
              "time %20.5lf\n", tol, (int)st.iters + (rc != SFR_OK ? 1 : 0), st.xKx,
              st.norm_x, st.time_s);
    else if (solver == SFR_GMRES)          /* src/Iters.c:778-786 */
      fprintf(fSolvLog, "GMRES: tol %20.2e  Cycles %d iteration %d  nKrylov %d  xKx %20.12e "
              " norma(x*x) %20.12e  time %20.5lf\n", tol, (int)st.outer, (int)st.iters,
              SFR_GM_K, st.xKx, st.norm_x, st.time_s);
    else
      fprintf(fSolvLog, "PBICGSTAB: tol %20.2e iteration %d xKx %20.12e ||x*x|| %20.12e "
              "time %20.5lf\n", tol, (int)st.iters, st.xKx, st.norm_x, st.time_s);
  }
}

int sfr_set_option(sfr_context *ctx, const char *name, double value)
{
  if (!ctx || !name) return SFR_BAD_ARG;
  if (!strcmp(name, "chunk")) {
    if (value < 1 || value > 256) return sfr_set_error(ctx, SFR_BAD_ARG, "chunk %g", value);
    ctx->chunk = (int)value;
    return SFR_OK;
  }
  if (!strcmp(name, "precision")) {
    if (value != SFR_PREC_FP64 && value != SFR_PREC_MIXED)
      return sfr_set_error(ctx, SFR_BAD_ARG, "precision %g", value);
    ctx->precision = (int)value;
    return SFR_OK;
  }
  if (!strcmp(name, "ru_delta")) {
    if (!(value > 0 && value < 1)) return sfr_set_error(ctx, SFR_BAD_ARG, "ru_delta %g", value);
    ctx->ru_delta = value;
    return SFR_OK;
  }
  if (!strcmp(name, "nt_loads")) {
    if (value != -1 && value != 0 && value != 1)
      return sfr_set_error(ctx, SFR_BAD_ARG, "nt_loads %g", value);
    ctx->nt_loads = (int)value;
    return SFR_OK;
  }
  if (!strcmp(name, "ell_pairs")) {
    if (value != 0 && value != 1) return sfr_set_error(ctx, SFR_BAD_ARG, "ell_pairs %g", value);
    ctx->ell_pairs = (int)value;
    return SFR_OK;
  }
  if (!strcmp(name, "ru_delta_bicg")) {
    if (!(value > 0 && value < 1))
      return sfr_set_error(ctx, SFR_BAD_ARG, "ru_delta_bicg %g", value);
    ctx->ru_delta_b = value;
    return SFR_OK;
  }
  return sfr_set_error(ctx, SFR_BAD_ARG, "unknown option %s", name);
}
