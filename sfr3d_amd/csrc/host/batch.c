/*
 * Uneven batch of independent systems: set-up, autotuner and solve driver
 * for the kernels of device/mixed_batch.hip.
 *
 * Each system is an ordinary device matrix (any storage, ELL or CSR device
 * format) with its FP32 Jacobi-scaled copy; the batch adds a descriptor
 * table, concatenated vectors (each system's rows start on a 64-row
 * boundary), per-system scalars and partial-sum slots, and the active list.
 * A solve is: active list of every system, two set-up kernels, then a
 * captured graph of `chunk` iterations plus a compaction of the active list,
 * replayed until no system is left running; then the final FP64 quantities
 * of every system.  The host polls one 8-byte copy (systems still running)
 * per chunk and never touches a system's state inside the loop.
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

#define NBUCKET  32
#define ROWALIGN 64
#define TUNE_SAMPLES 4
#define TUNE_REPS 8
#ifndef TUNE_ROWS
#define TUNE_ROWS ((int64_t)4 << 20)   /* rows a tuning trial fills          */
#endif
#define TUNE_MAXRPT 16

struct sfr_batch {
  int          nsys;
  sfr_matrix **A;
  sfr_ub_sys  *h_sys, *d_sys;
  int64_t      rows;          /* padded total rows of the batch vectors     */
  int64_t      items;         /* total work items (sum of g)                */
  int          grid;
  int         *d_act;
  double      *x, *b;
  float       *y, *r, *p, *q;
  sfr_scalars *d_sc, *h_sc;
  unsigned long long *d_ts, *h_ts;
  int         *h_poll;        /* pinned [2][2]: active systems, items       */
  double      *partA, *partB;
  hipGraphExec_t graph;
  int          graph_chunk;
  int          clock_khz;
  int          nbucket;
  sfr_bucket_choice bucket[NBUCKET];
};

static int64_t align_rows(int64_t n)
{
  return (n + ROWALIGN - 1) / ROWALIGN * ROWALIGN;
}

static int items_for(int64_t n, int rpt)
{
  const int64_t r = (int64_t)rpt * SFR_BLOCK;
  return (int)((n + r - 1) / r);
}

static int bucket_of(int32_t n)
{
  int k = 0;
  while (k < NBUCKET - 1 && ((int64_t)1 << (k + 1)) <= n) k++;
  return k;
}

/* rows per thread without measurement: the smallest that keeps a system at
 * or below 256 work items (each consumer reduces all of a system's
 * partials, so the redundant reduction stays small next to the rows) */
static int rpt_rule(int32_t n)
{
  int rpt = 1;
  while (rpt < 8 && items_for(n, rpt) > 256) rpt *= 2;
  return rpt;
}

static void describe(sfr_ub_sys *d, const sfr_matrix *A, int64_t row0,
                     int rpt, int64_t gfirst)
{
  memset(d, 0, sizeof *d);
  d->col = A->col;
  d->col16 = A->fmt == SFR_FMT_ELL ? A->col16 : NULL;
  d->exc_k = A->exc_k;
  d->exc_col = A->exc_col;
  d->nexc = A->nexc;
  d->rowptr = A->fmt == SFR_FMT_CSR ? A->rowptr : NULL;
  d->val = A->val; d->ad = A->ad; d->dinv = A->dinv;
  d->dscale = A->dscale; d->val32 = A->val32;
  d->row0 = row0;
  d->n = A->n; d->slots = A->slots; d->fmt = A->fmt; d->rpt = rpt;
  d->g = items_for(A->n, rpt);
  d->gfirst = (int32_t)gfirst;
}

/* device buffers of a (possibly temporary) batch of nsys systems */
typedef struct {
  sfr_ub_sys  *d_sys;
  int         *act;
  double      *x, *b;
  float       *y, *r, *p, *q;
  sfr_scalars *sc;
  unsigned long long *ts;
  double      *partA, *partB;
} ub_bufs;

static void free_bufs(ub_bufs *u)
{
  hipFree(u->d_sys); hipFree(u->act); hipFree(u->x); hipFree(u->b);
  hipFree(u->y); hipFree(u->r); hipFree(u->p); hipFree(u->q);
  hipFree(u->sc); hipFree(u->ts); hipFree(u->partA); hipFree(u->partB);
  memset(u, 0, sizeof *u);
}

static int alloc_bufs(sfr_context *ctx, ub_bufs *u, int nsys, int64_t rows,
                      int64_t items)
{
  memset(u, 0, sizeof *u);
  const size_t nr = (size_t)(rows > 0 ? rows : 1);
  const size_t np = (size_t)(4 * (items > 0 ? items : 1));
#define BTRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { free_bufs(u); \
    return sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s", #call, hipGetErrorString(e_)); } } while (0)
  BTRY(hipMalloc((void **)&u->d_sys, sizeof(sfr_ub_sys) * (size_t)nsys));
  BTRY(hipMalloc((void **)&u->act, sizeof(int) * (size_t)(2 + 2 * nsys)));
  BTRY(hipMalloc((void **)&u->x, sizeof(double) * nr));
  BTRY(hipMalloc((void **)&u->b, sizeof(double) * nr));
  BTRY(hipMalloc((void **)&u->y, sizeof(float) * nr));
  BTRY(hipMalloc((void **)&u->r, sizeof(float) * nr));
  BTRY(hipMalloc((void **)&u->p, sizeof(float) * nr));
  BTRY(hipMalloc((void **)&u->q, sizeof(float) * nr));
  BTRY(hipMalloc((void **)&u->sc, sizeof(sfr_scalars) * (size_t)nsys));
  BTRY(hipMalloc((void **)&u->ts, sizeof(unsigned long long) * (size_t)(nsys + 1)));
  BTRY(hipMalloc((void **)&u->partA, sizeof(double) * np));
  BTRY(hipMalloc((void **)&u->partB, sizeof(double) * np));
  /* every kernel of a solve starts from the active list; keep it valid */
  BTRY(hipMemsetAsync(u->act, 0, sizeof(int) * (size_t)(2 + 2 * nsys), ctx->stream));
  BTRY(hipMemsetAsync(u->sc, 0, sizeof(sfr_scalars) * (size_t)nsys, ctx->stream));
#undef BTRY
  return SFR_OK;
}

static int grid_of(const sfr_context *ctx, int64_t items)
{
  int64_t g = (int64_t)ctx->ncu * 8;
  const int64_t need = (items + 7) / 8 * 8;
  if (need < g) g = need;
  if (g < 8) g = 8;
  return (int)g;
}

static sfr_ub_view view_of(const ub_bufs *u, int nsys, int grid)
{
  sfr_ub_view v;
  memset(&v, 0, sizeof v);
  v.sys = u->d_sys; v.nsys = nsys; v.grid = grid; v.act = u->act;
  v.b = u->b; v.x = u->x; v.y = u->y; v.r = u->r; v.p = u->p; v.q = u->q;
  v.sc = u->sc; v.ts = u->ts; v.partA = u->partA; v.partB = u->partB;
  return v;
}

/* ---- autotuner ------------------------------------------------------------ */

/* Time of one FP32 CG iteration (the three slot kernels, every system in CG
 * state) per row, for `copies` systems cycling through the sample matrices
 * M[0..m), at rows per thread rpt.  u holds buffers for the copies at rows
 * per thread 1 (the most work items). */
static int time_iteration(sfr_context *ctx, const ub_bufs *u, sfr_matrix **M,
                          int m, int copies, int rpt, double *ns_per_row)
{
  sfr_ub_sys *h = (sfr_ub_sys *)malloc(sizeof(sfr_ub_sys) * (size_t)copies);
  if (!h) return sfr_set_error(ctx, SFR_BAD_ARG, "host allocation");
  int64_t rows = 0, items = 0, real = 0;
  for (int c = 0; c < copies; c++) {
    describe(&h[c], M[c % m], rows, rpt, items);
    rows += align_rows(M[c % m]->n);
    items += h[c].g;
    real += M[c % m]->n;
  }
  const int grid = grid_of(ctx, items);
  sfr_ub_view v = view_of(u, copies, grid);
  hipStream_t s = ctx->stream;
  int rc = SFR_OK;
  float ms = 0.0f;
#define UTRY(call) do { hipError_t e_ = (hipError_t)(call); if (e_ != hipSuccess) { \
    rc = sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s", #call, hipGetErrorString(e_)); goto out; } } while (0)
  UTRY(hipMemcpyAsync(u->d_sys, h, sizeof(sfr_ub_sys) * (size_t)copies, hipMemcpyHostToDevice, s));
  UTRY(sfr_k_fill(s, rows, 1.0, u->b));
  UTRY(sfr_k_ub_compact(s, &v, 1));
  /* tol 0, delta 0: no fold after the first, so every timed iteration is
   * a CG iteration of every copy */
  UTRY(sfr_k_ub_init(s, &v, 1, 0.0, 0x7fffffffu, 0.0));
  for (int it = 0; it < 2; it++) UTRY(sfr_k_ub_iter(s, &v));
  UTRY(hipEventRecord(ctx->ev_start, s));
  for (int it = 0; it < TUNE_REPS; it++) UTRY(sfr_k_ub_iter(s, &v));
  UTRY(hipEventRecord(ctx->ev_stop, s));
  UTRY(hipEventSynchronize(ctx->ev_stop));
  UTRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  *ns_per_row = (double)ms * 1e6 / TUNE_REPS / (double)(real > 0 ? real : 1);
out:
#undef UTRY
  hipStreamSynchronize(s);
  free(h);
  return rc;
}

/* Choose the device format and rows per thread of one bucket from up to
 * TUNE_SAMPLES of its members; the sample matrices in the chosen format are
 * kept in out[] (indexed by system), the others destroyed. */
static int tune_bucket(sfr_context *ctx, const sfr_system *sys, const int *mem,
                       int m, sfr_matrix **out, sfr_bucket_choice *ch)
{
  const int ns = m < TUNE_SAMPLES ? m : TUNE_SAMPLES;
  int pick[TUNE_SAMPLES];
  for (int k = 0; k < ns; k++) pick[k] = mem[(int64_t)k * m / ns];
  /* replicas of the samples up to TUNE_ROWS rows: the bulk of a batch solve
   * runs with the GPU full, and that is the regime the choice is made for */
  int64_t round_rows = 0, round_items = 0;
  for (int k = 0; k < ns; k++) {
    round_rows += align_rows(sys[pick[k]].neq);
    round_items += items_for(sys[pick[k]].neq, 1);
  }
  int64_t rounds = (TUNE_ROWS + round_rows - 1) / round_rows;
  if (rounds * ns > SFR_UB_MAXSYS) rounds = SFR_UB_MAXSYS / ns;
  if (rounds < 1) rounds = 1;
  const int copies = (int)(rounds * ns);
  ub_bufs u;
  CHECK(alloc_bufs(ctx, &u, copies, rounds * round_rows, rounds * round_items));
  sfr_matrix *M[2][TUNE_SAMPLES];
  memset(M, 0, sizeof M);
  const int fmts[2] = {SFR_FMT_ELL, SFR_FMT_CSR};
  double best = -1.0;
  int bf = 0, brpt = 1, rc = SFR_OK;
  for (int f = 0; f < 2 && rc == SFR_OK; f++) {
    for (int k = 0; k < ns && rc == SFR_OK; k++) {
      rc = sfr_matrix_create(ctx, &sys[pick[k]], fmts[f], &M[f][k]);
      if (rc == SFR_OK) rc = sfr_matrix_prepare_mixed(ctx, M[f][k]);
    }
    for (int rpt = 1; rpt <= TUNE_MAXRPT && rc == SFR_OK; rpt *= 2) {
      double t = 0.0;
      rc = time_iteration(ctx, &u, M[f], ns, copies, rpt, &t);
      if (rc == SFR_OK && (best < 0.0 || t < best)) { best = t; bf = f; brpt = rpt; }
    }
  }
  free_bufs(&u);
  for (int f = 0; f < 2; f++)
    for (int k = 0; k < ns; k++) {
      if (rc == SFR_OK && f == bf) {
        out[pick[k]] = M[f][k];
      } else if (M[f][k]) {
        sfr_matrix_destroy(ctx, M[f][k]);
      }
    }
  if (rc != SFR_OK) return rc;
  ch->format = fmts[bf];
  ch->rpt = brpt;
  ch->ns_per_row = best;
  return SFR_OK;
}

/* ---- create / destroy ----------------------------------------------------- */

void sfr_batch_destroy(sfr_context *ctx, sfr_batch *B)
{
  if (!B) return;
  if (ctx) hipStreamSynchronize(ctx->stream);
  if (B->graph) hipGraphExecDestroy(B->graph);
  if (B->A)
    for (int s = 0; s < B->nsys; s++)
      if (B->A[s]) sfr_matrix_destroy(ctx, B->A[s]);
  free(B->A);
  free(B->h_sys);
  ub_bufs u = {B->d_sys, B->d_act, B->x, B->b, B->y, B->r, B->p, B->q,
               B->d_sc, B->d_ts, B->partA, B->partB};
  free_bufs(&u);
  if (B->h_sc) hipHostFree(B->h_sc);
  if (B->h_ts) hipHostFree(B->h_ts);
  if (B->h_poll) hipHostFree(B->h_poll);
  free(B);
}

int sfr_batch_create(sfr_context *ctx, int nsys, const sfr_system *sys,
                     int tune, sfr_batch **out)
{
  if (!ctx || !sys || !out) return SFR_BAD_ARG;
  *out = NULL;
  if (nsys < 1 || nsys > SFR_UB_MAXSYS)
    return sfr_set_error(ctx, SFR_BAD_ARG, "batch of %d systems (1..%d)", nsys, SFR_UB_MAXSYS);
  for (int s = 0; s < nsys; s++)
    if (sys[s].neq <= 0)
      return sfr_set_error(ctx, SFR_BAD_ARG, "system %d: neq = %d", s, sys[s].neq);
  sfr_batch *B = (sfr_batch *)calloc(1, sizeof *B);
  if (!B) return sfr_set_error(ctx, SFR_BAD_ARG, "host allocation");
  B->nsys = nsys;
  B->A = (sfr_matrix **)calloc((size_t)nsys, sizeof(sfr_matrix *));
  B->h_sys = (sfr_ub_sys *)calloc((size_t)nsys, sizeof(sfr_ub_sys));
  int *rpt = (int *)calloc((size_t)nsys, sizeof(int));
  int *fmt = (int *)calloc((size_t)nsys, sizeof(int));
  int *mem = (int *)malloc(sizeof(int) * (size_t)nsys);
  int rc = SFR_OK;
  if (!B->A || !B->h_sys || !rpt || !fmt || !mem) {
    rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation");
    goto fail;
  }
  /* buckets and their choices */
  for (int k = 0; k < NBUCKET && rc == SFR_OK; k++) {
    int m = 0;
    for (int s = 0; s < nsys; s++)
      if (bucket_of(sys[s].neq) == k) mem[m++] = s;
    if (!m) continue;
    sfr_bucket_choice ch;
    memset(&ch, 0, sizeof ch);
    ch.bucket = k;
    ch.nsys = m;
    if (tune) {
      rc = tune_bucket(ctx, sys, mem, m, B->A, &ch);
    } else {
      ch.format = SFR_FMT_AUTO;
      ch.rpt = 0;
    }
    for (int j = 0; j < m && rc == SFR_OK; j++) {
      fmt[mem[j]] = ch.format;
      rpt[mem[j]] = ch.rpt ? ch.rpt : rpt_rule(sys[mem[j]].neq);
    }
    B->bucket[B->nbucket++] = ch;
  }
  /* the remaining matrices */
  for (int s = 0; s < nsys && rc == SFR_OK; s++) {
    if (!B->A[s]) rc = sfr_matrix_create(ctx, &sys[s], fmt[s], &B->A[s]);
    if (rc == SFR_OK) rc = sfr_matrix_prepare_mixed(ctx, B->A[s]);
  }
  if (rc != SFR_OK) goto fail;
  for (int i = 0; i < B->nbucket; i++)      /* report what AUTO resolved to */
    if (B->bucket[i].format == SFR_FMT_AUTO)
      for (int s = 0; s < nsys; s++)
        if (bucket_of(sys[s].neq) == B->bucket[i].bucket) {
          B->bucket[i].format = B->A[s]->fmt;
          B->bucket[i].rpt = rpt[s];
          break;
        }
  /* descriptors and layout */
  int64_t rows = 0, items = 0;
  for (int s = 0; s < nsys; s++) {
    describe(&B->h_sys[s], B->A[s], rows, rpt[s], items);
    rows += align_rows(B->A[s]->n);
    items += B->h_sys[s].g;
  }
  if (items > 0x7fffffff / 4) {
    rc = sfr_set_error(ctx, SFR_BAD_ARG, "batch too large (%lld work items)", (long long)items);
    goto fail;
  }
  B->rows = rows;
  B->items = items;
  B->grid = grid_of(ctx, items);
  {
    ub_bufs u;
    rc = alloc_bufs(ctx, &u, nsys, rows, items);
    if (rc != SFR_OK) goto fail;
    B->d_sys = u.d_sys; B->d_act = u.act; B->x = u.x; B->b = u.b;
    B->y = u.y; B->r = u.r; B->p = u.p; B->q = u.q; B->d_sc = u.sc;
    B->d_ts = u.ts; B->partA = u.partA; B->partB = u.partB;
  }
#define FTRY(call) do { hipError_t e_ = (hipError_t)(call); if (e_ != hipSuccess) { \
    rc = sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s", #call, hipGetErrorString(e_)); goto fail; } } while (0)
  FTRY(hipMemcpyAsync(B->d_sys, B->h_sys, sizeof(sfr_ub_sys) * (size_t)nsys,
                      hipMemcpyHostToDevice, ctx->stream));
  FTRY(hipHostMalloc((void **)&B->h_sc, sizeof(sfr_scalars) * (size_t)nsys, 0));
  FTRY(hipHostMalloc((void **)&B->h_ts, sizeof(unsigned long long) * (size_t)(nsys + 1), 0));
  FTRY(hipHostMalloc((void **)&B->h_poll, sizeof(int) * 4, 0));
  FTRY(hipDeviceGetAttribute(&B->clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
  FTRY(hipStreamSynchronize(ctx->stream));
#undef FTRY
  free(rpt); free(fmt); free(mem);
  *out = B;
  return SFR_OK;
fail:
  free(rpt); free(fmt); free(mem);
  sfr_batch_destroy(ctx, B);
  return rc;
}

int sfr_batch_set_values(sfr_context *ctx, sfr_batch *B, int s,
                         const double *al, const double *ad, const double *au)
{
  if (!ctx || !B || s < 0 || s >= B->nsys) return SFR_BAD_ARG;
  CHECK(sfr_matrix_set_values(ctx, B->A[s], al, ad, au));
  return sfr_matrix_prepare_mixed(ctx, B->A[s]);
}

int64_t sfr_batch_rows(const sfr_batch *B)
{
  return B ? B->rows : 0;
}

int sfr_batch_info(const sfr_batch *B, int s, int64_t *row0, int32_t *neq,
                   int *format, int *rpt)
{
  if (!B || s < 0 || s >= B->nsys) return SFR_BAD_ARG;
  if (row0) *row0 = B->h_sys[s].row0;
  if (neq) *neq = B->h_sys[s].n;
  if (format) *format = B->h_sys[s].fmt;
  if (rpt) *rpt = B->h_sys[s].rpt;
  return SFR_OK;
}

int sfr_batch_buckets(const sfr_batch *B, sfr_bucket_choice *out, int max)
{
  if (!B || !out) return 0;
  int n = B->nbucket < max ? B->nbucket : max;
  for (int i = 0; i < n; i++) out[i] = B->bucket[i];
  return n;
}

/* ---- solve ----------------------------------------------------------------- */

static sfr_ub_view batch_view(const sfr_batch *B)
{
  ub_bufs u = {B->d_sys, B->d_act, B->x, B->b, B->y, B->r, B->p, B->q,
               B->d_sc, B->d_ts, B->partA, B->partB};
  return view_of(&u, B->nsys, B->grid);
}

static int batch_graph(sfr_context *ctx, sfr_batch *B, const sfr_ub_view *v)
{
  if (B->graph && B->graph_chunk == ctx->chunk) return SFR_OK;
  if (B->graph) { hipGraphExecDestroy(B->graph); B->graph = NULL; }
  hipStream_t s = ctx->stream;
  TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int lrc = hipSuccess;
  for (int it = 0; it < ctx->chunk && lrc == hipSuccess; it++) lrc = sfr_k_ub_iter(s, v);
  if (lrc == hipSuccess) lrc = sfr_k_ub_compact(s, v, 0);
  hipGraph_t graph = NULL;
  const hipError_t erc = hipStreamEndCapture(s, &graph);
  if (lrc != hipSuccess || erc != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    return sfr_set_error(ctx, SFR_HIP_ERROR, "graph capture: %s",
                         hipGetErrorString(lrc != hipSuccess ? (hipError_t)lrc : erc));
  }
  const hipError_t irc = hipGraphInstantiate(&B->graph, graph, NULL, NULL, 0);
  hipGraphDestroy(graph);
  if (irc != hipSuccess) {
    B->graph = NULL;
    return sfr_set_error(ctx, SFR_HIP_ERROR, "graph instantiate: %s", hipGetErrorString(irc));
  }
  B->graph_chunk = ctx->chunk;
  return SFR_OK;
}

int sfr_batch_solve_device(sfr_context *ctx, sfr_batch *B, double tol,
                           unsigned maxit, const double *d_b, double *d_x,
                           int newX, sfr_stats *stats)
{
  if (!ctx || !B || !d_b || !d_x) return SFR_BAD_ARG;
  if (!(tol >= 0.0) || maxit < 1)
    return sfr_set_error(ctx, SFR_BAD_ARG, "tol %g maxit %u", tol, maxit);
  TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t vb = sizeof(double) * (size_t)B->rows;
  if (d_b != B->b) TRY(hipMemcpyAsync(B->b, d_b, vb, hipMemcpyDeviceToDevice, s));
  if (!newX && d_x != B->x) TRY(hipMemcpyAsync(B->x, d_x, vb, hipMemcpyDeviceToDevice, s));
  const sfr_ub_view v = batch_view(B);
  CHECK(batch_graph(ctx, B, &v));
  TRY(hipEventRecord(ctx->ev_start, s));
  TRY(sfr_k_ub_compact(s, &v, 1));
  TRY(sfr_k_ub_init(s, &v, newX, tol, maxit, ctx->ru_delta));
  /* every iteration of a system advances its inner or its outer count, and
   * a breakdown costs one extra iteration: 3 maxit + 2 iterations bound it */
  const uint64_t max_chunks = (3 * (uint64_t)maxit + 2) / (uint64_t)ctx->chunk + 3;
  uint64_t launched = 0, seen = 0;
  for (;;) {
    while (launched - seen < 2 && launched < max_chunks) {
      const int slot = (int)(launched & 1);
      TRY(hipGraphLaunch(B->graph, s));
      TRY(hipMemcpyAsync(B->h_poll + 2 * slot, B->d_act, 2 * sizeof(int),
                         hipMemcpyDeviceToHost, s));
      TRY(hipEventRecord(ctx->ev_poll[slot], s));
      launched++;
    }
    if (seen == launched) break;
    const int slot = (int)(seen & 1);
    TRY(hipEventSynchronize(ctx->ev_poll[slot]));
    seen++;
    if (B->h_poll[2 * slot] == 0) break;
  }
  TRY(sfr_k_ub_compact(s, &v, 2));
  TRY(sfr_k_ub_final(s, &v));
  TRY(hipEventRecord(ctx->ev_stop, s));
  TRY(hipMemcpyAsync(B->h_sc, B->d_sc, sizeof(sfr_scalars) * (size_t)B->nsys,
                     hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(B->h_ts, B->d_ts, sizeof(unsigned long long) * (size_t)(B->nsys + 1),
                     hipMemcpyDeviceToHost, s));
  if (d_x != B->x) TRY(hipMemcpyAsync(d_x, B->x, vb, hipMemcpyDeviceToDevice, s));
  TRY(hipStreamSynchronize(s));
  float ms = 0.0f;
  TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  int worst = SFR_OK;
  const double tick = B->clock_khz > 0 ? 1.0 / (1e3 * B->clock_khz) : 0.0;
  for (int c = 0; c < B->nsys; c++) {
    sfr_stats st;
    sfr_fill_stats(&B->h_sc[c], &st, ms);
    if (B->h_sc[c].done != SFR_DONE_RUN && B->h_ts[c] >= B->h_ts[B->nsys])
      st.time_s = (double)(B->h_ts[c] - B->h_ts[B->nsys]) * tick;
    if (st.status != SFR_OK && (worst == SFR_OK || st.status == SFR_BREAKDOWN)) worst = st.status;
    if (stats) stats[c] = st;
  }
  return worst;
}

int sfr_batch_solve(sfr_context *ctx, sfr_batch *B, double tol, unsigned maxit,
                    const double *const *b, double *const *x, int newX,
                    sfr_stats *stats)
{
  if (!ctx || !B || !b || !x) return SFR_BAD_ARG;
  TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  for (int c = 0; c < B->nsys; c++) {
    if (!b[c] || !x[c]) return sfr_set_error(ctx, SFR_BAD_ARG, "system %d: NULL b or x", c);
    const size_t nb = sizeof(double) * (size_t)B->h_sys[c].n;
    TRY(hipMemcpyAsync(B->b + B->h_sys[c].row0, b[c], nb, hipMemcpyHostToDevice, s));
    if (!newX) TRY(hipMemcpyAsync(B->x + B->h_sys[c].row0, x[c], nb, hipMemcpyHostToDevice, s));
  }
  const int rc = sfr_batch_solve_device(ctx, B, tol, maxit, B->b, B->x, newX, stats);
  if (rc == SFR_BAD_ARG || rc == SFR_HIP_ERROR) return rc;
  for (int c = 0; c < B->nsys; c++)
    TRY(hipMemcpyAsync(x[c], B->x + B->h_sys[c].row0, sizeof(double) * (size_t)B->h_sys[c].n,
                       hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  return rc;
}
