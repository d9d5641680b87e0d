/*
 * Context and device-resident matrices.
 *
 * A host matrix in one of the reference's storages (src/Datastruct.c:28-613)
 * is converted once, on the host, to a device format that holds complete
 * rows:
 *   ELL  slot-major val[f*n+i]/col[f*n+i], every row padded to `slots`
 *        entries with (col = i, value 0); lane per row, coalesced per slot;
 *   CSR  row pointers + columns, several lanes per row.
 * Symmetric storages (lower or upper triangle, src/HccaBlas.c:2620-2651,
 * 3297-3331) are expanded to full rows, so the device never scatters.
 * CSRC unsym (lower pattern, [al|au], src/HccaBlas.c:4223-4260) contributes
 * al to row i and au to row ja[k].  ELLPACK input (src/EllPack.c:26-79,
 * src/HccaBlas.c:4700-4752) keeps its slot order.
 * Each stored device entry remembers which raw host value feeds it (vmap), so
 * a re-assembly with the same pattern uploads the raw arrays and one gather
 * kernel rebuilds the device values.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../sfr_internal.h"

#define CHECK_RC(call) do { int rc_ = (call); if (rc_ != SFR_OK) return rc_; } while (0)

int sfr_set_error(sfr_context *ctx, int code, const char *fmt, ...)
{
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof ctx->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

const char *sfr_last_error(const sfr_context *ctx)
{
  return ctx ? ctx->err : "no context";
}

const char *sfr_status_string(int s)
{
  switch (s) {
  case SFR_OK: return "ok";
  case SFR_MAXIT: return "maximum iterations reached without convergence";
  case SFR_BREAKDOWN: return "Krylov breakdown (zero or non-finite denominator)";
  case SFR_BAD_ARG: return "invalid argument";
  case SFR_HIP_ERROR: return "HIP runtime error";
  case SFR_NO_DEVICE: return "no HIP device";
  case SFR_NOT_SUPPORTED: return "not supported";
  default: return "unknown status";
  }
}

#define TRY(call)                                                            \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s (%s:%d)", #call,      \
                           hipGetErrorString(e_), __FILE__, __LINE__);       \
  } while (0)

int sfr_init(sfr_context **out, int device)
{
  if (!out) return SFR_BAD_ARG;
  *out = NULL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SFR_NO_DEVICE;
  if (device < 0 || device >= ndev) return SFR_BAD_ARG;
  sfr_context *ctx = (sfr_context *)calloc(1, sizeof *ctx);
  if (!ctx) return SFR_BAD_ARG;
  ctx->device = device;
  ctx->chunk = 16;
  ctx->ru_delta = 0.1;
  ctx->ru_delta_b = 0.1;
  ctx->ell_pairs = 1;
  ctx->nt_loads = -1;
  if (hipSetDevice(device) != hipSuccess) { free(ctx); return SFR_NO_DEVICE; }
  int rc = SFR_OK;
#define INIT_TRY(call) do { if ((call) != hipSuccess) { rc = SFR_HIP_ERROR; goto fail; } } while (0)
  INIT_TRY(hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, device));
  INIT_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  INIT_TRY(hipMalloc((void **)&ctx->d_sc, sizeof(sfr_scalars) * SFR_MAXSYS));
  INIT_TRY(hipMemset(ctx->d_sc, 0, sizeof(sfr_scalars) * SFR_MAXSYS));
  INIT_TRY(hipHostMalloc((void **)&ctx->h_sc, sizeof(sfr_scalars) * SFR_MAXSYS, 0));
  INIT_TRY(hipHostMalloc((void **)&ctx->h_poll, sizeof(sfr_scalars) * 2 * SFR_MAXSYS, 0));
  INIT_TRY(hipMalloc((void **)&ctx->part, sizeof(double) * SFR_MAXDOT * SFR_MAXG * SFR_MAXSYS));
  INIT_TRY(hipEventCreate(&ctx->ev_start));
  INIT_TRY(hipEventCreate(&ctx->ev_stop));
  INIT_TRY(hipEventCreateWithFlags(&ctx->ev_poll[0], hipEventDisableTiming));
  INIT_TRY(hipEventCreateWithFlags(&ctx->ev_poll[1], hipEventDisableTiming));
  INIT_TRY(sfr_scratch_address(device, &ctx->scratch));
#undef INIT_TRY
  *out = ctx;
  return SFR_OK;
fail:
  sfr_finalize(ctx);
  return rc;
}

/* Loop graphs are destroyed only with the context stream drained: a graph
 * must not go while a launch of it may still be in flight. */
static void drop_graphs(sfr_context *ctx)
{
  if (ctx->ngraph && ctx->stream) hipStreamSynchronize(ctx->stream);
  for (int i = 0; i < ctx->ngraph; i++)
    if (ctx->graphs[i].exec) hipGraphExecDestroy(ctx->graphs[i].exec);
  ctx->ngraph = 0;
}

void sfr_drop_graphs_keyed(sfr_context *ctx, int slot, const void *key)
{
  int w = 0, synced = 0;
  for (int i = 0; i < ctx->ngraph; i++) {
    if (ctx->graphs[i].key[slot] == key) {
      if (!synced) { hipStreamSynchronize(ctx->stream); synced = 1; }
      hipGraphExecDestroy(ctx->graphs[i].exec);
    } else {
      ctx->graphs[w++] = ctx->graphs[i];
    }
  }
  ctx->ngraph = w;
}

void sfr_finalize(sfr_context *ctx)
{
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  drop_graphs(ctx);
  sfr_mcache_clear(ctx);
  for (int i = 0; i < 10; i++) hipFree(ctx->w64[i]);
  for (int i = 0; i < 6; i++) hipFree(ctx->w32[i]);
  hipFree(ctx->stage);
  hipFree(ctx->gm_buf);
  hipFree(ctx->nb_val);
  hipFree(ctx->nb_raw);
  hipFree(ctx->nb_val32);
  hipFree(ctx->mb32);
  hipFree(ctx->d_sc);
  hipFree(ctx->part);
  if (ctx->h_sc) hipHostFree(ctx->h_sc);
  if (ctx->h_poll) hipHostFree(ctx->h_poll);
  if (ctx->ev_start) hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) hipEventDestroy(ctx->ev_stop);
  if (ctx->ev_poll[0]) hipEventDestroy(ctx->ev_poll[0]);
  if (ctx->ev_poll[1]) hipEventDestroy(ctx->ev_poll[1]);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  free(ctx);
}

int sfr_synchronize(sfr_context *ctx)
{
  if (!ctx) return SFR_BAD_ARG;
  TRY(hipStreamSynchronize(ctx->stream));
  return SFR_OK;
}

int sfr_dev_alloc(sfr_context *ctx, size_t bytes, void **ptr)
{
  if (!ctx || !ptr) return SFR_BAD_ARG;
  TRY(hipMalloc(ptr, bytes ? bytes : 1));
  return SFR_OK;
}

int sfr_dev_free(sfr_context *ctx, void *ptr)
{
  if (!ctx) return SFR_BAD_ARG;
  TRY(hipStreamSynchronize(ctx->stream));
  TRY(hipFree(ptr));
  return SFR_OK;
}

int sfr_copy_h2d(sfr_context *ctx, void *dst, const void *src, size_t bytes)
{
  if (!ctx) return SFR_BAD_ARG;
  if (!bytes) return SFR_OK;
  TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  TRY(hipStreamSynchronize(ctx->stream));
  return SFR_OK;
}

int sfr_copy_d2h(sfr_context *ctx, void *dst, const void *src, size_t bytes)
{
  if (!ctx) return SFR_BAD_ARG;
  if (!bytes) return SFR_OK;
  TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  TRY(hipStreamSynchronize(ctx->stream));
  return SFR_OK;
}

/* Workspace vectors for systems of up to n rows (grow-only). */
int sfr_workspace(sfr_context *ctx, int64_t n)
{
  if (n <= ctx->ws_n) return SFR_OK;
  TRY(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);                /* captured graphs hold the old buffers */
  for (int i = 0; i < 10; i++) { hipFree(ctx->w64[i]); ctx->w64[i] = NULL; }
  for (int i = 0; i < 6; i++) { hipFree(ctx->w32[i]); ctx->w32[i] = NULL; }
  ctx->ws_n = 0;
  for (int i = 0; i < 10; i++) TRY(hipMalloc((void **)&ctx->w64[i], sizeof(double) * (size_t)n));
  for (int i = 0; i < 6; i++) TRY(hipMalloc((void **)&ctx->w32[i], sizeof(float) * (size_t)n));
  ctx->ws_n = n;
  return SFR_OK;
}

/* GMRES buffers: SFR_GM_K basis vectors and two raw vectors at a stride of
 * ld (n rounded up to 64 doubles), then the Hessenberg state. */
int sfr_gm_workspace(sfr_context *ctx, int64_t n)
{
  if (n <= ctx->gm_n) return SFR_OK;
  TRY(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  hipFree(ctx->gm_buf);
  ctx->gm_buf = NULL;
  ctx->gm_n = 0;
  const int64_t ld = (n + 63) & ~(int64_t)63;
  TRY(hipMalloc((void **)&ctx->gm_buf,
                sizeof(double) * ((size_t)(SFR_GM_K + 2) * (size_t)ld + SFR_GM_STATE)));
  ctx->gm_n = n;
  ctx->gm_ld = ld;
  return SFR_OK;
}

/* Species batch buffers: interleaved values of up to 4 systems (nnz each)
 * and one system's raw values.  Captured graphs hold the old pointers, so
 * they go when the buffers grow. */
int sfr_nb_buffers(sfr_context *ctx, int64_t nnz, int64_t nraw)
{
  const int64_t nv = 4 * (nnz + 1), nr = nraw + 1;
  if (nv <= ctx->nb_val_n && nr <= ctx->nb_raw_n) return SFR_OK;
  TRY(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  hipFree(ctx->nb_val);
  hipFree(ctx->nb_raw);
  hipFree(ctx->nb_val32);
  ctx->nb_val = ctx->nb_raw = NULL;
  ctx->nb_val32 = NULL;
  ctx->nb_val_n = ctx->nb_raw_n = ctx->nb_val32_n = 0;
  TRY(hipMalloc((void **)&ctx->nb_val, sizeof(double) * (size_t)nv));
  TRY(hipMalloc((void **)&ctx->nb_raw, sizeof(double) * (size_t)nr));
  TRY(hipMalloc((void **)&ctx->nb_val32, sizeof(float) * (size_t)nv));
  ctx->nb_val_n = nv;
  ctx->nb_raw_n = nr;
  ctx->nb_val32_n = nv;
  return SFR_OK;
}

sfr_mat_view sfr_view(const sfr_matrix *A)
{
  sfr_mat_view v;
  v.n = A->n; v.fmt = A->fmt; v.slots = A->slots; v.lanes = A->lanes;
  v.col = A->col; v.rowptr = A->rowptr; v.val = A->val; v.ad = A->ad;
  v.dinv = A->dinv; v.val32 = A->val32; v.dscale = A->dscale;
  v.col16 = A->col16;
  v.exc_k = A->exc_k; v.exc_col = A->exc_col; v.nexc = A->nexc;
  v.valnb = NULL;
  v.pair = A->pair_use ? A->pair_kind : SFR_PAIR_NONE;
  v.nt = 0;
  return v;
}

/* The face pairing of a slot-major ELL pattern (hcol[f*n + i]): the kind
 * whose pairs (f, g) satisfy, for every row i with a neighbour j across f,
 * that j's slot g faces back to i.  SFR3D's hex and quad face orders
 * (Adjcency.c's face tables) give the two kinds recognised here. */
static int ell_pair_kind(const int32_t *hcol, int32_t n, int s)
{
  static const int hex[6] = {1, 0, 4, 5, 2, 3}, quad[4] = {2, 3, 0, 1};
  const int *opp = s == 6 ? hex : (s == 4 ? quad : NULL);
  if (!opp) return SFR_PAIR_NONE;
  for (int f = 0; f < s; f++)
    for (int32_t i = 0; i < n; i++) {
      const int32_t j = hcol[(int64_t)f * n + i];
      if (j != i && hcol[(int64_t)opp[f] * n + j] != i) return SFR_PAIR_NONE;
    }
  return s == 6 ? SFR_PAIR_HEX : SFR_PAIR_QUAD;
}

/* ---- host conversion ------------------------------------------------------ */

typedef struct { int32_t col; int32_t raw; } entry;

static int cmp_entry(const void *a, const void *b)
{
  const entry *x = (const entry *)a, *y = (const entry *)b;
  if (x->col != y->col) return (x->col > y->col) - (x->col < y->col);
  return (x->raw > y->raw) - (x->raw < y->raw);
}

static int valid_storage(short s)
{
  return s >= SFR_CSR && s <= SFR_CSRCCOO;
}

/* Full rows of the system as (col, raw index) lists: off[rowptr[i]..]. */
static int full_rows(const sfr_system *sys, int64_t **rowptr_out,
                     entry **ent_out, int32_t **dmap_out, int64_t *nraw_out)
{
  const int32_t n = sys->neq;
  const int32_t *ia = sys->ia, *ja = sys->ja;
  const short st = sys->storage;
  const int unsym = sys->unsym != 0;
  int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int32_t *dmap = NULL;
  if (!cnt) return -1;
  int64_t nraw = 0;
  const int csrc = (st == SFR_CSRC || st == SFR_CSRCCOO);
  const int csrd = (st == SFR_CSRD || st == SFR_CSRDCOO);
  /* every raw value index must fall inside the staged arrays */
  if (st != SFR_ELLPACK && (ia[0] != 0 || ia[n] != sys->nad || sys->nad < 0)) {
    free(cnt);
    return -2;
  }
  if (st == SFR_CSR) {
    dmap = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!dmap) { free(cnt); return -1; }
    for (int32_t i = 0; i < n; i++) dmap[i] = -1;
  }
  /* pass 1: counts */
  for (int32_t i = 0; i < n; i++) {
    if (st == SFR_ELLPACK) { cnt[i + 1] += ia[0]; continue; }
    for (int32_t k = ia[i]; k < ia[i + 1]; k++) {
      const int32_t j = ja[k];
      if (j < 0 || j >= n) { free(cnt); free(dmap); return -2; }
      if (st == SFR_CSR && j == i) { dmap[i] = k; continue; }
      cnt[i + 1]++;
      if (st == SFR_CSR && !unsym) cnt[j + 1]++;
      else if (csrd && !unsym) cnt[j + 1]++;
      else if (csrc) cnt[j + 1]++;         /* CSRC: sym mirror or au entry */
    }
  }
  if (st == SFR_ELLPACK) nraw = (int64_t)n * ia[0];
  else if (csrc && unsym) nraw = 2 * (int64_t)sys->nad;
  else nraw = sys->nad;
  for (int32_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  entry *ent = (entry *)malloc(sizeof(entry) * (size_t)(cnt[n] > 0 ? cnt[n] : 1));
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  if (!ent || !pos) { free(cnt); free(dmap); free(ent); free(pos); return -1; }
  for (int32_t i = 0; i < n; i++) pos[i] = cnt[i];
  /* pass 2: fill */
  for (int32_t i = 0; i < n; i++) {
    if (st == SFR_ELLPACK) {
      const int s = ia[0];
      for (int f = 0; f < s; f++) {
        const int64_t k = (int64_t)i * s + f;
        const int32_t j = ja[k];
        if (j < 0 || j >= n) { free(cnt); free(dmap); free(ent); free(pos); return -2; }
        ent[pos[i]++] = (entry){j, (int32_t)k};
      }
      continue;
    }
    for (int32_t k = ia[i]; k < ia[i + 1]; k++) {
      const int32_t j = ja[k];
      if (st == SFR_CSR && j == i) continue;
      ent[pos[i]++] = (entry){j, k};
      if ((st == SFR_CSR || csrd) && !unsym) ent[pos[j]++] = (entry){i, k};
      else if (csrc) ent[pos[j]++] = (entry){i, unsym ? sys->nad + k : k};
    }
  }
  free(pos);
  if (st != SFR_ELLPACK)               /* ELLPACK keeps the face order */
    for (int32_t i = 0; i < n; i++)
      qsort(ent + cnt[i], (size_t)(cnt[i + 1] - cnt[i]), sizeof(entry), cmp_entry);
  if (st == SFR_CSR)
    for (int32_t i = 0; i < n; i++)
      if (dmap[i] < 0) { free(cnt); free(dmap); free(ent); return -3; }
  *rowptr_out = cnt;
  *ent_out = ent;
  *dmap_out = dmap;
  *nraw_out = nraw;
  return 0;
}

void sfr_matrix_destroy(sfr_context *ctx, sfr_matrix *A)
{
  if (!A) return;
  if (ctx) {
    hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < SFR_MCACHE; i++)
      if (ctx->mcache[i].A == A) ctx->mcache[i].A = NULL;
    /* graphs captured on this matrix are stale */
    sfr_drop_graphs_keyed(ctx, 0, A);
  }
  if (A->T) sfr_matrix_destroy(ctx, A->T);
  hipFree(A->col); hipFree(A->col16); hipFree(A->exc_k); hipFree(A->exc_col);
  hipFree(A->rowptr); hipFree(A->vmap); hipFree(A->val);
  hipFree(A->dscale); hipFree(A->val32); hipFree(A->val32r); hipFree(A->grad_rc);
  if (!A->borrowed) {
    hipFree(A->dmap); hipFree(A->raw); hipFree(A->raw_ad); hipFree(A->ad);
    hipFree(A->dinv);
  }
  free(A);
}

/* Build the device format of A (n rows, entries rp/ent) and upload the
 * index arrays.  fmt = SFR_FMT_ELL / SFR_FMT_CSR / SFR_FMT_AUTO. */
static int upload_format(sfr_context *ctx, sfr_matrix *A, int32_t n,
                         const int64_t *rp, const entry *ent, int fmt)
{
  int maxrow = 0;
  for (int32_t i = 0; i < n; i++) {
    const int64_t l = rp[i + 1] - rp[i];
    if (l > maxrow) maxrow = (int)l;
  }
  if (fmt == SFR_FMT_AUTO)
    fmt = (maxrow <= 16 && (int64_t)maxrow * n <= 2 * rp[n] + 16 * (int64_t)n)
              ? SFR_FMT_ELL : SFR_FMT_CSR;
  if (fmt != SFR_FMT_ELL && fmt != SFR_FMT_CSR)
    return sfr_set_error(ctx, SFR_BAD_ARG, "format %d", fmt);
  A->n = n;
  A->fmt = fmt;
  /* rows per thread: the smallest that keeps a streaming kernel at or below
   * 1024 workgroups, so a consumer reduces at most 1024 partials */
  A->rpt = 1;
  while (A->rpt < 8 && (n + (int64_t)SFR_BLOCK * A->rpt - 1) / ((int64_t)SFR_BLOCK * A->rpt) > 1024)
    A->rpt *= 2;
  A->lanes = 4;
  int32_t *hcol = NULL, *hmap = NULL, *hptr = NULL;
  int rc = SFR_OK;
  if (fmt == SFR_FMT_ELL) {
    const int s = maxrow > 0 ? maxrow : 1;
    A->slots = s;
    A->nnz = (int64_t)s * n;
    hcol = (int32_t *)malloc(sizeof(int32_t) * (size_t)A->nnz);
    hmap = (int32_t *)malloc(sizeof(int32_t) * (size_t)A->nnz);
    if (!hcol || !hmap) { rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation"); goto done; }
    for (int32_t i = 0; i < n; i++) {
      const int64_t b = rp[i], l = rp[i + 1] - rp[i];
      for (int f = 0; f < s; f++) {
        const int64_t k = (int64_t)f * n + i;
        if (f < l) { hcol[k] = ent[b + f].col; hmap[k] = ent[b + f].raw; }
        else { hcol[k] = i; hmap[k] = -1; }
      }
    }
  } else {
    A->slots = maxrow;
    A->nnz = rp[n];
    hcol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(A->nnz + 1));
    hmap = (int32_t *)malloc(sizeof(int32_t) * (size_t)(A->nnz + 1));
    hptr = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    if (!hcol || !hmap || !hptr) { rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation"); goto done; }
    for (int32_t i = 0; i <= n; i++) hptr[i] = (int32_t)rp[i];
    for (int64_t k = 0; k < A->nnz; k++) { hcol[k] = ent[k].col; hmap[k] = ent[k].raw; }
    A->lanes = maxrow <= 4 ? 2 : (maxrow <= 12 ? 4 : 8);
  }
#define MTRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
    rc = sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s", #call, hipGetErrorString(e_)); goto done; } } while (0)
  MTRY(hipMalloc((void **)&A->col, sizeof(int32_t) * (size_t)(A->nnz + 1)));
  MTRY(hipMalloc((void **)&A->vmap, sizeof(int32_t) * (size_t)(A->nnz + 1)));
  MTRY(hipMalloc((void **)&A->val, sizeof(double) * (size_t)(A->nnz + 1)));
  MTRY(hipMemcpyAsync(A->col, hcol, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyHostToDevice, ctx->stream));
  if (fmt == SFR_FMT_ELL) {
    /* 16-bit offsets from the row: half the index stream.  Offsets outside
     * +-32767 (e.g. a slot facing a cell without an equation, which SFR3D's
     * ELLPACK points at column 0) are escaped and listed, if they are few. */
    int64_t nexc = 0;
    for (int64_t k = 0; k < A->nnz && nexc <= SFR_D16_MAXEXC; k++) {
      const int64_t d = (int64_t)hcol[k] - (int64_t)(k % n);
      if (d < -32767 || d > 32767) nexc++;
    }
    if (nexc <= SFR_D16_MAXEXC && A->nnz > 0) {
      int16_t *h16 = (int16_t *)malloc(sizeof(int16_t) * (size_t)A->nnz);
      int64_t *hxk = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nexc + 1));
      int32_t *hxc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nexc + 1));
      if (!h16 || !hxk || !hxc) {
        free(h16); free(hxk); free(hxc);
        rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation");
        goto done;
      }
      int64_t e = 0;
      for (int64_t k = 0; k < A->nnz; k++) {
        const int64_t d = (int64_t)hcol[k] - (k % n);
        if (d < -32767 || d > 32767) {
          h16[k] = (int16_t)SFR_D16_ESC;
          hxk[e] = k;
          hxc[e] = hcol[k];
          e++;
        } else {
          h16[k] = (int16_t)d;
        }
      }
      hipError_t e1 = hipMalloc((void **)&A->col16, sizeof(int16_t) * (size_t)A->nnz);
      if (e1 == hipSuccess)
        e1 = hipMemcpy(A->col16, h16, sizeof(int16_t) * (size_t)A->nnz, hipMemcpyHostToDevice);
      if (e1 == hipSuccess) e1 = hipMalloc((void **)&A->exc_k, sizeof(int64_t) * (size_t)(nexc + 1));
      if (e1 == hipSuccess) e1 = hipMalloc((void **)&A->exc_col, sizeof(int32_t) * (size_t)(nexc + 1));
      if (e1 == hipSuccess && nexc)
        e1 = hipMemcpy(A->exc_k, hxk, sizeof(int64_t) * (size_t)nexc, hipMemcpyHostToDevice);
      if (e1 == hipSuccess && nexc)
        e1 = hipMemcpy(A->exc_col, hxc, sizeof(int32_t) * (size_t)nexc, hipMemcpyHostToDevice);
      free(h16); free(hxk); free(hxc);
      if (e1 != hipSuccess) {
        rc = sfr_set_error(ctx, SFR_HIP_ERROR, "col16 upload: %s", hipGetErrorString(e1));
        goto done;
      }
      A->nexc = (int)nexc;
      if (nexc == 0) A->pair_kind = ell_pair_kind(hcol, n, A->slots);
    }
  }
  MTRY(hipMemcpyAsync(A->vmap, hmap, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyHostToDevice, ctx->stream));
  if (hptr) {
    MTRY(hipMalloc((void **)&A->rowptr, sizeof(int32_t) * ((size_t)n + 1)));
    MTRY(hipMemcpyAsync(A->rowptr, hptr, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyHostToDevice, ctx->stream));
  }
  MTRY(hipStreamSynchronize(ctx->stream));
#undef MTRY
done:
  free(hcol); free(hmap); free(hptr);
  return rc;
}

int sfr_matrix_create(sfr_context *ctx, const sfr_system *sys, int format,
                      sfr_matrix **out)
{
  if (!ctx || !sys || !out) return SFR_BAD_ARG;
  *out = NULL;
  if (sys->neq <= 0 || !sys->ia || !valid_storage(sys->storage))
    return sfr_set_error(ctx, SFR_BAD_ARG, "bad system (neq=%d storage=%d)",
                         sys->neq, sys->storage);
  if (sys->storage != SFR_CSR && !sys->ad)
    return sfr_set_error(ctx, SFR_BAD_ARG, "ad is required for storage %d",
                         sys->storage);
  if (sys->storage == SFR_ELLPACK && (sys->ia[0] < 0 || sys->ia[0] > 64))
    return sfr_set_error(ctx, SFR_BAD_ARG, "ELLPACK width %d", sys->ia[0]);
  if (format != SFR_FMT_AUTO && format != SFR_FMT_ELL && format != SFR_FMT_CSR)
    return sfr_set_error(ctx, SFR_BAD_ARG, "format %d", format);
  int64_t *rp = NULL; entry *ent = NULL; int32_t *dmap = NULL; int64_t nraw = 0;
  const int frc = full_rows(sys, &rp, &ent, &dmap, &nraw);
  if (frc == -2)
    return sfr_set_error(ctx, SFR_BAD_ARG,
                         "index arrays inconsistent (ia[0] != 0, ia[neq] != nad or a column out of range)");
  if (frc == -3) return sfr_set_error(ctx, SFR_BAD_ARG, "CSR row without diagonal");
  if (frc) return sfr_set_error(ctx, SFR_BAD_ARG, "host allocation failed");
  const int32_t n = sys->neq;
  sfr_matrix *A = (sfr_matrix *)calloc(1, sizeof *A);
  if (!A) { free(rp); free(ent); free(dmap); return SFR_BAD_ARG; }
  A->storage = sys->storage; A->unsym = sys->unsym;
  A->nad = sys->nad; A->nraw = nraw; A->diag_in_al = sys->storage == SFR_CSR;
  A->h_ia = sys->ia; A->h_ja = sys->ja;
  int rc = upload_format(ctx, A, n, rp, ent, format);
  if (rc == SFR_OK) {
#define MTRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
    rc = sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s", #call, hipGetErrorString(e_)); goto done; } } while (0)
    MTRY(hipMalloc((void **)&A->ad, sizeof(double) * (size_t)n));
    MTRY(hipMalloc((void **)&A->dinv, sizeof(double) * (size_t)n));
    MTRY(hipMalloc((void **)&A->raw, sizeof(double) * (size_t)(nraw + 1)));
    MTRY(hipMalloc((void **)&A->raw_ad, sizeof(double) * (size_t)n));
    if (dmap) {
      MTRY(hipMalloc((void **)&A->dmap, sizeof(int32_t) * (size_t)n));
      MTRY(hipMemcpyAsync(A->dmap, dmap, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
      MTRY(hipStreamSynchronize(ctx->stream));
    }
#undef MTRY
    rc = sfr_matrix_set_values(ctx, A, sys->al, sys->ad, sys->au);
  }
done:
  free(rp); free(ent); free(dmap);
  if (rc != SFR_OK) { sfr_matrix_destroy(ctx, A); return rc; }
  *out = A;
  return SFR_OK;
}

/* Adjoint support, built on first use: the transposed operator (same raw
 * values, entries (j, i)) and, for each raw value, the matrix entries it
 * feeds.  The device index arrays are read back, so nothing of the host
 * system has to be kept. */
int sfr_matrix_prepare_adjoint(sfr_context *ctx, sfr_matrix *A)
{
  if (A->have_grad) return SFR_OK;
  const int32_t n = A->n;
  const int64_t nnz = A->nnz;
  int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
  int32_t *map = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
  int32_t *ptr = A->rowptr ? (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1)) : NULL;
  int32_t *dmap = A->dmap ? (int32_t *)malloc(sizeof(int32_t) * (size_t)n) : NULL;
  int32_t *rc4 = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)(A->nraw + 1));
  int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  entry *tent = NULL;
  int rc = SFR_OK;
  if (!col || !map || !rc4 || !cnt || (A->rowptr && !ptr) || (A->dmap && !dmap)) {
    rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation");
    goto out;
  }
#define ATRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
    rc = sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s", #call, hipGetErrorString(e_)); goto out; } } while (0)
  ATRY(hipMemcpy(col, A->col, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost));
  ATRY(hipMemcpy(map, A->vmap, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost));
  if (ptr) ATRY(hipMemcpy(ptr, A->rowptr, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost));
  if (dmap) ATRY(hipMemcpy(dmap, A->dmap, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < 4 * (A->nraw + 1); k++) rc4[k] = -1;
  /* entries (i, j, raw) of the device format */
  for (int32_t i = 0; i < n; i++) {
    const int64_t b = ptr ? ptr[i] : 0, e = ptr ? ptr[i + 1] : A->slots;
    for (int64_t q = b; q < e; q++) {
      const int64_t k = ptr ? q : q * n + i;
      const int32_t raw = map[k];
      if (raw < 0) continue;
      int32_t *slot = rc4 + 4 * (int64_t)raw;
      if (slot[0] < 0) { slot[0] = i; slot[1] = col[k]; }
      else if (slot[2] < 0) { slot[2] = i; slot[3] = col[k]; }
      else { rc = sfr_set_error(ctx, SFR_NOT_SUPPORTED, "raw value feeds > 2 entries"); goto out; }
      cnt[col[k] + 1]++;
    }
  }
  if (dmap)                             /* storage CSR: diagonal inside al */
    for (int32_t i = 0; i < n; i++) {
      int32_t *slot = rc4 + 4 * (int64_t)dmap[i];
      slot[0] = i; slot[1] = i;
    }
  for (int32_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  tent = (entry *)malloc(sizeof(entry) * (size_t)(cnt[n] + 1));
  if (!tent) { rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation"); goto out; }
  {
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!pos) { rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation"); goto out; }
    for (int32_t i = 0; i < n; i++) pos[i] = cnt[i];
    for (int32_t i = 0; i < n; i++) {
      const int64_t b = ptr ? ptr[i] : 0, e = ptr ? ptr[i + 1] : A->slots;
      for (int64_t q = b; q < e; q++) {
        const int64_t k = ptr ? q : q * n + i;
        if (map[k] < 0) continue;
        tent[pos[col[k]]++] = (entry){i, map[k]};   /* row col[k] of A^T */
      }
    }
    free(pos);
  }
  for (int32_t i = 0; i < n; i++)
    qsort(tent + cnt[i], (size_t)(cnt[i + 1] - cnt[i]), sizeof(entry), cmp_entry);
  {
    sfr_matrix *T = (sfr_matrix *)calloc(1, sizeof *T);
    if (!T) { rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation"); goto out; }
    T->storage = A->storage; T->unsym = A->unsym; T->nad = A->nad;
    T->nraw = A->nraw; T->diag_in_al = A->diag_in_al; T->borrowed = 1;
    rc = upload_format(ctx, T, n, cnt, tent, A->fmt);
    if (rc != SFR_OK) { sfr_matrix_destroy(ctx, T); goto out; }
    T->raw = A->raw; T->raw_ad = A->raw_ad; T->ad = A->ad; T->dinv = A->dinv;
    T->dmap = A->dmap;
    ATRY(sfr_k_gather(ctx->stream, T->nnz, T->vmap, T->raw, T->val));
    A->T = T;
  }
  ATRY(hipMalloc((void **)&A->grad_rc, sizeof(int32_t) * 4 * (size_t)(A->nraw + 1)));
  ATRY(hipMemcpyAsync(A->grad_rc, rc4, sizeof(int32_t) * 4 * (size_t)(A->nraw + 1),
                      hipMemcpyHostToDevice, ctx->stream));
  ATRY(hipStreamSynchronize(ctx->stream));
#undef ATRY
  A->have_grad = 1;
out:
  free(col); free(map); free(ptr); free(dmap); free(rc4); free(cnt); free(tent);
  return rc;
}

int sfr_matrix_set_values(sfr_context *ctx, sfr_matrix *A, const double *al,
                          const double *ad, const double *au)
{
  if (!ctx || !A) return SFR_BAD_ARG;
  if ((A->nraw > 0 && !al) || (!A->diag_in_al && !ad))
    return sfr_set_error(ctx, SFR_BAD_ARG, "missing values");
  const int csrc_unsym = (A->storage == SFR_CSRC || A->storage == SFR_CSRCCOO) && A->unsym;
  if (csrc_unsym) {
    TRY(hipMemcpyAsync(A->raw, al, sizeof(double) * (size_t)A->nad, hipMemcpyHostToDevice, ctx->stream));
    TRY(hipMemcpyAsync(A->raw + A->nad, au ? au : al + A->nad, sizeof(double) * (size_t)A->nad,
                       hipMemcpyHostToDevice, ctx->stream));
  } else if (A->nraw > 0) {
    TRY(hipMemcpyAsync(A->raw, al, sizeof(double) * (size_t)A->nraw, hipMemcpyHostToDevice, ctx->stream));
  }
  if (!A->diag_in_al)
    TRY(hipMemcpyAsync(A->raw_ad, ad, sizeof(double) * (size_t)A->n, hipMemcpyHostToDevice, ctx->stream));
  CHECK_RC(sfr_matrix_refresh(ctx, A));
  TRY(hipStreamSynchronize(ctx->stream));
  return SFR_OK;
}

/* The raw values (and raw diagonal) on the device changed: rebuild the
 * device-format values, the diagonal and its inverse, and mark the FP32
 * copy stale.  Enqueued on the context stream. */
int sfr_matrix_refresh(sfr_context *ctx, sfr_matrix *A)
{
  TRY(sfr_k_gather(ctx->stream, A->nnz, A->vmap, A->raw, A->val));
  if (A->T) TRY(sfr_k_gather(ctx->stream, A->T->nnz, A->T->vmap, A->raw, A->T->val));
  TRY(sfr_k_diag(ctx->stream, A->n, A->diag_in_al ? A->dmap : NULL, A->raw, A->raw_ad, A->ad, A->dinv));
  A->have32 = 0;
  A->have32r = 0;
  A->pair_ok = -1;
  A->pair_use = 0;
  return SFR_OK;
}

int sfr_matrix_pair_check(sfr_context *ctx, sfr_matrix *A)
{
  if (A->pair_kind == SFR_PAIR_NONE || !ctx->ell_pairs) {
    A->pair_use = 0;
    return SFR_OK;
  }
  if (A->pair_ok < 0) {
    sfr_mat_view v = sfr_view(A);
    int np = 0;
    TRY(sfr_k_pair_check(ctx->stream, &v, A->pair_kind, ctx->part, &np));
    double h[1024];
    TRY(hipMemcpyAsync(h, ctx->part, sizeof(double) * (size_t)np,
                       hipMemcpyDeviceToHost, ctx->stream));
    TRY(hipStreamSynchronize(ctx->stream));
    double bad = 0.0;
    for (int k = 0; k < np; k++) bad += h[k];
    A->pair_ok = bad == 0.0;
  }
  A->pair_use = A->pair_ok == 1;
  return SFR_OK;
}

int sfr_matrix_pair_planes(const sfr_matrix *A)
{
  if (!A) return SFR_BAD_ARG;
  return A->pair_use ? A->pair_kind : SFR_PAIR_NONE;
}

/* FP32 copy of the device-format values, unscaled (mixed BiCGSTAB) */
int sfr_matrix_prepare_f32(sfr_context *ctx, sfr_matrix *A)
{
  if (A->have32r) return SFR_OK;
  if (!A->val32r) TRY(hipMalloc((void **)&A->val32r, sizeof(float) * (size_t)(A->nnz + 1)));
  TRY(sfr_k_to_f32(ctx->stream, A->nnz, A->val, A->val32r));
  A->have32r = 1;
  return SFR_OK;
}

/* Mixed BiCGSTAB buffers: 9 FP32 vectors of up to n entries each (grow-only;
 * captured graphs hold the old pointers, so they go when it grows). */
int sfr_mb_buffers(sfr_context *ctx, int64_t n)
{
  if (n <= ctx->mb32_n) return SFR_OK;
  TRY(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  hipFree(ctx->mb32);
  ctx->mb32 = NULL;
  ctx->mb32_n = 0;
  TRY(hipMalloc((void **)&ctx->mb32, sizeof(float) * 9 * (size_t)n));
  ctx->mb32_n = n;
  return SFR_OK;
}

int sfr_matrix_prepare_mixed(sfr_context *ctx, sfr_matrix *A)
{
  if (A->have32) return SFR_OK;
  if (!A->val32) TRY(hipMalloc((void **)&A->val32, sizeof(float) * (size_t)(A->nnz + 1)));
  if (!A->dscale) TRY(hipMalloc((void **)&A->dscale, sizeof(double) * (size_t)A->n));
  sfr_mat_view v = sfr_view(A);
  TRY(sfr_k_scale32(ctx->stream, &v, A->dscale, A->val32));
  A->have32 = 1;
  return SFR_OK;
}

int sfr_matrix_tune(sfr_context *ctx, sfr_matrix *A, int rpt, int lanes)
{
  if (!ctx || !A) return SFR_BAD_ARG;
  if (rpt != 0 && rpt != 1 && rpt != 2 && rpt != 4 && rpt != 8)
    return sfr_set_error(ctx, SFR_BAD_ARG, "rows per thread %d", rpt);
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
    return sfr_set_error(ctx, SFR_BAD_ARG, "lanes per row %d", lanes);
  if (A->fmt == SFR_FMT_CSR && A->lanes != lanes) {
    A->lanes = lanes;
    /* the loop graphs of A were captured with the old lanes per row (the
     * graph key holds them too; this frees the stale ones at once) */
    sfr_drop_graphs_keyed(ctx, 0, A);
  }
  if (rpt == 0) return sfr_tune_pcg(ctx, A);
  A->rpt = rpt;
  A->rpt_pcg = 0;
  A->rpt_set = 1;
  return SFR_OK;
}

int sfr_matrix_rows_per_thread(const sfr_matrix *A, int *rpt, int *rpt_pcg)
{
  if (!A) return SFR_BAD_ARG;
  if (rpt) *rpt = A->rpt;
  if (rpt_pcg) *rpt_pcg = sfr_pcg_rpt(A);
  return SFR_OK;
}

int sfr_matrix_index_bytes(const sfr_matrix *A)
{
  if (!A) return 0;
  return A->col16 ? 2 : 4;
}

int sfr_matrix_info(const sfr_matrix *A, int32_t *neq, int *format, int *slots,
                    int64_t *nnz)
{
  if (!A) return SFR_BAD_ARG;
  if (neq) *neq = A->n;
  if (format) *format = A->fmt;
  if (slots) *slots = A->slots;
  if (nnz) *nnz = A->nnz;
  return SFR_OK;
}
