/*
 * Device assembly: the value scatter of sfr_assemble (the reference's
 * assbly / csr / ellPack, src/Assbly.c:43-95, src/Csr.c:549-975,
 * src/EllPack.c:113-149) replayed on the device into a resident matrix.
 *
 * The plan is made once per pattern by running the host assembler itself on
 * source indices instead of values (lA[i] = i + 1, lB[c] = c + 1, outputs
 * pre-set to -1): every stored value then holds the 1-based index of the
 * coefficient that lands in it (the last writer, as on the host), 0 where
 * the host writes 0.0 (an ELL slot facing a cell without an equation), or
 * -1 where the host leaves the value alone.  The device scatter therefore
 * writes bitwise what sfr_assemble would write, with no arithmetic.
 */
#include <stdint.h>
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

struct sfr_assembly {
  sfr_matrix *A;
  int64_t     nsrc;        /* numel * (maxViz + 1) coefficients             */
  int         numel;
  int32_t    *amap;        /* A->nraw: raw off-diagonal values              */
  int32_t    *dmap;        /* n: raw diagonal (NULL when inside al)         */
  int32_t    *bmap;        /* n: right-hand side                            */
};

static int32_t *to_map(const double *v, int64_t n)
{
  int32_t *m = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!m) return NULL;
  for (int64_t k = 0; k < n; k++) m[k] = v[k] < 0.0 ? -1 : (int32_t)v[k];
  return m;
}

void sfr_assembly_destroy(sfr_context *ctx, sfr_assembly *P)
{
  if (!P) return;
  if (ctx) hipStreamSynchronize(ctx->stream);
  hipFree(P->amap);
  hipFree(P->dmap);
  hipFree(P->bmap);
  free(P);
}

int sfr_assembly_create(sfr_context *ctx, sfr_matrix *A, const sfr_pattern *pat,
                        const int32_t *id, const int32_t *nelcon,
                        const short *nViz, int numel, int ell_quirk,
                        sfr_assembly **out)
{
  if (!ctx || !A || !pat || !id || !nelcon || !nViz || !out || numel <= 0)
    return SFR_BAD_ARG;
  *out = NULL;
  if (pat->neq != A->n || pat->storage != A->storage || pat->nad != A->nad ||
      (int)(pat->unsym != 0) != (int)(A->unsym != 0))
    return sfr_set_error(ctx, SFR_BAD_ARG,
                         "pattern (neq %d, storage %d, nad %d) is not the matrix's "
                         "(neq %d, storage %d, nad %d)", pat->neq, pat->storage,
                         pat->nad, A->n, A->storage, A->nad);
  const int w = pat->maxViz + 1;
  const int64_t nsrc = (int64_t)numel * w;
  if (nsrc >= INT32_MAX) return sfr_set_error(ctx, SFR_BAD_ARG, "too many cells");
  const int64_t nval = pat->nval > A->nraw ? pat->nval : A->nraw;
  const int32_t n = A->n;
  double *lA = (double *)malloc(sizeof(double) * (size_t)nsrc);
  double *lB = (double *)malloc(sizeof(double) * (size_t)numel);
  double *al = (double *)malloc(sizeof(double) * (size_t)(nval + 1));
  double *ad = (double *)malloc(sizeof(double) * ((size_t)n + 1));
  double *b = (double *)malloc(sizeof(double) * ((size_t)n + 1));
  int32_t *am = NULL, *dm = NULL, *bm = NULL;
  sfr_assembly *P = (sfr_assembly *)calloc(1, sizeof *P);
  int rc = SFR_OK;
  if (!lA || !lB || !al || !ad || !b || !P) {
    rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation");
    goto out;
  }
  for (int64_t i = 0; i < nsrc; i++) lA[i] = (double)(i + 1);
  for (int c = 0; c < numel; c++) lB[c] = (double)(c + 1);
  for (int64_t k = 0; k <= nval; k++) al[k] = -1.0;
  for (int32_t i = 0; i <= n; i++) { ad[i] = -1.0; b[i] = -1.0; }
  if (sfr_assemble(pat, id, nelcon, nViz, numel, lA, lB, al,
                   A->diag_in_al ? NULL : ad, b, ell_quirk) != 0) {
    rc = sfr_set_error(ctx, SFR_BAD_ARG, "assembly plan: bad arguments");
    goto out;
  }
  am = to_map(al, A->nraw);
  dm = to_map(ad, n);
  bm = to_map(b, n);
  if (!am || !dm || !bm) {
    rc = sfr_set_error(ctx, SFR_BAD_ARG, "host allocation");
    goto out;
  }
  P->A = A;
  P->nsrc = nsrc;
  P->numel = numel;
#define PTRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
    rc = sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s", #call, hipGetErrorString(e_)); goto out; } } while (0)
  PTRY(hipMalloc((void **)&P->amap, sizeof(int32_t) * (size_t)(A->nraw + 1)));
  PTRY(hipMalloc((void **)&P->bmap, sizeof(int32_t) * (size_t)n));
  PTRY(hipMemcpy(P->amap, am, sizeof(int32_t) * (size_t)A->nraw, hipMemcpyHostToDevice));
  PTRY(hipMemcpy(P->bmap, bm, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
  if (!A->diag_in_al) {
    PTRY(hipMalloc((void **)&P->dmap, sizeof(int32_t) * (size_t)n));
    PTRY(hipMemcpy(P->dmap, dm, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
  }
#undef PTRY
out:
  free(lA); free(lB); free(al); free(ad); free(b);
  free(am); free(dm); free(bm);
  if (rc != SFR_OK) { sfr_assembly_destroy(ctx, P); return rc; }
  *out = P;
  return SFR_OK;
}

int sfr_assemble_device(sfr_context *ctx, sfr_assembly *P, const double *d_lA,
                        const double *d_lB, double *d_b)
{
  if (!ctx || !P || !d_lA) return SFR_BAD_ARG;
  if ((d_b == NULL) != (d_lB == NULL))
    return sfr_set_error(ctx, SFR_BAD_ARG, "d_lB and d_b go together");
  TRY(hipSetDevice(ctx->device));
  sfr_matrix *A = P->A;
  hipStream_t s = ctx->stream;
  TRY(sfr_k_scatter(s, A->nraw, P->amap, d_lA, A->raw));
  if (P->dmap) TRY(sfr_k_scatter(s, A->n, P->dmap, d_lA, A->raw_ad));
  if (d_b) TRY(sfr_k_scatter(s, A->n, P->bmap, d_lB, d_b));
  const int rc = sfr_matrix_refresh(ctx, A);
  if (rc != SFR_OK) return rc;
  TRY(hipStreamSynchronize(s));
  return SFR_OK;
}
