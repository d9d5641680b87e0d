/*
 * MatrixMarket output and input of assembled systems.
 *
 * Output follows the reference's writeCoo / csrToCoo / writeCooB
 * (src/WriteMtx.c:26-176) and the file layout of the Matrix Market I/O
 * library it calls (mm_write_mtx_crd, mmio/lib/mmio.c): banner
 * "%%MatrixMarket matrix coordinate real general|symmetric", "M N nz", then
 * one "%d %d %30.14e" line per entry, 1-based; every row is written as its
 * diagonal followed by the row's stored entries, nz = neq + nad.  Values:
 * al for a column below the diagonal, above it (unsym) au when the caller
 * passes one -- the reference's argument -- else al itself, since CSRD keeps
 * full off-diagonal rows in al (dataStruct never sets au for CSRD); 0 above
 * the diagonal of a symmetric matrix; 1.0 everywhere in a "bin"
 * (pattern-only) file.  Storage CSRD (and its COO twin) only, as in the
 * reference.
 *
 * Input reads any real coordinate file (general or symmetric) into CSRD:
 * a symmetric file becomes strictly-lower sorted rows (unsym = 0), a general
 * one full off-diagonal sorted rows (unsym = 1); the diagonal goes to ad.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../sfr_internal.h"

int sfr_write_mtx(const char *path, const sfr_system *s, int bin)
{
  if (!path || !s || !s->ia || !s->ja || s->neq <= 0) return SFR_BAD_ARG;
  if (s->storage != SFR_CSRD && s->storage != SFR_CSRDCOO) return SFR_NOT_SUPPORTED;
  if (!bin && (!s->ad || (s->nad > 0 && !s->al))) return SFR_BAD_ARG;
  FILE *f = fopen(path, "w");
  if (!f) return SFR_BAD_ARG;
  const int32_t n = s->neq;
  const double *au = s->au ? s->au : s->al;
  fprintf(f, "%%%%MatrixMarket matrix coordinate real %s\n",
          s->unsym ? "general" : "symmetric");
  fprintf(f, "%d %d %d\n", n, n, n + s->nad);
  for (int32_t i = 0; i < n; i++) {
    fprintf(f, "%d %d %30.14e\n", i + 1, i + 1, bin ? 1.0 : s->ad[i]);
    for (int32_t k = s->ia[i]; k < s->ia[i + 1]; k++) {
      const int32_t j = s->ja[k];
      double v = 0.0;
      if (bin) v = 1.0;
      else if (j < i) v = s->al[k];
      else if (j > i && s->unsym) v = au[k];
      fprintf(f, "%d %d %30.14e\n", i + 1, j + 1, v);
    }
  }
  return fclose(f) == 0 ? SFR_OK : SFR_BAD_ARG;
}

int sfr_write_mtx_vector(const char *path, const double *b, int32_t neq)
{
  if (!path || !b || neq <= 0) return SFR_BAD_ARG;
  FILE *f = fopen(path, "w");
  if (!f) return SFR_BAD_ARG;
  fprintf(f, "%%%%MatrixMarket matrix array real general\n");
  fprintf(f, "%d %d\n", neq, 1);
  for (int32_t i = 0; i < neq; i++) fprintf(f, "%.14e\n", b[i]);
  return fclose(f) == 0 ? SFR_OK : SFR_BAD_ARG;
}

void sfr_mtx_free(sfr_mtx *m)
{
  if (!m) return;
  free(m->ia); free(m->ja); free(m->al); free(m->ad);
  memset(m, 0, sizeof *m);
}

typedef struct { int32_t i, j; double v; } coo;

static int cmp_coo(const void *a, const void *b)
{
  const coo *x = (const coo *)a, *y = (const coo *)b;
  if (x->i != y->i) return (x->i > y->i) - (x->i < y->i);
  return (x->j > y->j) - (x->j < y->j);
}

int sfr_read_mtx(const char *path, sfr_mtx *out)
{
  if (!path || !out) return SFR_BAD_ARG;
  memset(out, 0, sizeof *out);
  FILE *f = fopen(path, "r");
  if (!f) return SFR_BAD_ARG;
  char line[1100], obj[64], fmt[64], field[64], sym[64];
  int rc = SFR_BAD_ARG;
  coo *e = NULL;
  if (!fgets(line, sizeof line, f) ||
      sscanf(line, "%%%%MatrixMarket %63s %63s %63s %63s", obj, fmt, field, sym) != 4 ||
      strcmp(obj, "matrix") || strcmp(fmt, "coordinate") ||
      (strcmp(field, "real") && strcmp(field, "integer")) ||
      (strcmp(sym, "general") && strcmp(sym, "symmetric")))
    goto done;
  const int symmetric = !strcmp(sym, "symmetric");
  long long M = 0, N = 0, nz = 0;
  do {
    if (!fgets(line, sizeof line, f)) goto done;
  } while (line[0] == '%');
  if (sscanf(line, "%lld %lld %lld", &M, &N, &nz) != 3 || M != N || M <= 0 ||
      M > INT32_MAX || nz < 0)
    goto done;
  e = (coo *)malloc(sizeof(coo) * (size_t)(nz > 0 ? nz : 1));
  out->ad = (double *)calloc((size_t)M, sizeof(double));
  if (!e || !out->ad) goto done;
  long long m = 0;
  for (long long k = 0; k < nz; k++) {
    long long i, j;
    double v;
    if (fscanf(f, "%lld %lld %lf", &i, &j, &v) != 3 || i < 1 || j < 1 || i > M || j > M)
      goto done;
    if (i == j) { out->ad[i - 1] = v; continue; }
    if (symmetric && j > i) { long long t = i; i = j; j = t; }   /* keep lower */
    e[m].i = (int32_t)(i - 1); e[m].j = (int32_t)(j - 1); e[m].v = v;
    m++;
  }
  qsort(e, (size_t)m, sizeof(coo), cmp_coo);
  out->neq = (int32_t)M;
  out->nad = (int32_t)m;
  out->unsym = symmetric ? 0 : 1;
  out->ia = (int32_t *)calloc((size_t)M + 1, sizeof(int32_t));
  out->ja = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  out->al = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  if (!out->ia || !out->ja || !out->al) goto done;
  for (long long k = 0; k < m; k++) {
    out->ia[e[k].i + 1]++;
    out->ja[k] = e[k].j;
    out->al[k] = e[k].v;
  }
  for (int32_t i = 0; i < out->neq; i++) out->ia[i + 1] += out->ia[i];
  rc = SFR_OK;
done:
  free(e);
  fclose(f);
  if (rc != SFR_OK) sfr_mtx_free(out);
  return rc;
}
