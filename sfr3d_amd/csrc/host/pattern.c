/*
 * Equation numbering, sparse index builders and the per-cell value scatter
 * (ndf = 1, sequential branch of the reference):
 *   numEqV1 / numEqV2      reference src/Numeq.c:70-96, 109-135
 *   csrIa / csrJa + sort   reference src/Csr.c:30-96, 202-281, src/Graph.c:114
 *   ellPackJa              reference src/EllPack.c:26-79
 *   dataStruct             reference src/Datastruct.c:28-613
 *   csr / ellPack scatter  reference src/Csr.c:549-975, src/EllPack.c:113-149
 * Rows are counted first and prefix-summed, then filled and sorted; on the
 * reference's inputs (rows met in increasing order) this yields the same
 * ia/ja as its single-pass builder.
 */
#include <stdlib.h>
#include <string.h>
#include "sfr3d_setup.h"

int sfr_numeq_v1(int32_t *id, const int32_t *num, int numel)
{
  if (numel < 0 || !id || !num) return -1;
  int32_t neq = 0;
  for (int i = 0; i < numel; i++) id[num[i] - 1] = ++neq;
  return neq;
}

int sfr_numeq_v2(int32_t *id, const int32_t *num, const short *rt,
                 const short *nFace, int numel, short maxViz)
{
  if (numel < 0 || !id || !num || !rt || !nFace) return -1;
  const int w = maxViz + 1;
  int32_t neq = 0;
  for (int i = 0; i < numel; i++) {
    const int c = num[i] - 1;
    id[c] = (rt[(int64_t)c * w + nFace[c]] > 0) ? -1 : ++neq;
  }
  return neq;
}

static int cmp_i32(const void *a, const void *b)
{
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* Does column col of row `row` belong to the stored part? */
static inline int keep(int32_t col, int32_t row, int lower, int upper)
{
  return (lower && col < row) || (upper && col > row);
}

static int build_csr(sfr_pattern *p, const int32_t *id, const int32_t *num,
                     const int32_t *nelcon, const short *nViz, int numel,
                     short maxViz, int lower, int upper, int diag)
{
  const int32_t neq = p->neq;
  int64_t *cnt = (int64_t *)calloc((size_t)neq + 1, sizeof(int64_t));
  if (!cnt) return -2;
  for (int i = 0; i < numel; i++) {
    const int c = num[i] - 1;
    const int32_t row = id[c] - 1;
    if (row < 0) continue;
    int64_t n = diag ? 1 : 0;
    for (int f = 0; f < nViz[c]; f++) {
      const int32_t nb = nelcon[(int64_t)c * maxViz + f];
      if (nb == -1) continue;
      const int32_t col = id[nb - 1] - 1;
      if (col >= 0 && keep(col, row, lower, upper)) n++;
    }
    cnt[row + 1] = n;
  }
  for (int32_t r = 0; r < neq; r++) cnt[r + 1] += cnt[r];
  if (cnt[neq] > INT32_MAX) { free(cnt); return -3; }
  p->nia = neq + 1;
  p->nja = cnt[neq];
  p->nad = (int32_t)cnt[neq];
  p->ia = (int32_t *)malloc(sizeof(int32_t) * (size_t)(neq + 1));
  p->ja = (int32_t *)malloc(sizeof(int32_t) * (size_t)(p->nja > 0 ? p->nja : 1));
  if (!p->ia || !p->ja) { free(cnt); return -2; }
  for (int32_t r = 0; r <= neq; r++) p->ia[r] = (int32_t)cnt[r];
  for (int i = 0; i < numel; i++) {
    const int c = num[i] - 1;
    const int32_t row = id[c] - 1;
    if (row < 0) continue;
    int64_t k = cnt[row];
    if (diag) p->ja[k++] = row;
    for (int f = 0; f < nViz[c]; f++) {
      const int32_t nb = nelcon[(int64_t)c * maxViz + f];
      if (nb == -1) continue;
      const int32_t col = id[nb - 1] - 1;
      if (col >= 0 && keep(col, row, lower, upper)) p->ja[k++] = col;
    }
  }
  free(cnt);
  for (int32_t r = 0; r < neq; r++)
    qsort(p->ja + p->ia[r], (size_t)(p->ia[r + 1] - p->ia[r]), sizeof(int32_t),
          cmp_i32);
  return 0;
}

static int build_ell(sfr_pattern *p, const int32_t *id, const int32_t *num,
                     const int32_t *nelcon, const short *nViz, int numel,
                     short maxViz)
{
  const int64_t rows = (int64_t)p->neq + 1;       /* one spare row, as n1 */
  p->nia = 2;
  p->nja = rows * maxViz;
  p->ia = (int32_t *)malloc(sizeof(int32_t) * 2);
  p->ja = (int32_t *)calloc((size_t)p->nja, sizeof(int32_t));
  if (!p->ia || !p->ja) return -2;
  if (p->nja > INT32_MAX) return -3;
  p->ia[0] = maxViz;
  p->ia[1] = (int32_t)p->nja;
  int64_t nad = 0;
  for (int i = 0; i < numel; i++) {
    const int c = num[i] - 1;
    const int32_t row = id[c] - 1;
    if (row < 0) continue;
    int32_t *slot = p->ja + (int64_t)row * maxViz;
    for (int f = 0; f < nViz[c]; f++) {
      const int32_t nb = nelcon[(int64_t)c * maxViz + f];
      if (nb == -1) { slot[f] = row; continue; }       /* boundary face */
      const int32_t col = id[nb - 1] - 1;
      if (col >= 0) { slot[f] = col; nad++; }          /* else stays 0 */
    }
  }
  p->nad = (int32_t)nad;
  return 0;
}

int sfr_datastruct(sfr_pattern *p, short storage, int unsym,
                   const int32_t *id, const int32_t *num,
                   const int32_t *nelcon, const short *nViz,
                   int numel, short maxViz, int neq)
{
  if (!p || !id || !num || !nelcon || !nViz || neq < 0) return -1;
  memset(p, 0, sizeof *p);
  p->neq = neq;
  p->storage = storage;
  p->unsym = (short)(unsym != 0);
  p->maxViz = maxViz;
  int rc;
  switch (storage) {
  case SFR_CSR:                     /* diag + upper, + lower when unsym */
    rc = build_csr(p, id, num, nelcon, nViz, numel, maxViz, unsym, 1, 1);
    p->nval = p->nad;
    break;
  case SFR_CSRD: case SFR_CSRDCOO:  /* lower, + upper when unsym */
    rc = build_csr(p, id, num, nelcon, nViz, numel, maxViz, 1, unsym, 0);
    p->nval = p->nad;
    break;
  case SFR_CSRC: case SFR_CSRCCOO:  /* lower; values [al|au] when unsym */
    rc = build_csr(p, id, num, nelcon, nViz, numel, maxViz, 1, 0, 0);
    p->nval = unsym ? 2 * (int64_t)p->nad : p->nad;
    break;
  case SFR_ELLPACK:
    rc = build_ell(p, id, num, nelcon, nViz, numel, maxViz);
    p->nval = p->nja;
    break;
  default:
    return -4;
  }
  if (rc) sfr_pattern_free(p);
  return rc;
}

void sfr_pattern_free(sfr_pattern *p)
{
  if (!p) return;
  free(p->ia); free(p->ja);
  p->ia = p->ja = NULL;
}

/* write v at every entry of row r whose column is col (the reference scans
 * the whole row and overwrites every match) */
static void put_row(const int32_t *ia, const int32_t *ja, int32_t r,
                    int32_t col, double v, double *arr)
{
  for (int64_t k = ia[r]; k < ia[r + 1]; k++)
    if (ja[k] == col) arr[k] = v;
}

int sfr_assemble(const sfr_pattern *p, const int32_t *id,
                 const int32_t *nelcon, const short *nViz, int numel,
                 const double *lA, const double *lB,
                 double *al, double *ad, double *b, int ell_quirk)
{
  if (!p || !id || !nelcon || !nViz || !lA) return -1;
  const int s = p->maxViz, w = s + 1;
  for (int c = 0; c < numel; c++) {
    const int32_t row = id[c] - 1;
    if (row < 0) continue;
    const double *A = lA + (int64_t)c * w;
    if (b && lB) b[row] = lB[c];
    switch (p->storage) {
    case SFR_ELLPACK: {
      if (ad) ad[row] = A[s];
      if (!al) break;
      double *a = al + (int64_t)row * s;
      for (int f = 0; f < nViz[c]; f++) {
        const int32_t nb = nelcon[(int64_t)c * s + f];
        const int has_eq = nb != -1 && id[nb - 1] > 0;
        a[f] = (ell_quirk || has_eq) ? A[f] : 0.0;
      }
      break;
    }
    case SFR_CSR: {
      /* diagonal lives inside al; ad aliases al in the reference */
      if (!al) break;
      put_row(p->ia, p->ja, row, row, A[s], al);
      for (int f = 0; f < nViz[c]; f++) {
        const int32_t nb = nelcon[(int64_t)c * s + f];
        if (nb == -1) continue;
        const int32_t col = id[nb - 1] - 1;
        if (col < 0) continue;
        put_row(p->ia, p->ja, row, col, A[f], al);
      }
      break;
    }
    default: {                             /* CSRD / CSRC (+COO twins) */
      if (ad) ad[row] = A[s];
      if (!al) break;
      const int csrc_unsym =
          (p->storage == SFR_CSRC || p->storage == SFR_CSRCCOO) && p->unsym;
      double *au = al + p->nad;
      for (int f = 0; f < nViz[c]; f++) {
        const int32_t nb = nelcon[(int64_t)c * s + f];
        if (nb == -1) continue;
        const int32_t col = id[nb - 1] - 1;
        if (col < 0) continue;
        if (csrc_unsym && col > row) {
          /* upper coefficient kept in row col, at column row, of au */
          put_row(p->ia, p->ja, col, row, A[f], au);
        } else {
          put_row(p->ia, p->ja, row, col, A[f], al);
        }
      }
      break;
    }
    }
  }
  return 0;
}
