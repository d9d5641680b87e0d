/*
 * Cell reordering: the cell graph of reord (reference src/Reord.c:18-112,
 * convGraph src/Graph.c:24, rows sorted ascending as sortGraphCsr/bubblesort
 * src/Graph.c:114, src/HccaSort.c:18-35) and reverse Cuthill-McKee following
 * George & Liu's GENRCM (A. George, J. Liu, "Computer Solution of Large
 * Sparse Positive Definite Systems", 1981), the variant the reference ships
 * in src/Rcm.c:56-887.  The permutation must match that routine exactly, so
 * the same choices are made wherever the algorithm leaves one open:
 *   - components are started from the lowest-numbered unnumbered node;
 *   - the pseudo-peripheral root search (Gibbs-Poole-Stockmeyer style) takes,
 *     in the last level of the current rooted level structure, the FIRST node
 *     of minimum degree (degree counted over unnumbered nodes), and stops as
 *     soon as the number of levels does not grow (src/Rcm.c:217-320);
 *   - BFS visits neighbours in adjacency order; the children of each node are
 *     ordered by a stable sort on degree (src/Rcm.c:636-686);
 *   - each component's Cuthill-McKee order is reversed on its own.
 * The code below is a 0-based implementation with its own data layout.
 */
#include <stdlib.h>
#include <string.h>
#include "sfr3d_setup.h"

/* Rooted level structure over the unnumbered nodes (live[v] != 0).
 * Fills ls[] in BFS order and lstart[0..nlev]; returns the component size. */
static int level_structure(int root, const int64_t *xadj, const int32_t *adj,
                           unsigned char *live, int32_t *ls, int32_t *lstart,
                           int *nlev)
{
  int size = 1, lev = 0, begin = 0;
  ls[0] = root;
  live[root] = 0;
  for (;;) {
    const int end = size;
    lstart[lev++] = begin;
    for (int i = begin; i < end; i++) {
      const int v = ls[i];
      for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
        const int w = adj[e];
        if (live[w]) { live[w] = 0; ls[size++] = w; }
      }
    }
    begin = end;
    if (size == end) break;           /* the level just scanned added nothing */
  }
  lstart[lev] = size;
  *nlev = lev;
  for (int i = 0; i < size; i++) live[ls[i]] = 1;
  return size;
}

static int live_degree(int v, const int64_t *xadj, const int32_t *adj,
                       const unsigned char *live)
{
  int d = 0;
  for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) d += live[adj[e]] != 0;
  return d;
}

static int pseudo_peripheral(int root, const int64_t *xadj, const int32_t *adj,
                             unsigned char *live, int32_t *ls, int32_t *lstart)
{
  int nlev, nlev2;
  const int size = level_structure(root, xadj, adj, live, ls, lstart, &nlev);
  if (nlev == 1 || nlev == size) return root;
  for (;;) {
    const int first = lstart[nlev - 1];
    int best = ls[first], bestdeg = size;
    for (int j = first; j < size; j++) {
      const int d = live_degree(ls[j], xadj, adj, live);
      if (d < bestdeg) { bestdeg = d; best = ls[j]; }
    }
    root = best;
    level_structure(root, xadj, adj, live, ls, lstart, &nlev2);
    if (nlev2 <= nlev) break;
    nlev = nlev2;
    if (size <= nlev) break;
  }
  return root;
}

/* Cuthill-McKee order of root's component into out[]; numbers the nodes
 * (live = 0) and returns the component size.  deg is scratch. */
static int cuthill_mckee(int root, const int64_t *xadj, const int32_t *adj,
                         unsigned char *live, int32_t *deg, int32_t *out)
{
  /* degrees over unnumbered nodes for the whole component (BFS once) */
  int size = 1;
  out[0] = root;
  live[root] = 2;                        /* 2 = discovered in this pass */
  for (int i = 0; i < size; i++) {
    const int v = out[i];
    int d = 0;
    for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
      const int w = adj[e];
      if (live[w]) {
        d++;
        if (live[w] == 1) { live[w] = 2; out[size++] = w; }
      }
    }
    deg[v] = d;
  }
  for (int i = 0; i < size; i++) live[out[i]] = 1;

  live[root] = 0;
  if (size == 1) return 1;
  int tail = 1;
  for (int head = 0; head < tail; head++) {
    const int v = out[head];
    const int first = tail;
    for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
      const int w = adj[e];
      if (live[w]) { live[w] = 0; out[tail++] = w; }
    }
    /* stable insertion sort of the new children by degree */
    for (int k = first + 1; k < tail; k++) {
      const int w = out[k];
      int l = k - 1;
      while (l >= first && deg[out[l]] > deg[w]) { out[l + 1] = out[l]; l--; }
      out[l + 1] = w;
    }
  }
  return size;
}

int sfr_genrcm(int n, const int64_t *xadj, const int32_t *adj, int32_t *perm)
{
  if (n < 1 || !xadj || !adj || !perm) return -1;
  unsigned char *live = (unsigned char *)malloc((size_t)n);
  int32_t *ls = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  int32_t *lstart = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
  int32_t *deg = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  if (!live || !ls || !lstart || !deg) {
    free(live); free(ls); free(lstart); free(deg);
    return -2;
  }
  memset(live, 1, (size_t)n);
  int placed = 0;
  for (int v = 0; v < n && placed < n; v++) {
    if (!live[v]) continue;
    const int root = pseudo_peripheral(v, xadj, adj, live, ls, lstart);
    int32_t *out = perm + placed;
    const int size = cuthill_mckee(root, xadj, adj, live, deg, out);
    for (int i = 0, j = size - 1; i < j; i++, j--) {
      const int32_t t = out[i]; out[i] = out[j]; out[j] = t;
    }
    placed += size;
  }
  free(live); free(ls); free(lstart); free(deg);
  return 0;
}

static int cmp_i32(const void *a, const void *b)
{
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

int sfr_reord(int32_t *num, const int32_t *nelcon, const short *nViz,
              short maxViz, int numel, int flag)
{
  if (numel < 0 || !num) return -1;
  if (!flag) {
    for (int i = 0; i < numel; i++) num[i] = i + 1;
    return 0;
  }
  if (!nelcon || !nViz) return -1;
  /* cell graph: neighbours across faces, boundary faces dropped */
  int64_t *xadj = (int64_t *)malloc(sizeof(int64_t) * ((size_t)numel + 1));
  if (!xadj) return -2;
  xadj[0] = 0;
  for (int c = 0; c < numel; c++) {
    int64_t cnt = 0;
    for (int f = 0; f < nViz[c]; f++)
      cnt += nelcon[(int64_t)c * maxViz + f] != -1;
    xadj[c + 1] = xadj[c] + cnt;
  }
  int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(xadj[numel] + 1));
  int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)numel);
  if (!adj || !perm) { free(xadj); free(adj); free(perm); return -2; }
  for (int c = 0; c < numel; c++) {
    int64_t k = xadj[c];
    for (int f = 0; f < nViz[c]; f++) {
      const int32_t nb = nelcon[(int64_t)c * maxViz + f];
      if (nb != -1) adj[k++] = nb - 1;
    }
    qsort(adj + xadj[c], (size_t)(xadj[c + 1] - xadj[c]), sizeof(int32_t),
          cmp_i32);
  }
  int rc = numel ? sfr_genrcm(numel, xadj, adj, perm) : 0;
  if (rc == 0)
    for (int i = 0; i < numel; i++) num[i] = perm[i] + 1;
  free(xadj); free(adj); free(perm);
  return rc;
}
