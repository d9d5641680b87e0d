/*
 * Cell adjacency in SFR3D's face order.
 *
 * Structured generators write the neighbour across every face in closed form;
 * the general builder matches faces by their node sets, which is what
 * adj2d/adj3d (reference src/Adjcency.c:519-700, 702-) compute on a
 * conforming mesh.  Output convention (reference): 1-based neighbour, -1 when
 * the face is on the boundary.
 */
#include <stdlib.h>
#include <string.h>
#include "sfr3d_setup.h"

int sfr_mesh_quad_adj(int nx, int ny, int32_t *nelcon, short *nViz)
{
  if (nx < 1 || ny < 1 || !nelcon || !nViz) return -1;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int c = i + nx * j;
      int32_t *f = nelcon + 4 * (int64_t)c;
      f[0] = (j > 0)      ? c - nx + 1 : -1;   /* S: nodes SW-SE */
      f[1] = (i < nx - 1) ? c + 1 + 1  : -1;   /* E: nodes SE-NE */
      f[2] = (j < ny - 1) ? c + nx + 1 : -1;   /* N: nodes NE-NW */
      f[3] = (i > 0)      ? c - 1 + 1  : -1;   /* W: nodes NW-SW */
      nViz[c] = 4;
    }
  return 0;
}

int sfr_mesh_hex_adj(int nx, int ny, int nz, int32_t *nelcon, short *nViz)
{
  if (nx < 1 || ny < 1 || nz < 1 || !nelcon || !nViz) return -1;
  const int64_t sxy = (int64_t)nx * ny;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int64_t c = i + nx * (int64_t)j + sxy * k;
        int32_t *f = nelcon + 6 * c;
        f[0] = (k > 0)      ? (int32_t)(c - sxy + 1) : -1;  /* z- {0,3,2,1} */
        f[1] = (k < nz - 1) ? (int32_t)(c + sxy + 1) : -1;  /* z+ {4,5,6,7} */
        f[2] = (j > 0)      ? (int32_t)(c - nx + 1)  : -1;  /* y- {0,1,5,4} */
        f[3] = (i < nx - 1) ? (int32_t)(c + 1 + 1)   : -1;  /* x+ {1,2,6,5} */
        f[4] = (j < ny - 1) ? (int32_t)(c + nx + 1)  : -1;  /* y+ {2,3,7,6} */
        f[5] = (i > 0)      ? (int32_t)(c - 1 + 1)   : -1;  /* x- {3,0,4,7} */
        nViz[c] = 6;
      }
  return 0;
}

/* ---- general conforming meshes: faces matched through a hash table ---- */

typedef struct { int32_t key[4]; int32_t cell; short face; short used; } face_slot;

static const short tri_faces[3][2]  = {{0,1},{1,2},{2,0}};
static const short quad_faces[4][2] = {{0,1},{1,2},{2,3},{3,0}};
static const short hex_faces[6][4]  = {{0,3,2,1},{4,5,6,7},{0,1,5,4},
                                       {1,2,6,5},{2,3,7,6},{3,0,4,7}};

static void sort4(int32_t *k, int n)
{
  for (int a = 1; a < n; a++) {
    int32_t v = k[a]; int b = a - 1;
    while (b >= 0 && k[b] > v) { k[b + 1] = k[b]; b--; }
    k[b + 1] = v;
  }
}

static uint64_t hash_key(const int32_t *k)
{
  uint64_t h = 0x9E3779B97F4A7C15ull;
  for (int a = 0; a < 4; a++) {
    h ^= (uint64_t)(uint32_t)k[a] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  }
  h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 27;
  return h;
}

int sfr_mesh_adj_general(int ndm, int numel, int maxNo, const int32_t *el,
                         const short *nen, int maxViz, int32_t *nelcon,
                         short *nViz)
{
  if (numel < 0 || !el || !nen || !nelcon || !nViz) return -1;
  int64_t nfaces_total = 0;
  for (int c = 0; c < numel; c++) {
    int nf;
    if (ndm == 2) nf = nen[c];                 /* triangles 3, quads 4 */
    else if (nen[c] == 8) nf = 6;              /* hexahedra */
    else return -2;                            /* other 3-D cells: not needed */
    if (nf > maxViz) return -3;
    nViz[c] = (short)nf;
    nfaces_total += nf;
    for (int f = 0; f < maxViz; f++) nelcon[(int64_t)c * maxViz + f] = -1;
  }
  int64_t cap = 1;
  while (cap < 2 * nfaces_total + 16) cap <<= 1;
  face_slot *tab = (face_slot *)calloc((size_t)cap, sizeof(face_slot));
  if (!tab) return -4;
  for (int c = 0; c < numel; c++) {
    const int32_t *nodes = el + (int64_t)c * maxNo;
    for (int f = 0; f < nViz[c]; f++) {
      int32_t key[4] = {-1, -1, -1, -1};
      int nk;
      if (ndm == 2) {
        const short (*tab2)[2] = (nen[c] == 3) ? tri_faces : quad_faces;
        key[0] = nodes[tab2[f][0]]; key[1] = nodes[tab2[f][1]]; nk = 2;
      } else {
        for (int a = 0; a < 4; a++) key[a] = nodes[hex_faces[f][a]];
        nk = 4;
      }
      sort4(key, nk);
      uint64_t h = hash_key(key) & (uint64_t)(cap - 1);
      for (;;) {
        face_slot *s = &tab[h];
        if (!s->used) {
          memcpy(s->key, key, sizeof key); s->cell = c; s->face = (short)f;
          s->used = 1; break;
        }
        if (memcmp(s->key, key, sizeof key) == 0 && s->used == 1) {
          /* second owner of the face: link both sides (1-based) */
          nelcon[(int64_t)c * maxViz + f] = s->cell + 1;
          nelcon[(int64_t)s->cell * maxViz + s->face] = c + 1;
          s->used = 2; break;
        }
        h = (h + 1) & (uint64_t)(cap - 1);
      }
    }
  }
  free(tab);
  return 0;
}
