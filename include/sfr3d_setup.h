/*
 * sfr3d_setup.h -- host-side set-up of the linear systems that SFR3D's cell
 * loops hand to its solvers: mesh adjacency in SFR3D's face order, cell
 * reordering (reverse Cuthill-McKee), equation numbering, the CSR/CSRD/CSRC
 * and ELLPACK index builders and the per-cell value scatter.
 *
 * Every array follows SFR3D's own conventions so the results can be compared
 * bit for bit with the reference's builders:
 *   nelcon[c*maxViz+f]  1-based neighbour of cell c across face f, -1 = none
 *   num[i]              1-based cell placed at position i of the ordering
 *   id[c]               1-based equation of cell c, -1 = no equation
 *   ia/ja               0-based, int32
 * Only one degree of freedom per cell (ndf = 1) is supported; SFR3D's
 * solverC callers (pressure, momentum components, energy, species) all use it.
 */
#ifndef SFR3D_SETUP_H
#define SFR3D_SETUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* storage ids, identical to include/Define.h:407-412 of the reference */
#define SFR_CSR     1
#define SFR_CSRD    2
#define SFR_CSRC    3
#define SFR_ELLPACK 4
#define SFR_CSRDCOO 5
#define SFR_CSRCCOO 6

/* ---------------------------------------------------------------------
 * Mesh adjacency.
 * Quad cells: nodes (SW, SE, NE, NW) as script/retangular_mesh_mvf.py:53-58
 * writes them; face f joins nodes {f, f+1 mod 4}, i.e. faces S, E, N, W
 * (src/Adjcency.c:39).  Cell (i,j) of an nx*ny grid is number i + nx*j.
 * Hex cells: nodes 0-3 on the bottom (z-) face counter-clockwise from the
 * (x-,y-) corner, 4-7 above them; faces {0,3,2,1},{4,5,6,7},{0,1,5,4},
 * {1,2,6,5},{2,3,7,6},{3,0,4,7} (src/Adjcency.c:1336-1341), i.e. faces
 * z-, z+, y-, x+, y+, x-.  Cell (i,j,k) is number i + nx*(j + ny*k).
 * --------------------------------------------------------------------- */
int sfr_mesh_quad_adj(int nx, int ny, int32_t *nelcon, short *nViz);
int sfr_mesh_hex_adj(int nx, int ny, int nz, int32_t *nelcon, short *nViz);
/* General conforming mesh: cells given by 1-based node lists el[c*maxNo+..]
 * with nen[c] nodes (3/4 in 2-D, 8 for hexes in 3-D).  Neighbour across face
 * f is the other cell owning the same face; the result matches adj2d/adj3d
 * (src/Adjcency.c:519-700, 702-) on conforming meshes. */
int sfr_mesh_adj_general(int ndm, int numel, int maxNo, const int32_t *el,
                         const short *nen, int maxViz, int32_t *nelcon,
                         short *nViz);

/* ---------------------------------------------------------------------
 * Reordering: reord (src/Reord.c:18-112).  flag = 0 gives the identity
 * (num[i] = i+1); flag = 1 builds the cell graph (convGraph, sorted rows)
 * and runs George & Liu's GENRCM with its tie-breaks (src/Rcm.c:56-887).
 * --------------------------------------------------------------------- */
int sfr_reord(int32_t *num, const int32_t *nelcon, const short *nViz,
              short maxViz, int numel, int flag);
/* GENRCM on a 0-based CSR graph; perm receives 0-based node ids. */
int sfr_genrcm(int node_num, const int64_t *xadj, const int32_t *adj,
               int32_t *perm);

/* Numbering: numEqV1 (src/Numeq.c:70-96) and numEqV2 (src/Numeq.c:109-135).
 * rt is the (maxViz+1)-wide restriction table; a cell with
 * rt[c*(maxViz+1)+nFace[c]] > 0 (a pinned pressure cell) gets no equation. */
int sfr_numeq_v1(int32_t *id, const int32_t *num, int numel);
int sfr_numeq_v2(int32_t *id, const int32_t *num, const short *rt,
                 const short *nFace, int numel, short maxViz);

/* ---------------------------------------------------------------------
 * Index builders (dataStruct, src/Datastruct.c:28-613, sequential branch).
 * The function fills an sfr_pattern whose arrays it allocates with malloc;
 * release them with sfr_pattern_free.
 *   CSR  : diag + upper (sym) or diag + lower + upper (unsym), sorted rows;
 *          the reference aliases ad to al for this storage (Datastruct.c:85)
 *   CSRD : strictly lower (sym) or lower + upper (unsym), sorted rows
 *   CSRC : strictly lower, sorted rows; values [al | au] when unsym
 *   ELLPACK: ia = {maxViz, (neq+1)*maxViz}, ja row-major (neq+1)*maxViz,
 *          slot = face index, boundary slot -> own row, slot facing a cell
 *          without equation left 0 (EllPack.c:26-79)
 * --------------------------------------------------------------------- */
typedef struct sfr_pattern {
  int32_t  neq;        /* equations */
  int32_t  nad;        /* stored off-diagonal (CSR: stored) entries      */
  int32_t  nia;        /* length of ia                                   */
  int64_t  nja;        /* length of ja                                   */
  int64_t  nval;       /* length of the value array the storage needs    */
  short    storage;
  short    unsym;
  short    maxViz;
  int32_t *ia;
  int32_t *ja;
} sfr_pattern;

int  sfr_datastruct(sfr_pattern *pat, short storage, int unsym,
                    const int32_t *id, const int32_t *num,
                    const int32_t *nelcon, const short *nViz,
                    int numel, short maxViz, int neq);
void sfr_pattern_free(sfr_pattern *pat);

/* Value scatter of one batch of cells, with the semantics of assbly/csr/
 * ellPack (src/Assbly.c:43-95, src/Csr.c:549-975, src/EllPack.c:113-149),
 * sequential branch, ndf = 1.
 *   lA[c*(maxViz+1) + f]  coefficient of cell c for face f, f < nViz[c]
 *   lA[c*(maxViz+1) + maxViz] diagonal of cell c  (stored at index nFace in
 *                          the reference, nFace == maxViz here)
 *   lB[c]                 right-hand side of cell c
 * al receives the off-diagonal values in the storage's layout (for CSRC
 * unsym, al has 2*nad entries: [al | au]); ad and b have neq (+1 for ELL)
 * entries.  Cells with id == -1 are skipped; a face whose neighbour has no
 * equation is dropped (csr) or written with ja = 0 (ellPack, ell_quirk = 1)
 * or written as 0 (ell_quirk = 0). */
int sfr_assemble(const sfr_pattern *pat, const int32_t *id,
                 const int32_t *nelcon, const short *nViz, int numel,
                 const double *lA, const double *lB,
                 double *al, double *ad, double *b, int ell_quirk);

/* Solver configuration file (setSolverConfig, src/Solv.c:547-608):
 * keys name|tol|maxit|storage up to "end"; tol 0 -> machine epsilon. */
int    sfr_read_solver_config(const char *path, short *solver, double *tol,
                              unsigned *maxit, short *storage);
short  sfr_solver_id(const char *word);   /* setSolver, Solv.c:507-528     */
short  sfr_storage_id(const char *word);  /* setDataStruct, Datastruct.c:1707 */
double sfr_smachn(void);                  /* smachn, Solv.c:614-622        */

/* ---- MatrixMarket --------------------------------------------------------
 * sfr_write_mtx: the reference's writeCoo (src/WriteMtx.c:26-161) -- CSRD
 * storage, every row as its diagonal then its stored entries, 1-based,
 * "%d %d %30.14e" lines under a "coordinate real symmetric|general" banner;
 * bin = 1 writes 1.0 for every value (the pattern files of the reference's
 * test/output/coo).  sfr_write_mtx_vector: writeCooB (src/WriteMtx.c:165-176).
 * sfr_read_mtx: any real coordinate file into CSRD (symmetric file: strictly
 * lower sorted rows, unsym = 0; general: full off-diagonal sorted rows,
 * unsym = 1; diagonal in ad); release with sfr_mtx_free.  Status codes are
 * those of sfr3d.h (0 = success). */
typedef struct sfr_mtx {
  int32_t  neq, nad;
  short    unsym;
  int32_t *ia, *ja;
  double  *al, *ad;
} sfr_mtx;

struct sfr_system;
int  sfr_write_mtx(const char *path, const struct sfr_system *sys, int bin);
int  sfr_write_mtx_vector(const char *path, const double *b, int32_t neq);
int  sfr_read_mtx(const char *path, sfr_mtx *out);
void sfr_mtx_free(sfr_mtx *m);

#ifdef __cplusplus
}
#endif
#endif /* SFR3D_SETUP_H */
