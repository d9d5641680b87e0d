/*
 * sfr3d.h -- MI355X (gfx950) linear-solve path for SFR3D-assembled systems:
 * Jacobi-preconditioned CG, BiCGSTAB, BiCGSTAB(2), GMRES(15) and MINRES
 * over CSR-family and ELLPACK matrices.
 *
 * The boundary keeps the contract of the reference's solver dispatch
 * (solverC, reference include/Solv.h:59-70, src/Solv.c:64-501): a matrix in
 * one of SFR3D's storages, a solver id, a tolerance, a maximum iteration
 * count, b and x.  Solver ids are those of include/Define.h:392-397 and
 * storage ids those of include/Define.h:407-412 (see sfr3d_setup.h).
 *
 * Stopping tests are the reference's:
 *   PCG      sqrt|r.D^-1 r| < tol * sqrt|b.D^-1 b|   (src/Cg.c:73-78,141-142)
 *   BiCGSTAB ||r||_2 < tol * ||b||_2, checked after each half step
 *            (src/Bicgstab.c:79-81,129-131,153-154)
 *   BiCGSTAB(2) ||r||_2 < tol * ||b||_2 after each iteration of two BiCG
 *            steps and a GCR(2) step (src/Bicgstab.c:1202-1205,1326-1327)
 *   GMRES    |e_k| < tol * ||D^-1 b||_2 after each restart cycle of k = 15
 *            Arnoldi steps, left Jacobi preconditioning
 *            (src/Iters.c:625-628,740); maxit counts cycles
 *   MINRES   the recurrence ||r|| (|s| times the previous) < tol *
 *            ||D^-1/2 b||_2 (src/Iters.c:360-364,476-477), Jacobi
 *            preconditioning, A symmetric (definite or not)
 * and the iteration count reported is the reference's (PCG, BiCGSTAB: j+1;
 * GMRES: Arnoldi steps + 1, with the cycles in stats.outer).
 *
 * Every entry point returns an SFR_* status.  All arrays are caller-owned and
 * never freed or reallocated; x is read (newX = 0) and written, b is read.
 * One sfr_context owns one HIP stream and its device buffers; calls on one
 * context must come from one host thread at a time, distinct contexts are
 * independent.
 */
#ifndef SFR3D_H
#define SFR3D_H

#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include "sfr3d_setup.h"

#ifdef __cplusplus
extern "C" {
#endif

/* solver ids (include/Define.h:392-397) */
#define SFR_PCG        1
#define SFR_PBICGSTAB  2
#define SFR_PBICGSTABL2 3
#define SFR_GMRES      4
#define SFR_MINRES     5

/* status codes */
#define SFR_OK             0
#define SFR_MAXIT          1   /* j == maxIt: reference prints + exit()    */
#define SFR_BREAKDOWN      2   /* zero / non-finite denominator            */
#define SFR_BAD_ARG        3
#define SFR_HIP_ERROR      4
#define SFR_NO_DEVICE      5
#define SFR_NOT_SUPPORTED  6

/* arithmetic of a device solve */
#define SFR_PREC_FP64   0      /* FP64 throughout, the reference algorithm */
#define SFR_PREC_MIXED  1      /* FP32 inner Krylov, FP64 residual/update  */

/* device matrix formats (chosen at upload, independent of the storage) */
#define SFR_FMT_AUTO   0
#define SFR_FMT_ELL    1       /* slot-major (column-major) ELL, lane/row  */
#define SFR_FMT_CSR    2       /* full-row CSR, several lanes per row      */

/* A matrix in one of the reference's host storages (SistEq, reference
 * include/Sisteq.h:27-45).  For CSRC with unsym, the upper part is au, or
 * al + nad when au is NULL (the reference keeps [al | au] contiguous). */
typedef struct sfr_system {
  int32_t        neq;
  int32_t        nad;
  const int32_t *ia;
  const int32_t *ja;
  const double  *al;
  const double  *ad;
  const double  *au;
  short          storage;
  short          unsym;
} sfr_system;

typedef struct sfr_stats {
  int      status;
  unsigned iters;     /* Krylov iterations (reference j+1); MIXED: sum of inner */
  unsigned outer;     /* MIXED: FP64 refinement steps; GMRES: restart cycles   */
  double   tol;
  double   norm_b;    /* PCG: sqrt|b.D^-1 b|  BiCGSTAB: ||b||  GMRES: ||D^-1 b|| */
  double   conv;      /* tol * norm_b                                          */
  double   xKx;       /* x.Ax            (src/Cg.c:155-158)                    */
  double   norm_x;    /* ||x||           (src/Cg.c:162)                        */
  double   norm_r;    /* ||b - Ax||, FP64                                      */
  double   norm_r_m;  /* sqrt|r.D^-1 r|, FP64; GMRES: |e_k| of the last cycle  */
  double   relres;    /* ||b - Ax|| / ||b||                                    */
  double   time_s;    /* solve time on the device (events on the context stream) */
} sfr_stats;

typedef struct sfr_context sfr_context;
typedef struct sfr_matrix  sfr_matrix;

/* ---- context -------------------------------------------------------- */
int         sfr_init(sfr_context **ctx, int device);
void        sfr_finalize(sfr_context *ctx);
const char *sfr_status_string(int status);
const char *sfr_last_error(const sfr_context *ctx);
int         sfr_synchronize(sfr_context *ctx);

/* device memory helpers on the context's stream (plain hipMalloc'd memory) */
int sfr_dev_alloc(sfr_context *ctx, size_t bytes, void **ptr);
int sfr_dev_free(sfr_context *ctx, void *ptr);
int sfr_copy_h2d(sfr_context *ctx, void *dst, const void *src, size_t bytes);
int sfr_copy_d2h(sfr_context *ctx, void *dst, const void *src, size_t bytes);

/* ---- device-resident matrices -------------------------------------------
 * sfr_matrix_create uploads the pattern once (conversion to the device
 * format, value gather maps) and the values; sfr_matrix_set_values uploads
 * new values for the same pattern, as SFR3D's SIMPLE loop re-assembles every
 * outer iteration.  format = SFR_FMT_* (AUTO lets the tuner decide). */
int  sfr_matrix_create(sfr_context *ctx, const sfr_system *sys, int format,
                       sfr_matrix **A);
int  sfr_matrix_set_values(sfr_context *ctx, sfr_matrix *A, const double *al,
                           const double *ad, const double *au);
void sfr_matrix_destroy(sfr_context *ctx, sfr_matrix *A);
/* launch shape of the solve kernels: rows per thread (1, 2, 4, 8) and, for
 * the CSR format, lanes per row (1, 2, 4, 8).  An explicit rpt applies to
 * every loop; without one, a matrix of 2^21 rows or more runs its mixed
 * loop at 2 rows per thread on at most 1,024 workgroups and its FP64 PCG
 * loop at 1 row per thread on at most 1,024 workgroups.  rpt = 0
 * measures:
 * short FP64 PCG runs at each rows-per-thread choice pick the FP64 PCG
 * loop's (the other loops keep theirs).  It runs solves on the context's
 * stream, allocates two temporary n-vectors and takes tens of
 * milliseconds at 2 M rows.  The choice changes the order of the device
 * sums, so solutions can differ in the last bits from untuned ones. */
int  sfr_matrix_tune(sfr_context *ctx, sfr_matrix *A, int rpt, int lanes);
/* neq, device format, ELL slots, stored nnz of the device format */
int  sfr_matrix_info(const sfr_matrix *A, int32_t *neq, int *format,
                     int *slots, int64_t *nnz);
/* bytes per stored column index the solve kernels read: 2 when the device
 * ELL's columns are kept as 16-bit offsets from the row (every |j - i| <
 * 32768, e.g. banded after RCM), else 4 */
int  sfr_matrix_index_bytes(const sfr_matrix *A);
/* Face pairs the FP64 PCG SpMV reads, as of the last FP64 PCG solve on A:
 * 1 (hex rows, slots (0,1) (2,4) (3,5)), 2 (quad rows, (0,2) (1,3)) or 0.
 * A symmetric ELLPACK system in SFR3D's face order (src/EllPack.c:26-79
 * with the face tables of src/Adjcency.c) stores each off-diagonal twice,
 * in the two rows it couples; when the values mirror bit for bit (checked
 * on the device once after they change), the FP64 PCG SpMV reads the first
 * slot of each pair from the row and the second from the neighbour, half
 * the value stream, with the same products and sums as the full read.  No
 * counterpart in the reference (a performance option; context option
 * "ell_pairs" = 0 turns it off). */
int  sfr_matrix_pair_planes(const sfr_matrix *A);
/* rows per thread of the solve kernels and of the FP64 PCG loop (the
 * latter differs from 2^21 rows up and after sfr_matrix_tune with
 * rpt = 0) */
int  sfr_matrix_rows_per_thread(const sfr_matrix *A, int *rpt, int *rpt_pcg);

/* The host-pointer entry points (sfr_solve, sfr_solverC, sfr_solve_shared,
 * sfr_solve_batched) keep up to 4 device matrices per context, one per
 * pattern, recognised by the identity of ia/ja with neq, nad, storage and
 * unsym; a later call on a cached pattern uploads values only, and the least
 * recently used pattern is dropped for a fifth.  SFR3D alternates at least
 * the velocity and the pressure patterns every SIMPLE step
 * (src/Simple.c:2983-3136).  A caller that frees a pattern and builds
 * another at the same addresses must call sfr_matrix_cache_clear first.
 * sfr_cache_info: matrices held, matrices created and lookups that hit. */
int  sfr_cache_info(const sfr_context *ctx, int *entries, int64_t *created,
                    int64_t *hits);
void sfr_matrix_cache_clear(sfr_context *ctx);

/* ---- solves ---------------------------------------------------------------
 * sfr_solve: host arrays in and out, the drop-in for solverC's PCG,
 * PBICGSTAB, PBICGSTABL2, GMRES and MINRES branches (reference
 * src/Solv.c:93-445).
 * sfr_solve_device: b and x already on the device (d_b, d_x: neq doubles).
 * precision SFR_PREC_MIXED (PCG and PBICGSTAB) runs the Krylov iterations
 * in FP32 -- vectors and matrix values, dot products accumulated in FP64 --
 * and folds their correction into the FP64 solution with reliable updates:
 * the residual, the solution update and the stopping test are FP64, the
 * test is the FP64 solver's own.  That test bounds the residual, not the
 * error: on an ill-conditioned system the mixed solution meets it with a
 * larger error than the FP64 one (8.5e-7 against 1.0e-7 on a 128^3
 * variable-coefficient Poisson system at tol 1e-8), since FP32 rounding
 * leaves error in the lowest modes, which the test weighs least; a caller
 * who needs the solution itself to FP64 accuracy keeps SFR_PREC_FP64.
 * BiCGSTAB(2), GMRES and MINRES run in FP64
 * only (SFR_PREC_MIXED returns SFR_NOT_SUPPORTED).  The host-pointer and
 * batch entry points (sfr_solve, sfr_solverC, sfr_solve_shared*,
 * sfr_solve_batched) take the precision from the context option
 * "precision" (default SFR_PREC_FP64), for PCG and PBICGSTAB. */
int sfr_solve(sfr_context *ctx, const sfr_system *sys, int solver, double tol,
              unsigned maxit, const double *b, double *x, int newX,
              sfr_stats *stats);
int sfr_solve_device(sfr_context *ctx, sfr_matrix *A, int solver,
                     int precision, double tol, unsigned maxit,
                     const double *d_b, double *d_x, int newX,
                     sfr_stats *stats);

/* Allocation-free solve: sfr_solve_device for SFR_PCG or SFR_PBICGSTAB in
 * FP64 that allocates nothing -- no hipMalloc, hipFree, graph or other
 * allocating runtime call.  No counterpart in the reference: it is for
 * callers that cannot allocate (code around HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  The caller passes the
 * resident matrix, d_b and d_x; the solve's scalars, partial sums and work
 * vectors live in the library's static device scratch, 64 MiB that each GPU
 * the library runs on holds until the process ends.  Limit: 2,708,640 rows
 * for PCG, 1,354,304 for PBICGSTAB; a larger system is refused with
 * SFR_BAD_ARG.  The kernels are those of sfr_solve_device, launched directly
 * instead of from captured graphs, so x, the iteration count and the
 * statistics are bitwise those of sfr_solve_device.  Every call on one GPU
 * uses that GPU's one scratch and finds the previous call's contents there:
 * calls that overlap on one GPU (other contexts, other threads) race, and
 * the caller must serialise them. */
int sfr_solve_device_noalloc(sfr_context *ctx, sfr_matrix *A, int solver,
                             double tol, unsigned maxit, const double *d_b,
                             double *d_x, int newX, sfr_stats *stats);

/* Shared-pattern batch (SFR3D's u/v/w momentum solves, reference
 * src/Simple.c:2738-2750, 2980-3036): nsys (1..4) systems with the same
 * ia/ja and off-diagonal values, each with its own diagonal, right-hand side
 * and solution, solved by one set of batched launches in which every SpMV
 * reads the shared pattern once.  Each system keeps its own scalars and stops
 * on its own test; stats[c] holds system c's results.  Solver: PBICGSTAB (in FP32
 * with FP64 reliable updates under the context option "precision").
 * Host form: ad[c] (NULL: sys->ad for every system), b[c], x[c] are arrays of
 * neq doubles.  Device form: d_ad (or NULL), d_b, d_x hold nsys*neq doubles,
 * system c at offset c*neq (the reference's b + c*neq layout). */
int sfr_solve_shared(sfr_context *ctx, const sfr_system *sys, int nsys,
                     const double *const *ad, const double *const *b,
                     double *const *x, int solver, double tol, unsigned maxit,
                     int newX, sfr_stats *stats);
int sfr_solve_shared_device(sfr_context *ctx, sfr_matrix *A, int nsys,
                            const double *d_ad, const double *d_b,
                            double *d_x, int solver, double tol,
                            unsigned maxit, int newX, sfr_stats *stats);

/* Species batch (SFR3D's transported species, reference
 * src/Combustion.c:76-92, 222-240: one solverC call per species on the same
 * ia/ja, with al + i*(2*nad+nadr), ad + i*neq, au, b and x of its own):
 * nsys (1..SFR_BATCH_MAX) systems on sys's pattern, system c with al[c],
 * ad[c], au[c] (au or au[c] may be NULL: CSRC unsym then reads al[c] + nad),
 * b[c] and x[c].  sys supplies neq, nad, ia, ja, storage and unsym; its
 * value pointers are ignored.  With solver PBICGSTAB the systems run in
 * groups of 2-4 through one set of batched launches per group, every SpMV
 * reading the pattern once for the group and each system's own values
 * (in FP32 with FP64 reliable updates under the "precision" option);
 * other solver ids and storage CSR solve the systems one after another.
 * Each system keeps its own stop test; stats[c] holds system c's results.
 * (The reference skips a species whose ||b|| < SZERO before calling the
 * solver; pass only the systems to solve, a zero b converges at once.) */
#define SFR_BATCH_MAX 10       /* MAXSPECIES, include/Define.h:92           */
int sfr_solve_batched(sfr_context *ctx, const sfr_system *sys, int nsys,
                      const double *const *al, const double *const *ad,
                      const double *const *au, const double *const *b,
                      double *const *x, int solver, double tol,
                      unsigned maxit, int newX, sfr_stats *stats);

/* Adjoint of a solve x = A^-1 b, for the gradient of a loss L(x): solves
 * A^T lambda = g with g = dL/dx (solver SFR_PCG when A is symmetric,
 * A^T = A; SFR_PBICGSTAB runs on the transposed operator, built on first
 * use), then
 *   dL/db = lambda                          (d_lambda, neq)
 *   dL/dal[k] = -sum lambda_i x_j over the entries (i, j) raw value k stands
 *     for, in the caller's storage order ([al | au] for CSRC unsym, the
 *     row-major ELLPACK array, ...); a value kept once for (i,j) and (j,i)
 *     gets both terms                        (d_grad_al, NULL to skip)
 *   dL/dad_i = -lambda_i x_i                 (d_grad_ad, neq, NULL to skip)
 * d_x is the forward solution.  No counterpart in the reference. */
int sfr_solve_adjoint(sfr_context *ctx, sfr_matrix *A, int solver, double tol,
                      unsigned maxit, const double *d_x, const double *d_g,
                      double *d_lambda, double *d_grad_al, double *d_grad_ad,
                      sfr_stats *stats);

/* ---- uneven batch ----------------------------------------------------------
 * Many independent SPD systems of any sizes, patterns and storages (up to
 * 1024), solved together by mixed-precision Jacobi-PCG (FP32 inner CG with
 * FP64 reliable updates, each system stopping on the reference PCG's own
 * test, src/Cg.c:73-78,141-142): one fixed set of launches per iteration
 * serves every system still running, replayed from one captured graph.  The
 * reference solves such systems one solverC call at a time; this entry has
 * no counterpart there.
 *
 * sfr_batch_create uploads the systems.  tune = 1 runs the autotuner: the
 * systems are grouped in size buckets (2^k <= neq < 2^(k+1)), and for each
 * bucket the device format (ELL or CSR) and the work-item size (rows per
 * thread) are chosen by timing the batch's own iteration kernels on up to
 * four of the bucket's systems.  tune = 0 takes the format heuristic of
 * sfr_matrix_create and a size rule.
 * stats[s] (nsys entries) holds system s's results; its time_s is the time
 * from the start of the batch solve to the moment system s converged,
 * measured on the device clock. */
typedef struct sfr_batch sfr_batch;

typedef struct sfr_bucket_choice {
  int    bucket;       /* k: systems with 2^k <= neq < 2^(k+1)              */
  int    nsys;         /* systems in the bucket                             */
  int    format;       /* SFR_FMT_ELL / SFR_FMT_CSR                         */
  int    rpt;          /* rows per thread (work item = rpt * 256 rows)      */
  double ns_per_row;   /* measured FP32 iteration time per row (0: untuned) */
} sfr_bucket_choice;

int  sfr_batch_create(sfr_context *ctx, int nsys, const sfr_system *sys,
                      int tune, sfr_batch **B);
int  sfr_batch_set_values(sfr_context *ctx, sfr_batch *B, int s,
                          const double *al, const double *ad,
                          const double *au);
/* host form: b[s], x[s] are arrays of system s's neq doubles */
int  sfr_batch_solve(sfr_context *ctx, sfr_batch *B, double tol,
                     unsigned maxit, const double *const *b, double *const *x,
                     int newX, sfr_stats *stats);
/* device form: d_b, d_x hold every system's vector, system s at the row
 * offset sfr_batch_info reports (total rows: sfr_batch_rows) */
int  sfr_batch_solve_device(sfr_context *ctx, sfr_batch *B, double tol,
                            unsigned maxit, const double *d_b, double *d_x,
                            int newX, sfr_stats *stats);
int64_t sfr_batch_rows(const sfr_batch *B);
int  sfr_batch_info(const sfr_batch *B, int s, int64_t *row0, int32_t *neq,
                    int *format, int *rpt);
/* tuner choices, one per non-empty bucket; returns how many were written */
int  sfr_batch_buckets(const sfr_batch *B, sfr_bucket_choice *out, int max);
void sfr_batch_destroy(sfr_context *ctx, sfr_batch *B);

/* ---- device assembly -------------------------------------------------------
 * The value scatter of sfr_assemble (the reference's assbly / csr / ellPack,
 * src/Assbly.c:43-95, src/Csr.c:549-975, src/EllPack.c:113-149) done on the
 * device into a resident matrix, for callers whose cell loops produce the
 * per-cell coefficients on the GPU: no values cross the host link between
 * SIMPLE iterations.  sfr_assembly_create makes the plan once per pattern
 * (pat, id, nelcon, nViz, numel, ell_quirk as for sfr_assemble; pat must be
 * the pattern A was created from).  sfr_assemble_device then takes d_lA
 * (numel*(maxViz+1) doubles, layout of sfr_assemble's lA) and optionally
 * d_lB -> d_b (neq doubles); the stored values are bitwise those
 * sfr_assemble writes, and the next solve on A uses them. */
typedef struct sfr_assembly sfr_assembly;
int  sfr_assembly_create(sfr_context *ctx, sfr_matrix *A,
                         const sfr_pattern *pat, const int32_t *id,
                         const int32_t *nelcon, const short *nViz, int numel,
                         int ell_quirk, sfr_assembly **plan);
int  sfr_assemble_device(sfr_context *ctx, sfr_assembly *plan,
                         const double *d_lA, const double *d_lB, double *d_b);
void sfr_assembly_destroy(sfr_context *ctx, sfr_assembly *plan);

/* ---- cell coefficients on the device ---------------------------------------
 * The finite-volume cell loop of SFR3D's SIMPLE pressure correction
 * (systFormSimplePres, src/CellLoopInc.c:471-679, called at
 * src/Simple.c:3053-3080; cell kernel cellSimplePres3D,
 * src/CellPres.c:903-1149) run on the GPU: the mesh's geometry is uploaded
 * once, and each call turns the current fields into the per-cell
 * coefficients lA / lB (layout of sfr_assemble: numel*(maxViz+1) doubles,
 * the centre at index maxViz) and the mass residual rCell, for
 * sfr_assemble_device.  Together with the solve, nothing crosses the host
 * link between SIMPLE iterations but the fields the caller chooses to read.
 *
 * sfr_cell_mesh holds SFR3D's own mesh arrays (struct Mesh), in its layout
 * and 1-based conventions: nelcon (elm.adj.nelcon, -1 on the boundary),
 * nFace (elm.adj.nViz), cellFace (elm.cellFace), owner (face.owner, two
 * per face, the owner first), the cell geometry elm.geom.volume / dcca /
 * xmcc and the face geometry face.mksi / ksi / area / normal / vSkew, the
 * material mat (elm.mat, 1-based) with the property table prop
 * (elm.material.prop, MAXPROP = 10 per material, density at index 1), and
 * the boundary markers faceRvel / faceRpres (elm.faceRvel / faceRpres,
 * maxViz+1 per cell, 1-based load index or 0; a cell whose own entry
 * faceRpres[c*(maxViz+1)+nFace] is set has no pressure equation and is
 * skipped, its lA / lB left as they are).  ndm must be 3; nFace <= maxViz
 * <= 6.
 *
 * sfr_load is a constant SFR3D load (struct Loads with nTypeVar constant):
 * type (Define.h: DIRICHLETBC 1, INLET 6, OUTLET 7, MOVEWALL 8,
 * INLETSTAICTPRES 9, OPEN 10, INLETTOTALPRES 13, STATICWALL 14,
 * FLUXPRES 15), its np parameters par and density.  loadsVel and loadsPres
 * are the tables faceRvel / faceRpres index (loadsVel and loadsPresC at
 * Simple.c:3053).  A load type the pressure cell does not use adds
 * nothing, as in the reference (src/Loads.c:265-727, 728-845); load
 * functions of position or time are refused.
 *
 * sfr_cells_pres_device: diffusion = the scheme of difusionSchemeAnisotropic
 * (src/CellLib.c:6054-6190; 1 ORTHOGONAL or 4 OVERRELAXED: the reference
 * leaves MINIMAL and ORTHOGONALC undefined); device fields in the
 * reference's layout: pres[numel], gradPres[numel*3], vel[numel*3],
 * dField[numel*3] (the SIMPLE d field); outputs d_lA, d_lB, and d_rCell
 * when not NULL.  The arithmetic keeps the reference's operation order
 * without fused multiply-adds, so lA / lB are bitwise those of the
 * reference's cell loop on the same inputs. */
typedef struct {
  int32_t numel;            /* cells (numelNov)                           */
  int32_t nface;            /* faces: length of the face arrays          */
  short maxViz;             /* face slots per cell                       */
  short ndm;                /* 3                                          */
  const short   *nFace;     /* [numel]                                   */
  const int32_t *nelcon;    /* [numel*maxViz] 1-based, -1 boundary       */
  const int32_t *cellFace;  /* [numel*maxViz] 1-based face index         */
  const int32_t *owner;     /* [nface*2] 1-based owner cell first        */
  const double  *volume;    /* [numel]                                   */
  const double  *dcca;      /* [numel*maxViz]                            */
  const double  *xmcc;      /* [numel*maxViz*3]                          */
  const double  *mksi;      /* [nface]                                   */
  const double  *ksi;       /* [nface*3]                                 */
  const double  *area;      /* [nface]                                   */
  const double  *normal;    /* [nface*3]                                 */
  const double  *vSkew;     /* [nface*3]                                 */
  const short   *mat;       /* [numel] 1-based material                  */
  const short   *faceRvel;  /* [numel*(maxViz+1)]                        */
  const short   *faceRpres; /* [numel*(maxViz+1)]                        */
  const double  *cc;        /* [numel*3] centroids (momentum only)       */
  const double  *xm;        /* [nface*3] face midpoints (momentum only)  */
} sfr_cell_mesh;

/* Options of the momentum cell loop (cellSimpleVel3D): advection 1 FoUp,
 * 2 Cd, 3 SoUp (advVel->iCod1); diffusion 1-4 (diffVel->iCod); turb /
 * wall_model (tModel->fTurb / fWall; tModel->fTurbStruct is not supported);
 * rhie, so_pressure, viscosity, residual (momentumModel->fRhieChowInt,
 * fSoPressure, fViscosity, fRes); simplec (typeSimple SIMPLEC, else
 * SIMPLE); time / backward (ddt->flag, ddt->type BACKWARD, else EULER);
 * dt, dt0 (ddt->dt[0], dt[1]); underU (the velocity under-relaxation). */
typedef struct {
  int    advection, diffusion, turb, wall_model, rhie, so_pressure,
         viscosity, residual, simplec, time, backward;
  double dt, dt0, underU;
} sfr_vel_opts;

typedef struct {
  short  type;              /* load type (Define.h)                      */
  short  np;                /* parameters in par                         */
  double par[8];
  double density;           /* Loads.density                             */
} sfr_load;

typedef struct sfr_cells sfr_cells;
int  sfr_cells_create(sfr_context *ctx, const sfr_cell_mesh *mesh,
                      const double *prop, int nmat,
                      const sfr_load *loadsVel, int nloadsVel,
                      const sfr_load *loadsPres, int nloadsPres,
                      sfr_cells **cells);
int  sfr_cells_pres_device(sfr_context *ctx, sfr_cells *cells, int diffusion,
                           const double *d_pres, const double *d_gradPres,
                           const double *d_vel, const double *d_dField,
                           double *d_lA, double *d_lB, double *d_rCell);
/* The momentum cell loop of SIMPLE (systFormSimpleVel,
 * src/CellLoopInc.c:98-385; cellSimpleVel3D, src/CellSimple.c:106-638),
 * u, v and w together, for a mesh created with cc and xm.  Inputs in the
 * reference's layout: d_id (the momentum equations' numbers, -1 for a cell
 * of prescribed velocity), pres, gradPres (3/cell), vel (3), gradVel (9,
 * d u_i / d x_j at i*3+j), eddyVisc (1; NULL unless turb), wallPar (4; NULL
 * unless wall_model), dField (3).  Outputs: d_lA in sfr_assemble's layout
 * (the shared off-diagonal coefficients, the u diagonal at maxViz) for
 * sfr_assemble_device; d_ad and d_b (3*neq, component c at c*neq, the
 * layout csrSimple writes and sfr_solve_shared_device reads); d_rCell
 * (3*numel, component-major); d_dOut (3/cell) the new d field, which must
 * not alias d_dField.  Velocity loads: MOVEWALL, STATICWALL, INLET,
 * OUTLET (INLETSTAICTPRES, OPEN and INLETTOTALPRES are refused).
 * Coefficients are bitwise the reference's loop, except that every cell
 * reads its neighbours' input d: the reference's serial loop writes d in
 * place, so with Rhie-Chow interpolation a cell there sees the new d of
 * lower-numbered neighbours (an order its OpenMP loop does not keep
 * either). */
int  sfr_cells_vel_device(sfr_context *ctx, sfr_cells *cells,
                          const sfr_vel_opts *opts, const int32_t *d_id,
                          int neq, const double *d_pres,
                          const double *d_gradPres, const double *d_vel,
                          const double *d_gradVel, const double *d_eddyVisc,
                          const double *d_wallPar, const double *d_dField,
                          double *d_dOut, double *d_lA, double *d_ad,
                          double *d_b, double *d_rCell);
void sfr_cells_destroy(sfr_context *ctx, sfr_cells *cells);
/* The two field updates around SIMPLE's pressure-correction solve, on
 * device arrays: sfr_cells_update_value is updateCellValueBlock
 * (src/CellLoop.c:4087-4130, ndf 1, one process): u[c] = x[id[c]-1], or
 * u[c] += x[id[c]-1] with add, for every cell with an equation (id > 0),
 * others untouched.  sfr_simple_update is simpleUpdate
 * (src/Simple.c:1615-1668, ndm 3): vel -= dField * gradPresC per component,
 * pres += alphaPres * presC, in the reference's operation order. */
int  sfr_cells_update_value(sfr_context *ctx, int numel, const int32_t *d_id,
                            const double *d_x, double *d_u, int add);
int  sfr_simple_update(sfr_context *ctx, int numel, double alphaPres,
                       double *d_vel, double *d_pres, const double *d_presC,
                       const double *d_gradPresC, const double *d_dField);

/* Measurement hook: average device time of one launch of a loop kernel,
 * HIP events on the context stream.  which = 0: FP64 PCG SpMV + p.q;
 * which = 1: the first kernel of a mixed-precision FP32 iteration (SpMV of
 * the updated direction, direction and correction updates, p.q). */
int sfr_kernel_time(sfr_context *ctx, sfr_matrix *A, int which, int reps,
                    double *ms_per_launch);
/* Measurement hook, no counterpart in the reference: average device time of
 * one copy of n doubles d_src -> d_dst (16-byte aligned, not overlapping)
 * by the library's streaming copy kernel (16-byte non-temporal loads and
 * stores), HIP events on the context stream.  2 n 8 bytes / time is the
 * HBM rate a pure stream attains, the ceiling the loop kernels are held
 * against; on MI355X it is above what hipMemcpy's device copy reaches. */
int sfr_copy_time(sfr_context *ctx, const double *d_src, double *d_dst,
                  int64_t n, int reps, double *ms_per_copy);
/* Options of a context: "chunk" = iterations per full captured graph
 * chunk (1..256; solves start on short chunks; GMRES captures chunk/16
 * restart cycles, at least one); "precision" = SFR_PREC_FP64 or
 * SFR_PREC_MIXED for the host-pointer and batch entry points (PCG and
 * PBICGSTAB; the other solver ids stay FP64); "ru_delta" / "ru_delta_bicg"
 * = MIXED PCG / BiCGSTAB: drop of the FP32 residual norm (relative to its
 * largest value since the last fold) that triggers an FP64 reliable update
 * (both default to 0.1); "ell_pairs" = 1 (default) lets the FP64 PCG SpMV
 * read mirrored ELL values through face pairs (sfr_matrix_pair_planes), 0 reads
 * every slot (the results are the same bits either way); "nt_loads" = -1
 * (default) lets the PCG loops' SpMV kernels (FP64 and the mixed FP32 loop)
 * stream an ELL matrix with non-temporal loads once the loop's footprint,
 * about 84 B/row in FP64 with face pairs and 56 B/row in FP32, exceeds the
 * 256 MiB last-level cache, 0 / 1 force plain / non-temporal loads (a cache
 * policy only: the same bits either way).  A mixed BiCGSTAB
 * solve whose FP64 residual stops
 * improving (or breaks down) is finished in FP64 from its last x, within
 * the same call; stats.iters then counts both parts. */
int sfr_set_option(sfr_context *ctx, const char *name, double value);

/* Drop-in twin of solverC (reference include/Solv.h:59-70), same prototype
 * with the reference's opaque Memoria/Interface/BufferOmp pointers passed
 * through untouched.  Storage CSR is refused (exit with a message): the
 * reference's solverC does not solve a CSR system as stored (ad aliases al,
 * src/Datastruct.c:85; matVecCsrSym drops the lower triangle,
 * src/HccaBlas.c:3315-3325), so no faithful answer exists; sfr_solve takes
 * CSR and solves the stored operator.  Non-convergence and bad ids end the
 * process with the
 * reference's messages and exit codes (src/Cg.c:195-200,
 * src/Bicgstab.c:225-231, 1362-1366, src/Iters.c:770-774, src/Solv.c:496;
 * MINRES only warns, src/Iters.c:540-544); with fLog one line in the
 * reference's format goes to fSolvLog (src/Cg.c:201-208,
 * src/Bicgstab.c:235-242, 1368-1376, src/Iters.c:545-552, 778-786). */
void sfr_solverC(void *m, int nEq, int nEqNov, int nAd, int nAdR,
                 int *ia, int *ja, double *al, double *ad, double *au,
                 double *b, double *x, void *iNeq, void *bOmp,
                 double tol, unsigned int maxIt, short storage, short solver,
                 FILE *fSolvLog, bool fLog, bool newX, bool unSym,
                 const char *name);
/* The context sfr_solverC solves on (device $SFR_DEVICE, default 0),
 * created on first use; for options and sfr_cache_info. */
int sfr_solverC_context(sfr_context **ctx);

#ifdef __cplusplus
}
#endif
#endif /* SFR3D_H */
