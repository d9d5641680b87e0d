/*
 * Internal interface between the C host layer (drivers, matrix conversion,
 * public ABI) and the HIP kernel layer (kernels + extern "C" launchers).
 * Plain C: included by gcc-compiled host files and by hipcc-compiled .hip
 * files alike.
 */
#ifndef SFR_INTERNAL_H
#define SFR_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>
#include "sfr3d.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SFR_BLOCK   256          /* threads per workgroup (4 wave64)        */
#define SFR_MAXG    16384        /* partial slots per reduction region      */
#define SFR_GRIDCAP 2048         /* max workgroups of a streaming kernel    */
#define SFR_MAXDOT  8            /* partial sums one kernel may reduce      */
#define SFR_MAXSYS  8            /* systems of one shared-pattern batch     */
#define SFR_D16_ESC (-32768)     /* 16-bit column offset: see the exceptions */
#define SFR_D16_MAXEXC 256       /* most exceptions a 16-bit ELL may have   */
#define SFR_GM_K    15           /* GMRES Krylov dimension (src/Solv.c:82)  */
#define SFR_GM_STATE 512         /* doubles of GMRES state: h, c, s, e      */
/* Static device scratch of the allocation-free solve (device/scratch.hip):
 * scalars at 0, the four partial-sum regions of a solve (4 x 4 x SFR_MAXG
 * doubles) from SFR_SCRATCH_PART, work vectors from SFR_SCRATCH_VEC, each
 * on a 256-byte boundary. */
#define SFR_SCRATCH_BYTES ((size_t)64 << 20)
#define SFR_SCRATCH_PART  ((size_t)4096)
#define SFR_SCRATCH_VEC   (SFR_SCRATCH_PART + (size_t)16 * SFR_MAXG * sizeof(double))

/* Scalars of one Krylov solve, resident on the device.  The last workgroup
 * of each reducing kernel writes them; later kernels read them. */
typedef struct sfr_scalars {
  double   d;          /* PCG: r.D^-1 r   BiCGSTAB: r.r0   inner CG: r.r     */
  double   d_prev;     /* value of d the previous kernel started from      */
  double   alpha, beta, omega;
  double   norm_b;     /* reference norm of b                               */
  double   conv;       /* tol * norm_b                                      */
  double   bb;         /* b.b                                               */
  double   tol;
  double   rho_max;    /* mixed: largest inner r.r since the last update    */
  double   delta2;     /* mixed: reliable-update threshold, squared         */
  double   res[6];     /* final / outer dots                                */
  double   dd[2];      /* mixed: inner r.r, written alternately by parity   */
  /* polled by the host (8 words from iter) */
  unsigned iter;       /* completed iterations                              */
  unsigned done;       /* SFR_DONE_*                                        */
  unsigned maxit;
  unsigned outer;      /* mixed: reliable (FP64) updates taken              */
  unsigned ru;         /* mixed: reliable update requested                  */
  unsigned run2, run3, run4;  /* run flags handed from kernel to kernel     */
  unsigned run5;
  unsigned fresh;      /* mixed: the direction was just rebuilt by a fold   */
  unsigned pbuf;       /* mixed: direction buffer current at a fold request */
  unsigned pad[5];
} sfr_scalars;

enum { SFR_DONE_RUN = 0, SFR_DONE_CONV = 1, SFR_DONE_BREAK = 2,
       SFR_DONE_MAXIT = 3 };

/* Read-only view of a device matrix handed to the launchers. */
typedef struct sfr_mat_view {
  int64_t        n;
  int            fmt;       /* SFR_FMT_ELL / SFR_FMT_CSR                 */
  int            slots;     /* ELL slots                                 */
  int            lanes;     /* CSR lanes per row                         */
  const int32_t *col;
  const int32_t *rowptr;
  const double  *val;
  const double  *ad;
  const double  *dinv;
  const float   *val32;     /* D^-1/2 A D^-1/2 off-diagonals (unit diag) */
  const double  *dscale;    /* 1/sqrt(ad)                                */
  const int16_t *col16;     /* ELL columns as 16-bit row offsets, or NULL */
  const int64_t *exc_k;     /* entries whose offset is SFR_D16_ESC: their  */
  const int32_t *exc_col;   /*   slot index f*n+i and column               */
  int            nexc;
  const double  *valnb;     /* batched kernels: per-system values          */
                            /*   interleaved, valnb[k*nb + c], or NULL     */
  int            pair;      /* ELL face pairs the FP64 PCG SpMV may read   */
                            /*   (SFR_PAIR_HEX / SFR_PAIR_QUAD), or 0      */
  int            nt;        /* the PCG loops' SpMV kernels stream the      */
                            /*   matrix with non-temporal loads            */
} sfr_mat_view;

/* ELL face pairs: SFR3D's hex face order pairs slots (0,1) (2,4) (3,5),
 * its quad order (0,2) (1,3); a row's second slot of a pair holds the value
 * of the neighbour's first slot facing back */
enum { SFR_PAIR_NONE = 0, SFR_PAIR_HEX = 1, SFR_PAIR_QUAD = 2 };

/* launch geometry: rows per thread, workgroups (derived), and a cap on the
 * workgroups below SFR_GRIDCAP (0: SFR_GRIDCAP) */
typedef struct sfr_geom { int rpt; int grid; int cap; } sfr_geom;

struct sfr_matrix {
  int32_t  n;
  int      fmt, slots, lanes;
  int64_t  nnz;          /* stored entries of the device format            */
  short    storage, unsym;
  int32_t  nad;
  int64_t  nraw;         /* raw off-diagonal values of the host storage    */
  int      diag_in_al;   /* storage CSR: diagonal inside al                */
  int      have32;       /* FP32 scaled copy is current                    */
  int      rpt;          /* tuned rows per thread                          */
  int      rpt_pcg;      /* FP64 PCG loop's rows per thread (0: rpt)       */
  int      rpt_set;      /* rpt was chosen by the caller (sfr_matrix_tune) */
  int32_t *col, *rowptr, *vmap, *dmap;
  int16_t *col16;        /* ELL: col - row, SFR_D16_ESC for the few that   */
  int64_t *exc_k;        /*   do not fit 16 bits, listed in exc_k/exc_col  */
  int32_t *exc_col;
  int      nexc;
  int      pair_kind;    /* face pairing of the ELL pattern (SFR_PAIR_*)   */
  int      pair_ok;      /* values mirror across the pairs: 1 yes, 0 no,   */
                         /*   -1 not checked since they last changed       */
  int      pair_use;     /* the FP64 PCG SpMV reads them (last check)      */
  double  *raw, *raw_ad, *val, *ad, *dinv, *dscale;
  float   *val32;
  float   *val32r;       /* FP32 copy of val (unscaled), mixed BiCGSTAB     */
  int      have32r;      /*   and whether it is current                    */
  const int32_t *h_ia, *h_ja;   /* host pattern identity (value-only uploads) */
  /* adjoint support, built on first use */
  struct sfr_matrix *T;  /* transposed operator (shares raw, ad, dinv)     */
  int      borrowed;     /* raw/ad/dinv belong to another matrix           */
  int32_t *grad_rc;      /* nraw x 4: (row, col) of up to two entries fed  */
  int      have_grad;    /*   by each raw value, -1 = none                 */
};

typedef struct sfr_graph {
  hipGraphExec_t exec;
  const void    *key[8];
  int            kind, chunk;
} sfr_graph;

/* Host-pointer paths (sfr_solve, sfr_solverC, sfr_solve_shared,
 * sfr_solve_batched): device matrices kept per pattern, recognised by the
 * identity of ia/ja and the sizes, least recently used evicted.  SFR3D
 * alternates at least two patterns per SIMPLE step (velocity and pressure,
 * plus energy / turbulence / species), each re-assembled on a fixed
 * structure. */
#define SFR_MCACHE 4
typedef struct sfr_mcache {
  sfr_matrix *A;
  uint64_t    used;             /* tick of the last lookup that hit       */
} sfr_mcache;

struct sfr_context {
  int          device, ncu;
  hipStream_t  stream;
  char         err[512];
  sfr_scalars *d_sc, *h_sc;     /* SFR_MAXSYS each; h_sc is pinned        */
  double      *part;            /* SFR_MAXDOT * SFR_MAXG * SFR_MAXSYS      */
  hipEvent_t   ev_start, ev_stop, ev_poll[2];
  sfr_scalars *h_poll;          /* pinned [2][SFR_MAXSYS] status copies    */
  double      *stage;           /* system-major staging of batched inputs  */
  int64_t      stage_n;
  int64_t      ws_n;
  double      *w64[10];
  float       *w32[6];
  sfr_graph    graphs[16];
  int          ngraph;
  sfr_mcache   mcache[SFR_MCACHE]; /* patterns of the host-pointer paths */
  uint64_t     mtick;             /* lookups so far (LRU clock)          */
  int64_t      mcreated, mhits;   /* matrices created / lookups that hit */
  int          chunk;           /* iterations per captured chunk           */
  double      *gm_buf;          /* GMRES: basis, raw vectors, state        */
  int64_t      gm_n, gm_ld;
  double      *nb_val;          /* species batch: interleaved values       */
  double      *nb_raw;          /*   and one system's raw values           */
  int64_t      nb_val_n, nb_raw_n;
  double       ru_delta;        /* MIXED PCG: reliable-update residual drop */
  double       ru_delta_b;      /* MIXED BiCGSTAB: the same                */
  int          ell_pairs;       /* FP64 PCG SpMV may read ELL face pairs   */
  int          nt_loads;        /* -1 by size, 0 / 1: non-temporal matrix   */
                                /*   loads in the PCG loops' SpMV kernels   */
  void        *scratch;         /* device address of the static scratch    */
  int          precision;       /* host-pointer and batch entry points     */
  float       *mb32;            /* mixed BiCGSTAB: 7 vectors, diagonal and */
  int64_t      mb32_n;          /*   inverse, 9 x mb32_n floats            */
  float       *nb_val32;        /* species batch: FP32 interleaved values  */
  int64_t      nb_val32_n;
};

int sfr_set_error(sfr_context *ctx, int code, const char *fmt, ...);
/* destroy every cached host-path matrix (context teardown) */
void sfr_mcache_clear(sfr_context *ctx);
/* destroy the loop graphs whose key slot `slot` equals `key` */
void sfr_drop_graphs_keyed(sfr_context *ctx, int slot, const void *key);
/* solve statistics from a scalar block (ms = device time of the solve) */
void sfr_fill_stats(const sfr_scalars *sc, sfr_stats *st, float ms);
int  sfr_matrix_prepare_mixed(sfr_context *ctx, sfr_matrix *A);
/* whether the FP64 PCG SpMV of A may read its ELL face pairs: checks the
 * values once after they change (a device pass and a sync), sets pair_use */
int  sfr_matrix_pair_check(sfr_context *ctx, sfr_matrix *A);
/* raw values changed on the device: device-format values, diagonal */
int  sfr_matrix_refresh(sfr_context *ctx, sfr_matrix *A);
sfr_mat_view sfr_view(const sfr_matrix *A);
sfr_geom     sfr_geom_for(const sfr_context *ctx, int64_t n, int rpt);

/* device address of the static scratch on `device` (the current device),
 * looked up once per device; returns a hipError_t */
int sfr_scratch_address(int device, void **ptr);

/* ---- launchers (the .hip files); each enqueues on `s`, returns a hipError_t */
/* matrix values */
int sfr_k_gather(hipStream_t s, int64_t nnz, const int32_t *vmap,
                 const double *raw, double *val);
int sfr_k_diag(hipStream_t s, int64_t n, const int32_t *dmap,
               const double *raw, const double *raw_ad, double *ad,
               double *dinv);
/* mismatches of A's values across its face pairs (kind = SFR_PAIR_*):
 * per-workgroup counts in part[0..*nparts) */
int sfr_k_pair_check(hipStream_t s, const sfr_mat_view *A, int kind,
                     double *part, int *nparts);
int sfr_k_scale32(hipStream_t s, const sfr_mat_view *A, double *dscale,
                  float *val32);
int sfr_k_fill(hipStream_t s, int64_t n, double v, double *x);
int sfr_k_spmv(hipStream_t s, const sfr_mat_view *A, const double *x,
               double *y);
int sfr_k_fill32(hipStream_t s, int64_t n, float v, float *x);
int sfr_k_stream_copy(hipStream_t s, int64_t n, const double *src,
                      double *dst);
/* single SpMV kernels of the PCG / inner loops (measurement hooks) */
int sfr_k_pcg_spmv_only(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                        const double *p, double *q, sfr_scalars *sc,
                        double *part);

/* PCG, FP64 (reference src/Cg.c:47-209) */
int sfr_k_pcg_setup(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                    const double *b, double *x, double *r, double *p,
                    int newX, double tol, unsigned maxit, sfr_scalars *sc,
                    double *part);
int sfr_k_pcg_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                   double *x, double *r, double *p, double *q,
                   sfr_scalars *sc, double *part);
/* partials of b.D^-1 b and b.b (region 2), x = 0 when newX */
int sfr_k_pcg_norm(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                   const double *b, double *x, int newX, double *part,
                   int *nparts);
/* final quantities: x.Ax, x.x, r.D^-1 r, r.r with r = b - Ax -> sc->res */
int sfr_k_final(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                const double *b, const double *x, sfr_scalars *sc,
                double *part);

/* SIMPLE pressure-correction cell loop, one work-item per cell (the
 * per-slot planes of host/cells.c) */
int sfr_k_cell_pres(hipStream_t s, int numel, int maxViz, int diffusion,
                    const signed char *nface, const int32_t *nb,
                    const int32_t *bc, const double *g, const double *rho,
                    const sfr_load *lv, const sfr_load *lp, const double *pres,
                    const double *gp, const double *vel, const double *df,
                    double *lA, double *lB, double *rc);

/* SIMPLE momentum cell loop, one work-item per cell (planes 0-14) */
int sfr_k_cell_vel(hipStream_t s, int numel, int maxViz, const sfr_vel_opts *o,
                   const signed char *nfacev, const int32_t *nb,
                   const int32_t *bc, const double *g, const double *rho,
                   const double *visc, const double *vol, const sfr_load *lv,
                   const int32_t *id, int neq, const double *pres,
                   const double *gp, const double *vel, const double *gv,
                   const double *eddy, const double *wall, const double *din,
                   double *dout, double *lA, double *ad, double *b, double *rc);
/* u[c] (+)= x[id[c]-1] for the cells with an equation */
int sfr_k_cell_value(hipStream_t s, int64_t n, const int32_t *id,
                     const double *x, double *u, int add);
/* w -= d grad(p'), p += alpha p' (ndm 3) */
int sfr_k_simple_update(hipStream_t s, int64_t n, double alpha, double *w,
                        double *p, const double *pc, const double *gpc,
                        const double *d);

/* measured rows per thread of the FP64 PCG loop -> A->rpt_pcg */
int sfr_tune_pcg(sfr_context *ctx, sfr_matrix *A);
/* Rows per thread of the FP64 PCG loop: the measured choice when there is
 * one; else, unless the caller fixed the shape, 1 from 2^21 rows up (with
 * the SpMV capped at 1,024 workgroups there, pcg.hip); else the matrix's
 * default.  Measured on the 128^3 / 160^3 / 200^3 Poisson loops (ms per
 * solve at 1, 2, 4, 8 rows per thread): 63.2 / 64.8 / 65.4 / 69.2,
 * 180.0 / 183.7 / 183.8 / 187.4 and 415.9 / 424.7 / 430.4 / 439.9; the
 * tuner picks 1 at all three.  The shape does not depend on whether the
 * face pairs are read, so a solve that has not checked them (the
 * allocation-free entry point) sums in the same order as one that has. */
static inline int sfr_pcg_rpt(const sfr_matrix *A)
{
  if (A->rpt_pcg) return A->rpt_pcg;
  if (!A->rpt_set && A->n >= ((int64_t)1 << 21)) return 1;
  return A->rpt;
}

/* Non-temporal matrix loads in the PCG loops' SpMV kernels (FP64
 * k_pcg_spmv, FP32 k_mx_a): the context option "nt_loads" when it is set,
 * else once the loop's footprint -- the matrix bytes its SpMV reads plus
 * its five vectors -- exceeds the 256 MiB last-level cache.  Below that
 * the plain loads keep the matrix cached from one iteration to the next;
 * above it they only displace the vectors, which the vector kernels then
 * fetch from HBM.  Measured on MI355X, solves/s with the loads plain /
 * non-temporal, 128^3 / 160^3 / 200^3 Poisson (84 B/row FP64 with face
 * pairs, 56 B/row FP32): FP64 15.6 / 14.5, 5.29 / 5.79, 2.38 / 2.41;
 * mixed 27.2 / 23.3, 12.0 / 10.5, 4.69 / 5.08.  The values are the same
 * bits either way. */
#define SFR_LLC_BYTES ((int64_t)256 << 20)
static inline int sfr_loop_nt(const sfr_context *ctx, const sfr_matrix *A,
                              int fp32)
{
  if (ctx->nt_loads >= 0) return ctx->nt_loads;
  const int64_t s = A->slots;
  const int64_t row = fp32 ? s * 4 + s * 2 + 5 * 4
                           : 8 + (A->pair_use ? s / 2 : s) * 8 + s * 2 + 5 * 8;
  return (int64_t)A->n * row > SFR_LLC_BYTES;
}

/* GMRES(SFR_GM_K), FP64 (reference src/Iters.c:587-790) */
int sfr_gm_workspace(sfr_context *ctx, int64_t n);
int sfr_nb_buffers(sfr_context *ctx, int64_t nnz, int64_t nraw);
/* mixed BiCGSTAB: FP32 values of A (unscaled); 9 FP32 vectors of n entries */
int sfr_matrix_prepare_f32(sfr_context *ctx, sfr_matrix *A);
int sfr_mb_buffers(sfr_context *ctx, int64_t n);
int sfr_k_gm_setup(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                   const double *b, double *x, int newX, double tol,
                   unsigned maxit, sfr_scalars *sc, double *part);
int sfr_k_gm_cycle(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                   const double *b, double *x, double *basis, int64_t ld,
                   double *const w[2], double *st, sfr_scalars *sc,
                   double *part);

/* BiCGSTAB(2), FP64 (reference src/Bicgstab.c:1172-1378); W: r0, r, u, v,
 * w, s, p, z/t */
int sfr_k_b2_setup(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                   const double *b, double *x, double *const W[8], int newX,
                   double tol, unsigned maxit, sfr_scalars *sc, double *part);
int sfr_k_b2_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g, double *x,
                  double *const W[8], sfr_scalars *sc, double *part);

/* MINRES, FP64 (reference src/Iters.c:309-555); V, W: rings of 3 vectors,
 * Z: ring of 2; one iteration per call at ring phase q = j % 6 */
int sfr_k_mr_setup(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                   const double *b, double *x, double *const V[3],
                   double *const W[3], double *const Z[2], int newX,
                   double tol, unsigned maxit, sfr_scalars *sc, double *part);
int sfr_k_mr_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g, double *x,
                  double *const V[3], double *const W[3], double *const Z[2],
                  int q, sfr_scalars *sc, double *part);

/* BiCGSTAB, FP64 (reference src/Bicgstab.c:50-245) */
int sfr_k_bicg_setup(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                     const double *b, double *x, double *r, double *r0,
                     double *p, double *z, int newX, double tol,
                     unsigned maxit, sfr_scalars *sc, double *part);
int sfr_k_bicg_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                    double *x, double *r, const double *r0, double *p,
                    double *z, double *v, double *t, sfr_scalars *sc,
                    double *part);

/* shared-pattern batch of nb (2..4) BiCGSTAB systems, vectors interleaved
 * per row (x[i*nb + c]), per-system diagonal ad/dinv interleaved alike,
 * scalars sc[0..nb) */
int sfr_k_bb_setup(hipStream_t s, const sfr_mat_view *A, sfr_geom g, int nb,
                   const double *ad, const double *dinv, const double *b,
                   double *x, double *r, double *r0, double *p, double *z,
                   int newX, double tol, unsigned maxit, sfr_scalars *sc,
                   double *part);
int sfr_k_bb_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g, int nb,
                  const double *ad, const double *dinv, double *x, double *r,
                  const double *r0, double *p, double *z, double *v,
                  double *t, sfr_scalars *sc, double *part);
int sfr_k_bb_final(hipStream_t s, const sfr_mat_view *A, sfr_geom g, int nb,
                   const double *ad, const double *dinv, const double *b,
                   const double *x, sfr_scalars *sc, double *part);
/* system-major src[c*n+i] -> interleaved dst[i*nb+c] (inv: 1/value, or NULL) */
/* valnb[k*nb + c] = raw[vmap[k]] (0 where vmap[k] < 0) for k < nnz */
int sfr_k_gather_nb(hipStream_t s, int64_t nnz, const int32_t *vmap,
                    const double *raw, double *valnb, int nb, int c);
int sfr_k_interleave(hipStream_t s, int64_t n, int nb, const double *src,
                     double *dst, double *inv);
int sfr_k_deinterleave(hipStream_t s, int64_t n, int nb, const double *src,
                       double *dst);

/* adjoint: dL/d raw[k] = -(lam[r0] x[c0] + lam[r1] x[c1]) over the entries
 * raw value k feeds; dL/d ad_i = -lam_i x_i */
int sfr_k_adj_grad(hipStream_t s, int64_t nraw, const int32_t *rc,
                   const double *lam, const double *x, double *grad_raw,
                   int64_t n, double *grad_ad);

/* mixed precision: FP32 CG on D^-1/2 A D^-1/2 with FP64 reliable updates */
int sfr_k_mx_start(hipStream_t s, const double *part, int nparts, double tol,
                   unsigned maxit, double delta, sfr_scalars *sc);
/* fold y into x, true FP64 residual, stop test, new direction; init = first
 * call (p = r, y = 0), zero_x = x is known to be 0 */
int sfr_k_mx_fold(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                  const double *b, double *x, float *y, float *r,
                  const float *p_old, float *p_new, int init, int zero_x,
                  sfr_scalars *sc, double *part);
/* one FP32 iteration at position parity par of a captured chunk: the
 * direction is read from pb[par] and written to pb[par ^ 1] */
int sfr_k_mx_inner_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                        float *y, float *r, float *const pb[2], float *q,
                        int par, sfr_scalars *sc, double *part);
/* the fused SpMV kernel of the FP32 iteration alone (measurement hook) */
int sfr_k_mx_a_only(hipStream_t s, const sfr_mat_view *A, sfr_geom g,
                    const float *r, const float *po, float *pn, float *q,
                    float *y, sfr_scalars *sc, double *part);

/* mixed precision BiCGSTAB for nb (1..4) systems on one pattern, vectors
 * interleaved per row (nb = 1: plain), FP32 iterations with FP64 reliable
 * updates.  w: y, r, r0, p, z, v, t (nb*n floats each); ad32, dinv32: the
 * FP32 diagonal and inverse (nb*n).  ad, b, x: FP64, interleaved.  The
 * FP32 operator is val32 (the shared values) or valnb32 (per-system values
 * interleaved per entry) when not NULL; the view's valnb likewise for the
 * FP64 products of the folds. */
int sfr_k_mb_start(hipStream_t s, const sfr_mat_view *A, sfr_geom g, int nb,
                   const double *ad, const double *b, double *x, float *ad32,
                   float *dinv32, int newX, double tol, unsigned maxit,
                   double delta, sfr_scalars *sc, double *part);
int sfr_k_mb_fold(hipStream_t s, const sfr_mat_view *A, sfr_geom g, int nb,
                  const double *ad, const double *b, double *x,
                  float *const w[7], const float *ad32, const float *dinv32,
                  int init, int zero_x, sfr_scalars *sc, double *part);
int sfr_k_mb_iter(hipStream_t s, const sfr_mat_view *A, sfr_geom g, int nb,
                  const float *val32, const float *valnb32, float *const w[7],
                  const float *ad32, const float *dinv32, sfr_scalars *sc,
                  double *part);
/* dst[k] = (float) src[k], k < n */
int sfr_k_to_f32(hipStream_t s, int64_t n, const double *src, float *dst);

/* device assembly: dst[k] = src[map[k] - 1] (map > 0), 0 (map == 0),
 * unchanged (map < 0) */
int sfr_k_scatter(hipStream_t s, int64_t n, const int32_t *map,
                  const double *src, double *dst);

/* ---- uneven batch: independent systems, mixed-precision PCG ------------- */
#define SFR_UB_MAXSYS 1024       /* systems of one uneven batch             */

/* One system of an uneven batch as the kernels see it: its device matrix
 * (either format) and where its rows sit in the batch vectors. */
typedef struct sfr_ub_sys {
  const int32_t *col, *rowptr;
  const int16_t *col16;     /* ELL: 16-bit row offsets, or NULL             */
  const int64_t *exc_k;     /* escaped offsets: slot index and column       */
  const int32_t *exc_col;
  int64_t        nexc;
  const double  *val, *ad, *dinv, *dscale;
  const float   *val32;
  int64_t        row0;      /* first row in the batch vectors               */
  int32_t        n, slots, fmt, rpt;
  int32_t        g;         /* work items of rpt*SFR_BLOCK rows             */
  int32_t        gfirst;    /* first partial slot (4 per work item)         */
} sfr_ub_sys;

typedef struct sfr_ub_view {
  const sfr_ub_sys *sys;
  int           nsys, grid;
  int          *act;        /* [0] active systems, [1] their work items,
                               [2, 2+nsys) ids, [2+nsys, 2+2nsys) prefix     */
  const double *b;
  double       *x;
  float        *y, *r, *p, *q;
  sfr_scalars  *sc;         /* nsys                                         */
  unsigned long long *ts;   /* [s] completion stamp, [nsys] start stamp     */
  double       *partA, *partB;
} sfr_ub_view;

/* all = 1: list every system and stamp the start; 2: list every system;
 * 0: list the systems still running */
int sfr_k_ub_compact(hipStream_t s, const sfr_ub_view *v, int all);
int sfr_k_ub_init(hipStream_t s, const sfr_ub_view *v, int newX, double tol,
                  unsigned maxit, double delta);
int sfr_k_ub_iter(hipStream_t s, const sfr_ub_view *v);
int sfr_k_ub_final(hipStream_t s, const sfr_ub_view *v);

#ifdef __cplusplus
}
#endif
#endif /* SFR_INTERNAL_H */
