/*
 * ORACLE -- test infrastructure only.
 *
 * Thin harness around the reference's own C sources (compiled where they
 * lie by oracle/Makefile into oracle/_ref/libsfr_ref.so).  It sets the
 * reference's process globals the way its main program does and calls its
 * functions unchanged:
 *   reord          src/Reord.c:18          numEqV1/V2  src/Numeq.c:70,109
 *   dataStruct     src/Datastruct.c:28     assbly      src/Assbly.c:43
 *   solverC        src/Solv.c:64 (PCG / PBICGSTAB, serial, log line captured;
 *                  with -fopenmp also its OpenMP solvers, ref_solve_omp)
 *   genrcm         src/Rcm.c:56            matvecs     src/HccaBlas.c
 *   systFormSimplePres  src/CellLoopInc.c:471 (the SIMPLE pressure-correction
 *                  cell loop, with cellSimplePres3D, src/CellPres.c:903)
 * Used only to pin the oracle and to generate golden vectors.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <Solv.h>
#include <Reord.h>
#include <Sisteq.h>
#include <Assbly.h>
#include <CellLoop.h>

static Memoria g_mem;
static int g_init;
static SistEq g_sys;
static int g_have_sys;
static int g_sys_ad_alloc;

int ref_init(long arena_bytes)
{
  if (g_init) finalizeMem(&g_mem, false);
  nmax = arena_bytes;
  initMem(&g_mem, nmax, false);
  fileLogExc = fopen("/dev/null", "w");
  fileLogDebug = fileLogExc;
  mpiVar.nPrcs = 1;
  mpiVar.myId = 0;
  ompVar.fSolver = false;
  memset(&tm, 0, sizeof tm);
  /* dalloc (src/Memoria.c:280-330) refuses to free the arena's only block,
   * so one small block stays allocated for the life of the harness */
  static char *sentinel;
  HccaAlloc(char, &g_mem, sentinel, 64, "refSentinel", false);
  g_init = 1;
  g_have_sys = 0;
  return 0;
}

int ref_reord(int *num, const int *nelcon, const short *nViz, short maxViz,
              int numel, int flag)
{
  reord(&g_mem, num, nelcon, nViz, maxViz, numel, numel, flag != 0, 1);
  return 0;
}

int ref_genrcm(int node_num, int *adj_row_1based, int *adj_1based, int *perm)
{
  genrcm(node_num, adj_row_1based, adj_1based, perm);
  return 0;
}

int ref_numeq_v1(int *id, int *num, int numel) { return numEqV1(id, num, numel); }

int ref_numeq_v2(int *id, int *num, short *rt, short *nFace, int numel,
                 short maxViz)
{
  return numEqV2(id, num, rt, nFace, numel, maxViz);
}

static void drop_sys(void)
{
  if (!g_have_sys) return;
  HccaDealloc(&g_mem, g_sys.al, "refA", false);
  if (g_sys_ad_alloc) HccaDealloc(&g_mem, g_sys.ad, "refAd", false);
  HccaDealloc(&g_mem, g_sys.ja, "refJa", false);
  HccaDealloc(&g_mem, g_sys.ia, "refIa", false);
  g_have_sys = 0;
}

/* Builds the structure with dataStruct; returns nad, sizes in out[0..2] =
 * {len(ia), len(ja), len(al)}. */
int ref_datastruct(short storage, int unsym, int *id, int *num, int *nelcon,
                   short *nViz, int numel, short maxViz, int neq, long *out)
{
  drop_sys();
  memset(&g_sys, 0, sizeof g_sys);
  g_sys.storage = storage;
  g_sys.unsym = unsym != 0;
  g_sys.neq = neq;
  g_sys.neqNov = neq;
  dataStruct(&g_mem, id, num, nelcon, nViz, numel, maxViz, 1,
             "refIa", "refJa", "refAd", "refA", &g_sys);
  g_have_sys = 1;
  g_sys_ad_alloc = storage != CSR;
  long nia, nja, nal;
  if (storage == ELLPACK) {
    nia = 2; nja = (long)(neq + 1) * maxViz; nal = nja;
  } else {
    nia = neq + 1; nja = g_sys.ia[neq]; nal = nja;
    if ((storage == CSRC || storage == CSRCCOO) && unsym) nal = 2 * nja;
  }
  out[0] = nia; out[1] = nja; out[2] = nal;
  return g_sys.nad;
}

int ref_datastruct_copy(int *ia, int *ja, long nia, long nja)
{
  if (!g_have_sys) return -1;
  memcpy(ia, g_sys.ia, sizeof(int) * (size_t)nia);
  memcpy(ja, g_sys.ja, sizeof(int) * (size_t)nja);
  return 0;
}

/* Scatter per-cell coefficients with assbly into the held structure.
 * lA[c*(maxViz+1)+f], diagonal at f = maxViz; lB[c]. */
int ref_assemble(int *id, int *nelcon, short *nViz, int numel, short maxViz,
                 const double *lA, const double *lB, double *al, double *ad,
                 double *b, long nal, int neqad)
{
  if (!g_have_sys) return -1;
  const short storage = g_sys.storage;
  zero(g_sys.al, nal, DOUBLEC);
  if (storage != CSR) zero(g_sys.ad, neqad, DOUBLEC);
  double *bb = calloc((size_t)neqad, sizeof(double));
  INT lId[16];
  DOUBLE la[16], lb[1];
  for (int c = 0; c < numel; c++) {
    if (id[c] < 1) continue;
    const short nFace = nViz[c];
    for (int f = 0; f < nFace; f++) {
      const int viz = nelcon[c * maxViz + f] - 1;
      lId[f] = -1;
      if (viz != -2) lId[f] = id[viz] - 1;
      la[f] = lA[(long)c * (maxViz + 1) + f];
    }
    lId[nFace] = id[c] - 1;
    la[nFace] = lA[(long)c * (maxViz + 1) + maxViz];
    lb[0] = lB[c];
    assbly(g_sys.ia, g_sys.ja, g_sys.al, g_sys.ad, bb, lId, la, lb,
           g_sys.neq, g_sys.neq, g_sys.nad, 0, nFace, 1, storage, true, true,
           g_sys.unsym);
  }
  memcpy(al, g_sys.al, sizeof(double) * (size_t)nal);
  if (storage != CSR) memcpy(ad, g_sys.ad, sizeof(double) * (size_t)neqad);
  memcpy(b, bb, sizeof(double) * (size_t)neqad);
  free(bb);
  return 0;
}

/* solverC with the given arrays; out[0] = iterations, out[1] = xKx,
 * out[2] = ||x||, out[3] = time, and for GMRES out[4] = cycles (parsed back
 * from the reference's own log line). */
int ref_solve(int neq, int nad, int *ia, int *ja, double *al, double *ad,
              double *au, double *b, double *x, double tol, unsigned maxit,
              short storage, short solver, int newX, int unsym, double *out,
              char *line, int linelen)
{
  FILE *f = tmpfile();
  if (!f) return -1;
  solverC(&g_mem, neq, neq, nad, 0, ia, ja, al, ad, au, b, x, NULL, NULL,
          tol, maxit, storage, solver, f, true, newX != 0, unsym != 0, "ref");
  fflush(f);
  rewind(f);
  char buf[512] = {0};
  if (!fgets(buf, sizeof buf, f)) { fclose(f); return -2; }
  fclose(f);
  if (line && linelen > 0) { strncpy(line, buf, (size_t)linelen - 1); line[linelen - 1] = 0; }
  double t, xkx, nx, tm_;
  int it;
  if (solver == PCG) {
    if (sscanf(buf, "PCG: tol %lf iteration %d xKx %lf norma(x*x) %lf time %lf",
               &t, &it, &xkx, &nx, &tm_) != 5) return -3;
  } else if (solver == PBICGSTABL2) {
    if (sscanf(buf, "PBICGSTAB(2): tol %lf iteration %d xKx %lf norma(x*x) %lf time %lf",
               &t, &it, &xkx, &nx, &tm_) != 5) return -3;
  } else if (solver == MINRES) {
    if (sscanf(buf, "PMINRES: tol %lf iteration %d xKx %lf norma(x*x) %lf time %lf",
               &t, &it, &xkx, &nx, &tm_) != 5) return -3;
  } else if (solver == GMRES) {
    int cyc, nk;
    if (sscanf(buf, "GMRES: tol %lf Cycles %d iteration %d nKrylov %d xKx %lf "
               "norma(x*x) %lf time %lf", &t, &cyc, &it, &nk, &xkx, &nx,
               &tm_) != 7) return -3;
    out[4] = cyc;                       /* GMRES callers pass out[5] */
  } else {
    if (sscanf(buf, "PBICGSTAB: tol %lf iteration %d xKx %lf ||x*x|| %lf time %lf",
               &t, &it, &xkx, &nx, &tm_) != 5) return -3;
  }
  out[0] = it; out[1] = xkx; out[2] = nx; out[3] = tm_;
  return 0;
}

#ifdef _OPENMP
#include <omp.h>
#include <Csr.h>
/* solverC with the reference's OpenMP solvers (pcgOmp, src/Cg.c:490-715;
 * pbicgstabOmp, src/Bicgstab.c:554-803) on `nthreads` threads: the process
 * globals set as the reference's openMpSet and ReadFile do for the solver
 * (ompVar.fSolver, nThreadsSolver, src/ReadFile.c:5570-5600), the rows
 * split among the threads by the reference's partitionCsrByNonzeros
 * (src/Csr.c:1827, called from pMatrixSolverOmp), and the scatter buffer
 * thY of nthreads * neq doubles.  out and line as ref_solve. */
int ref_solve_omp(int neq, int nad, int *ia, int *ja, double *al, double *ad,
                  double *b, double *x, double tol, unsigned maxit,
                  short storage, short solver, int unsym, int nthreads,
                  double *out, char *line, int linelen)
{
  /* computeEffectiveWork (src/Csr.c) opens a parallel region of the default
   * size and writes one entry per thread: the default is set to nthreads
   * and the arrays sized for whichever is larger */
  const int nmax_th = omp_get_max_threads() > nthreads ? omp_get_max_threads() : nthreads;
  omp_set_num_threads(nthreads);
  BufferOmp bo;
  bo.thBegin = calloc((size_t)nmax_th, sizeof(INT));
  bo.thEnd = calloc((size_t)nmax_th, sizeof(INT));
  bo.thSize = calloc((size_t)nmax_th, sizeof(INT));
  bo.thHeight = calloc((size_t)nmax_th, sizeof(INT));
  bo.thY = calloc((size_t)nthreads * (size_t)neq, sizeof(DOUBLE));
  FILE *f = tmpfile();
  int rc = 0;
  if (!bo.thBegin || !bo.thEnd || !bo.thSize || !bo.thHeight || !bo.thY || !f) {
    rc = -1;
    goto out;
  }
  partitionCsrByNonzeros(ia, ja, neq, nthreads, bo.thBegin, bo.thEnd, bo.thSize,
                         bo.thHeight, storage);
  ompVar.fSolver = true;
  ompVar.nThreadsSolver = (unsigned short)nthreads;
  solverC(&g_mem, neq, neq, nad, 0, ia, ja, al, ad, NULL, b, x, NULL, &bo,
          tol, maxit, storage, solver, f, true, true, unsym != 0, "ref");
  ompVar.fSolver = false;
  ompVar.nThreadsSolver = 1;
  fflush(f);
  rewind(f);
  char buf[512] = {0};
  if (!fgets(buf, sizeof buf, f)) { rc = -2; goto out; }
  if (line && linelen > 0) { strncpy(line, buf, (size_t)linelen - 1); line[linelen - 1] = 0; }
  double t, xkx, nx, tm_;
  int it;
  /* the OpenMP solvers' own log lines (src/Cg.c:706-713,
   * src/Bicgstab.c:794-801) */
  const int ok = solver == PCG
      ? sscanf(buf, "PCGOMP: tol %lf iteration %d xKx %lf norma(x*x) %lf time %lf",
               &t, &it, &xkx, &nx, &tm_)
      : sscanf(buf, "PBICGSTABOMP: tol %lf iteration %d xKx %lf norma(x*x) %lf time %lf",
               &t, &it, &xkx, &nx, &tm_);
  if (ok != 5) { rc = -3; goto out; }
  out[0] = it; out[1] = xkx; out[2] = nx; out[3] = tm_;
out:
  if (f) fclose(f);
  free(bo.thBegin); free(bo.thEnd); free(bo.thSize); free(bo.thHeight); free(bo.thY);
  return rc;
}
#endif

/* y = A x with the reference's serial matvec for the storage. */
int ref_matvec(short storage, int unsym, int neq, int *ia, int *ja,
               double *al, double *ad, double *x, double *y)
{
  void (*mv)() = NULL;
  const bool o = ompVar.fSolver;
  ompVar.fSolver = false;
  setMatVec(&mv, storage, unsym != 0, false);
  ompVar.fSolver = o;
  mv(neq, ia, ja, al, ad, x, y);
  return 0;
}

/* Constant loads from flat arrays: type[i], np[i], par[i*8 + k], density[i]. */
static void fill_loads(Loads *ld, int n, const short *type, const short *np,
                       const double *par, const double *density)
{
  memset(ld, 0, sizeof(Loads) * MAXLOAD);
  for (int i = 0; i < n && i < MAXLOAD; i++) {
    ld[i].fUse = true;
    ld[i].nTypeVar = LVARCONST;
    ld[i].type = type[i];
    ld[i].np = np[i];
    for (int k = 0; k < 8; k++) ld[i].par[k] = par[i * 8 + k];
    ld[i].density = density[i];
  }
}

/* The reference's SIMPLE pressure-correction cell loop, systFormSimplePres
 * (src/CellLoopInc.c:471-679), unchanged, assembling into the structure held
 * by ref_datastruct (which must have been built for these cells).  Mesh and
 * field arrays in the reference's layout (see Simple.c:3053-3080 for which
 * Mesh member each one is); ndm = 3, calType 1 for every material.  Writes
 * al (nal), ad and b (neq) and rCell (numel). */
int ref_simple_pres(int numel, short maxViz, short *nFace, int *nelcon,
                    int *cellFace, int *fOwner, double *volume, double *dcca,
                    double *xmcc, double *mksi, double *ksi, double *eta,
                    double *area, double *normal, double *xm, double *mvSkew,
                    double *vSkew, short *mat, double *prop, int nmat,
                    short *faceRvel, short *faceRpres, int *id,
                    int nlv, short *lv_type, short *lv_np, double *lv_par,
                    double *lv_density,
                    int nlp, short *lp_type, short *lp_np, double *lp_par,
                    double *lp_density,
                    short diffusion, double *pres, double *gradPres,
                    double *vel, double *dField, double *al, double *ad,
                    double *b, double *rCell, long nal)
{
  if (!g_have_sys || nmat > MAXMAT) return -1;
  /* cellSimplePres3D reads the velocity loads from the global loadsVel
   * (src/CellPres.c:1119, `&loadsVel[nCarg]`), whatever table its caller
   * passes; Simple.c:3053 passes that same global */
  static Loads lpres[MAXLOAD];
  fill_loads(loadsVel, nlv, lv_type, lv_np, lv_par, lv_density);
  fill_loads(lpres, nlp, lp_type, lp_np, lp_par, lp_density);
  Loads *lvel = loadsVel;
  Diffusion dif = {diffusion};
  Turbulence tm_;
  memset(&tm_, 0, sizeof tm_);
  Temporal ddt;
  memset(&ddt, 0, sizeof ddt);
  short *calType = calloc((size_t)nmat, sizeof(short));
  short *geomType = malloc(sizeof(short) * (size_t)numel);
  short *nen = malloc(sizeof(short) * (size_t)numel);
  double *wallPar = calloc((size_t)numel * NWALLPAR, sizeof(double));
  double *gCc = calloc((size_t)numel * 3, sizeof(double));
  for (int m = 0; m < nmat; m++) calType[m] = 1;
  for (int c = 0; c < numel; c++) { geomType[c] = HEXACELL; nen[c] = 8; }
  const int neq = g_sys.neq;
  zero(g_sys.al, nal, DOUBLEC);
  zero(g_sys.ad, neq, DOUBLEC);
  double *bb = calloc((size_t)neq + 1, sizeof(double));
  ompVar.fCell = false;
  ompVar.nThreadsCell = 1;
  systFormSimplePres(lvel, lpres, &dif, &tm_, NULL, nelcon, nen, nFace,
                     cellFace, fOwner, volume, dcca, xmcc, gCc, mksi, ksi,
                     eta, area, normal, xm, mvSkew, vSkew, geomType, prop,
                     calType, mat, g_sys.ia, g_sys.ja, g_sys.al, g_sys.ad,
                     bb, id, faceRvel, faceRpres, pres, gradPres, vel,
                     dField, wallPar, rCell, &ddt, neq, neq, g_sys.nad, 0,
                     8, maxViz, 3, numel, 1, g_sys.storage, true, true, true,
                     g_sys.unsym);
  memcpy(al, g_sys.al, sizeof(double) * (size_t)nal);
  memcpy(ad, g_sys.ad, sizeof(double) * (size_t)neq);
  memcpy(b, bb, sizeof(double) * (size_t)neq);
  free(bb); free(calType); free(geomType); free(nen); free(wallPar); free(gCc);
  return 0;
}

/* The reference's SIMPLE momentum cell loop, systFormSimpleVel
 * (src/CellLoopInc.c:98-385, with cellSimpleVel3D,
 * src/CellSimple.c:106-638), unchanged, assembling u/v/w (ndf 3) into the
 * CSRD unsym structure held by ref_datastruct: al (nal, shared), ad and b
 * (3*neq, component-major), rCell (3*numel), and dField updated in place as
 * the loop does.  opt = {advection, diffusion, fTurb, fWall, fRhieChowInt,
 * fSoPressure, fViscosity, fRes, SIMPLEC, fTime, BACKWARD}; optd = {dt, dt0,
 * underU}. */
int ref_simple_vel(int numel, short maxViz, short *nFace, int *nelcon,
                   int *cellFace, int *fOwner, double *volume, double *dcca,
                   double *xmcc, double *cc, double *mksi, double *ksi,
                   double *eta, double *area, double *normal, double *xm,
                   double *mvSkew, double *vSkew, short *mat, double *prop,
                   int nmat, short *faceRvel, short *faceRpres, int *id,
                   int nlv, short *lv_type, short *lv_np, double *lv_par,
                   double *lv_density, int *opt, double *optd, double *pres,
                   double *gradPres, double *vel, double *gradVel,
                   double *eddyVisc, double *wallPar, double *dField,
                   double *al, double *ad, double *b, double *rCell, long nal)
{
  if (!g_have_sys || nmat > MAXMAT) return -1;
  static Loads lpres[MAXLOAD];
  fill_loads(loadsVel, nlv, lv_type, lv_np, lv_par, lv_density);
  memset(lpres, 0, sizeof lpres);
  Advection adv;
  memset(&adv, 0, sizeof adv);
  adv.iCod1 = (short)opt[0];
  Diffusion dif = {(short)opt[1]};
  Turbulence tm_;
  memset(&tm_, 0, sizeof tm_);
  tm_.fTurb = opt[2] != 0;
  tm_.fWall = opt[3] != 0;
  MomentumModel mm;
  memset(&mm, 0, sizeof mm);
  mm.fRhieChowInt = opt[4] != 0;
  mm.fSoPressure = opt[5] != 0;
  mm.fViscosity = opt[6] != 0;
  mm.fRes = opt[7] != 0;
  Temporal ddt;
  memset(&ddt, 0, sizeof ddt);
  ddt.flag = opt[9] != 0;
  ddt.type = opt[10] ? BACKWARD : EULER;
  ddt.dt[0] = optd[0];
  ddt.dt[1] = optd[1];
  short *calType = calloc((size_t)nmat, sizeof(short));
  short *geomType = malloc(sizeof(short) * (size_t)numel);
  short *nen = malloc(sizeof(short) * (size_t)numel);
  double *stressR = calloc((size_t)numel * 6, sizeof(double));
  for (int m = 0; m < nmat; m++) calType[m] = 1;
  for (int c = 0; c < numel; c++) { geomType[c] = HEXACELL; nen[c] = 8; }
  const int neq = g_sys.neq;
  zero(g_sys.al, nal, DOUBLEC);
  double *ad3 = calloc((size_t)3 * neq, sizeof(double));
  double *b3 = calloc((size_t)3 * neq, sizeof(double));
  ompVar.fCell = false;
  ompVar.nThreadsCell = 1;
  systFormSimpleVel(loadsVel, lpres, &adv, &dif, &tm_, &mm, NULL,
                    opt[8] ? SIMPLEC : SIMPLE, NULL, nelcon, nen, nFace,
                    cellFace, fOwner, volume, dcca, xmcc, cc, mksi, ksi, eta,
                    area, normal, xm, mvSkew, vSkew, geomType, prop, calType,
                    mat, g_sys.ia, g_sys.ja, g_sys.al, ad3, b3, id, faceRvel,
                    faceRpres, pres, gradPres, vel, gradVel, dField, optd[2],
                    rCell, stressR, eddyVisc, wallPar, &ddt, neq, neq,
                    g_sys.nad, 0, 8, maxViz, 3, numel, 3, 6, g_sys.storage,
                    true, true, true, g_sys.unsym);
  memcpy(al, g_sys.al, sizeof(double) * (size_t)nal);
  memcpy(ad, ad3, sizeof(double) * (size_t)3 * neq);
  memcpy(b, b3, sizeof(double) * (size_t)3 * neq);
  free(ad3); free(b3); free(calType); free(geomType); free(nen); free(stressR);
  return 0;
}
