/*
 * ORACLE -- test infrastructure only.  Nothing in the product links it.
 *
 * CPU restatement of SFR3D's SIMPLE pressure-correction cell loop for 3-D
 * cells: the per-cell gather of systFormSimplePres (src/CellLoopInc.c:
 * 471-679) and the cell kernel cellSimplePres3D (src/CellPres.c:903-1149)
 * with the pieces it calls:
 *   interpolFace, linear        src/CellLib.c:7174-7200
 *   difusionSchemeAnisotropic   src/CellLib.c:6054-6190 (ORTHOGONAL,
 *                               OVERRELAXED; MINIMAL / ORTHOGONALC leave e
 *                               undefined in the reference and are refused)
 *   interpolFaceVel, ndm 3      src/CellLib.c:6853-6940
 *   pLoadSimplePres             src/Loads.c:728-845 (iCodPres 0, not
 *                               buoyant, fCal true)
 *   pLoadSimple, pressure part  src/Loads.c:265-727 (fCalVel false,
 *                               fCalPres true)
 *   getLoads, constant loads    src/Loads.c:122-166
 * Expressions keep the reference's operation order; with no contraction of
 * multiplies and adds (this file is built without FMA) the results are
 * bitwise the reference's.  Output lA in the layout of sfr_assemble
 * (numel*(maxViz+1), the centre at maxViz), lB[numel], rCell[numel].
 */
#include <math.h>
#include <string.h>

#define ORC_DENSITY 1          /* Define.h:203 */
#define ORC_MAXPROP 10         /* Define.h:349 */
enum { L_DIRICHLET = 1, L_INLET = 6, L_OUTLET = 7, L_MOVEWALL = 8,
       L_INLETSTATIC = 9, L_OPEN = 10, L_INLETTOTAL = 13, L_STATICWALL = 14,
       L_FLUXPRES = 15 };      /* Define.h:317-331 */

typedef struct {
  short type, np;
  double par[8];
  double density;
} orc_load;

static void load_at(orc_load *ld, int i, const short *type, const short *np,
                    const double *par, const double *dens)
{
  ld->type = type[i];
  ld->np = np[i];
  for (int k = 0; k < 8; k++) ld->par[k] = par[i * 8 + k];
  ld->density = dens[i];
}

/* e of the implicit part of the diffusive flux through s */
static int diffusion_e(int code, const double *s, const double *k, double *e)
{
  if (code == 1) {                                /* ORTHOGONAL */
    const double ss = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    e[0] = ss * k[0];
    e[1] = ss * k[1];
    e[2] = ss * k[2];
    return 0;
  }
  if (code == 4) {                                /* OVERRELAXED */
    const double ss = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const double nk = s[0] * k[0] + s[1] * k[1] + s[2] * k[2];
    const double tmp = ss / nk;
    e[0] = tmp * k[0];
    e[1] = tmp * k[1];
    e[2] = tmp * k[2];
    return 0;
  }
  return -1;
}

/* pressure load of a boundary face: adds to sP (no other effect here) */
static void pres_load(const orc_load *ld, double *sP, const double *e,
                      const double *n, const double *velC, double densityC,
                      double fArea, double dd)
{
  double t = e[0] * e[0] + e[1] * e[1];
  t += e[2] * e[2];
  const double modE = sqrt(t);
  if (ld->type == L_DIRICHLET) {
    *sP += densityC * modE / dd;
  } else if (ld->type == L_INLETTOTAL) {
    const double rhoE = ld->par[0];
    const double ev[3] = {ld->par[2], ld->par[3], ld->par[4]};
    const double dc = modE / dd;
    const double vn = velC[0] * n[0] + velC[1] * n[1] + velC[2] * n[2];
    const double en = (ev[0] * n[0] + ev[1] * n[1] + ev[2] * n[2]) * fArea;
    const double m = rhoE * vn * fArea;
    const double modVel = m / (rhoE * en);
    const double rv = rhoE * modVel;
    const double den = m - dc * rv * rv;
    *sP += rhoE * m * dc / den;
  } else if (ld->type == L_OPEN) {
    double vn = velC[0] * n[0] + velC[1] * n[1];
    vn += velC[2] * n[2];
    if (vn > 0.0) {                      /* the reference adds it twice */
      *sP += densityC * modE / dd;
      *sP += densityC * modE / dd;
    }
  }
  /* FLUXPRES without buoyancy and every other type: nothing */
}

/* mass flux of a velocity load through a boundary face: subtracted from p */
static void vel_load(const orc_load *ld, double *p, const double *n,
                     const double *velC, double densityC, double fArea)
{
  double wfn, m;
  switch (ld->type) {
  case L_INLETSTATIC:
    wfn = velC[0] * n[0] + velC[1] * n[1] + velC[2] * n[2];
    m = densityC * wfn * fArea;
    *p -= m;
    break;
  case L_INLET: {
    wfn = ld->par[0] * n[0] + ld->par[1] * n[1];
    wfn += ld->par[2] * n[2];
    const double rhoE = ld->par[ld->np - 1];
    m = rhoE * wfn * fArea;
    *p -= m;
    break;
  }
  case L_OUTLET:
    wfn = velC[0] * n[0] + velC[1] * n[1];
    wfn += velC[2] * n[2];
    m = densityC * wfn * fArea;
    *p -= m;
    break;
  case L_OPEN:
    wfn = velC[0] * n[0] + velC[1] * n[1];
    wfn += velC[2] * n[2];
    m = (wfn >= 0.0 ? densityC : ld->density) * wfn * fArea;
    *p -= m;
    break;
  case L_INLETTOTAL: {
    const double vn = velC[0] * n[0] + velC[1] * n[1] + velC[2] * n[2];
    m = ld->par[0] * vn * fArea;
    *p -= m;
    break;
  }
  default:                              /* walls and the rest: no flux */
    break;
  }
}

int orc_simple_pres3d(int numel, short maxViz, const short *nFace,
                      const int *nelcon, const int *cellFace,
                      const int *fOwner, const double *dcca,
                      const double *xmcc, const double *mksi,
                      const double *ksi, const double *area,
                      const double *normal, const double *vSkew,
                      const short *mat, const double *prop,
                      const short *faceRvel, const short *faceRpres,
                      int nlv, const short *lv_type, const short *lv_np,
                      const double *lv_par, const double *lv_density,
                      int nlp, const short *lp_type, const short *lp_np,
                      const double *lp_par, const double *lp_density,
                      short diffusion, const double *pres,
                      const double *gradPres, const double *vel,
                      const double *dField, double *lA, double *lB,
                      double *rCell)
{
  if (diffusion != 1 && diffusion != 4) return -1;
  const int w = maxViz + 1;
  for (int c = 0; c < numel; c++) {
    if (rCell) rCell[c] = 0.0;
    const int nf = nFace[c];
    if (faceRpres[c * w + nf] > 0) continue;         /* no equation */
    const double rhoC = prop[(mat[c] - 1) * ORC_MAXPROP + ORC_DENSITY];
    const double presC = pres[c];
    double velC[3], gpC[3], dC[3];
    for (int j = 0; j < 3; j++) {
      velC[j] = vel[c * 3 + j];
      gpC[j] = gradPres[c * 3 + j];
      dC[j] = dField[c * 3 + j];
    }
    double p = 0.0, sP = 0.0, la[8];
    for (int f = 0; f < nf; f++) {
      const int face = cellFace[c * maxViz + f] - 1;
      const double ch = (fOwner[face * 2] - 1 == c) ? 1.0 : -1.0;
      double n[3], k[3], xc[3], vs[3], s[3], e[3];
      for (int j = 0; j < 3; j++) {
        n[j] = ch * normal[face * 3 + j];
        k[j] = ch * ksi[face * 3 + j];
        xc[j] = xmcc[(c * maxViz + f) * 3 + j];
        vs[j] = vSkew[face * 3 + j];
      }
      const double A = area[face], mk = mksi[face];
      const int nb = nelcon[c * maxViz + f] - 1;
      if (nb > -1) {
        double v[3];
        v[0] = vs[0] + xc[0];
        v[1] = vs[1] + xc[1];
        v[2] = vs[2] + xc[2];
        const double alpha = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / mk;
        const double am1 = 1.0 - alpha;
        const double rhoV = prop[(mat[nb] - 1) * ORC_MAXPROP + ORC_DENSITY];
        const double density = am1 * rhoC + alpha * rhoV;
        double dF[3], wf[3], gf[3];
        for (int j = 0; j < 3; j++)
          dF[j] = am1 * dC[j] + alpha * dField[nb * 3 + j];
        for (int j = 0; j < 3; j++) s[j] = dF[j] * A * n[j];
        diffusion_e(diffusion, s, k, e);
        const double modE = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        la[f] = density * modE / mk;
        /* interpolFaceVel: face velocity with the pressure-gradient
         * (Rhie-Chow style) correction */
        for (int j = 0; j < 3; j++) {
          wf[j] = am1 * velC[j] + alpha * vel[nb * 3 + j];
          gf[j] = am1 * gpC[j] + alpha * gradPres[nb * 3 + j];
        }
        double wfn = wf[0] * n[0] + wf[1] * n[1] + wf[2] * n[2];
        const double g0 = (pres[nb] - presC) / mk;
        const double nk = dF[0] * k[0] * n[0] + dF[1] * k[1] * n[1]
                        + dF[2] * k[2] * n[2];
        double g1 = gf[0] * k[0] + gf[1] * k[1] + gf[2] * k[2];
        g1 = (g1 - g0) * nk;
        wfn += g1;
        p -= density * wfn * A;
      } else {
        la[f] = 0.0;
        const int lp = faceRpres[c * w + f];
        if (lp) {
          orc_load ld;
          load_at(&ld, lp - 1, lp_type, lp_np, lp_par, lp_density);
          for (int j = 0; j < 3; j++) s[j] = dC[j] * A * n[j];
          diffusion_e(diffusion, s, k, e);
          pres_load(&ld, &sP, e, n, velC, rhoC, A, dcca[c * maxViz + f]);
        }
        const int lv = faceRvel[c * w + f];
        if (lv) {
          orc_load ld;
          load_at(&ld, lv - 1, lv_type, lv_np, lv_par, lv_density);
          vel_load(&ld, &p, n, velC, rhoC, A);
        }
      }
    }
    double ap = sP;
    for (int f = 0; f < nf; f++) ap += la[f];
    double *out = lA + (long)c * w;
    for (int f = 0; f < nf; f++) out[f] = -la[f];
    out[maxViz] = ap;
    lB[c] = p;
    if (rCell) rCell[c] = p;
  }
  (void)nlv; (void)nlp;
  return 0;
}

/* ---- the SIMPLE momentum cell loop --------------------------------------
 * systFormSimpleVel (src/CellLoopInc.c:98-385) with cellSimpleVel3D
 * (src/CellSimple.c:106-638) and the pieces it calls:
 *   difusionSchemeNew          src/CellLib.c:5824-5990 (all four schemes)
 *   mediaHarmonica             src/CellLib.c:7879-7884
 *   advectiveScheme            src/CellLib.c:6394-6640 (FoUp, Cd, SoUp;
 *                              deferredCd, upwindLinearV1)
 *   facePressure               src/CellLib.c:7499-7560 (first or second
 *                              order, not buoyant)
 *   viscosityPartExp           src/CellLib.c:7439-7470 (fDiv false)
 *   pLoadSimple, momentum part src/Loads.c:265-520 (MOVEWALL, STATICWALL
 *                              through moveWall, INLET and OUTLET through
 *                              stress and gradFaceNull)
 * ndf = ndm = 3; no structural turbulence model.  Output per cell c:
 * lA[c*(maxViz+3) + f] for its faces, the three diagonals at maxViz..+2;
 * lB[c*3 + k]; rCell[k*numel + c]; dField: the new d, written to dOut.
 * in_place = 1 writes dOut = dField as the reference's loop does, so a cell
 * reads the already updated d of lower-numbered neighbours (the result then
 * depends on the order of the cells); 0 reads the input d for every
 * neighbour, the order-free form a parallel loop computes. */
enum { A_FOUP = 1, A_CD = 2, A_SOUP = 3 };           /* Define.h:267-269 */
#define ORC_VISC 0                                   /* DYNAMICVISCOSITY */

typedef struct {
  int advection, diffusion, turb, wall_model, rhie, so_pressure, viscosity,
      residual, simplec, time, backward;
  double dt, dt0, underU;
} orc_vel_opts;

/* e and t of the diffusive flux through s (difusionSchemeNew, ndm 3) */
static int diffusion_et(int code, const double *s, const double *k, double *e,
                        double *t)
{
  if (code == 1 || code == 3) {                      /* ORTHOGONAL(C) */
    const double ss = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int j = 0; j < 3; j++) e[j] = ss * k[j];
  } else if (code == 2) {                            /* MINIMAL */
    const double nk = s[0] * k[0] + s[1] * k[1] + s[2] * k[2];
    for (int j = 0; j < 3; j++) e[j] = nk * k[j];
  } else if (code == 4) {                            /* OVERRELAXED */
    const double ss = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const double nk = s[0] * k[0] + s[1] * k[1] + s[2] * k[2];
    const double tmp = ss / nk;
    for (int j = 0; j < 3; j++) e[j] = tmp * k[j];
  } else {
    return -1;
  }
  for (int j = 0; j < 3; j++) t[j] = code == 1 ? 0.0 : s[j] - e[j];
  return 0;
}

/* velocity load of a boundary face: diagonal parts sPc, sources p */
static void vel_load_mom(const orc_load *ld, double *sPc, double *p,
                         const double *velC, const double *gv, double viscC,
                         double effViscC, const double *s, const double *n,
                         const double *xmcc, double rhoC, const double *wallPar,
                         int wallModel, double fArea, double dcca)
{
  double tA[3], ss[6];
  if (ld->type == L_MOVEWALL || ld->type == L_STATICWALL) {
    for (int j = 0; j < 3; j++)
      tA[j] = ld->type == L_MOVEWALL ? ld->par[j] : 0.0;
    double vw;
    if (wallModel) {
      const double yPlus = wallPar[0], uPlus = wallPar[1];
      vw = yPlus > 0.01 ? viscC * yPlus / uPlus : viscC;
    } else {
      vw = effViscC;
    }
    /* moveWall (src/CellLib.c:7715-7760), not Dirichlet */
    const double aP = vw * fArea / dcca;
    sPc[0] += aP * (1.0 - n[0] * n[0]);
    sPc[1] += aP * (1.0 - n[1] * n[1]);
    sPc[2] += aP * (1.0 - n[2] * n[2]);
    p[0] += aP * (tA[0] * (1.0 - n[0] * n[0]) + (velC[1] - tA[1]) * n[1] * n[0]
                  + (velC[2] - tA[2]) * n[2] * n[0]);
    p[1] += aP * (tA[1] * (1.0 - n[1] * n[1]) + (velC[0] - tA[0]) * n[0] * n[1]
                  + (velC[2] - tA[2]) * n[2] * n[1]);
    p[2] += aP * (tA[2] * (1.0 - n[2] * n[2]) + (velC[0] - tA[0]) * n[0] * n[2]
                  + (velC[1] - tA[1]) * n[1] * n[2]);
    return;
  }
  if (ld->type != L_INLET && ld->type != L_OUTLET) return;   /* checked */
  /* stress (src/CellLib.c:7013-7040), lambda 0 */
  const double tr = gv[0] + gv[4] + gv[8];
  ss[0] = 2.0 * viscC * gv[0] + 0.0 * tr;
  ss[1] = 2.0 * viscC * gv[4] + 0.0 * tr;
  ss[2] = 2.0 * viscC * gv[8] + 0.0 * tr;
  ss[3] = viscC * (gv[3] + gv[1]);
  ss[4] = viscC * (gv[7] + gv[5]);
  ss[5] = viscC * (gv[6] + gv[2]);
  if (ld->type == L_INLET) {
    tA[0] = ld->par[0];
    tA[1] = ld->par[1];
    double wfn = tA[0] * n[0] + tA[1] * n[1];
    tA[2] = ld->par[2];
    wfn += tA[2] * n[2];
    const double m = ld->par[ld->np - 1] * wfn * fArea;
    p[0] += ss[0] * s[0] + ss[3] * s[1] + ss[5] * s[2];
    p[1] += ss[3] * s[0] + ss[1] * s[1] + ss[4] * s[2];
    p[2] += ss[5] * s[0] + ss[4] * s[1] + ss[2] * s[2];
    p[0] -= m * tA[0];
    p[1] -= m * tA[1];
    p[2] -= m * tA[2];
    return;
  }
  /* OUTLET: zero normal gradient of the velocity at the face */
  double wfn = velC[0] * n[0] + velC[1] * n[1];
  wfn += velC[2] * n[2];
  const double m = rhoC * wfn * fArea;
  p[0] += ss[0] * s[0] + ss[3] * s[1] + ss[5] * s[2];
  p[1] += ss[3] * s[0] + ss[1] * s[1] + ss[4] * s[2];
  p[2] += ss[5] * s[0] + ss[4] * s[1] + ss[2] * s[2];
  /* gradFaceNull (src/CellLib.c:5477-5600), ndm 3 */
  const double mod = sqrt(xmcc[0] * xmcc[0] + xmcc[1] * xmcc[1] + xmcc[2] * xmcc[2]);
  const double ef[3] = {xmcc[0] / mod, xmcc[1] / mod, xmcc[2] / mod};
  double gf[9], tmp[3];
  for (int i = 0; i < 3; i++) {
    const double pr = gv[i * 3] * ef[0] + gv[i * 3 + 1] * ef[1] + gv[i * 3 + 2] * ef[2];
    for (int j = 0; j < 3; j++) gf[i * 3 + j] = gv[i * 3 + j] - pr * ef[j];
  }
  for (int i = 0; i < 3; i++)
    tmp[i] = gf[i * 3] * xmcc[0] + gf[i * 3 + 1] * xmcc[1] + gf[i * 3 + 2] * xmcc[2];
  p[0] -= m * tmp[0];
  p[1] -= m * tmp[1];
  p[2] -= m * tmp[2];
  sPc[0] += m;
  sPc[1] += m;
  sPc[2] += m;
}

int orc_simple_vel3d(int numel, short maxViz, const short *nFace,
                     const int *nelcon, const int *cellFace, const int *fOwner,
                     const double *volume, const double *dcca,
                     const double *xmcc, const double *cc, const double *mksi,
                     const double *ksi, const double *area,
                     const double *normal, const double *xm,
                     const double *vSkew, const short *mat, const double *prop,
                     const short *faceRvel, const int *id, int nlv,
                     const short *lv_type, const short *lv_np,
                     const double *lv_par, const double *lv_density,
                     const int *opt, const double *optd, const double *pres,
                     const double *gradPres, const double *vel,
                     const double *gradVel, const double *eddyVisc,
                     const double *wallPar, const double *dField, double *dOut,
                     int in_place, double *lA, double *lB, double *rCell)
{
  const orc_vel_opts o = {opt[0], opt[1], opt[2], opt[3], opt[4], opt[5],
                          opt[6], opt[7], opt[8], opt[9], opt[10],
                          optd[0], optd[1], optd[2]};
  if (o.advection != A_FOUP && o.advection != A_CD && o.advection != A_SOUP)
    return -1;
  if (o.diffusion < 1 || o.diffusion > 4) return -2;
  /* momentum terms of these velocity loads are not restated */
  for (int i = 0; i < nlv; i++)
    if (lv_type[i] == L_INLETSTATIC || lv_type[i] == L_OPEN ||
        lv_type[i] == L_INLETTOTAL)
      return -3;
  const int w = maxViz + 1, wa = maxViz + 3;
  for (int c = 0; c < numel; c++) {
    for (int k = 0; k < 3; k++) rCell[k * numel + c] = 0.0;
    const double *dIn = in_place ? dOut : dField;
    const int nf = nFace[c];
    if (faceRvel[c * w + nf] > 0) {                  /* no equations */
      if (!in_place)
        for (int j = 0; j < 3; j++) dOut[c * 3 + j] = dField[c * 3 + j];
      continue;
    }
    const double *pc = prop + (mat[c] - 1) * ORC_MAXPROP;
    const double rhoC = pc[ORC_DENSITY], viscC = pc[ORC_VISC];
    const double eddyC = o.turb ? eddyVisc[c] : 0.0;
    const double effC = viscC + eddyC;
    const double presC = pres[c], V = volume[c];
    double velC[3], gpC[3], dC[3], gvC[9];
    for (int j = 0; j < 3; j++) {
      velC[j] = vel[c * 3 + j];
      gpC[j] = gradPres[c * 3 + j];
      dC[j] = dIn[c * 3 + j];
    }
    for (int j = 0; j < 9; j++) gvC[j] = gradVel[c * 9 + j];
    double sP = 0.0, p[3] = {0.0, 0.0, 0.0}, pf[3] = {0.0, 0.0, 0.0},
           sPc[3] = {0.0, 0.0, 0.0}, la[8];
    for (int f = 0; f < nf; f++) {
      const int face = cellFace[c * maxViz + f] - 1;
      const double ch = (fOwner[face * 2] - 1 == c) ? 1.0 : -1.0;
      double n[3], k[3], xc[3], s[3], e[3], t[3];
      for (int j = 0; j < 3; j++) {
        n[j] = ch * normal[face * 3 + j];
        k[j] = ch * ksi[face * 3 + j];
        xc[j] = xmcc[(c * maxViz + f) * 3 + j];
      }
      const double A = area[face], mk = mksi[face];
      const int nb = nelcon[c * maxViz + f] - 1;
      if (nb > -1) {
        const double *pv = prop + (mat[nb] - 1) * ORC_MAXPROP;
        const double rhoV = pv[ORC_DENSITY], viscV = pv[ORC_VISC];
        const double eddyV = o.turb ? eddyVisc[nb] : 0.0;
        const double presV = pres[nb];
        double velV[3], gpV[3], dV[3], vs[3], duDksi[3], ccV[3], xmf[3];
        const double *gvV = gradVel + nb * 9;
        for (int j = 0; j < 3; j++) {
          velV[j] = vel[nb * 3 + j];
          gpV[j] = gradPres[nb * 3 + j];
          dV[j] = dIn[nb * 3 + j];
          vs[j] = vSkew[face * 3 + j];
          duDksi[j] = (velV[j] - velC[j]) / mk;
          ccV[j] = cc[nb * 3 + j];
          xmf[j] = xm[face * 3 + j];
        }
        for (int j = 0; j < 3; j++) s[j] = A * n[j];
        diffusion_et(o.diffusion, s, k, e, t);
        double v[3];
        v[0] = vs[0] + xc[0];
        v[1] = vs[1] + xc[1];
        v[2] = vs[2] + xc[2];
        const double alpha = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / mk;
        const double am1 = 1.0 - alpha;
        const double effV = viscV + eddyV;
        /* mediaHarmonica(effV, effC, am1, alpha) */
        const double visc = (effC == 0.0 || effV == 0.0) ? 0.0
                            : 1.0 / (alpha / effC + am1 / effV);
        const double coef = visc;
        const double modE = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        const double dfd = coef * modE / mk;
        double gf[9], wf[3];
        for (int i = 0; i < 3; i++) {
          for (int j = 0; j < 3; j++)
            gf[i * 3 + j] = am1 * gvC[i * 3 + j] + alpha * gvV[i * 3 + j];
          wf[i] = am1 * velC[i] + alpha * velV[i];
        }
        const double density = am1 * rhoC + alpha * rhoV;
        double wfn;
        if (o.rhie) {
          double dF[3], g[3];
          for (int j = 0; j < 3; j++) dF[j] = am1 * dC[j] + alpha * dV[j];
          for (int j = 0; j < 3; j++) g[j] = am1 * gpC[j] + alpha * gpV[j];
          wfn = wf[0] * n[0] + wf[1] * n[1] + wf[2] * n[2];
          const double g0 = (presV - presC) / mk;
          const double nk = dF[0] * k[0] * n[0] + dF[1] * k[1] * n[1]
                          + dF[2] * k[2] * n[2];
          double g1 = g[0] * k[0] + g[1] * k[1] + g[2] * k[2];
          g1 = (g1 - g0) * nk;
          wfn += g1;
        } else {
          wfn = wf[0] * n[0] + wf[1] * n[1] + wf[2] * n[2];
        }
        double gcomp[9], dfdc[3];
        for (int i = 0; i < 3; i++) {
          double gk = gf[i * 3] * k[0] + gf[i * 3 + 1] * k[1] + gf[i * 3 + 2] * k[2];
          const double du = duDksi[i] - gk;
          for (int j = 0; j < 3; j++) gcomp[i * 3 + j] = gf[i * 3 + j] + du * k[j];
          gk = gcomp[i * 3] * t[0] + gcomp[i * 3 + 1] * t[1] + gcomp[i * 3 + 2] * t[2];
          dfdc[i] = coef * gk;
        }
        const double cv = density * wfn * A;
        double rv[3], cvc[3];
        rv[0] = xmf[0] - ccV[0];
        rv[1] = xmf[1] - ccV[1];
        rv[2] = xmf[2] - ccV[2];
        for (int i = 0; i < 3; i++) {
          if (o.advection == A_FOUP) {
            cvc[i] = 0.0;
          } else if (o.advection == A_CD) {
            const double cd = am1 * velC[i] + alpha * velV[i];
            cvc[i] = cd - (cv < 0.0 ? velV[i] : velC[i]);
          } else {
            const double *g = cv < 0.0 ? gvV + i * 3 : gvC + i * 3;
            const double *r = cv < 0.0 ? rv : xc;
            double u = g[0] * r[0] + g[1] * r[1];
            u += g[2] * r[2];
            cvc[i] = u;
          }
        }
        la[f] = dfd - (cv < 0.0 ? cv : 0.0);
        sP += cv;
        for (int i = 0; i < 3; i++) p[i] += dfdc[i] - cv * cvc[i];
        /* facePressure */
        double pFace;
        if (o.so_pressure) {
          double p1 = presC + gpC[0] * xc[0] + gpC[1] * xc[1] + gpC[2] * xc[2];
          double p2 = presV + gpV[0] * rv[0] + gpV[1] * rv[1] + gpV[2] * rv[2];
          pFace = 0.5 * (p1 + p2);
        } else {
          pFace = am1 * presC + alpha * presV;
        }
        const double tp = pFace * A;
        for (int j = 0; j < 3; j++) pf[j] += tp * n[j];
        if (o.viscosity) {                 /* viscosityPartExp, fDiv false */
          const double aP = viscC * A;
          for (int j = 0; j < 3; j++)
            p[j] += aP * (gcomp[0 * 3 + j] * n[0] + gcomp[1 * 3 + j] * n[1]
                          + gcomp[2 * 3 + j] * n[2]);
        }
      } else {
        la[f] = 0.0;
        const double tp = presC * A;                /* facePressure, boundary */
        for (int j = 0; j < 3; j++) pf[j] += tp * n[j];
        for (int j = 0; j < 3; j++) s[j] = A * n[j];
        diffusion_et(o.diffusion, s, k, e, t);
        const int lv = faceRvel[c * w + f];
        if (lv) {
          orc_load ld;
          load_at(&ld, lv - 1, lv_type, lv_np, lv_par, lv_density);
          vel_load_mom(&ld, sPc, p, velC, gvC, viscC, effC, s, n, xc, rhoC,
                       o.wall_model ? wallPar + c * 4 : NULL, o.wall_model, A,
                       dcca[c * maxViz + f]);
        }
      }
    }
    if (o.time) {
      if (!o.backward) {
        sP += rhoC * V / o.dt;
      } else {
        const double tmp = 1.0 / o.dt + 1.0 / (o.dt + o.dt0);
        sP += tmp * rhoC * V;
      }
    }
    double lAn = 0.0;
    for (int f = 0; f < nf; f++) lAn += la[f];
    double ap[3];
    for (int k = 0; k < 3; k++) ap[k] = sP + sPc[k] + lAn;
    for (int k = 0; k < 3; k++) ap[k] = ap[k] / o.underU;
    for (int k = 0; k < 3; k++) p[k] += (1 - o.underU) * ap[k] * velC[k];
    for (int k = 0; k < 3; k++) p[k] -= pf[k];
    double rc[3] = {0.0, 0.0, 0.0};
    for (int f = 0; f < nf; f++) {
      const int nb = nelcon[c * maxViz + f] - 1;
      if (nb > -1) {
        const int lid = id[nb] - 1;
        for (int k = 0; k < 3; k++) {
          if (lid == -2) p[k] += la[f] * vel[nb * 3 + k];
          else rc[k] += la[f] * vel[nb * 3 + k];
        }
      }
    }
    for (int k = 0; k < 3; k++) rc[k] += p[k] - ap[k] * velC[k];
    for (int k = 0; k < 3; k++)
      dOut[c * 3 + k] = o.simplec ? V / (ap[k] - lAn) : V / ap[k];
    double *out = lA + (long)c * wa;
    for (int f = 0; f < nf; f++) out[f] = -la[f];
    for (int k = 0; k < 3; k++) out[maxViz + k] = ap[k];
    for (int k = 0; k < 3; k++) {
      lB[c * 3 + k] = o.residual ? rc[k] : p[k];
      rCell[k * numel + c] = rc[k];
    }
  }
  return 0;
}
