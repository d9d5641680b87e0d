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
