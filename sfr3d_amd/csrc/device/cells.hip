// Finite-volume cell coefficients on the device: SFR3D's SIMPLE
// pressure-correction cell loop (systFormSimplePres,
// src/CellLoopInc.c:471-679, with cellSimplePres3D, src/CellPres.c:903-1149)
// for 3-D cells.
//
// One work-item per cell.  The geometry is stored per face slot,
// slot-major (entry f * numel + c) like the solver's ELL, so the six slots
// of a wavefront's cells are six coalesced sweeps; the signs of the owner's
// face vectors and the linear interpolation weight are folded in at
// upload.  The neighbours' fields are gathers that hit in L2 for cells
// close in the numbering (the XCD-aware chunk order keeps a contiguous range
// of cells per XCD).
//
// Arithmetic keeps the reference's operation order with contraction off,
// so each coefficient is the reference's to the last bit (division and
// square root are correctly rounded in double precision).
#include "dev_common.hpp"

// Read-once streams (the face geometry planes, the face neighbour and load
// indices) are non-temporal loads, so that they do not push out of L2 the
// cell fields (velocity, gradients, pressure, d) that the neighbour gathers
// of nearby cells read again: a z-plane of 128^2 cells streams about 12 MB
// of geometry, three times an XCD's L2.  Measured at 128^3: -6 % pressure
// loop, -3 % momentum loop.  Non-temporal stores of the outputs (rows of
// 7 doubles, partial lines) measured 60 % and 14 % slower and are not used.
template <class T>
__device__ __forceinline__ T lds(const T *p)
{
  return __builtin_nontemporal_load(p);
}

namespace {

constexpr int L_DIRICHLET = 1, L_INLET = 6, L_OUTLET = 7, L_INLETSTATIC = 9,
              L_OPEN = 10, L_INLETTOTAL = 13;

// e of the implicit part of the diffusive flux through s
// (difusionSchemeAnisotropic, src/CellLib.c:6054-6190, ndm 3)
__device__ __forceinline__ void diffusion_e(int code, const double *s,
                                            const double *k, double *e)
{
#pragma clang fp contract(off)
  if (code == 1) {
    const double ss = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    e[0] = ss * k[0];
    e[1] = ss * k[1];
    e[2] = ss * k[2];
  } else {
    const double ss = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const double nk = s[0] * k[0] + s[1] * k[1] + s[2] * k[2];
    const double t = ss / nk;
    e[0] = t * k[0];
    e[1] = t * k[1];
    e[2] = t * k[2];
  }
}

// boundary pressure load (pLoadSimplePres, src/Loads.c:728-845, not
// buoyant): its contribution to the diagonal sP
__device__ __forceinline__ void pres_load(const sfr_load &ld, double *sP,
                                          const double *e, const double *n,
                                          const double *vc, double rhoC,
                                          double A, double dd)
{
#pragma clang fp contract(off)
  double t = e[0] * e[0] + e[1] * e[1];
  t += e[2] * e[2];
  const double modE = sqrt(t);
  if (ld.type == L_DIRICHLET) {
    *sP += rhoC * modE / dd;
  } else if (ld.type == L_INLETTOTAL) {
    const double rhoE = ld.par[0];
    const double dc = modE / dd;
    const double vn = vc[0] * n[0] + vc[1] * n[1] + vc[2] * n[2];
    const double en = (ld.par[2] * n[0] + ld.par[3] * n[1] + ld.par[4] * n[2]) * A;
    const double m = rhoE * vn * A;
    const double mv = m / (rhoE * en);
    const double rv = rhoE * mv;
    const double den = m - dc * rv * rv;
    *sP += rhoE * m * dc / den;
  } else if (ld.type == L_OPEN) {
    double vn = vc[0] * n[0] + vc[1] * n[1];
    vn += vc[2] * n[2];
    if (vn > 0.0) {               // outflow: the reference adds the term twice
      *sP += rhoC * modE / dd;
      *sP += rhoC * modE / dd;
    }
  }
  // FLUXPRES without buoyancy and every other type: nothing
}

// boundary velocity load's mass flux (pLoadSimple, src/Loads.c:265-727,
// pressure equation only)
__device__ __forceinline__ bool vel_flux(const sfr_load &ld, const double *n,
                                         const double *vc, double rhoC,
                                         double A, double *m)
{
#pragma clang fp contract(off)
  if (ld.type == L_INLETSTATIC || ld.type == L_INLETTOTAL) {
    const double vn = vc[0] * n[0] + vc[1] * n[1] + vc[2] * n[2];
    *m = (ld.type == L_INLETSTATIC ? rhoC : ld.par[0]) * vn * A;
    return true;
  }
  if (ld.type == L_INLET) {
    double wfn = ld.par[0] * n[0] + ld.par[1] * n[1];
    wfn += ld.par[2] * n[2];
    *m = ld.par[ld.np - 1] * wfn * A;
    return true;
  }
  if (ld.type == L_OUTLET || ld.type == L_OPEN) {
    double wfn = vc[0] * n[0] + vc[1] * n[1];
    wfn += vc[2] * n[2];
    const double rho = (ld.type == L_OPEN && !(wfn >= 0.0)) ? ld.density : rhoC;
    *m = rho * wfn * A;
    return true;
  }
  return false;                     // walls: no flux
}

struct CellPres {
  int numel, maxViz, diffusion;
  const signed char *nface;         // -1: no pressure equation
  const int32_t *nb;                // slot-major; -1 boundary
  const int32_t *bc;                // velocity load + 1 | (pressure + 1) << 16
  const double *g;                  // 9 slot-major planes (see cells.c)
  const double *rho;
  const sfr_load *lv, *lp;
  const double *pres, *gp, *vel, *df;
  double *lA, *lB, *rc;
};

__global__ __launch_bounds__(SFR_BLOCK) void k_cell_pres(CellPres a)
{
#pragma clang fp contract(off)
  const int64_t c = (int64_t)xcd_chunk(blockIdx.x, gridDim.x) * SFR_BLOCK + threadIdx.x;
  if (c >= a.numel) return;
  const int nf = a.nface[c];
  if (a.rc) a.rc[c] = 0.0;
  if (nf < 0) return;
  const int64_t S = (int64_t)a.maxViz * a.numel;   // plane stride
  const double rhoC = a.rho[c], presC = a.pres[c];
  double vc[3], gc[3], dc[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    vc[j] = a.vel[c * 3 + j];
    gc[j] = a.gp[c * 3 + j];
    dc[j] = a.df[c * 3 + j];
  }
  double p = 0.0, sP = 0.0, la[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    if (f >= nf) break;
    const int64_t q = (int64_t)f * a.numel + c;
    const double n[3] = {lds(a.g + (q)), lds(a.g + (S + q)), lds(a.g + (2 * S + q))};
    const double k[3] = {lds(a.g + (3 * S + q)), lds(a.g + (4 * S + q)), lds(a.g + (5 * S + q))};
    const double A = lds(a.g + (6 * S + q)), mk = lds(a.g + (7 * S + q)), w = lds(a.g + (8 * S + q));
    const int nb = lds(a.nb + q);
    double s[3], e[3];
    if (nb > -1) {
      const double alpha = w, am1 = 1.0 - alpha;
      const double density = am1 * rhoC + alpha * a.rho[nb];
      double dF[3], wf[3], gf[3];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        dF[j] = am1 * dc[j] + alpha * a.df[nb * 3 + j];
        wf[j] = am1 * vc[j] + alpha * a.vel[nb * 3 + j];
        gf[j] = am1 * gc[j] + alpha * a.gp[nb * 3 + j];
      }
#pragma unroll
      for (int j = 0; j < 3; j++) s[j] = dF[j] * A * n[j];
      diffusion_e(a.diffusion, s, k, e);
      const double modE = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      la[f] = density * modE / mk;
      // face velocity with the pressure-gradient correction
      // (interpolFaceVel, src/CellLib.c:6853-6940)
      double wfn = wf[0] * n[0] + wf[1] * n[1] + wf[2] * n[2];
      const double g0 = (a.pres[nb] - presC) / mk;
      const double nk = dF[0] * k[0] * n[0] + dF[1] * k[1] * n[1]
                      + dF[2] * k[2] * n[2];
      double g1 = gf[0] * k[0] + gf[1] * k[1] + gf[2] * k[2];
      g1 = (g1 - g0) * nk;
      wfn += g1;
      p -= density * wfn * A;
    } else {
      la[f] = 0.0;
      const int32_t bc = lds(a.bc + q);
      const int lvi = (bc & 0xffff) - 1, lpi = (bc >> 16) - 1;
      if (lpi >= 0) {
#pragma unroll
        for (int j = 0; j < 3; j++) s[j] = dc[j] * A * n[j];
        diffusion_e(a.diffusion, s, k, e);
        pres_load(a.lp[lpi], &sP, e, n, vc, rhoC, A, w);
      }
      if (lvi >= 0) {
        double m;
        if (vel_flux(a.lv[lvi], n, vc, rhoC, A, &m)) p -= m;
      }
    }
  }
  double ap = sP;
#pragma unroll
  for (int f = 0; f < 6; f++)
    if (f < nf) ap += la[f];
  double *out = a.lA + c * (a.maxViz + 1);
#pragma unroll
  for (int f = 0; f < 6; f++)
    if (f < nf) out[f] = -la[f];
  out[a.maxViz] = ap;
  a.lB[c] = p;
  if (a.rc) a.rc[c] = p;
}

// ---- the momentum cell loop ------------------------------------------------
// systFormSimpleVel (src/CellLoopInc.c:98-385) with cellSimpleVel3D
// (src/CellSimple.c:106-638) for u, v, w together, one work-item per cell.
// Every neighbour's d is the input one (the reference's serial loop writes
// d in place; see sfr3d.h).  Planes 9-11 hold xmcc, 12-14 the vector from
// the neighbour's centroid to the face midpoint.

constexpr int L_MOVEWALL = 8, L_STATICWALL = 14;

// difusionSchemeNew (src/CellLib.c:5824-5990), ndm 3: e and t
__device__ __forceinline__ void diffusion_et(int code, const double *s,
                                             const double *k, double *e,
                                             double *t)
{
#pragma clang fp contract(off)
  if (code == 2) {
    const double nk = s[0] * k[0] + s[1] * k[1] + s[2] * k[2];
    e[0] = nk * k[0];
    e[1] = nk * k[1];
    e[2] = nk * k[2];
  } else if (code == 4) {
    const double ss = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const double nk = s[0] * k[0] + s[1] * k[1] + s[2] * k[2];
    const double tmp = ss / nk;
    e[0] = tmp * k[0];
    e[1] = tmp * k[1];
    e[2] = tmp * k[2];
  } else {                                   // 1 ORTHOGONAL, 3 ORTHOGONALC
    const double ss = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    e[0] = ss * k[0];
    e[1] = ss * k[1];
    e[2] = ss * k[2];
  }
#pragma unroll
  for (int j = 0; j < 3; j++) t[j] = code == 1 ? 0.0 : s[j] - e[j];
}

// pLoadSimple, momentum part (src/Loads.c:265-520): MOVEWALL / STATICWALL
// through moveWall (src/CellLib.c:7715-7760), INLET and OUTLET with stress
// (src/CellLib.c:7013-7040) and gradFaceNull (src/CellLib.c:5477-5600)
__device__ __forceinline__ void vel_load_mom(const sfr_load &ld, double *sPc,
    double *p, const double *vc, const double *gv, double viscC,
    double effC, const double *s, const double *n, const double *xmcc,
    double rhoC, const double *wall, int wallModel, double A, double dcca)
{
#pragma clang fp contract(off)
  if (ld.type == L_MOVEWALL || ld.type == L_STATICWALL) {
    double tA[3];
#pragma unroll
    for (int j = 0; j < 3; j++) tA[j] = ld.type == L_MOVEWALL ? ld.par[j] : 0.0;
    double vw;
    if (wallModel) {
      const double yPlus = wall[0], uPlus = wall[1];
      vw = yPlus > 0.01 ? viscC * yPlus / uPlus : viscC;
    } else {
      vw = effC;
    }
    const double aP = vw * A / dcca;
    sPc[0] += aP * (1.0 - n[0] * n[0]);
    sPc[1] += aP * (1.0 - n[1] * n[1]);
    sPc[2] += aP * (1.0 - n[2] * n[2]);
    p[0] += aP * (tA[0] * (1.0 - n[0] * n[0]) + (vc[1] - tA[1]) * n[1] * n[0]
                  + (vc[2] - tA[2]) * n[2] * n[0]);
    p[1] += aP * (tA[1] * (1.0 - n[1] * n[1]) + (vc[0] - tA[0]) * n[0] * n[1]
                  + (vc[2] - tA[2]) * n[2] * n[1]);
    p[2] += aP * (tA[2] * (1.0 - n[2] * n[2]) + (vc[0] - tA[0]) * n[0] * n[2]
                  + (vc[1] - tA[1]) * n[1] * n[2]);
    return;
  }
  if (ld.type != L_INLET && ld.type != L_OUTLET) return;
  const double tr = gv[0] + gv[4] + gv[8];
  double ss[6];
  ss[0] = 2.0 * viscC * gv[0] + 0.0 * tr;
  ss[1] = 2.0 * viscC * gv[4] + 0.0 * tr;
  ss[2] = 2.0 * viscC * gv[8] + 0.0 * tr;
  ss[3] = viscC * (gv[3] + gv[1]);
  ss[4] = viscC * (gv[7] + gv[5]);
  ss[5] = viscC * (gv[6] + gv[2]);
  if (ld.type == L_INLET) {
    double wfn = ld.par[0] * n[0] + ld.par[1] * n[1];
    wfn += ld.par[2] * n[2];
    const double m = ld.par[ld.np - 1] * wfn * A;
    p[0] += ss[0] * s[0] + ss[3] * s[1] + ss[5] * s[2];
    p[1] += ss[3] * s[0] + ss[1] * s[1] + ss[4] * s[2];
    p[2] += ss[5] * s[0] + ss[4] * s[1] + ss[2] * s[2];
    p[0] -= m * ld.par[0];
    p[1] -= m * ld.par[1];
    p[2] -= m * ld.par[2];
    return;
  }
  double wfn = vc[0] * n[0] + vc[1] * n[1];
  wfn += vc[2] * n[2];
  const double m = rhoC * wfn * A;
  p[0] += ss[0] * s[0] + ss[3] * s[1] + ss[5] * s[2];
  p[1] += ss[3] * s[0] + ss[1] * s[1] + ss[4] * s[2];
  p[2] += ss[5] * s[0] + ss[4] * s[1] + ss[2] * s[2];
  const double mod = sqrt(xmcc[0] * xmcc[0] + xmcc[1] * xmcc[1] + xmcc[2] * xmcc[2]);
  const double ef[3] = {xmcc[0] / mod, xmcc[1] / mod, xmcc[2] / mod};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double pr = gv[i * 3] * ef[0] + gv[i * 3 + 1] * ef[1] + gv[i * 3 + 2] * ef[2];
    const double g0 = gv[i * 3] - pr * ef[0], g1 = gv[i * 3 + 1] - pr * ef[1],
                 g2 = gv[i * 3 + 2] - pr * ef[2];
    const double tmp = g0 * xmcc[0] + g1 * xmcc[1] + g2 * xmcc[2];
    p[i] -= m * tmp;
  }
  sPc[0] += m;
  sPc[1] += m;
  sPc[2] += m;
}

struct CellVel {
  sfr_vel_opts o;
  int numel, maxViz, neq;
  const signed char *nface;
  const int32_t *nb, *bc;
  const double *g, *rho, *visc, *vol;
  const sfr_load *lv;
  const int32_t *id;
  const double *pres, *gp, *vel, *gv, *eddy, *wall, *din;
  double *dout, *lA, *ad, *b, *rc;
};

__global__ __launch_bounds__(SFR_BLOCK) void k_cell_vel(CellVel a)
{
#pragma clang fp contract(off)
  const int64_t c = (int64_t)xcd_chunk(blockIdx.x, gridDim.x) * SFR_BLOCK + threadIdx.x;
  if (c >= a.numel) return;
  const sfr_vel_opts &o = a.o;
  const int nf = a.nface[c];
#pragma unroll
  for (int k = 0; k < 3; k++) a.rc[k * a.numel + c] = 0.0;
  if (nf < 0) {                              // prescribed velocity: d kept
#pragma unroll
    for (int j = 0; j < 3; j++) a.dout[c * 3 + j] = a.din[c * 3 + j];
    return;
  }
  const int64_t S = (int64_t)a.maxViz * a.numel;
  const double rhoC = a.rho[c], viscC = a.visc[c];
  const double effC = viscC + (o.turb ? a.eddy[c] : 0.0);
  const double presC = a.pres[c], V = a.vol[c];
  double vc[3], gpC[3], dC[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    vc[j] = a.vel[c * 3 + j];
    gpC[j] = a.gp[c * 3 + j];
    dC[j] = a.din[c * 3 + j];
  }
  // the velocity gradients are read a row at a time where they are used
  // (from cache), which keeps the live set small enough not to spill
  const double *gvC = a.gv + c * 9;
  double sP = 0.0, p[3] = {0.0, 0.0, 0.0}, pf[3] = {0.0, 0.0, 0.0},
         sPc[3] = {0.0, 0.0, 0.0}, la[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    if (f < nf) {
    const int64_t q = (int64_t)f * a.numel + c;
    const double n[3] = {lds(a.g + (q)), lds(a.g + (S + q)), lds(a.g + (2 * S + q))};
    const double k[3] = {lds(a.g + (3 * S + q)), lds(a.g + (4 * S + q)), lds(a.g + (5 * S + q))};
    const double A = lds(a.g + (6 * S + q)), mk = lds(a.g + (7 * S + q)), w = lds(a.g + (8 * S + q));
    const double xc[3] = {lds(a.g + (9 * S + q)), lds(a.g + (10 * S + q)), lds(a.g + (11 * S + q))};
    const int nb = lds(a.nb + q);
    double s[3], e[3], t[3];
    if (nb > -1) {
      const double rv[3] = {lds(a.g + (12 * S + q)), lds(a.g + (13 * S + q)), lds(a.g + (14 * S + q))};
      const double rhoV = a.rho[nb], viscV = a.visc[nb];
      const double effV = viscV + (o.turb ? a.eddy[nb] : 0.0);
      const double presV = a.pres[nb];
      const double *gvV = a.gv + (int64_t)nb * 9;
#pragma unroll
      for (int j = 0; j < 3; j++) s[j] = A * n[j];
      diffusion_et(o.diffusion, s, k, e, t);
      const double alpha = w, am1 = 1.0 - alpha;
      // mediaHarmonica (src/CellLib.c:7879-7884)
      const double coef = (effC == 0.0 || effV == 0.0) ? 0.0
                          : 1.0 / (alpha / effC + am1 / effV);
      const double modE = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      const double dfd = coef * modE / mk;
      double wf[3];
#pragma unroll
      for (int i = 0; i < 3; i++) wf[i] = am1 * vc[i] + alpha * a.vel[nb * 3 + i];
      const double density = am1 * rhoC + alpha * rhoV;
      double wfn = wf[0] * n[0] + wf[1] * n[1] + wf[2] * n[2];
      if (o.rhie) {                          // interpolFaceVel
        double dF[3], gfp[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
          dF[j] = am1 * dC[j] + alpha * a.din[nb * 3 + j];
          gfp[j] = am1 * gpC[j] + alpha * a.gp[nb * 3 + j];
        }
        const double g0 = (presV - presC) / mk;
        const double nk = dF[0] * k[0] * n[0] + dF[1] * k[1] * n[1]
                        + dF[2] * k[2] * n[2];
        double g1 = gfp[0] * k[0] + gfp[1] * k[1] + gfp[2] * k[2];
        g1 = (g1 - g0) * nk;
        wfn += g1;
      }
      const double cv = density * wfn * A;
      la[f] = dfd - (cv < 0.0 ? cv : 0.0);
      sP += cv;
      // per velocity component i: the interpolated gradient row, its
      // compact (Darwish-Moukalled) correction, the non-orthogonal flux and
      // the deferred advection term; the explicit viscous flux accumulates
      // the rows in the reference's order (row 0, 1, 2)
      double vis[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double vVi = a.vel[nb * 3 + i];
        const double duDksi = (vVi - vc[i]) / mk;
        double gf[3], gc[3];
#pragma unroll
        for (int j = 0; j < 3; j++)
          gf[j] = am1 * gvC[i * 3 + j] + alpha * gvV[i * 3 + j];
        double gk = gf[0] * k[0] + gf[1] * k[1] + gf[2] * k[2];
        const double du = duDksi - gk;
#pragma unroll
        for (int j = 0; j < 3; j++) gc[j] = gf[j] + du * k[j];
        gk = gc[0] * t[0] + gc[1] * t[1] + gc[2] * t[2];
        const double dfdc = coef * gk;
        double cvc;
        if (o.advection == 1) {
          cvc = 0.0;
        } else if (o.advection == 2) {       // deferredCd
          const double cd = am1 * vc[i] + alpha * vVi;
          cvc = cd - (cv < 0.0 ? vVi : vc[i]);
        } else {                             // upwindLinearV1
          const double *gg = cv < 0.0 ? gvV + i * 3 : gvC + i * 3;
          const double *r = cv < 0.0 ? rv : xc;
          double u = gg[0] * r[0] + gg[1] * r[1];
          u += gg[2] * r[2];
          cvc = u;
        }
        p[i] += dfdc - cv * cvc;
#pragma unroll
        for (int j = 0; j < 3; j++) vis[j] = i == 0 ? gc[j] * n[0] : vis[j] + gc[j] * n[i];
      }
      double pFace;                          // facePressure
      if (o.so_pressure) {
        const double p1 = presC + gpC[0] * xc[0] + gpC[1] * xc[1] + gpC[2] * xc[2];
        const double p2 = presV + a.gp[nb * 3] * rv[0] + a.gp[nb * 3 + 1] * rv[1]
                          + a.gp[nb * 3 + 2] * rv[2];
        pFace = 0.5 * (p1 + p2);
      } else {
        pFace = am1 * presC + alpha * presV;
      }
      const double tp = pFace * A;
#pragma unroll
      for (int j = 0; j < 3; j++) pf[j] += tp * n[j];
      if (o.viscosity) {                     // viscosityPartExp, no divergence
        const double aP = viscC * A;
#pragma unroll
        for (int j = 0; j < 3; j++) p[j] += aP * vis[j];
      }
    } else {
      la[f] = 0.0;
      const double tp = presC * A;
#pragma unroll
      for (int j = 0; j < 3; j++) pf[j] += tp * n[j];
#pragma unroll
      for (int j = 0; j < 3; j++) s[j] = A * n[j];
      diffusion_et(o.diffusion, s, k, e, t);
      const int lvi = (lds(a.bc + q) & 0xffff) - 1;
      if (lvi >= 0)
        vel_load_mom(a.lv[lvi], sPc, p, vc, gvC, viscC, effC, s, n, xc, rhoC,
                     o.wall_model ? a.wall + c * 4 : nullptr, o.wall_model, A, w);
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
#pragma unroll
  for (int f = 0; f < 6; f++)
    if (f < nf) lAn += la[f];
  double ap[3], rcl[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ap[k] = sP + sPc[k] + lAn;
    ap[k] = ap[k] / o.underU;
    p[k] += (1 - o.underU) * ap[k] * vc[k];
    p[k] -= pf[k];
  }
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int nb = f < nf ? lds(a.nb + (int64_t)f * a.numel + c) : -1;
    if (nb > -1) {
      const bool fixed = a.id[nb] - 1 == -2;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (fixed) p[k] += la[f] * a.vel[nb * 3 + k];
        else rcl[k] += la[f] * a.vel[nb * 3 + k];
      }
    }
  }
  double *out = a.lA + c * (a.maxViz + 1);
#pragma unroll
  for (int f = 0; f < 6; f++)
    if (f < nf) out[f] = -la[f];
  out[a.maxViz] = ap[0];
  const int32_t row = a.id[c] - 1;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    rcl[k] += p[k] - ap[k] * vc[k];
    a.dout[c * 3 + k] = o.simplec ? V / (ap[k] - lAn) : V / ap[k];
    a.rc[k * a.numel + c] = rcl[k];
    if (row > -1 && row < a.neq) {
      a.ad[(int64_t)k * a.neq + row] = ap[k];
      a.b[(int64_t)k * a.neq + row] = o.residual ? rcl[k] : p[k];
    }
  }
}

// updateCellValueBlock (src/CellLoop.c:4087-4130), ndf 1: cell values from
// a solution vector through the equation numbers
__global__ __launch_bounds__(SFR_BLOCK) void k_cell_value(int64_t n,
    const int32_t *__restrict__ id, const double *__restrict__ x,
    double *__restrict__ u, int add)
{
  const int64_t c = (int64_t)xcd_chunk(blockIdx.x, gridDim.x) * SFR_BLOCK + threadIdx.x;
  if (c >= n) return;
  const int32_t q = id[c] - 1;
  if (q > -1) u[c] = add ? u[c] + x[q] : x[q];
}

// simpleUpdate (src/Simple.c:1615-1668), ndm 3: w -= d grad(p'); p += a p'
__global__ __launch_bounds__(SFR_BLOCK) void k_simple_update(int64_t n,
    double alpha, double *__restrict__ w, double *__restrict__ p,
    const double *__restrict__ pc, const double *__restrict__ gpc,
    const double *__restrict__ d)
{
#pragma clang fp contract(off)
  const int64_t c = (int64_t)xcd_chunk(blockIdx.x, gridDim.x) * SFR_BLOCK + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int j = 0; j < 3; j++) w[c * 3 + j] -= d[c * 3 + j] * gpc[c * 3 + j];
  p[c] += alpha * pc[c];
}

static unsigned cell_grid(int64_t n)
{
  int64_t g = (n + SFR_BLOCK - 1) / SFR_BLOCK;
  return (unsigned)((g + 7) / 8 * 8);
}

}  // namespace

extern "C" int sfr_k_cell_value(hipStream_t s, int64_t n, const int32_t *id,
                                const double *x, double *u, int add)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_value, dim3(cell_grid(n)), dim3(SFR_BLOCK), 0, s, n,
                     id, x, u, add);
  return hipGetLastError();
}

extern "C" int sfr_k_simple_update(hipStream_t s, int64_t n, double alpha,
                                   double *w, double *p, const double *pc,
                                   const double *gpc, const double *d)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_simple_update, dim3(cell_grid(n)), dim3(SFR_BLOCK), 0, s,
                     n, alpha, w, p, pc, gpc, d);
  return hipGetLastError();
}

extern "C" int sfr_k_cell_vel(hipStream_t s, int numel, int maxViz,
                              const sfr_vel_opts *o, const signed char *nfacev,
                              const int32_t *nb, const int32_t *bc,
                              const double *g, const double *rho,
                              const double *visc, const double *vol,
                              const sfr_load *lv, const int32_t *id, int neq,
                              const double *pres, const double *gp,
                              const double *vel, const double *gv,
                              const double *eddy, const double *wall,
                              const double *din, double *dout, double *lA,
                              double *ad, double *b, double *rc)
{
  if (numel <= 0) return hipSuccess;
  const CellVel a = {*o, numel, maxViz, neq, nfacev, nb, bc, g, rho, visc, vol,
                     lv, id, pres, gp, vel, gv, eddy, wall, din, dout, lA, ad,
                     b, rc};
  hipLaunchKernelGGL(k_cell_vel, dim3(cell_grid(numel)), dim3(SFR_BLOCK), 0, s, a);
  return hipGetLastError();
}

extern "C" int sfr_k_cell_pres(hipStream_t s, int numel, int maxViz,
                               int diffusion, const signed char *nface,
                               const int32_t *nb, const int32_t *bc,
                               const double *g, const double *rho,
                               const sfr_load *lv, const sfr_load *lp,
                               const double *pres, const double *gp,
                               const double *vel, const double *df,
                               double *lA, double *lB, double *rc)
{
  if (numel <= 0) return hipSuccess;
  const CellPres a = {numel, maxViz, diffusion, nface, nb, bc, g, rho, lv, lp,
                      pres, gp, vel, df, lA, lB, rc};
  hipLaunchKernelGGL(k_cell_pres, dim3(cell_grid(numel)), dim3(SFR_BLOCK), 0, s, a);
  return hipGetLastError();
}
