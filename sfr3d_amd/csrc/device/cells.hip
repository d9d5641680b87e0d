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
    const double n[3] = {a.g[q], a.g[S + q], a.g[2 * S + q]};
    const double k[3] = {a.g[3 * S + q], a.g[4 * S + q], a.g[5 * S + q]};
    const double A = a.g[6 * S + q], mk = a.g[7 * S + q], w = a.g[8 * S + q];
    const int nb = a.nb[q];
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
      const int32_t bc = a.bc[q];
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
