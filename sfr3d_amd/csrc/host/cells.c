/*
 * Cell coefficients on the device: set-up of the mesh geometry for the
 * SIMPLE pressure-correction cell loop (systFormSimplePres,
 * src/CellLoopInc.c:471-679; cell kernel cellSimplePres3D,
 * src/CellPres.c:903-1149) and its launch.
 *
 * The caller's arrays are SFR3D's Mesh members (face arrays indexed through
 * cellFace, seen from the face's owner).  They are rearranged once into
 * per-slot planes, slot-major (entry f * numel + c):
 *   0-2  normal as the cell sees it (the owner's, negated for the other cell)
 *   3-5  ksi, the same way
 *   6    area       7  mksi
 *   8    interior: the linear interpolation weight alpha = |vSkew + xmcc| /
 *        mksi (interpolFace, src/CellLib.c:7174-7200, computed here in the
 *        reference's operation order); boundary: dcca
 * plus the neighbour (0-based, -1 on the boundary), the slot's load indices
 * and each cell's density (prop of its material).  Sign flips and the
 * weight are exact restatements of what the reference computes per cell,
 * so the device sees bitwise the operands the reference's loop forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../sfr_internal.h"

#define TRY(call)                                                            \
  do {                                                                       \
    hipError_t e_ = (hipError_t)(call);                                      \
    if (e_ != hipSuccess) {                                                  \
      rc = sfr_set_error(ctx, SFR_HIP_ERROR, "%s: %s (%s:%d)", #call,        \
                         hipGetErrorString(e_), __FILE__, __LINE__);         \
      goto done;                                                             \
    }                                                                        \
  } while (0)

#define PROP_STRIDE 10   /* MAXPROP, Define.h:349 */
#define PROP_DENSITY 1   /* DENSITY, Define.h:203 */
#define PROP_VISC 0      /* DYNAMICVISCOSITY, Define.h:202 */
#define NPLANES 9        /* the pressure loop's planes                   */
#define NPLANES_VEL 15   /* + xmcc (9-11), xm - cc of the neighbour (12-14) */

struct sfr_cells {
  int32_t      numel;
  int          maxViz;
  signed char *nface;    /* faces per cell, -1: no pressure equation */
  int32_t     *nb;       /* [maxViz * numel] slot-major                */
  int32_t     *bc;       /* velocity load + 1 | (pressure load + 1) << 16 */
  double      *g;        /* NPLANES (NPLANES_VEL) planes of maxViz*numel */
  double      *rho;      /* [numel]                                    */
  sfr_load    *lv, *lp;  /* load tables (at least one entry each)      */
  /* momentum (mesh given with cc and xm) */
  int          vel;      /* the momentum planes exist                  */
  int          vel_loads;/* every velocity load is one the loop restates */
  signed char *nfacev;   /* faces per cell, -1: prescribed velocity    */
  double      *visc;     /* [numel] dynamic viscosity of the material  */
  double      *vol;      /* [numel]                                    */
};

void sfr_cells_destroy(sfr_context *ctx, sfr_cells *C)
{
  if (!C) return;
  if (ctx) hipStreamSynchronize(ctx->stream);
  hipFree(C->nface);
  hipFree(C->nb);
  hipFree(C->bc);
  hipFree(C->g);
  hipFree(C->rho);
  hipFree(C->lv);
  hipFree(C->lp);
  hipFree(C->nfacev);
  hipFree(C->visc);
  hipFree(C->vol);
  free(C);
}

static int check_loads(sfr_context *ctx, const sfr_load *ld, int n, const char *what)
{
  for (int i = 0; i < n; i++) {
    if (ld[i].np < 0 || ld[i].np > 8)
      return sfr_set_error(ctx, SFR_BAD_ARG, "cells: %s load %d has %d parameters",
                           what, i + 1, ld[i].np);
    if (ld[i].type == 6 && ld[i].np < 4)             /* INLET: u v w density */
      return sfr_set_error(ctx, SFR_BAD_ARG,
                           "cells: %s load %d (INLET) needs 4 parameters", what, i + 1);
  }
  return SFR_OK;
}

int sfr_cells_create(sfr_context *ctx, const sfr_cell_mesh *m,
                     const double *prop, int nmat,
                     const sfr_load *loadsVel, int nlv,
                     const sfr_load *loadsPres, int nlp, sfr_cells **out)
{
  if (!ctx || !m || !prop || !out || nmat <= 0 || nlv < 0 || nlp < 0 ||
      (nlv && !loadsVel) || (nlp && !loadsPres))
    return SFR_BAD_ARG;
  *out = NULL;
  if (m->ndm != 3)
    return sfr_set_error(ctx, SFR_NOT_SUPPORTED, "cells: ndm %d (3-D cells only)", m->ndm);
  if (m->numel <= 0 || m->nface <= 0 || m->maxViz < 1 || m->maxViz > 6 ||
      !m->nFace || !m->nelcon || !m->cellFace || !m->owner || !m->dcca ||
      !m->xmcc || !m->mksi || !m->ksi || !m->area || !m->normal ||
      !m->vSkew || !m->mat || !m->faceRvel || !m->faceRpres)
    return sfr_set_error(ctx, SFR_BAD_ARG, "cells: incomplete mesh");
  int rc = check_loads(ctx, loadsVel, nlv, "velocity");
  if (rc == SFR_OK) rc = check_loads(ctx, loadsPres, nlp, "pressure");
  if (rc != SFR_OK) return rc;
  const int64_t n = m->numel, mv = m->maxViz, S = mv * n, w = mv + 1;
  const int vel = m->cc && m->xm && m->volume;
  const int np = vel ? NPLANES_VEL : NPLANES;
  signed char *hnf = malloc((size_t)n), *hnfv = malloc((size_t)n);
  int32_t *hnb = malloc(sizeof(int32_t) * (size_t)S);
  int32_t *hbc = malloc(sizeof(int32_t) * (size_t)S);
  double *hg = calloc((size_t)(np * S), sizeof(double));
  double *hrho = malloc(sizeof(double) * (size_t)n);
  double *hvisc = malloc(sizeof(double) * (size_t)n);
  sfr_cells *C = calloc(1, sizeof *C);
  if (!hnf || !hnfv || !hnb || !hbc || !hg || !hrho || !hvisc || !C) {
    rc = sfr_set_error(ctx, SFR_BAD_ARG, "cells: host allocation");
    goto done;
  }
  for (int64_t c = 0; c < n; c++) {
    const int nf = m->nFace[c];
    const int mat = m->mat[c];
    if (nf < 1 || nf > mv || mat < 1 || mat > nmat) {
      rc = sfr_set_error(ctx, SFR_BAD_ARG, "cells: cell %lld has %d faces, material %d",
                         (long long)c + 1, nf, mat);
      goto done;
    }
    hrho[c] = prop[(int64_t)(mat - 1) * PROP_STRIDE + PROP_DENSITY];
    hvisc[c] = prop[(int64_t)(mat - 1) * PROP_STRIDE + PROP_VISC];
    hnf[c] = m->faceRpres[c * w + nf] > 0 ? (signed char)-1 : (signed char)nf;
    hnfv[c] = m->faceRvel[c * w + nf] > 0 ? (signed char)-1 : (signed char)nf;
    for (int64_t f = 0; f < mv; f++) {
      const int64_t q = f * n + c;
      hnb[q] = -1;
      hbc[q] = 0;
      if (f >= nf) continue;
      const int32_t face = m->cellFace[c * mv + f] - 1;
      if (face < 0 || face >= m->nface) {
        rc = sfr_set_error(ctx, SFR_BAD_ARG, "cells: cell %lld slot %lld names face %d",
                           (long long)c + 1, (long long)f, face + 1);
        goto done;
      }
      const double ch = (m->owner[(int64_t)face * 2] - 1 == c) ? 1.0 : -1.0;
      for (int j = 0; j < 3; j++) {
        hg[(0 + j) * S + q] = ch * m->normal[(int64_t)face * 3 + j];
        hg[(3 + j) * S + q] = ch * m->ksi[(int64_t)face * 3 + j];
      }
      hg[6 * S + q] = m->area[face];
      hg[7 * S + q] = m->mksi[face];
      if (vel)
        for (int j = 0; j < 3; j++) hg[(9 + j) * S + q] = m->xmcc[(c * mv + f) * 3 + j];
      const int32_t nb = m->nelcon[c * mv + f] - 1;
      if (nb > -1) {
        if (nb >= n) {
          rc = sfr_set_error(ctx, SFR_BAD_ARG, "cells: neighbour %d out of range", nb + 1);
          goto done;
        }
        hnb[q] = nb;
        const double *vs = m->vSkew + (int64_t)face * 3;
        const double *xc = m->xmcc + (c * mv + f) * 3;
        double v[3];
        v[0] = vs[0] + xc[0];
        v[1] = vs[1] + xc[1];
        v[2] = vs[2] + xc[2];
        hg[8 * S + q] = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / m->mksi[face];
        if (vel)   /* xm of the face minus the neighbour's centroid */
          for (int j = 0; j < 3; j++)
            hg[(12 + j) * S + q] = m->xm[(int64_t)face * 3 + j] - m->cc[(int64_t)nb * 3 + j];
      } else {
        hg[8 * S + q] = m->dcca[c * mv + f];
        const int lv = m->faceRvel[c * w + f], lp = m->faceRpres[c * w + f];
        if (lv < 0 || lv > nlv || lp < 0 || lp > nlp) {
          rc = sfr_set_error(ctx, SFR_BAD_ARG,
                             "cells: cell %lld face %lld names load %d / %d beyond "
                             "the tables (%d / %d)", (long long)c + 1,
                             (long long)f + 1, lv, lp, nlv, nlp);
          goto done;
        }
        hbc[q] = lv | (lp << 16);
      }
    }
  }
  C->numel = (int32_t)n;
  C->maxViz = (int)mv;
  C->vel = vel;
  C->vel_loads = 1;
  for (int i = 0; i < nlv; i++)   /* INLETSTAICTPRES, OPEN, INLETTOTALPRES */
    if (loadsVel[i].type == 9 || loadsVel[i].type == 10 || loadsVel[i].type == 13)
      C->vel_loads = 0;
  TRY(hipSetDevice(ctx->device));
  TRY(hipMalloc((void **)&C->nface, (size_t)n));
  TRY(hipMalloc((void **)&C->nb, sizeof(int32_t) * (size_t)S));
  TRY(hipMalloc((void **)&C->bc, sizeof(int32_t) * (size_t)S));
  TRY(hipMalloc((void **)&C->g, sizeof(double) * (size_t)(np * S)));
  TRY(hipMalloc((void **)&C->rho, sizeof(double) * (size_t)n));
  TRY(hipMalloc((void **)&C->lv, sizeof(sfr_load) * (size_t)(nlv > 0 ? nlv : 1)));
  TRY(hipMalloc((void **)&C->lp, sizeof(sfr_load) * (size_t)(nlp > 0 ? nlp : 1)));
  TRY(hipMemcpy(C->nface, hnf, (size_t)n, hipMemcpyHostToDevice));
  TRY(hipMemcpy(C->nb, hnb, sizeof(int32_t) * (size_t)S, hipMemcpyHostToDevice));
  TRY(hipMemcpy(C->bc, hbc, sizeof(int32_t) * (size_t)S, hipMemcpyHostToDevice));
  TRY(hipMemcpy(C->g, hg, sizeof(double) * (size_t)(np * S), hipMemcpyHostToDevice));
  if (vel) {
    TRY(hipMalloc((void **)&C->nfacev, (size_t)n));
    TRY(hipMalloc((void **)&C->visc, sizeof(double) * (size_t)n));
    TRY(hipMalloc((void **)&C->vol, sizeof(double) * (size_t)n));
    TRY(hipMemcpy(C->nfacev, hnfv, (size_t)n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(C->visc, hvisc, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(C->vol, m->volume, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  }
  TRY(hipMemcpy(C->rho, hrho, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  if (nlv) TRY(hipMemcpy(C->lv, loadsVel, sizeof(sfr_load) * (size_t)nlv, hipMemcpyHostToDevice));
  if (nlp) TRY(hipMemcpy(C->lp, loadsPres, sizeof(sfr_load) * (size_t)nlp, hipMemcpyHostToDevice));
  *out = C;
  C = NULL;
done:
  free(hnf); free(hnfv); free(hnb); free(hbc); free(hg); free(hrho); free(hvisc);
  if (C) sfr_cells_destroy(ctx, C);
  return rc;
}

int sfr_cells_pres_device(sfr_context *ctx, sfr_cells *C, int diffusion,
                          const double *d_pres, const double *d_gradPres,
                          const double *d_vel, const double *d_dField,
                          double *d_lA, double *d_lB, double *d_rCell)
{
  if (!ctx || !C || !d_pres || !d_gradPres || !d_vel || !d_dField || !d_lA || !d_lB)
    return SFR_BAD_ARG;
  if (diffusion != 1 && diffusion != 4)
    return sfr_set_error(ctx, SFR_NOT_SUPPORTED,
                         "cells: diffusion scheme %d (1 ORTHOGONAL or 4 OVERRELAXED; the "
                         "reference leaves e undefined for MINIMAL and ORTHOGONALC)",
                         diffusion);
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess)
    e = (hipError_t)sfr_k_cell_pres(ctx->stream, C->numel, C->maxViz, diffusion,
                                    C->nface, C->nb, C->bc, C->g, C->rho, C->lv, C->lp,
                                    d_pres, d_gradPres, d_vel, d_dField, d_lA, d_lB,
                                    d_rCell);
  if (e != hipSuccess)
    return sfr_set_error(ctx, SFR_HIP_ERROR, "cells: %s", hipGetErrorString(e));
  return SFR_OK;
}

int sfr_cells_update_value(sfr_context *ctx, int numel, const int32_t *d_id,
                           const double *d_x, double *d_u, int add)
{
  if (!ctx || numel < 0 || (numel && (!d_id || !d_x || !d_u))) return SFR_BAD_ARG;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess)
    e = (hipError_t)sfr_k_cell_value(ctx->stream, numel, d_id, d_x, d_u, add != 0);
  if (e != hipSuccess)
    return sfr_set_error(ctx, SFR_HIP_ERROR, "cells: %s", hipGetErrorString(e));
  return SFR_OK;
}

int sfr_simple_update(sfr_context *ctx, int numel, double alphaPres,
                      double *d_vel, double *d_pres, const double *d_presC,
                      const double *d_gradPresC, const double *d_dField)
{
  if (!ctx || numel < 0 ||
      (numel && (!d_vel || !d_pres || !d_presC || !d_gradPresC || !d_dField)))
    return SFR_BAD_ARG;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess)
    e = (hipError_t)sfr_k_simple_update(ctx->stream, numel, alphaPres, d_vel, d_pres,
                                        d_presC, d_gradPresC, d_dField);
  if (e != hipSuccess)
    return sfr_set_error(ctx, SFR_HIP_ERROR, "cells: %s", hipGetErrorString(e));
  return SFR_OK;
}

int sfr_cells_vel_device(sfr_context *ctx, sfr_cells *C, const sfr_vel_opts *o,
                         const int32_t *d_id, int neq, const double *d_pres,
                         const double *d_gradPres, const double *d_vel,
                         const double *d_gradVel, const double *d_eddyVisc,
                         const double *d_wallPar, const double *d_dField,
                         double *d_dOut, double *d_lA, double *d_ad, double *d_b,
                         double *d_rCell)
{
  if (!ctx || !C || !o || !d_id || neq < 1 || !d_pres || !d_gradPres || !d_vel ||
      !d_gradVel || !d_dField || !d_dOut || !d_lA || !d_ad || !d_b || !d_rCell ||
      (o->turb && !d_eddyVisc) || (o->wall_model && !d_wallPar))
    return SFR_BAD_ARG;
  if (d_dOut == d_dField)
    return sfr_set_error(ctx, SFR_BAD_ARG, "cells: the new d field must not alias the old");
  if (!C->vel)
    return sfr_set_error(ctx, SFR_BAD_ARG, "cells: mesh created without cc / xm");
  if (!C->vel_loads)
    return sfr_set_error(ctx, SFR_NOT_SUPPORTED,
                         "cells: the momentum loop restates MOVEWALL, STATICWALL, INLET "
                         "and OUTLET velocity loads only");
  if (o->advection < 1 || o->advection > 3)
    return sfr_set_error(ctx, SFR_NOT_SUPPORTED,
                         "cells: advection scheme %d (1 FoUp, 2 Cd, 3 SoUp)", o->advection);
  if (o->diffusion < 1 || o->diffusion > 4)
    return sfr_set_error(ctx, SFR_BAD_ARG, "cells: diffusion scheme %d", o->diffusion);
  if (!(o->underU > 0.0) || (o->time && !(o->dt > 0.0)))
    return sfr_set_error(ctx, SFR_BAD_ARG, "cells: underU %g, dt %g", o->underU, o->dt);
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess)
    e = (hipError_t)sfr_k_cell_vel(ctx->stream, C->numel, C->maxViz, o, C->nfacev,
                                   C->nb, C->bc, C->g, C->rho, C->visc, C->vol, C->lv,
                                   d_id, neq, d_pres, d_gradPres, d_vel, d_gradVel,
                                   d_eddyVisc, d_wallPar, d_dField, d_dOut, d_lA, d_ad,
                                   d_b, d_rCell);
  if (e != hipSuccess)
    return sfr_set_error(ctx, SFR_HIP_ERROR, "cells: %s", hipGetErrorString(e));
  return SFR_OK;
}
