"""Synthetic SFR3D-style systems for the solve path (the workloads the
benchmarks and parity tests run).

Every system goes through the same pipeline as in the reference: adjacency in
the reference's face order -> reord (identity or GENRCM) -> numEqV1/numEqV2 ->
dataStruct -> per-cell coefficients lA/lB -> assbly.  Only the per-cell
stencils are written here, in numpy; the integer structure and the scatter
are the C set-up layer.

Random numbers come from splitmix64 keyed by (seed, stream, index), so a
value depends only on what it belongs to (a face, a cell) and every size and
ordering of the same mesh sees the same field.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import structure as S
from ._lib import ELLPACK, CSRD
from .solver import System

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def u01(seed, stream, idx):
    """Uniform [0,1) doubles for integer indices idx."""
    idx = np.asarray(idx, np.uint64)
    with np.errstate(over="ignore"):
        key = (splitmix64(np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)
                          + np.uint64(stream)) + idx)
    return (splitmix64(key) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class Built:
    """An assembled system plus what produced it."""
    system: System
    b: np.ndarray
    id: np.ndarray
    num: np.ndarray
    nelcon: np.ndarray
    nviz: np.ndarray
    pattern: object
    extra: dict = field(default_factory=dict)


def assemble_system(nelcon, nviz, num, id_, neq, storage, unsym, lA, lB,
                    ell_quirk=False):
    pat = S.datastruct(storage, unsym, id_, num, nelcon, nviz, neq)
    al, ad, b = S.assemble(pat, id_, nelcon, nviz, lA, lB, ell_quirk)
    sysm = System(neq, pat.nad, pat.ia, pat.ja, al, ad, storage, unsym)
    return sysm, b[:neq].copy(), pat


# ---------------------------------------------------------------------------
# 3-D variable-coefficient Poisson (hex, 7-point): per-face conductance
# k_f = exp(U[ln kmin, 0]); Dirichlet p = 0 on the z = 0 face (half-cell
# distance: 2 k_f on the diagonal), homogeneous Neumann elsewhere; RHS
# U[-1, 1].  Optionally all-Neumann with one pinned cell (numEqV2).
# Face streams: 1 x-faces, 2 y-faces, 3 z-faces, 4 RHS.

def face_conductances(nx, ny, nz, seed, kmin, theta=None):
    """(kx, ky, kz) on all faces including the boundary ones:
    kx[(i+1) + (nx+1)*(j + ny*k)] for i = -1..nx-1, etc."""
    nfx, nfy, nfz = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    lk = math.log(kmin)
    if theta is None:
        tx = lk * (1.0 - u01(seed, 1, np.arange(nfx)))
        ty = lk * (1.0 - u01(seed, 2, np.arange(nfy)))
        tz = lk * (1.0 - u01(seed, 3, np.arange(nfz)))
    else:
        tx, ty, tz = theta
    return np.exp(tx), np.exp(ty), np.exp(tz), (tx, ty, tz)


def poisson_cells(nx, ny, nz, kx, ky, kz, dirichlet_z0=True):
    """lA (numel x 7, faces z-,z+,y-,x+,y+,x-, diagonal last)."""
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    shp = (nz, ny, nx)
    kxm = kx[(i + 0) + (nx + 1) * (j + ny * k)]     # face x- of cell (i,j,k)
    kxp = kx[(i + 1) + (nx + 1) * (j + ny * k)]
    kym = ky[i + nx * ((j + 0) + (ny + 1) * k)]
    kyp = ky[i + nx * ((j + 1) + (ny + 1) * k)]
    kzm = kz[i + nx * (j + ny * (k + 0))]
    kzp = kz[i + nx * (j + ny * (k + 1))]
    inner = np.ones(shp, bool)
    fz_m = np.broadcast_to(k > 0, shp)
    fz_p = np.broadcast_to(k < nz - 1, shp)
    fy_m = np.broadcast_to(j > 0, shp)
    fy_p = np.broadcast_to(j < ny - 1, shp)
    fx_p = np.broadcast_to(i < nx - 1, shp)
    fx_m = np.broadcast_to(i > 0, shp)
    coef = [np.where(fz_m, kzm, 0.0), np.where(fz_p, kzp, 0.0),
            np.where(fy_m, kym, 0.0), np.where(fx_p, kxp, 0.0),
            np.where(fy_p, kyp, 0.0), np.where(fx_m, kxm, 0.0)]
    diag = sum(coef)
    if dirichlet_z0:
        diag = diag + np.where(~fz_m & inner, 2.0 * kzm, 0.0)
    lA = np.empty((nx * ny * nz, 7))
    for f in range(6):
        lA[:, f] = -coef[f].ravel()
    lA[:, 6] = diag.ravel()
    return lA


def poisson3d(n, seed=1, kmin=1e-2, reorder=True, storage=ELLPACK,
              pinned_cell=None, theta=None, ny=None, nz=None):
    nx = n
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    numel = nx * ny * nz
    nelcon, nviz = S.hex_adjacency(nx, ny, nz)
    num = S.reord(nelcon, nviz, reorder)
    kx, ky, kz, th = face_conductances(nx, ny, nz, seed, kmin, theta)
    lA = poisson_cells(nx, ny, nz, kx, ky, kz,
                       dirichlet_z0=pinned_cell is None)
    lB = 2.0 * u01(seed, 4, np.arange(numel)) - 1.0
    if pinned_cell is None:
        id_, neq = S.numeq_v1(num)
    else:
        rt = np.zeros((numel, 7), np.int16)
        rt[pinned_cell, 6] = 1
        id_, neq = S.numeq_v2(num, rt, nviz, 6)
    sysm, b, pat = assemble_system(nelcon, nviz, num, id_, neq, storage,
                                   False, lA, lB)
    return Built(sysm, b, id_, num, nelcon, nviz, pat,
                 {"theta": th, "lA": lA, "lB": lB, "shape": (nx, ny, nz)})


# ---------------------------------------------------------------------------
# 2-D variable-coefficient Poisson on quads (faces S,E,N,W), the 5-point
# pressure-type system; one pinned reference cell (numEqV2) by default.

def poisson2d(nx, ny, seed=1, kmin=1e-2, reorder=False, storage=CSRD,
              pinned_cell=0):
    numel = nx * ny
    nelcon, nviz = S.quad_adjacency(nx, ny)
    num = S.reord(nelcon, nviz, reorder)
    lk = math.log(kmin)
    kxf = np.exp(lk * (1.0 - u01(seed, 1, np.arange((nx + 1) * ny))))
    kyf = np.exp(lk * (1.0 - u01(seed, 2, np.arange(nx * (ny + 1)))))
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    shp = (ny, nx)
    kS = kyf[i + nx * j]
    kN = kyf[i + nx * (j + 1)]
    kW = kxf[i + (nx + 1) * j]
    kE = kxf[(i + 1) + (nx + 1) * j]
    coef = [np.where(np.broadcast_to(j > 0, shp), kS, 0.0),
            np.where(np.broadcast_to(i < nx - 1, shp), kE, 0.0),
            np.where(np.broadcast_to(j < ny - 1, shp), kN, 0.0),
            np.where(np.broadcast_to(i > 0, shp), kW, 0.0)]
    lA = np.empty((numel, 5))
    for f in range(4):
        lA[:, f] = -coef[f].ravel()
    lA[:, 4] = sum(coef).ravel()
    if pinned_cell is None:                    # Dirichlet on the south wall
        lA[:, 4] += np.where(np.broadcast_to(j == 0, shp), 2.0 * kS, 0.0).ravel()
    lB = 2.0 * u01(seed, 4, np.arange(numel)) - 1.0
    if pinned_cell is None:
        id_, neq = S.numeq_v1(num)
    else:
        rt = np.zeros((numel, 5), np.int16)
        rt[pinned_cell, 4] = 1
        id_, neq = S.numeq_v2(num, rt, nviz, 4)
    sysm, b, pat = assemble_system(nelcon, nviz, num, id_, neq, storage,
                                   False, lA, lB)
    return Built(sysm, b, id_, num, nelcon, nviz, pat, {"lA": lA, "lB": lB})


# ---------------------------------------------------------------------------
# Momentum (u, v, w) on a 3-D box [0, 2 pi]^3 with walls: first-order upwind
# convection-diffusion of a Taylor-Green field, implicit Euler.
#   u = sin x cos y cos z, v = -cos x sin y cos z, w = 0  (amplitude 1)
#   nu = h/10 (cell Peclet 10), dt = 0.5 h (CFL 0.5)
# Face flux F_f = (u_f . n_f) h^2 from the analytic field at the face centre.
# a_nb = D + max(-F, 0), a_P = sum_f (D + max(F, 0)) + V/dt  (D = nu h);
# walls are Dirichlet with the analytic velocity (2D, half-cell distance).
# RHS_c = V/dt u_c - V dp/dx_c + wall terms, with the Taylor-Green pressure
# p = (cos 2x + cos 2y)(cos 2z + 2)/16.  One shared pattern, three RHS.

def _tg(x, y, z):
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros(np.broadcast_shapes(np.shape(x), np.shape(y), np.shape(z)))
    return np.broadcast_to(u, w.shape), np.broadcast_to(v, w.shape), w


def momentum3d(n, seed=1, storage=CSRD, reorder=False):
    h = 2.0 * math.pi / n
    nu, dt, V, A = h / 10.0, 0.5 * h, h ** 3, h * h
    D = nu * A / h
    c = (np.arange(n) + 0.5) * h
    z = c[:, None, None]
    y = c[None, :, None]
    x = c[None, None, :]
    shp = (n, n, n)
    numel = n ** 3
    # face centres and outward normals in face order z-, z+, y-, x+, y+, x-
    faces = [((x, y, z - h / 2), (0, 0, -1)), ((x, y, z + h / 2), (0, 0, 1)),
             ((x, y - h / 2, z), (0, -1, 0)), ((x + h / 2, y, z), (1, 0, 0)),
             ((x, y + h / 2, z), (0, 1, 0)), ((x - h / 2, y, z), (-1, 0, 0))]
    k = np.arange(n)[:, None, None]
    j = np.arange(n)[None, :, None]
    i = np.arange(n)[None, None, :]
    interior = [np.broadcast_to(k > 0, shp), np.broadcast_to(k < n - 1, shp),
                np.broadcast_to(j > 0, shp), np.broadcast_to(i < n - 1, shp),
                np.broadcast_to(j < n - 1, shp), np.broadcast_to(i > 0, shp)]
    lA = np.zeros((numel, 7))
    diag = np.full(shp, V / dt)
    rhs = [np.zeros(shp) for _ in range(3)]
    for f, ((fx, fy, fz), nrm) in enumerate(faces):
        uf = _tg(fx, fy, fz)
        F = (uf[0] * nrm[0] + uf[1] * nrm[1] + uf[2] * nrm[2]) * A
        F = np.broadcast_to(F, shp)
        anb = D + np.maximum(-F, 0.0)
        diag = diag + np.where(interior[f], D + np.maximum(F, 0.0),
                               2.0 * D + np.maximum(F, 0.0))
        lA[:, f] = np.where(interior[f], -anb, 0.0).ravel()
        wall = np.where(interior[f], 0.0, 2.0 * D + np.maximum(-F, 0.0))
        for comp in range(3):
            rhs[comp] = rhs[comp] + wall * np.broadcast_to(uf[comp], shp)
    lA[:, 6] = diag.ravel()
    uc = _tg(x, y, z)
    g2 = np.cos(2 * z) + 2.0
    dpdx = -np.sin(2 * x) * g2 / 8.0
    dpdy = -np.sin(2 * y) * g2 / 8.0
    dpdz = -(np.cos(2 * x) + np.cos(2 * y)) * np.sin(2 * z) / 8.0
    for comp, dp in enumerate((dpdx, dpdy, dpdz)):
        rhs[comp] = rhs[comp] + V / dt * np.broadcast_to(uc[comp], shp) - V * np.broadcast_to(dp, shp)
    nelcon, nviz = S.hex_adjacency(n, n, n)
    num = S.reord(nelcon, nviz, reorder)
    id_, neq = S.numeq_v1(num)
    sysm, b0, pat = assemble_system(nelcon, nviz, num, id_, neq, storage,
                                    True, lA, rhs[0].ravel())
    bs = [b0]
    for comp in (1, 2):
        _, _, bc = S.assemble(pat, id_, nelcon, nviz, lA, rhs[comp].ravel())
        bs.append(bc[:neq].copy())
    return Built(sysm, bs[0], id_, num, nelcon, nviz, pat,
                 {"b_all": np.stack(bs), "lA": lA})


# ---------------------------------------------------------------------------
# Uneven batch: nsys independent all-Neumann 3-D Poisson systems, cube side
# n_s = round(smin * (smax/smin)^u), u ~ U[0,1) (log-uniform), conductances
# exp(U[ln kmin, 0]), RHS U[-1, 1], one seeded cell pinned (numEqV2 drops it,
# as the reference does for the pressure reference cell).  System s draws
# its field from seed (seed << 16) + s; the batch-level draws (sides, pinned
# cells) use stream 5 of `seed`.

def uneven_sides(nsys, seed=1, smin=16, smax=96):
    u = u01(seed, 5, np.arange(nsys))
    return np.rint(smin * (smax / smin) ** u).astype(np.int64)


def uneven_batch(nsys=256, seed=1, smin=16, smax=96, kmin=1e-3,
                 storage=ELLPACK, sides=None):
    sides = uneven_sides(nsys, seed, smin, smax) if sides is None else sides
    pin = u01(seed, 6, np.arange(len(sides)))
    out = []
    for s, n in enumerate(sides):
        n = int(n)
        cell = int(pin[s] * n ** 3)
        B = poisson3d(n, seed=(seed << 16) + s, kmin=kmin, reorder=False,
                      storage=storage, pinned_cell=cell)
        B.extra["side"] = n
        B.extra["pinned"] = cell
        out.append(B)
    return out
