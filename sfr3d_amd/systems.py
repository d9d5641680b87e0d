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


def species3d(n, nsys, seed=1, storage=CSRD, reorder=False):
    """nsys transported scalars (SFR3D's species, src/Combustion.c:222-240)
    on the 96^3-style periodic Taylor-Green box of momentum3d: the same
    pattern and convective fluxes, FOUP, implicit Euler, but each species
    with its own diffusivity nu_c = h*Sc_c/10 (Sc_c log-uniform in
    [0.5, 2]), so its own off-diagonals and diagonal, Dirichlet walls at
    phi = 0, source U[-1, 1] and old value U[0, 1].  Returns (system of
    species 0, list of (al, ad, b) per species)."""
    rng = np.random.default_rng(seed)
    h = 2.0 * math.pi / n
    dt, V, A = 0.5 * h, h ** 3, h * h
    c = (np.arange(n) + 0.5) * h
    z = c[:, None, None]
    y = c[None, :, None]
    x = c[None, None, :]
    shp = (n, n, n)
    numel = n ** 3
    faces = [((x, y, z - h / 2), (0, 0, -1)), ((x, y, z + h / 2), (0, 0, 1)),
             ((x, y - h / 2, z), (0, -1, 0)), ((x + h / 2, y, z), (1, 0, 0)),
             ((x, y + h / 2, z), (0, 1, 0)), ((x - h / 2, y, z), (-1, 0, 0))]
    k = np.arange(n)[:, None, None]
    j = np.arange(n)[None, :, None]
    i = np.arange(n)[None, None, :]
    interior = [np.broadcast_to(k > 0, shp), np.broadcast_to(k < n - 1, shp),
                np.broadcast_to(j > 0, shp), np.broadcast_to(i < n - 1, shp),
                np.broadcast_to(j < n - 1, shp), np.broadcast_to(i > 0, shp)]
    flux = []
    for (fx, fy, fz), nrm in faces:
        uf = _tg(fx, fy, fz)
        flux.append(np.broadcast_to((uf[0] * nrm[0] + uf[1] * nrm[1] + uf[2] * nrm[2]) * A, shp))
    nelcon, nviz = S.hex_adjacency(n, n, n)
    num = S.reord(nelcon, nviz, reorder)
    id_, neq = S.numeq_v1(num)
    sysm, pat, out = None, None, []
    for sp in range(nsys):
        nu = h * float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))) / 10.0
        D = nu * A / h
        lA = np.zeros((numel, 7))
        diag = np.full(shp, V / dt)
        for f in range(6):
            F = flux[f]
            anb = D + np.maximum(-F, 0.0)
            diag = diag + np.where(interior[f], D + np.maximum(F, 0.0),
                                   2.0 * D + np.maximum(F, 0.0))
            lA[:, f] = np.where(interior[f], -anb, 0.0).ravel()
        lA[:, 6] = diag.ravel()
        rhs = (V * rng.uniform(-1.0, 1.0, numel)
               + V / dt * rng.uniform(0.0, 1.0, numel))
        if sysm is None:
            sysm, b0, pat = assemble_system(nelcon, nviz, num, id_, neq, storage,
                                            True, lA, rhs)
            out.append((sysm.al, sysm.ad, b0))
        else:
            al, ad, bc = S.assemble(pat, id_, nelcon, nviz, lA, rhs)
            out.append((al, ad[:neq].copy(), bc[:neq].copy()))
    return sysm, out


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


# ---------------------------------------------------------------------------
# 2-D lid-driven cavity (unit square, n x n quads numbered row-major from the
# bottom-left, faces S,E,N,W): SIMPLE on a collocated grid with first-order
# upwind convection and Rhie-Chow face fluxes, the pressure-correction system
# of every outer iteration recorded in SFR3D's storages.  The pressure-
# correction matrix has off-diagonals -rho d_f A_f/|xi| (d_f = V/a_P averaged
# to the face), its diagonal the sum over interior faces, its right-hand side
# minus the cell's net mass outflow, and one reference cell pinned (numEqV2
# drops it); walls carry no correction.

def _momentum_matrix(aS, aE, aN, aW, aP):
    """CSR (scipy) of the n x n five-point operator aP x_P - sum a_nb x_nb."""
    import scipy.sparse as sp
    ny, nx = aP.shape
    idx = np.arange(nx * ny).reshape(ny, nx)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [aP.ravel()]
    for a, dj, di in ((aS, -1, 0), (aE, 0, 1), (aN, 1, 0), (aW, 0, -1)):
        jj, ii = np.nonzero(a)
        rows.append(idx[jj, ii])
        cols.append(idx[jj + dj, ii + di])
        vals.append(-a[jj, ii])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows),
                                                 np.concatenate(cols))),
                         shape=(nx * ny, nx * ny))


def cavity2d(n=128, outer=20, nu=0.01, rho=1.0, ulid=1.0, alpha_u=0.7,
             alpha_p=0.3, pinned_cell=0, storages=(CSRD, ELLPACK),
             p_solve=None, record=True):
    """Run `outer` SIMPLE iterations; return (systems, fields) where
    systems[k][storage] is the Built pressure-correction system of outer
    iteration k and fields = {u, v, p, mass} (mass = max |net outflow| per
    iteration).  p_solve(system, b) -> x solves the pressure correction (CSRD
    storage); by default a direct sparse solve."""
    import scipy.sparse.linalg as spla
    h = 1.0 / n
    A = h                                   # face area (unit depth)
    V = h * h
    D = nu * A / h                          # interior face diffusion
    Dw = nu * A / (0.5 * h)                 # wall face diffusion
    u = np.zeros((n, n))
    v = np.zeros((n, n))
    p = np.zeros((n, n))
    Fx = np.zeros((n, n + 1))               # outward +x mass flux, x-faces
    Fy = np.zeros((n + 1, n))               # +y mass flux, y-faces
    nelcon, nviz = S.quad_adjacency(n, n)
    num = S.reord(nelcon, nviz, False)
    rt = np.zeros((n * n, 5), np.int16)
    rt[pinned_cell, 4] = 1
    id_, neq = S.numeq_v2(num, rt, nviz, 4)
    pats = {st: S.datastruct(st, False, id_, num, nelcon, nviz, neq)
            for st in storages}
    inE = np.zeros((n, n), bool); inE[:, :-1] = True
    inW = np.zeros((n, n), bool); inW[:, 1:] = True
    inN = np.zeros((n, n), bool); inN[:-1, :] = True
    inS = np.zeros((n, n), bool); inS[1:, :] = True

    def grad(f):
        """cell-centre gradient, face values averaged, boundary = cell."""
        fe = np.where(inE, 0.5 * (f + np.roll(f, -1, 1)), f)
        fw = np.where(inW, 0.5 * (f + np.roll(f, 1, 1)), f)
        fn = np.where(inN, 0.5 * (f + np.roll(f, -1, 0)), f)
        fs = np.where(inS, 0.5 * (f + np.roll(f, 1, 0)), f)
        return (fe - fw) / h, (fn - fs) / h

    systems_out, mass = [], []
    for it in range(outer):
        Fe, Fw = Fx[:, 1:], Fx[:, :-1]
        Fn, Fs = Fy[1:, :], Fy[:-1, :]
        aE = np.where(inE, D + np.maximum(-Fe, 0.0), 0.0)
        aW = np.where(inW, D + np.maximum(Fw, 0.0), 0.0)
        aN = np.where(inN, D + np.maximum(-Fn, 0.0), 0.0)
        aS = np.where(inS, D + np.maximum(Fs, 0.0), 0.0)
        wall = (np.where(inE, 0.0, Dw) + np.where(inW, 0.0, Dw)
                + np.where(inN, 0.0, Dw) + np.where(inS, 0.0, Dw))
        aP = aE + aW + aN + aS + wall + (Fe - Fw + Fn - Fs)
        aP = np.maximum(aP, aE + aW + aN + aS + wall) / alpha_u
        dpx, dpy = grad(p)
        lid = np.where(inN, 0.0, Dw * ulid)
        bu = -V * dpx + lid + (1.0 - alpha_u) * aP * u
        bv = -V * dpy + (1.0 - alpha_u) * aP * v
        M = _momentum_matrix(aS, aE, aN, aW, aP)
        u = spla.spsolve(M.tocsc(), bu.ravel()).reshape(n, n)
        v = spla.spsolve(M.tocsc(), bv.ravel()).reshape(n, n)
        # Rhie-Chow face fluxes
        d = V / aP
        dpx, dpy = grad(p)
        de = 0.5 * (d[:, :-1] + d[:, 1:])
        ue = (0.5 * (u[:, :-1] + u[:, 1:])
              - de * ((p[:, 1:] - p[:, :-1]) / h - 0.5 * (dpx[:, :-1] + dpx[:, 1:])))
        dn = 0.5 * (d[:-1, :] + d[1:, :])
        vn = (0.5 * (v[:-1, :] + v[1:, :])
              - dn * ((p[1:, :] - p[:-1, :]) / h - 0.5 * (dpy[:-1, :] + dpy[1:, :])))
        Fx = np.zeros((n, n + 1)); Fx[:, 1:-1] = rho * ue * A
        Fy = np.zeros((n + 1, n)); Fy[1:-1, :] = rho * vn * A
        # pressure correction: faces S,E,N,W; coefficient rho d_f A / h
        cE = np.zeros((n, n)); cE[:, :-1] = rho * de * A / h
        cW = np.zeros((n, n)); cW[:, 1:] = rho * de * A / h
        cN = np.zeros((n, n)); cN[:-1, :] = rho * dn * A / h
        cS = np.zeros((n, n)); cS[1:, :] = rho * dn * A / h
        net = Fx[:, 1:] - Fx[:, :-1] + Fy[1:, :] - Fy[:-1, :]
        mass.append(float(np.abs(net).max()))
        lA = np.empty((n * n, 5))
        lA[:, 0], lA[:, 1], lA[:, 2], lA[:, 3] = -cS.ravel(), -cE.ravel(), -cN.ravel(), -cW.ravel()
        lA[:, 4] = (cS + cE + cN + cW).ravel()
        lB = -net.ravel()
        rec = {}
        for st, pat in pats.items():
            al, ad, b = S.assemble(pat, id_, nelcon, nviz, lA, lB)
            sysm = System(neq, pat.nad, pat.ia, pat.ja, al, ad, st, False)
            rec[st] = Built(sysm, b[:neq].copy(), id_, num, nelcon, nviz, pat,
                            {"outer": it})
        if record:
            systems_out.append(rec)
        first = rec[storages[0]]
        if p_solve is None:
            pr = spla.spsolve(_pressure_csr(first.system).tocsc(), first.b)
        else:
            pr = p_solve(first.system, first.b)
        pc = np.zeros(n * n)
        cells = np.nonzero(id_ > 0)[0]
        pc[cells] = pr[id_[cells] - 1]
        pc = pc.reshape(n, n)
        # corrections
        p = p + alpha_p * pc
        gx, gy = grad(pc)
        u = u - d * gx
        v = v - d * gy
        Fx[:, 1:-1] -= rho * de * A * (pc[:, 1:] - pc[:, :-1]) / h
        Fy[1:-1, :] -= rho * dn * A * (pc[1:, :] - pc[:-1, :]) / h
    return systems_out, {"u": u, "v": v, "p": p, "mass": np.array(mass)}


def _pressure_csr(sysm):
    """scipy CSR of a CSRD symmetric (lower) system, for the default direct
    pressure solve of cavity2d."""
    import scipy.sparse as sp
    n = sysm.neq
    rows = np.repeat(np.arange(n), np.diff(sysm.ia))
    L = sp.csr_matrix((sysm.al, (rows, sysm.ja)), shape=(n, n))
    return L + L.T + sp.diags(sysm.ad[:n])


# ---- hex meshes with SFR3D's geometry arrays --------------------------------

HEX_FACE_NODES = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4],
                           [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]])
HEX_OPPOSITE = np.array([1, 0, 4, 5, 2, 3])      # z-/z+, y-/y+, x+/x-


def _graded(n, ratio):
    """n+1 node coordinates on [0, 1] with a geometric cell-size ratio."""
    if ratio == 1.0:
        return np.linspace(0.0, 1.0, n + 1)
    w = ratio ** np.arange(n)
    return np.concatenate([[0.0], np.cumsum(w) / w.sum()])


def hex_mesh(nx, ny, nz, ratio=(1.0, 1.0, 1.0), shear=(0.0, 0.0, 0.0),
             warp=0.0):
    """An nx*ny*nz hex mesh of the unit box, graded by `ratio` per axis,
    sheared by x += s0*y + s1*z, y += s2*z (faces stay planar) and warped by
    x += w sin(pi y) sin(pi z), y += w sin(pi x) sin(pi z) (interior faces
    bend, and the line between two centroids misses the face's midpoint),
    with the
    arrays SFR3D's Mesh holds for the finite-volume cell loops, in its layout
    and 1-based conventions:

      nelcon [numel*6]  neighbour across each face (faces z-, z+, y-, x+, y+,
                        x-; -1 on the boundary), nFace [numel] = 6
      cellFace [numel*6], owner [nface*2]  face of each cell slot; owner cell
                        first, the other cell second (-1 on the boundary)
      volume [numel], dcca [numel*6], xmcc [numel*6*3], cc [numel*3]
      mksi, area [nface], ksi, normal, xm, vSkew, eta [nface*3], mvSkew
                        face quantities as the owner sees them

    The definitions follow cellGeom3D (reference src/CellLib.c:1901-2290):
    the centroid is the mean of the nodes; a face's normal is the normalised
    mean of its four corner cross products and its area the length of that
    mean; xm is the mean of its nodes, xmcc = xm - cc, dcca = xmcc . normal;
    ksi is the unit vector to the neighbour's centroid (to xm on the
    boundary) and mksi that distance; vSkew = xi - xm, xi the point where the
    line along ksi meets the face (vectorKm3d, src/CellLib.c:4247-4310); the
    volume is the Green-Gauss sum over faces.  eta holds the face's first
    edge (the reference's per-face eta is not used by the cell loops here)."""
    gx, gy, gz = (_graded(n, r) for n, r in zip((nx, ny, nz), ratio))
    Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
    px = X + shear[0] * Y + shear[1] * Z + warp * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    py = Y + shear[2] * Z + warp * np.sin(np.pi * X) * np.sin(np.pi * Z)
    P = np.stack([px, py, Z], axis=-1)              # (nz+1, ny+1, nx+1, 3)
    numel = nx * ny * nz
    # node o of every cell: offsets (di, dj, dk) of the reference's hex order
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
           (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    cn = np.stack([P[dk:dk + nz, dj:dj + ny, di:di + nx].reshape(numel, 3)
                   for di, dj, dk in off], axis=1)  # (numel, 8, 3)
    cc = cn.sum(axis=1) * (1.0 / 8.0)
    nelcon, nviz = S.hex_adjacency(nx, ny, nz)
    nelcon = np.ascontiguousarray(nelcon, np.int32)
    nb = nelcon - 1                                    # -2 on the boundary
    normal = np.empty((numel, 6, 3))
    area = np.empty((numel, 6))
    xm = np.empty((numel, 6, 3))
    eta = np.empty((numel, 6, 3))
    for f in range(6):
        x1, x2, x3, x4 = (cn[:, HEX_FACE_NODES[f, q]] for q in range(4))
        v1, v2, v3, v4 = x2 - x1, x3 - x2, x4 - x3, x1 - x4
        nv = 0.25 * (np.cross(v4, v1) + np.cross(v1, v2) + np.cross(v2, v3)
                     + np.cross(v3, v4))
        a = np.sqrt((nv * nv).sum(axis=1))
        normal[:, f] = nv / a[:, None]
        area[:, f] = a
        xm[:, f] = 0.25 * (x1 + x2 + x3 + x4)
        eta[:, f] = v1
    xmcc = xm - cc[:, None, :]
    dcca = (xmcc * normal).sum(axis=2)
    bnd = nb < 0
    cnb = np.where(bnd[..., None], xm, cc[np.where(bnd, 0, nb)])
    kv = np.where(bnd[..., None], xmcc, cnb - cc[:, None, :])
    mksi = np.sqrt((kv * kv).sum(axis=2))
    ksi = kv / mksi[..., None]
    t = (normal * xmcc).sum(axis=2) / (normal * ksi).sum(axis=2)
    xi = cc[:, None, :] + t[..., None] * ksi
    vskew = xi - xm
    mvskew = np.sqrt((vskew * vskew).sum(axis=2))
    volume = ((xm * normal).sum(axis=2) * area).sum(axis=1) / 3.0
    # faces in cell order: a slot owns its face on the boundary or when the
    # neighbour comes later
    cells = np.arange(numel)[:, None]
    own = bnd | (nb > cells)
    fid = np.full((numel, 6), -1, np.int64)
    fid[own] = np.arange(int(own.sum()))
    oc, of = np.nonzero(~own)
    fid[oc, of] = fid[nb[oc, of], HEX_OPPOSITE[of]]
    nface = int(own.sum())
    ocell, oslot = np.nonzero(own)                    # in face order
    owner = np.empty((nface, 2), np.int32)
    owner[:, 0] = ocell + 1
    owner[:, 1] = np.where(bnd[ocell, oslot], -1, nb[ocell, oslot] + 1)

    def fsel(a):
        return np.ascontiguousarray(a[ocell, oslot])

    return {
        "nx": nx, "ny": ny, "nz": nz, "numel": numel, "nface": nface,
        "maxViz": 6, "nelcon": nelcon.reshape(-1),
        "nFace": np.ascontiguousarray(nviz, np.int16),
        "cellFace": np.ascontiguousarray((fid + 1).reshape(-1), np.int32),
        "owner": owner.reshape(-1), "volume": volume, "cc": cc.reshape(-1),
        "dcca": np.ascontiguousarray(dcca.reshape(-1)),
        "xmcc": np.ascontiguousarray(xmcc.reshape(-1)),
        "mksi": fsel(mksi), "ksi": fsel(ksi).reshape(-1),
        "area": fsel(area), "normal": fsel(normal).reshape(-1),
        "xm": fsel(xm).reshape(-1), "vSkew": fsel(vskew).reshape(-1),
        "mvSkew": fsel(mvskew), "eta": fsel(eta).reshape(-1),
    }


# SFR3D load types (reference include/Define.h:317-331)
DIRICHLETBC, INLET, OUTLET, MOVEWALL = 1, 6, 7, 8
INLETSTAICTPRES, OPEN, INLETTOTALPRES, STATICWALL, FLUXPRES = 9, 10, 13, 14, 15


def simple_pres_case(mesh, seed=1, diffusion=1, boundary="mixed",
                     skip_cell=None):
    """Fields, materials, boundary markers and loads for the SIMPLE
    pressure-correction cell loop on a hex_mesh, as SFR3D holds them:
    mat (1-based), prop (10 per material, density at index 1), faceRvel /
    faceRpres (maxViz+1 per cell, 1-based load index or 0), the load tables,
    pres, gradPres, vel and the SIMPLE d field (positive), equation numbers
    id (all cells, natural order) and num.

    boundary "cavity": a moving lid (z+) and static walls, no pressure
    loads.  "mixed": every load type the pressure cell reads: z+ moving wall;
    z- static-pressure inlet and total-pressure inlet in alternate cells;
    x- velocity inlet; x+ outlet with prescribed pressure; y- static wall
    with a pressure-flux load; y+ open boundary (flow both ways).
    skip_cell: a cell marked as having no pressure equation."""
    nx, ny, nz = mesh["nx"], mesh["ny"], mesh["nz"]
    numel, mv = mesh["numel"], mesh["maxViz"]
    c = np.arange(numel)
    i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
    mat = np.where(k < max(nz // 2, 1), 1, 2).astype(np.int16)
    prop = np.zeros(20)
    prop[1], prop[11] = 1.0, 1.25
    prop[0], prop[10] = 1e-3, 2e-3
    u = lambda s, n: u01(seed, s, np.arange(n))
    pres = 2.0 * u(1, numel) - 1.0
    gradPres = 2.0 * u(2, 3 * numel) - 1.0
    vel = 2.0 * u(3, 3 * numel) - 1.0
    dField = 1e-3 * (0.5 + u(4, 3 * numel))
    fv = np.zeros((numel, mv + 1), np.int16)
    fp = np.zeros((numel, mv + 1), np.int16)
    # faces z-, z+, y-, x+, y+, x- (boundary where the cell is at the side)
    side = [k == 0, k == nz - 1, j == 0, i == nx - 1, j == ny - 1, i == 0]
    if boundary == "cavity":
        loadsVel = [{"type": MOVEWALL, "par": [1.0, 0.0, 0.0, 0.0]},
                    {"type": STATICWALL}]
        loadsPres = []
        for f in range(6):
            fv[side[f], f] = 1 if f == 1 else 2
    else:
        loadsVel = [{"type": MOVEWALL, "par": [1.0, 0.0, 0.0, 0.0]},
                    {"type": STATICWALL},
                    {"type": INLETSTAICTPRES},
                    {"type": INLETTOTALPRES, "par": [1.1, 2.0, 0.0, 0.0, 1.0]},
                    {"type": INLET, "par": [0.8, 0.1, -0.05, 1.05]},
                    {"type": OUTLET},
                    {"type": OPEN, "par": [0.3], "density": 1.15}]
        loadsPres = [{"type": DIRICHLETBC, "par": [0.5]},
                     {"type": INLETTOTALPRES, "par": [1.1, 2.0, 0.0, 0.0, 1.0]},
                     {"type": FLUXPRES},
                     {"type": OPEN, "par": [0.3]}]
        fv[side[1], 1] = 1
        zm = side[0]
        fv[zm & (i % 2 == 0), 0] = 3
        fv[zm & (i % 2 == 1), 0] = 4
        fp[zm & (i % 2 == 1), 0] = 2
        fv[side[5], 5] = 5
        fv[side[3], 3] = 6
        fp[side[3], 3] = 1
        fv[side[2], 2] = 2
        fp[side[2], 2] = 3
        fv[side[4], 4] = 7
        fp[side[4], 4] = 4
    if skip_cell is not None:
        fp[skip_cell, mesh["nFace"][skip_cell]] = 1
    return {"mat": mat, "prop": prop, "pres": pres, "gradPres": gradPres,
            "vel": vel, "dField": dField,
            "faceRvel": np.ascontiguousarray(fv.reshape(-1)),
            "faceRpres": np.ascontiguousarray(fp.reshape(-1)),
            "loadsVel": loadsVel, "loadsPres": loadsPres,
            "diffusion": diffusion,
            "id": (c + 1).astype(np.int32), "num": (c + 1).astype(np.int32)}


def simple_vel_case(mesh, seed=1, advection=3, diffusion=4, boundary="cavity",
                    rhie=True, so_pressure=True, viscosity=True, residual=True,
                    simplec=False, time="backward", turb=True, wall_model=False,
                    fixed_cell=None):
    """Fields, materials, markers, loads and model options for the SIMPLE
    momentum cell loop on a hex_mesh, as SFR3D holds them: vel, gradVel
    (9 per cell, d u_i / d x_j at i*3 + j), pres, gradPres, the d field,
    eddy viscosity and wall parameters (4 per cell), mat / prop (viscosity
    at index 0, density at 1), faceRvel / faceRpres, the velocity load
    table, id / num and the options of cellSimpleVel3D (advection 1 FoUp,
    2 Cd, 3 SoUp; diffusion 1-4; time "steady", "euler" or "backward").

    boundary "cavity": moving lid (z+) and static walls.  "channel": inlet
    at x-, outlet at x+, moving lid, static walls elsewhere.  fixed_cell: a
    cell with prescribed velocity (no equation, id -1): its neighbours move
    its coefficient to the right-hand side."""
    nx, ny, nz = mesh["nx"], mesh["ny"], mesh["nz"]
    numel, mv = mesh["numel"], mesh["maxViz"]
    c = np.arange(numel)
    i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
    mat = np.where(k < max(nz // 2, 1), 1, 2).astype(np.int16)
    prop = np.zeros(20)
    prop[0], prop[1] = 1e-2, 1.0
    prop[10], prop[11] = 1.5e-2, 1.2
    u = lambda s, n: u01(seed, s, np.arange(n))
    vel = 2.0 * u(11, 3 * numel) - 1.0
    gradVel = 2.0 * u(12, 9 * numel) - 1.0
    pres = 2.0 * u(13, numel) - 1.0
    gradPres = 2.0 * u(14, 3 * numel) - 1.0
    dField = 1e-3 * (0.5 + u(15, 3 * numel))
    eddy = 1e-3 * u(16, numel)
    wall = np.zeros(4 * numel)
    wall[0::4] = 0.02 * u(17, numel)          # y+ either side of 0.01
    wall[1::4] = 1.0 + u(18, numel)           # u+
    fv = np.zeros((numel, mv + 1), np.int16)
    fp = np.zeros((numel, mv + 1), np.int16)
    side = [k == 0, k == nz - 1, j == 0, i == nx - 1, j == ny - 1, i == 0]
    loads = [{"type": MOVEWALL, "par": [1.0, 0.2, 0.0, 0.0]},
             {"type": STATICWALL}]
    for f in range(6):
        fv[side[f], f] = 2
    fv[side[1], 1] = 1
    if boundary == "channel":
        loads += [{"type": INLET, "par": [0.8, 0.05, -0.02, 1.1]},
                  {"type": OUTLET}]
        fv[side[5], 5] = 3
        fv[side[3], 3] = 4
    id_ = (c + 1).astype(np.int32)
    if fixed_cell is not None:
        fv[fixed_cell, mesh["nFace"][fixed_cell]] = 1
        id_[fixed_cell] = -1
        id_[fixed_cell + 1:] -= 1
    t = {"steady": (0, 0), "euler": (1, 0), "backward": (1, 1)}[time]
    opts = {"advection": advection, "diffusion": diffusion, "turb": int(turb),
            "wall_model": int(wall_model), "rhie": int(rhie),
            "so_pressure": int(so_pressure), "viscosity": int(viscosity),
            "residual": int(residual), "simplec": int(simplec),
            "time": t[0], "backward": t[1], "dt": 0.05, "dt0": 0.04,
            "underU": 0.7}
    return {"mat": mat, "prop": prop, "vel": vel, "gradVel": gradVel,
            "pres": pres, "gradPres": gradPres, "dField": dField,
            "eddyVisc": eddy, "wallPar": wall,
            "faceRvel": np.ascontiguousarray(fv.reshape(-1)),
            "faceRpres": np.ascontiguousarray(fp.reshape(-1)),
            "loadsVel": loads, "id": id_, "num": (c + 1).astype(np.int32),
            "opts": opts}
