"""Adjoint solve (sfr_solve_adjoint) on the 64^3 variable-coefficient
diffusion workload: log-conductances theta_f ~ U[ln 1e-2, 0] per face,
b ~ U[-1, 1], loss L = sum over 512 seeded probe cells of (x_i - x*_i)^2 with
x* from a second theta.  The gradient of L w.r.t. theta and b from one
forward PCG + one adjoint solve, wrapped in a torch.autograd.Function, is
checked against
  - FP64 central finite differences (h = 1e-6) along 32 random directions,
    to 1e-5 relative;
  - an FP64 CPU transpose solve (oracle PCG on A^T = A) with the closed form
    dL/dtheta_f = -k_f (lam_i - lam_j)(x_i - x_j) (interior face),
    -2 k_f lam_i x_i (Dirichlet face).
The nonsymmetric path (BiCGSTAB on the transposed operator) is checked on a
momentum matrix against a direct transpose solve and finite differences of
individual stored values.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

import oracle
import sfr3d_amd as S
from sfr3d_amd import structure as ST, systems
from sfr3d_amd.solver import System

pytestmark = pytest.mark.gpu

N = 64
TOL = 1e-13


class Diffusion:
    """Fixed structure of the 64^3 system; values from theta."""

    def __init__(self, ctx, n=N, seed=11):
        self.ctx, self.n = ctx, n
        B = systems.poisson3d(n, seed=seed)
        self.B = B
        self.shape = (n, n, n)
        self.nf = [(n + 1) * n * n, n * (n + 1) * n, n * n * (n + 1)]
        self.theta0 = np.concatenate(B.extra["theta"])
        self.A = ctx.matrix(B.system)
        self.d_b = ctx.empty(B.system.neq)
        self.d_x = ctx.empty(B.system.neq)
        self.d_g = ctx.empty(B.system.neq)
        self.d_lam = ctx.empty(B.system.neq)
        self.d_gal = ctx.empty(len(B.system.al))
        self.d_gad = ctx.empty(B.system.neq)
        self._faces()

    def _faces(self):
        """face id and kind (0 interior, 1 Dirichlet z=0, 2 Neumann) for
        every (cell, slot) in the reference face order z-,z+,y-,x+,y+,x-."""
        n = self.n
        k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        ox, oy, oz = 0, self.nf[0], self.nf[0] + self.nf[1]
        fx = lambda ii: ox + ii + (n + 1) * (j + n * k)
        fy = lambda jj: oy + i + n * (jj + (n + 1) * k)
        fz = lambda kk: oz + i + n * (j + n * kk)
        ids = [fz(k), fz(k + 1), fy(j), fx(i + 1), fy(j + 1), fx(i)]
        inner = [k > 0, k < n - 1, j > 0, i < n - 1, j < n - 1, i > 0]
        self.face = np.stack([a.ravel() for a in ids], 1)
        kind = np.stack([np.where(m, 0, 2).ravel() for m in inner], 1)
        kind[:, 0] = np.where((k == 0).ravel(), 1, kind[:, 0])
        self.kind = kind

    def assemble(self, theta, b_cells):
        n = self.n
        th = np.split(theta, np.cumsum(self.nf)[:2])
        kx, ky, kz = np.exp(th[0]), np.exp(th[1]), np.exp(th[2])
        lA = systems.poisson_cells(n, n, n, kx, ky, kz)
        B = self.B
        al, ad, b = ST.assemble(B.pattern, B.id, B.nelcon, B.nviz, lA, b_cells)
        return al, ad, b[:B.system.neq]

    def solve(self, theta, b_cells):
        al, ad, b = self.assemble(theta, b_cells)
        self.A.set_values(al, ad)
        self.d_b.set(b)
        st = self.ctx.solve_device(self.A, S.PCG, TOL, 100000, self.d_b, self.d_x)
        assert st["status"] == 0
        return self.d_x.get(), (al, ad, b)

    def cells(self, x_eq):
        """equation order -> cell order"""
        return x_eq[self.B.id - 1]

    def grad_theta(self, theta, gal, gad):
        k = np.exp(theta)
        row = self.B.id - 1                       # cell -> equation row
        s = 6
        g_k = np.zeros_like(theta)
        raw = row[:, None] * s + np.arange(s)[None, :]
        inner = self.kind == 0
        np.add.at(g_k, self.face[inner], -gal[raw[inner]] + gad[np.repeat(row[:, None], s, 1)[inner]])
        dir_ = self.kind == 1
        np.add.at(g_k, self.face[dir_], 2.0 * gad[np.repeat(row[:, None], s, 1)[dir_]])
        return k * g_k


class SolveFn(torch.autograd.Function):
    """x(theta, b) through the GPU forward solve; backward = adjoint solve."""

    @staticmethod
    def forward(fctx, theta, b_cells, env):
        x_eq, _ = env.solve(theta.detach().numpy(), b_cells.detach().numpy())
        fctx.env = env
        fctx.save_for_backward(theta)
        fctx.x_eq = x_eq
        return torch.from_numpy(env.cells(x_eq))

    @staticmethod
    def backward(fctx, g_cells):
        env = fctx.env
        (theta,) = fctx.saved_tensors
        g_eq = np.zeros(env.B.system.neq)
        g_eq[env.B.id - 1] = g_cells.numpy()
        env.d_x.set(fctx.x_eq)
        env.d_g.set(g_eq)
        st = env.ctx.solve_adjoint(env.A, S.PCG, TOL, 100000, env.d_x, env.d_g,
                                   env.d_lam, env.d_gal, env.d_gad)
        assert st["status"] == 0
        lam, gal, gad = env.d_lam.get(), env.d_gal.get(), env.d_gad.get()
        g_theta = env.grad_theta(theta.numpy(), gal, gad)
        return torch.from_numpy(g_theta), torch.from_numpy(env.cells(lam)), None


@pytest.fixture(scope="module")
def env(ctx):
    e = Diffusion(ctx)
    yield e
    e.A.destroy()


def loss_setup(env, seed=5):
    rng = np.random.default_rng(seed)
    n3 = env.n ** 3
    probes = rng.choice(n3, 512, replace=False)
    theta_t = np.concatenate(systems.face_conductances(env.n, env.n, env.n, 99, 1e-2)[3])
    b0 = env.B.extra["lB"]
    xt, _ = env.solve(theta_t, b0)
    xstar = env.cells(xt)[probes]
    return probes, xstar, b0


def test_adjoint_gradient_vs_finite_differences(env):
    probes, xstar, b0 = loss_setup(env)
    theta = torch.tensor(env.theta0, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(b0, dtype=torch.float64, requires_grad=True)
    x = SolveFn.apply(theta, b, env)
    L = ((x[probes] - torch.from_numpy(xstar)) ** 2).sum()
    L.backward()
    gt, gb = theta.grad.numpy(), b.grad.numpy()

    def Lval(th, bb):
        xe, _ = env.solve(th, bb)
        return float(((env.cells(xe)[probes] - xstar) ** 2).sum())

    rng = np.random.default_rng(17)
    h = 1e-6
    worst = 0.0
    for d in range(32):
        vt = rng.standard_normal(len(env.theta0))
        vb = rng.standard_normal(len(b0))
        fd = (Lval(env.theta0 + h * vt, b0 + h * vb) -
              Lval(env.theta0 - h * vt, b0 - h * vb)) / (2 * h)
        ad = gt @ vt + gb @ vb
        rel = abs(fd - ad) / max(abs(fd), abs(ad))
        worst = max(worst, rel)
        print(f"direction {d}: fd={fd:.12e} adjoint={ad:.12e} rel={rel:.2e}", flush=True)
    assert worst < 1e-5


def test_adjoint_vs_cpu_transpose_solve(env):
    probes, xstar, b0 = loss_setup(env, seed=6)
    x_eq, (al, ad, b) = env.solve(env.theta0, b0)
    g_cells = np.zeros(env.n ** 3)
    g_cells[probes] = 2.0 * (env.cells(x_eq)[probes] - xstar)
    g_eq = np.zeros(env.B.system.neq)
    g_eq[env.B.id - 1] = g_cells
    env.d_x.set(x_eq)
    env.d_g.set(g_eq)
    env.ctx.solve_adjoint(env.A, S.PCG, TOL, 100000, env.d_x, env.d_g,
                          env.d_lam, env.d_gal, env.d_gad)
    lam = env.d_lam.get()
    g_gpu = env.grad_theta(env.theta0, env.d_gal.get(), env.d_gad.get())
    # FP64 CPU transpose solve (A symmetric) and the closed form per face
    sysm = System(env.B.system.neq, env.B.system.nad, env.B.system.ia,
                  env.B.system.ja, al, ad, S.ELLPACK, False)
    lam_cpu, so = oracle.pcg(sysm, g_eq, TOL, 100000)
    assert np.linalg.norm(lam - lam_cpu) / np.linalg.norm(lam_cpu) < 1e-9
    lc, xc = env.cells(lam_cpu), env.cells(x_eq)
    k = np.exp(env.theta0)
    g_cf = np.zeros_like(env.theta0)
    n3 = env.n ** 3
    for f in range(6):
        kind = env.kind[:, f]
        nb = env.B.nelcon[:, f] - 1
        inner = kind == 0
        c = np.arange(n3)[inner]
        j = nb[inner]
        # each interior face is met from both sides: half the term each time
        np.add.at(g_cf, env.face[inner, f], -0.5 * k[env.face[inner, f]] * (lc[c] - lc[j]) * (xc[c] - xc[j]))
        d = kind == 1
        np.add.at(g_cf, env.face[d, f], -2.0 * k[env.face[d, f]] * lc[d] * xc[d])
    assert np.linalg.norm(g_gpu - g_cf) / np.linalg.norm(g_cf) < 1e-8


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
def test_nonsymmetric_adjoint(ctx, storage):
    M = systems.momentum3d(8, storage=storage)
    sysm = M.system
    A = ctx.matrix(sysm)
    n = sysm.neq
    rng = np.random.default_rng(3)
    b = M.b
    g = rng.standard_normal(n)
    d_b, d_x, d_g, d_l = ctx.array(b), ctx.empty(n), ctx.array(g), ctx.empty(n)
    nval = len(sysm.al) if storage != S.ELLPACK else n * sysm.ia[0]
    d_gal, d_gad = ctx.empty(nval), ctx.empty(n)
    ctx.solve_device(A, S.PBICGSTAB, 1e-13, 5000, d_b, d_x)
    st = ctx.solve_adjoint(A, S.PBICGSTAB, 1e-13, 5000, d_x, d_g, d_l, d_gal, d_gad)
    assert st["status"] == 0
    from tests.test_oracle import dense
    Ad = dense(sysm)
    lam_ref = np.linalg.solve(Ad.T, g)
    lam = d_l.get()
    assert np.linalg.norm(lam - lam_ref) / np.linalg.norm(lam_ref) < 1e-9
    x = d_x.get()
    gal, gad = d_gal.get(), d_gad.get()
    assert np.allclose(gad, -lam_ref * x, rtol=1e-8, atol=1e-12)
    # L = g.x: dL/dal[k] by central differences on a few stored values
    al0 = sysm.al.copy()
    idx = [k for k in np.flatnonzero(al0[:nval]) if True][:6]
    for k in idx:
        h = 1e-6 * abs(al0[k])
        vals = []
        for sgn in (1, -1):
            al = al0.copy()
            al[k] += sgn * h
            s2 = System(n, sysm.nad, sysm.ia, sysm.ja, al, sysm.ad, storage, True)
            vals.append(g @ np.linalg.solve(dense(s2), b))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - gal[k]) <= 1e-6 * max(abs(fd), 1e-12), (k, fd, gal[k])
    A.destroy()
