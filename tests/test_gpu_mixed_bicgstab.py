"""GPU parity of mixed-precision BiCGSTAB (FP32 iterations, FP64 reliable
updates; device/bicgstab_mixed.hip) against the oracle's FP64 BiCGSTAB
(the reference's pbicgstab, src/Bicgstab.c:50-245, restated and pinned
bit-exact in tests/test_oracle.py).

Bar: every system meets the reference's own stop test ||b - Ax|| <
tol ||b|| on the true FP64 residual, and x agrees with the oracle's FP64
solution to a relative 1e-6 at tol 1e-8 -- for one system (sfr_solve_device
with SFR_PREC_MIXED), the u/v/w batch with distinct per-component diagonals
(the reference's ad + c*n, src/Simple.c:2748-2750) at 12^3 and at the full
96^3, and the species batch, the last two through the context's "precision"
option."""
import contextlib

import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems
from sfr3d_amd.solver import System

pytestmark = pytest.mark.gpu
TOL = 1e-8


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


@contextlib.contextmanager
def mixed(ctx):
    ctx.set_option("precision", S.PREC_MIXED)
    try:
        yield
    finally:
        ctx.set_option("precision", S.PREC_FP64)


def with_ad(sysm, ad):
    return System(sysm.neq, sysm.nad, sysm.ia, sysm.ja, sysm.al, ad, sysm.storage,
                  sysm.unsym, au=sysm.au)


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
def test_mixed_bicgstab_single(ctx, storage):
    M = systems.momentum3d(12, storage=storage)
    print(f"mixed bicgstab storage={storage}", flush=True)
    A = ctx.matrix(M.system)
    for c in range(3):
        b = M.extra["b_all"][c]
        db, dx = ctx.array(b), ctx.empty(M.system.neq)
        st = ctx.solve_device(A, S.PBICGSTAB, TOL, 5000, db, dx, precision=S.PREC_MIXED)
        xo, so = oracle.pbicgstab(M.system, b, TOL, 5000)
        x = dx.get()
        assert st["status"] == S._lib.OK
        assert st["relres"] < TOL
        assert np.linalg.norm(b - oracle.matvec(M.system, x)) < TOL * np.linalg.norm(b)
        assert relerr(x, xo) < 1e-6
        db.free(); dx.free()
    A.destroy()


def test_mixed_bicgstab_zero_rhs_and_warm_start(ctx):
    M = systems.momentum3d(10)
    n = M.system.neq
    A = ctx.matrix(M.system)
    db, dx = ctx.array(np.zeros(n)), ctx.array(np.ones(n))
    st = ctx.solve_device(A, S.PBICGSTAB, TOL, 100, db, dx, newX=False,
                          precision=S.PREC_MIXED)
    assert st["status"] == S._lib.OK and st["iters"] == 0 and not dx.get().any()
    # warm start from the FP64 solution: converges at once
    b = M.extra["b_all"][1]
    xo, _ = oracle.pbicgstab(M.system, b, 1e-12, 5000)
    db2, dx2 = ctx.array(b), ctx.array(xo)
    st = ctx.solve_device(A, S.PBICGSTAB, TOL, 100, db2, dx2, newX=False,
                          precision=S.PREC_MIXED)
    assert st["status"] == S._lib.OK and st["iters"] <= 1
    assert relerr(dx2.get(), xo) < 1e-9
    A.destroy(); db.free(); dx.free(); db2.free(); dx2.free()


@pytest.mark.parametrize("n", [12, 96])
def test_mixed_uvw_distinct_diagonals(ctx, n):
    """u/v/w on one pattern with the per-component diagonals the reference
    builds (ad + c*n), batched in one set of launches, FP32 inner."""
    M = systems.momentum3d(n)
    s = M.system
    ads = [s.ad * (1.0 + 0.05 * c) for c in range(3)]
    bs = list(M.extra["b_all"])
    print(f"mixed uvw n={n}", flush=True)
    with mixed(ctx):
        xs, sts = ctx.solve_shared(s, S.PBICGSTAB, TOL, 5000, bs, ads=ads)
    for c in range(3):
        sc = with_ad(s, ads[c])
        xo, so = oracle.pbicgstab(sc, bs[c], TOL, 5000)
        assert sts[c]["status"] == S._lib.OK
        assert np.linalg.norm(bs[c] - oracle.matvec(sc, xs[c])) < TOL * np.linalg.norm(bs[c])
        assert relerr(xs[c], xo) < 1e-6, (c, relerr(xs[c], xo), sts[c], so["iters"])


@pytest.mark.parametrize("delta", [1e-2, 1e-4])
def test_mixed_uvw_fold_threshold(ctx, delta):
    """The fold threshold ru_delta_bicg below its default on the 96^3 u/v/w
    batch: each system still meets the reference's stop test and agrees with
    the oracle, with fewer folds than at the default 0.1 (one per iteration
    there: these systems drop the residual about tenfold per iteration)."""
    M = systems.momentum3d(96)
    s = M.system
    ads = [s.ad * (1.0 + 0.05 * c) for c in range(3)]
    bs = list(M.extra["b_all"])
    print(f"mixed uvw 96^3 ru_delta_bicg {delta:g}", flush=True)
    with mixed(ctx):
        _, st_def = ctx.solve_shared(s, S.PBICGSTAB, TOL, 5000, bs, ads=ads)
        ctx.set_option("ru_delta_bicg", delta)
        try:
            xs, sts = ctx.solve_shared(s, S.PBICGSTAB, TOL, 5000, bs, ads=ads)
        finally:
            ctx.set_option("ru_delta_bicg", 0.1)
    print(f"  folds {[st['outer'] for st in sts]} against "
          f"{[st['outer'] for st in st_def]} at 0.1", flush=True)
    for c in range(3):
        sc = with_ad(s, ads[c])
        xo, so = oracle.pbicgstab(sc, bs[c], TOL, 5000)
        assert sts[c]["status"] == S._lib.OK
        assert np.linalg.norm(bs[c] - oracle.matvec(sc, xs[c])) < TOL * np.linalg.norm(bs[c])
        assert relerr(xs[c], xo) < 1e-6, (c, relerr(xs[c], xo), sts[c], so["iters"])
        assert sts[c]["outer"] < st_def[c]["outer"]


def test_mixed_species_batch(ctx):
    sysm, sp = systems.species3d(16, 5, seed=2)
    als = [a for a, _, _ in sp]
    ads = [d for _, d, _ in sp]
    bs = [b for _, _, b in sp]
    with mixed(ctx):
        xs, sts = ctx.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
    for c in range(len(sp)):
        sc = System(sysm.neq, sysm.nad, sysm.ia, sysm.ja, als[c], ads[c], sysm.storage,
                    sysm.unsym)
        xo, so = oracle.pbicgstab(sc, bs[c], TOL, 5000)
        assert sts[c]["status"] == S._lib.OK
        assert np.linalg.norm(bs[c] - oracle.matvec(sc, xs[c])) < TOL * np.linalg.norm(bs[c])
        assert relerr(xs[c], xo) < 1e-6


def test_mixed_option_reaches_the_dropin_path(ctx):
    """sfr_solve (the host-pointer path behind sfr_solverC) runs PCG and
    BiCGSTAB in mixed precision under the option; GMRES stays FP64."""
    M = systems.momentum3d(10)
    B = systems.poisson3d(10, seed=4)
    with mixed(ctx):
        x, st = ctx.solve(M.system, S.PBICGSTAB, TOL, 5000, M.b)
        assert st["outer"] >= 1 and relerr(x, oracle.pbicgstab(M.system, M.b, TOL, 5000)[0]) < 1e-6
        x, st = ctx.solve(B.system, S.PCG, TOL, 5000, B.b)
        assert st["outer"] >= 1 and relerr(x, oracle.pcg(B.system, B.b, TOL, 5000)[0]) < 1e-6
        x, st = ctx.solve(M.system, S.GMRES, TOL, 5000, M.b)
        xo, so = oracle.gmres(M.system, M.b, TOL, 5000)
        assert st["outer"] == so["cycles"] and relerr(x, xo) < 1e-6


@pytest.mark.parametrize("n,kmin", [(48, 1e-2), (64, 1e-3)])
def test_mixed_bicgstab_long_solve(ctx, n, kmin):
    """A long solve (a variable-coefficient Poisson system, hundreds of
    iterations): the FP32 loop restarts at each fold and, should its FP64
    residual stop improving, the solve is finished in FP64 within the same
    call -- either way the result meets the reference's stop test and agrees
    with the FP64 BiCGSTAB solution."""
    B = systems.poisson3d(n, seed=2, kmin=kmin, reorder=True, storage=S.ELLPACK)
    print(f"mixed bicgstab long {n}^3 kmin {kmin}", flush=True)
    A = ctx.matrix(B.system)
    db, dx = ctx.array(B.b), ctx.empty(B.system.neq)
    st64 = ctx.solve_device(A, S.PBICGSTAB, TOL, 20000, db, dx)
    x64 = dx.get()
    st = ctx.solve_device(A, S.PBICGSTAB, TOL, 20000, db, dx, precision=S.PREC_MIXED)
    x = dx.get()
    print(f"  fp64 {st64['iters']} iterations; mixed {st['iters']} iterations, "
          f"{st['outer']} folds, status {st['status']}", flush=True)
    assert st64["status"] == st["status"] == S._lib.OK
    assert st["relres"] < TOL
    # both meet ||b - Ax|| < tol ||b||; their distance is bounded by the
    # condition number times tol, far below this on these systems
    assert relerr(x, x64) < 1e-4
    db.free(); dx.free(); A.destroy()
