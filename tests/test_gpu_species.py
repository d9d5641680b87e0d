"""GPU parity of the species batch (sfr_solve_batched): transported scalars
on one pattern with their own values, as the reference's per-species
solverC loop (src/Combustion.c:76-92, 222-240) solves them one by one.

Bar: each system's solution agrees with the oracle's BiCGSTAB (the
reference's pbicgstab restated, pinned in test_oracle.py) to a relative
1e-6 at tol 1e-8, and is bitwise the solution of the single-system GPU
solve of the same values (the batched kernels keep each system's
recurrences, summation order included).
"""
import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems
from sfr3d_amd.solver import System

pytestmark = pytest.mark.gpu

TOL = 1e-8
XTOL = 1e-6


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


def as_system(sysm, al, ad):
    return System(sysm.neq, sysm.nad, sysm.ia, sysm.ja, al, ad, sysm.storage,
                  sysm.unsym)


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
@pytest.mark.parametrize("nsys", [3, 5, 10])
def test_species_batch_matches_oracle_and_single_solves(ctx, storage, nsys):
    sysm, sp = systems.species3d(12, nsys, seed=nsys, storage=storage)
    print(f"species storage={storage} nsys={nsys}", flush=True)
    als, ads, bs = [a for a, _, _ in sp], [d for _, d, _ in sp], [b for _, _, b in sp]
    xs, sts = ctx.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
    for c in range(nsys):
        sc = as_system(sysm, als[c], ads[c])
        xo, so = oracle.pbicgstab(sc, bs[c], TOL, 5000)
        assert sts[c]["status"] == S._lib.OK, (c, sts[c])
        assert abs(int(sts[c]["iters"]) - int(so["iters"])) <= 2
        assert relerr(xs[c], xo) < XTOL, (c, relerr(xs[c], xo))
        x1, s1 = ctx.solve(sc, S.PBICGSTAB, TOL, 5000, bs[c])
        assert s1["iters"] == sts[c]["iters"]
        assert relerr(xs[c], x1) < 1e-12, (c, relerr(xs[c], x1))


def test_species_zero_rhs_maxit_and_warm_start(ctx):
    sysm, sp = systems.species3d(10, 4, seed=7)
    als, ads, bs = [a for a, _, _ in sp], [d for _, d, _ in sp], [b.copy() for _, _, b in sp]
    bs[1][:] = 0.0
    xs, sts = ctx.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
    assert sts[1]["status"] == S._lib.OK and sts[1]["iters"] == 0 and not xs[1].any()
    for c in (0, 2, 3):
        xo, _ = oracle.pbicgstab(as_system(sysm, als[c], ads[c]), bs[c], TOL, 5000)
        assert relerr(xs[c], xo) < XTOL
    xw, stw = ctx.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs, xs=xs,
                                newX=False)
    for c in (0, 2, 3):
        assert stw[c]["status"] == S._lib.OK and stw[c]["iters"] <= 2
        assert relerr(xw[c], xs[c]) < 1e-7
    xm, stm = ctx.solve_batched(sysm, S.PBICGSTAB, 1e-14, 2, als, ads, bs)
    for c in (0, 2, 3):
        assert stm[c]["status"] == S._lib.MAXIT and stm[c]["iters"] == 2
        x2, _ = ctx.solve(as_system(sysm, als[c], ads[c]), S.PBICGSTAB, 1e-14, 2, bs[c])
        assert relerr(xm[c], x2) < 1e-12


def test_species_other_solvers_run_one_after_another(ctx):
    """Another solver id (the species solver is read from the solver
    configuration, src/ReadFile.c:5421-5444) solves the systems in turn."""
    sysm, sp = systems.species3d(10, 3, seed=3)
    als, ads, bs = [a for a, _, _ in sp], [d for _, d, _ in sp], [b for _, _, b in sp]
    for solver, orc in ((S.PBICGSTABL2, oracle.pbicgstabl2), (S.GMRES, oracle.gmres)):
        xs, sts = ctx.solve_batched(sysm, solver, TOL, 5000, als, ads, bs)
        for c in range(3):
            xo, so = orc(as_system(sysm, als[c], ads[c]), bs[c], TOL, 5000)
            assert sts[c]["status"] == S._lib.OK
            assert relerr(xs[c], xo) < XTOL, (solver, c)


def test_species_batch_then_single_solve_same_pattern(ctx):
    """The batch leaves the cached pattern usable: a host-pointer solve of
    one species afterwards uploads its values and matches the oracle."""
    sysm, sp = systems.species3d(9, 4, seed=11, storage=S.ELLPACK)
    als, ads, bs = [a for a, _, _ in sp], [d for _, d, _ in sp], [b for _, _, b in sp]
    ctx.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
    sc = as_system(sysm, als[2], ads[2])
    x, st = ctx.solve(sc, S.PBICGSTAB, TOL, 5000, bs[2])
    xo, _ = oracle.pbicgstab(sc, bs[2], TOL, 5000)
    assert relerr(x, xo) < XTOL


def test_species_batch_survives_other_workspaces():
    """Regression: growing another solver's workspace (GMRES's basis) between
    two species batches must leave the batch's own buffers valid."""
    with S.Context(0) as c2:
        sysm, sp = systems.species3d(9, 3, seed=5, storage=S.ELLPACK)
        als, ads, bs = [a for a, _, _ in sp], [d for _, d, _ in sp], [b for _, _, b in sp]
        x1, s1 = c2.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
        B = systems.momentum3d(11, seed=2)
        c2.solve(B.system, S.GMRES, TOL, 5000, B.b)
        x2, s2 = c2.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
        for c in range(3):
            assert s2[c]["status"] == S._lib.OK
            assert np.array_equal(x1[c], x2[c])
