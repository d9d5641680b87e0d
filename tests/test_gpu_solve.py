"""GPU parity of the solve path against the oracle (the reference's pcg /
pbicgstab restated in C, itself checked bit-exact against the reference's own
compiled solverC in test_oracle.py).

Bar: solutions agree with the oracle's to a relative 1e-6 at the
reference tolerance 1e-8; iteration counts agree to within 2 (device dot
products are summed in a different, fixed order, so the last bits differ).
"""
import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems

pytestmark = pytest.mark.gpu

TOL = 1e-8
XTOL = 1e-6          # relative solution agreement with the oracle


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


def true_relres(sysm, x, b):
    return np.linalg.norm(b - oracle.matvec(sysm, x)) / np.linalg.norm(b)


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
@pytest.mark.parametrize("n", [9, 20])
def test_pcg_fp64_matches_oracle(ctx, storage, n):
    B = systems.poisson3d(n, seed=11, storage=storage)
    print(f"pcg storage={storage} n={n}", flush=True)
    x, st = ctx.solve(B.system, S.PCG, TOL, 10000, B.b)
    xo, so = oracle.pcg(B.system, B.b, TOL, 10000)
    assert st["status"] == S._lib.OK
    assert abs(st["iters"] - so["iters"]) <= 2
    assert relerr(x, xo) < XTOL
    assert st["norm_r_m"] < 1.0001 * st["conv"] or abs(st["iters"] - so["iters"]) <= 1
    assert abs(st["xKx"] - so["xKx"]) <= 1e-8 * abs(so["xKx"])
    assert np.isclose(st["relres"], true_relres(B.system, x, B.b), rtol=1e-6, atol=1e-15)


@pytest.mark.parametrize("fmt", [S.FMT_ELL, S.FMT_CSR])
def test_pcg_device_formats(ctx, fmt):
    B = systems.poisson3d(14, seed=2, storage=S.CSRD)
    A = ctx.matrix(B.system, fmt)
    assert A.info()["format"] == fmt
    db = ctx.array(B.b)
    dx = ctx.empty(B.system.neq)
    st = ctx.solve_device(A, S.PCG, TOL, 10000, db, dx)
    xo, so = oracle.pcg(B.system, B.b, TOL, 10000)
    assert abs(st["iters"] - so["iters"]) <= 2
    assert relerr(dx.get(), xo) < XTOL
    A.destroy(); db.free(); dx.free()


def test_pcg_pinned_cell_2d(ctx):
    B = systems.poisson2d(40, 33, seed=4, storage=S.CSRD, pinned_cell=0)
    assert B.system.neq == 40 * 33 - 1            # numEqV2 drops the cell
    x, st = ctx.solve(B.system, S.PCG, TOL, 20000, B.b)
    xo, so = oracle.pcg(B.system, B.b, TOL, 20000)
    assert abs(st["iters"] - so["iters"]) <= 2
    assert relerr(x, xo) < XTOL


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
def test_bicgstab_matches_oracle(ctx, storage):
    M = systems.momentum3d(12, storage=storage)
    for comp in range(3):
        b = M.extra["b_all"][comp]
        x, st = ctx.solve(M.system, S.PBICGSTAB, TOL, 5000, b)
        xo, so = oracle.pbicgstab(M.system, b, TOL, 5000)
        assert st["status"] == S._lib.OK
        assert abs(st["iters"] - so["iters"]) <= 2
        assert relerr(x, xo) < XTOL


def test_bicgstab_harder_system(ctx):
    # weaker diagonal (large dt): more iterations, exercises both half steps
    B = systems.poisson3d(16, seed=9, storage=S.CSRD)
    x, st = ctx.solve(B.system, S.PBICGSTAB, TOL, 10000, B.b)
    xo, so = oracle.pbicgstab(B.system, B.b, TOL, 10000)
    assert abs(st["iters"] - so["iters"]) <= 3
    assert relerr(x, xo) < 1e-5


def test_mixed_precision_pcg(ctx):
    B = systems.poisson3d(24, seed=5)
    A = ctx.matrix(B.system)
    db, dx = ctx.array(B.b), ctx.empty(B.system.neq)
    st = ctx.solve_device(A, S.PCG, TOL, 20000, db, dx, precision=S.PREC_MIXED)
    xo, so = oracle.pcg(B.system, B.b, TOL, 20000)
    x = dx.get()
    assert st["status"] == S._lib.OK
    assert st["outer"] >= 1
    # same stopping test as the FP64 solver, on the true FP64 residual
    assert st["norm_r_m"] < st["conv"]
    assert relerr(x, xo) < XTOL
    A.destroy(); db.free(); dx.free()


def test_warm_start_and_reuse_pattern(ctx):
    B = systems.poisson3d(10, seed=3, storage=S.CSRC)
    x0 = np.linspace(-1, 1, B.system.neq)
    x, st = ctx.solve(B.system, S.PCG, TOL, 5000, B.b, x=x0, newX=False)
    xo, so = oracle.pcg(B.system, B.b, TOL, 5000, x=x0, newX=False)
    assert abs(st["iters"] - so["iters"]) <= 2 and relerr(x, xo) < XTOL
    # new values, same pattern arrays: value-only upload path
    B.system.al *= 1.5
    B.system.ad *= 1.5
    x2, st2 = ctx.solve(B.system, S.PCG, TOL, 5000, B.b)
    xo2, _ = oracle.pcg(B.system, B.b, TOL, 5000)
    assert relerr(x2, xo2) < XTOL


def test_zero_rhs_and_maxit(ctx):
    B = systems.poisson3d(8, seed=1)
    x, st = ctx.solve(B.system, S.PCG, TOL, 100, np.zeros(B.system.neq))
    assert st["status"] == S._lib.OK and st["iters"] == 0 and not x.any()
    x, st = ctx.solve(B.system, S.PCG, TOL, 5, B.b)
    assert st["status"] == S._lib.MAXIT and st["iters"] == 5
    xo = np.zeros(B.system.neq)
    # the first 5 iterates agree with the oracle's 5th iterate
    orc = oracle.orc()
    import ctypes
    sto = oracle.OrcStats()
    orc.orc_pcg(B.system.neq, B.system.nad, B.system.ia, B.system.ja,
                B.system.al, B.system.ad, B.b, xo, TOL, 5, 1, B.system.storage,
                0, ctypes.byref(sto))
    assert sto.status == 1 and relerr(x, xo) < 1e-10


def test_deterministic(ctx):
    B = systems.poisson3d(18, seed=8)
    x1, s1 = ctx.solve(B.system, S.PCG, TOL, 5000, B.b)
    x2, s2 = ctx.solve(B.system, S.PCG, TOL, 5000, B.b)
    assert s1["iters"] == s2["iters"] and np.array_equal(x1, x2)


def test_retune_and_graph_cache_churn():
    """Rows per thread re-tuned and the chunk length changed between solves
    of several solver ids: many more loop graphs are captured than the
    context's cache holds (16), so entries are evicted, least recently used
    first, and captured again.  Every solve matches the oracle."""
    B = systems.poisson3d(10, seed=4, storage=S.ELLPACK)
    ref = {S.PCG: oracle.pcg(B.system, B.b, TOL, 5000)[0],
           S.PBICGSTAB: oracle.pbicgstab(B.system, B.b, TOL, 5000)[0],
           S.MINRES: oracle.minres(B.system, B.b, TOL, 5000)[0]}
    cases = ((S.PCG, S.PREC_FP64), (S.PCG, S.PREC_MIXED),
             (S.PBICGSTAB, S.PREC_FP64), (S.MINRES, S.PREC_FP64))
    with S.Context(0) as c2:
        A = c2.matrix(B.system)
        db, dx = c2.array(B.b), c2.empty(B.system.neq)
        for chunk in (16, 4, 8, 16):
            c2.set_option("chunk", chunk)
            for rpt in (8, 2, 1):
                print(f"churn chunk={chunk} rpt={rpt}", flush=True)
                A.tune(rpt)
                for solver, prec in cases:
                    st = c2.solve_device(A, solver, TOL, 5000, db, dx, precision=prec)
                    assert st["status"] == S._lib.OK, (chunk, rpt, solver, prec)
                    assert relerr(dx.get(), ref[solver]) < XTOL, (chunk, rpt, solver, prec)
        A.destroy(); db.free(); dx.free()


def test_measured_pcg_tuning(ctx):
    """sfr_matrix_tune with rpt = 0 times short FP64 PCG runs at 1, 2, 4 and
    8 rows per thread and keeps the fastest for the FP64 PCG loop; solves
    afterwards (FP64 and mixed, ELL and CSR) still match the oracle, and an
    explicit rows-per-thread choice afterwards still applies."""
    B = systems.poisson3d(24, seed=5, storage=S.CSRD)
    xo, so = oracle.pcg(B.system, B.b, TOL, 5000)
    for fmt in (S.FMT_ELL, S.FMT_CSR):
        print(f"measured tuning fmt={fmt}", flush=True)
        A = ctx.matrix(B.system, fmt)
        A.tune(0)
        assert A.info()["rows_per_thread_pcg"] in (1, 2, 4, 8)
        db, dx = ctx.array(B.b), ctx.empty(B.system.neq)
        for prec in (S.PREC_FP64, S.PREC_MIXED):
            st = ctx.solve_device(A, S.PCG, TOL, 5000, db, dx, precision=prec)
            assert st["status"] == S._lib.OK, (fmt, prec)
            assert relerr(dx.get(), xo) < XTOL, (fmt, prec)
        st = ctx.solve_device(A, S.PCG, TOL, 5000, db, dx)
        assert abs(int(st["iters"]) - int(so["iters"])) <= 2
        x_tuned = dx.get()
        A.tune(8)
        assert A.info()["rows_per_thread"] == A.info()["rows_per_thread_pcg"] == 8
        st8 = ctx.solve_device(A, S.PCG, TOL, 5000, db, dx)
        assert st8["status"] == S._lib.OK and relerr(dx.get(), x_tuned) < 1e-9
        A.destroy(); db.free(); dx.free()
    with pytest.raises(S.SolverError):
        A = ctx.matrix(B.system)
        try:
            A.tune(3)
        finally:
            A.destroy()


def test_ell_16bit_offsets_with_escaped_far_columns(ctx):
    """SFR3D's ELLPACK points a slot facing a cell without an equation at
    column 0 (src/EllPack.c:54-60), and with the reference's assembly
    (ell_quirk) that slot carries the face coefficient.  In rows far from 0
    the offset does not fit 16 bits: those entries are escaped to an
    exception list while the rest of the matrix keeps 16-bit offsets.  The
    solve must see the reference's operator, column-0 couplings included."""
    from sfr3d_amd import structure as ST
    from sfr3d_amd.solver import System
    B = systems.poisson3d(34, seed=6, storage=S.ELLPACK, reorder=False,
                          pinned_cell=34 ** 3 - 500)
    al, ad, b = ST.assemble(B.pattern, B.id, B.nelcon, B.nviz, B.extra["lA"],
                            B.extra["lB"], ell_quirk=True)
    s0 = B.system
    # 5 % heavier diagonal: the extra column-0 couplings leave the
    # all-Neumann system near-singular otherwise
    sq = System(s0.neq, s0.nad, s0.ia, s0.ja, al, 1.05 * ad, S.ELLPACK, True)
    far = [k for k in range(len(s0.ja)) if s0.ja[k] == 0 and k // 6 > 32767]
    assert far and np.any(al[far] != 0.0)         # far, nonzero couplings exist
    A = ctx.matrix(sq)
    assert A.info()["index_bytes"] == 2
    bb = b[:s0.neq].copy()
    db, dx = ctx.array(bb), ctx.empty(s0.neq)
    st = ctx.solve_device(A, S.PBICGSTAB, TOL, 5000, db, dx)
    xo, so = oracle.pbicgstab(sq, bb, TOL, 5000)
    assert st["status"] == S._lib.OK
    assert abs(st["iters"] - so["iters"]) <= 2
    assert relerr(dx.get(), xo) < XTOL
    A.destroy(); db.free(); dx.free()
    # the plain assembly (zero in those slots) through the mixed path
    A = ctx.matrix(s0)
    assert A.info()["index_bytes"] == 2
    db, dx = ctx.array(B.b), ctx.empty(s0.neq)
    st = ctx.solve_device(A, S.PCG, TOL, 20000, db, dx, precision=S.PREC_MIXED)
    xo, _ = oracle.pcg(s0, B.b, TOL, 20000)
    assert st["status"] == S._lib.OK and relerr(dx.get(), xo) < XTOL
    A.destroy(); db.free(); dx.free()
