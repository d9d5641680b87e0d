"""GPU parity of BiCGSTAB(2) (solver id 3) against the oracle's
pbicgstabl2, the reference's src/Bicgstab.c:1172-1378 restated in C and
pinned bit for bit to the reference's compiled solverC
(test_oracle_bicgl2.py, golden vectors tests/golden/ref_bicgl2*).

Bar: the solution agrees with the oracle's to a relative 1e-6 at the
reference tolerance 1e-8, and the iteration counts (of two BiCG steps each)
agree to within 2 or 10 %, whichever is larger.  Device dot products are
summed in a different, fixed order, and BiCGSTAB's irregular convergence
amplifies the last bits: the oracle itself, given b scaled by 1 + 1e-15
noise, stops the 12^3 Poisson case (48 iterations) after 44 to 51.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems

pytestmark = pytest.mark.gpu

TOL = 1e-8
XTOL = 1e-6
GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


def iters_agree(a, b):
    return abs(int(a) - int(b)) <= max(2, 0.1 * b)


def check(st, so, x, xo):
    assert st["status"] == S._lib.OK, st
    assert iters_agree(st["iters"], so["iters"]), (st["iters"], so["iters"])
    assert relerr(x, xo) < XTOL, relerr(x, xo)
    assert np.isclose(st["conv"], so["conv"], rtol=1e-12)
    # the true residual stays under the reference's own warning threshold,
    # 3.16 conv (src/Bicgstab.c:1345-1346)
    assert st["norm_r"] < 3.16 * st["conv"]


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
def test_bicgl2_momentum_matches_oracle(ctx, storage):
    B = systems.momentum3d(12, seed=5, storage=storage)
    for comp in range(3):
        b = B.extra["b_all"][comp]
        print(f"bicgl2 momentum storage={storage} comp={comp}", flush=True)
        x, st = ctx.solve(B.system, S.PBICGSTABL2, TOL, 5000, b)
        xo, so = oracle.pbicgstabl2(B.system, b, TOL, 5000)
        check(st, so, x, xo)


def test_bicgl2_golden_vectors(ctx):
    import test_oracle_bicgl2 as T
    sols = np.load(os.path.join(GOLD, "ref_bicgl2_solutions.npz"))
    for g in T.bicgl2_cases():
        B, b = T.make_case(g["name"])
        print("bicgl2 golden", g["name"], flush=True)
        x, st = ctx.solve(B.system, S.PBICGSTABL2, g["tol"], 5000, b)
        assert st["status"] == S._lib.OK
        assert iters_agree(st["iters"], g["iters"]), (g["name"], st["iters"])
        assert relerr(x, sols[g["name"]]) < XTOL
        assert abs(st["xKx"] - g["xKx"]) <= 1e-6 * abs(g["xKx"])


def test_bicgl2_device_formats_and_larger(ctx):
    B = systems.momentum3d(40, seed=2)
    b = B.extra["b_all"][2]
    xo, so = oracle.pbicgstabl2(B.system, b, TOL, 5000)
    for fmt in (S.FMT_ELL, S.FMT_CSR):
        print(f"bicgl2 40^3 fmt={fmt}", flush=True)
        A = ctx.matrix(B.system, fmt)
        db, dx = ctx.array(b), ctx.empty(B.system.neq)
        st = ctx.solve_device(A, S.PBICGSTABL2, TOL, 5000, db, dx)
        check(st, so, dx.get(), xo)
        A.destroy(); db.free(); dx.free()


def test_bicgl2_poisson_pinned_and_escaped_far_columns(ctx):
    B = systems.poisson2d(24, 18, seed=5, storage=S.CSRD, pinned_cell=3)
    x, st = ctx.solve(B.system, S.PBICGSTABL2, TOL, 5000, B.b)
    xo, so = oracle.pbicgstabl2(B.system, B.b, TOL, 5000)
    check(st, so, x, xo)
    from sfr3d_amd import structure as ST
    from sfr3d_amd.solver import System
    B = systems.poisson3d(34, seed=6, storage=S.ELLPACK, reorder=False,
                          pinned_cell=34 ** 3 - 500)
    al, ad, b = ST.assemble(B.pattern, B.id, B.nelcon, B.nviz, B.extra["lA"],
                            B.extra["lB"], ell_quirk=True)
    s0 = B.system
    sq = System(s0.neq, s0.nad, s0.ia, s0.ja, al, 1.05 * ad, S.ELLPACK, True)
    bq = b[:s0.neq].copy()
    A = ctx.matrix(sq)
    assert A.info()["index_bytes"] == 2
    db, dx = ctx.array(bq), ctx.empty(sq.neq)
    st = ctx.solve_device(A, S.PBICGSTABL2, TOL, 5000, db, dx)
    xo, so = oracle.pbicgstabl2(sq, bq, TOL, 5000)
    check(st, so, dx.get(), xo)
    A.destroy(); db.free(); dx.free()


def test_bicgl2_maxit_zero_rhs_warm_start(ctx):
    B = systems.momentum3d(10, seed=8)
    b = B.extra["b_all"][0]
    for k in (1, 2, 3):
        x, st = ctx.solve(B.system, S.PBICGSTABL2, 1e-14, k, b)
        xo, so = oracle.pbicgstabl2(B.system, b, 1e-14, k)
        assert st["status"] == S._lib.MAXIT and so["status"] == 1
        assert st["iters"] == k
        assert relerr(x, xo) < 1e-10, k          # the same k iterates
    x, st = ctx.solve(B.system, S.PBICGSTABL2, TOL, 100, np.zeros(B.system.neq))
    assert st["status"] == S._lib.OK and st["iters"] == 0 and not x.any()
    x0 = np.linspace(-1.0, 1.0, B.system.neq)
    xw, stw = ctx.solve(B.system, S.PBICGSTABL2, TOL, 5000, b, x=x0, newX=False)
    xwo, swo = oracle.pbicgstabl2(B.system, b, TOL, 5000, x=x0, newX=False)
    check(stw, swo, xw, xwo)
    A = ctx.matrix(B.system)
    db, dx = ctx.array(b), ctx.empty(B.system.neq)
    with pytest.raises(S.SolverError):
        ctx.solve_device(A, S.PBICGSTABL2, TOL, 100, db, dx, precision=S.PREC_MIXED)
    A.destroy(); db.free(); dx.free()


def test_bicgl2_deterministic_and_chunk_sizes(ctx):
    B = systems.momentum3d(16, seed=9)
    b = B.extra["b_all"][1]
    x1, s1 = ctx.solve(B.system, S.PBICGSTABL2, 1e-10, 5000, b)
    x2, s2 = ctx.solve(B.system, S.PBICGSTABL2, 1e-10, 5000, b)
    assert s1["iters"] == s2["iters"] and np.array_equal(x1, x2)
    ctx.set_option("chunk", 1)
    x3, s3 = ctx.solve(B.system, S.PBICGSTABL2, 1e-10, 5000, b)
    ctx.set_option("chunk", 16)
    assert s3["iters"] == s1["iters"] and np.array_equal(x1, x3)


def test_bicgl2_solverC_log_line():
    from test_gpu_minres import _solverC
    B = systems.momentum3d(8, seed=3)
    b = B.extra["b_all"][0]
    x, line = _solverC(3, B.system, b, TOL, 5000)
    xo, so = oracle.pbicgstabl2(B.system, b, TOL, 5000)
    assert line.startswith("PBICGSTAB(2): tol")
    it = int(line.split("iteration")[1].split()[0])
    assert iters_agree(it, so["iters"]) and relerr(x, xo) < XTOL
