"""GPU parity of MINRES (solver id 5) against the oracle's pminres, the
reference's src/Iters.c:309-555 restated in C and pinned bit for bit to the
reference's compiled solverC (test_oracle_minres.py, golden vectors
tests/golden/ref_minres*).

Bar: the solution agrees with the oracle's to a relative 1e-6 at the
reference tolerance 1e-8, and the iteration counts agree to within 2 or 2 %
(device dot products are summed in a different, fixed order).
"""
import ctypes
import json
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


def check(st, so, x, xo):
    assert st["status"] == S._lib.OK, st
    assert abs(int(st["iters"]) - int(so["iters"])) <= max(2, 0.02 * so["iters"]), \
        (st["iters"], so["iters"])
    assert relerr(x, xo) < XTOL, relerr(x, xo)
    assert np.isclose(st["conv"], so["conv"], rtol=1e-12)
    # the true M-norm residual stays under the reference's own warning
    # threshold, 3.16 conv (src/Iters.c:529-530)
    assert st["norm_r_m"] < 3.16 * st["conv"]


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
def test_minres_poisson_matches_oracle(ctx, storage):
    B = systems.poisson3d(14, seed=4, storage=storage)
    print(f"minres storage={storage}", flush=True)
    x, st = ctx.solve(B.system, S.MINRES, TOL, 5000, B.b)
    xo, so = oracle.minres(B.system, B.b, TOL, 5000)
    check(st, so, x, xo)


def test_minres_golden_vectors(ctx):
    import test_oracle_minres as T
    sols = np.load(os.path.join(GOLD, "ref_minres_solutions.npz"))
    for g in T.minres_cases():
        B = T.make_case(g["name"])
        print("minres golden", g["name"], flush=True)
        x, st = ctx.solve(B.system, S.MINRES, g["tol"], g["maxit"], B.b)
        if g["log_iters"] > g["maxit"]:
            assert st["status"] == S._lib.MAXIT and st["iters"] == g["maxit"]
            assert relerr(x, sols[g["name"]]) < 1e-10      # the same iterates
        else:
            assert st["status"] == S._lib.OK
            assert abs(st["iters"] - g["log_iters"]) <= 2
            assert relerr(x, sols[g["name"]]) < XTOL
        assert abs(st["xKx"] - g["xKx"]) <= 1e-6 * abs(g["xKx"])


def test_minres_symmetric_indefinite(ctx):
    """MINRES's reason to exist next to CG: a symmetric indefinite system
    (the Poisson matrix shifted past its smallest eigenvalue, diagonal still
    positive for the Jacobi preconditioner)."""
    from test_oracle_minres import shifted
    B = systems.poisson3d(10, seed=5, storage=S.CSRD)
    sysm, sigma = shifted(B, 0.3)
    x, st = ctx.solve(sysm, S.MINRES, TOL, 20000, B.b)
    xo, so = oracle.minres(sysm, B.b, TOL, 20000)
    check(st, so, x, xo)


def test_minres_device_formats_and_larger(ctx):
    B = systems.poisson3d(40, seed=2)
    xo, so = oracle.minres(B.system, B.b, TOL, 20000)
    for fmt in (S.FMT_ELL, S.FMT_CSR):
        print(f"minres 40^3 fmt={fmt}", flush=True)
        A = ctx.matrix(B.system, fmt)
        db, dx = ctx.array(B.b), ctx.empty(B.system.neq)
        st = ctx.solve_device(A, S.MINRES, TOL, 20000, db, dx)
        check(st, so, dx.get(), xo)
        r = B.b - oracle.matvec(B.system, dx.get())
        assert np.isclose(st["norm_r"], np.linalg.norm(r), rtol=1e-6)
        A.destroy(); db.free(); dx.free()


def test_minres_maxit_zero_rhs_warm_start(ctx):
    B = systems.poisson3d(10, seed=7, storage=S.CSRD)
    for k in (1, 2, 7, 13):
        x, st = ctx.solve(B.system, S.MINRES, 1e-12, k, B.b)
        xo, so = oracle.minres(B.system, B.b, 1e-12, k)
        assert st["status"] == S._lib.MAXIT and so["status"] == 1
        assert st["iters"] == k
        assert relerr(x, xo) < 1e-10, k          # the same k iterates
    x, st = ctx.solve(B.system, S.MINRES, TOL, 100, np.zeros(B.system.neq))
    assert st["status"] == S._lib.OK and st["iters"] == 0 and not x.any()
    x0 = np.linspace(-1.0, 1.0, B.system.neq)
    xw, stw = ctx.solve(B.system, S.MINRES, TOL, 5000, B.b, x=x0, newX=False)
    xwo, swo = oracle.minres(B.system, B.b, TOL, 5000, x=x0, newX=False)
    check(stw, swo, xw, xwo)
    A = ctx.matrix(B.system)
    db, dx = ctx.array(B.b), ctx.empty(B.system.neq)
    with pytest.raises(S.SolverError):
        ctx.solve_device(A, S.MINRES, TOL, 100, db, dx, precision=S.PREC_MIXED)
    A.destroy(); db.free(); dx.free()


def test_minres_deterministic_and_chunk_sizes(ctx):
    B = systems.poisson3d(16, seed=9)
    x1, s1 = ctx.solve(B.system, S.MINRES, TOL, 5000, B.b)
    x2, s2 = ctx.solve(B.system, S.MINRES, TOL, 5000, B.b)
    assert s1["iters"] == s2["iters"] and np.array_equal(x1, x2)
    ctx.set_option("chunk", 6)
    x3, s3 = ctx.solve(B.system, S.MINRES, TOL, 5000, B.b)
    ctx.set_option("chunk", 40)
    x4, s4 = ctx.solve(B.system, S.MINRES, TOL, 5000, B.b)
    ctx.set_option("chunk", 16)
    assert s3["iters"] == s1["iters"] == s4["iters"]
    assert np.array_equal(x1, x3) and np.array_equal(x1, x4)


def _solverC(solver, sysm, b, tol, maxit):
    lib = S._lib.lib
    libc = ctypes.CDLL(None)
    libc.tmpfile.restype = ctypes.c_void_p
    libc.fflush.argtypes = [ctypes.c_void_p]
    libc.rewind.argtypes = [ctypes.c_void_p]
    libc.fgets.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    vp = ctypes.c_void_p
    lib.sfr_solverC.restype = None
    lib.sfr_solverC.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, vp,
                                vp, ctypes.c_double, ctypes.c_uint,
                                ctypes.c_short, ctypes.c_short, vp,
                                ctypes.c_bool, ctypes.c_bool, ctypes.c_bool,
                                ctypes.c_char_p]
    x = np.zeros(sysm.neq)
    b = np.ascontiguousarray(b)
    f = libc.tmpfile()
    lib.sfr_solverC(None, sysm.neq, sysm.neq, sysm.nad, 0, sysm.ia.ctypes.data,
                    sysm.ja.ctypes.data, sysm.al.ctypes.data, sysm.ad.ctypes.data,
                    None, b.ctypes.data, x.ctypes.data, None, None, tol, maxit,
                    sysm.storage, solver, f, True, True, bool(sysm.unsym), b"pres")
    libc.fflush(f)
    libc.rewind(f)
    buf = ctypes.create_string_buffer(512)
    libc.fgets(buf, 512, f)
    libc.fclose(f)
    return x, buf.value.decode()


def test_minres_solverC_log_line_and_no_exit_at_maxit():
    """The reference's pminres only warns at maxIt and logs maxIt + 1
    iterations (src/Iters.c:540-552); the drop-in does the same and
    returns."""
    B = systems.poisson3d(8, seed=3, storage=S.CSRD)
    x, line = _solverC(5, B.system, B.b, TOL, 5000)
    xo, so = oracle.minres(B.system, B.b, TOL, 5000)
    assert line.startswith("PMINRES: tol") and "xKx" in line
    it = int(line.split("iteration")[1].split()[0])
    assert abs(it - so["iters"]) <= 2 and relerr(x, xo) < XTOL
    x, line = _solverC(5, B.system, B.b, 1e-14, 4)
    assert line.startswith("PMINRES: tol") and " iteration 5 " in line
