"""Face-pair reads of symmetric ELLPACK systems on the GPU.

SFR3D's ELLPACK (src/EllPack.c:26-79) stores every off-diagonal of a
symmetric system twice, in the two rows it couples, in the face order of
src/Adjcency.c.  When the values mirror bit for bit, the FP64 PCG SpMV
reads the first slot of each face pair from the row and the second from the
neighbour (sfr_matrix_pair_planes).  Bar: the same bits as the full read --
x and the iteration count of FP64 PCG equal with the context option
"ell_pairs" on and off, and the mixed loop (which keeps the full read)
unchanged by it -- on the 128^3 bench system (hex rows, RCM) and a
128 x 128 quad system; a value that stops mirroring turns the pairs off and
mirroring values turn them on again, with the solution of the full read
each time.  The same bar holds for the context option "nt_loads" (the
matrix streamed by non-temporal loads in the loops' SpMV kernels, a cache
policy that must not change a bit).
"""
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
def pairs(ctx, on):
    ctx.set_option("ell_pairs", 1 if on else 0)
    try:
        yield
    finally:
        ctx.set_option("ell_pairs", 1)


def solve_both(ctx, A, db, dx, prec):
    out = []
    for on in (True, False):
        with pairs(ctx, on):
            st = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx, precision=prec)
            out.append((st, dx.get(), A.info()["pair_planes"]))
    return out


@pytest.mark.parametrize("case", ["hex128", "quad128"])
def test_pairs_bitwise_fp64_and_mixed(ctx, case):
    if case == "hex128":
        B = systems.poisson3d(128, seed=1, kmin=1e-2, reorder=True, storage=S.ELLPACK)
        kind = 1
    else:
        B = systems.poisson2d(128, 128, seed=2, storage=S.ELLPACK, pinned_cell=None)
        kind = 2
    n = B.system.neq
    print(f"{case}: {n} rows", flush=True)
    A = ctx.matrix(B.system)
    # one launch shape both ways: without a fixed one, the FP64 loop's
    # default rows per thread depends on whether the read is in use, and a
    # different shape sums in a different order
    A.tune(2)
    db, dx = ctx.array(B.b), ctx.empty(n)
    try:
        for prec in (S.PREC_FP64, S.PREC_MIXED):
            (s1, x1, p1), (s0, x0, p0) = solve_both(ctx, A, db, dx, prec)
            print(f"  precision {prec}: iters {s1['iters']} / {s0['iters']}, "
                  f"planes {p1} / {p0}", flush=True)
            assert s1["status"] == S._lib.OK and s0["status"] == S._lib.OK
            if prec == S.PREC_FP64:
                assert p1 == kind and p0 == 0
            assert s1["iters"] == s0["iters"] and s1["outer"] == s0["outer"]
            assert np.array_equal(x1, x0)
            assert s1["norm_r_m"] < s1["conv"]
    finally:
        A.destroy(); db.free(); dx.free()


def test_pairs_follow_the_values(ctx):
    """A value that stops mirroring turns the pair reads off at the next
    solve; restoring it turns them on; each solve is the oracle's."""
    B = systems.poisson3d(24, seed=5, kmin=1e-2, reorder=True, storage=S.ELLPACK)
    n = B.system.neq
    A = ctx.matrix(B.system)
    db, dx = ctx.array(B.b), ctx.empty(n)
    try:
        st = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx)
        x_pair = dx.get()
        assert A.info()["pair_planes"] == 1
        al = B.system.al.copy()
        k = int(np.flatnonzero(al)[len(np.flatnonzero(al)) // 3])
        al[k] *= 1.0 + 2.0 ** -40
        A.set_values(al, B.system.ad)
        sysm = System(n, B.system.nad, B.system.ia, B.system.ja, al, B.system.ad,
                      B.system.storage, B.system.unsym)
        st2 = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx)
        x2 = dx.get()
        assert A.info()["pair_planes"] == 0
        xo2, so2 = oracle.pcg(sysm, B.b, TOL, 100000)
        assert st2["status"] == S._lib.OK and relerr(x2, xo2) < 1e-9
        A.set_values(B.system.al, B.system.ad)
        st3 = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx)
        assert A.info()["pair_planes"] == 1
        assert st3["iters"] == st["iters"] and np.array_equal(dx.get(), x_pair)
        xo, so = oracle.pcg(B.system, B.b, TOL, 100000)
        assert abs(int(st["iters"]) - int(so["iters"])) <= 1 and relerr(x_pair, xo) < 1e-9
    finally:
        A.destroy(); db.free(); dx.free()


@pytest.mark.parametrize("case", ["hex64", "quad128"])
def test_nt_loads_same_bits(ctx, case):
    """FP64 (face-pair read) and mixed PCG with the matrix loads forced
    plain and non-temporal: the same iterations and the same x."""
    if case == "hex64":
        B = systems.poisson3d(64, seed=3, kmin=1e-2, reorder=True, storage=S.ELLPACK)
    else:
        B = systems.poisson2d(128, 128, seed=4, storage=S.ELLPACK, pinned_cell=None)
    n = B.system.neq
    print(f"nt_loads {case}: {n} rows", flush=True)
    A = ctx.matrix(B.system)
    db, dx = ctx.array(B.b), ctx.empty(n)
    try:
        for prec in (S.PREC_FP64, S.PREC_MIXED):
            out = []
            for nt in (0, 1):
                ctx.set_option("nt_loads", nt)
                try:
                    st = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx, precision=prec)
                finally:
                    ctx.set_option("nt_loads", -1)
                out.append((st, dx.get()))
            (s0, x0), (s1, x1) = out
            print(f"  precision {prec}: iters {s0['iters']} / {s1['iters']}", flush=True)
            assert s0["status"] == s1["status"] == S._lib.OK
            assert s0["iters"] == s1["iters"] and s0["outer"] == s1["outer"]
            assert np.array_equal(x0, x1)
        assert A.info()["pair_planes"] in (1, 2)
        with pytest.raises(S.SolverError):
            ctx.set_option("nt_loads", 2)
    finally:
        A.destroy(); db.free(); dx.free()
