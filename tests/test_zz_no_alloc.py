"""GPU test of the allocation-free solve (sfr_solve_device_noalloc), called
through the C ABI: its scalars, partials and work vectors live in the
library's static device scratch, its kernels are launched without graphs.
Bar: x, the iteration count and x.Ax bitwise those of sfr_solve_device on
the same resident matrix (same kernels, same launch shape, same summation
order), and x within 1e-6 of the oracle at tol 1e-8 -- on two calls in a
row, the second with a smaller system, for PCG and BiCGSTAB; a system
beyond the scratch's limit is refused with SFR_BAD_ARG."""
import ctypes as C

import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems
from sfr3d_amd._lib import lib, SfrStats

pytestmark = pytest.mark.gpu


def noalloc(ctx, A, solver, tol, maxit, db, dx):
    st = SfrStats()
    rc = lib.sfr_solve_device_noalloc(ctx.h, A.h, solver, tol, maxit, db.ptr, dx.ptr, 1,
                                      C.byref(st))
    return rc, st.as_dict()


def relerr(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("solver,make,orc", [
    (S.PCG, lambda n: systems.poisson3d(n, seed=5), oracle.pcg),
    (S.PBICGSTAB, lambda n: systems.momentum3d(n), oracle.pbicgstab),
])
def test_noalloc_matches_device_solve(ctx, solver, make, orc):
    for n in (40, 18):                       # the second call is smaller
        B = make(n)
        print(f"noalloc solver={solver} n={n}", flush=True)
        A = ctx.matrix(B.system)
        db = ctx.array(B.b)
        dx, dy = ctx.empty(B.system.neq), ctx.empty(B.system.neq)
        rc, st = noalloc(ctx, A, solver, 1e-8, 10000, db, dx)
        assert rc == S._lib.OK, lib.sfr_last_error(ctx.h)
        ref = ctx.solve_device(A, solver, 1e-8, 10000, db, dy)
        x, y = dx.get(), dy.get()
        assert np.array_equal(x, y)
        assert st["iters"] == ref["iters"] and st["xKx"] == ref["xKx"]
        assert st["norm_r"] == ref["norm_r"]
        xo, so = orc(B.system, B.b, 1e-8, 10000)
        assert relerr(x, xo) < 1e-6 and abs(st["iters"] - so["iters"]) <= 2
        # a second allocation-free call on the same scratch: same answer
        rc, st2 = noalloc(ctx, A, solver, 1e-8, 10000, db, dx)
        assert rc == S._lib.OK and np.array_equal(dx.get(), x) and st2["iters"] == st["iters"]
        A.destroy(); db.free(); dx.free(); dy.free()


def test_noalloc_full_size_pcg(ctx):
    """The 128^3 Poisson system (2,097,152 rows, within the PCG limit)."""
    B = systems.poisson3d(128, seed=1)
    print("noalloc 128^3 PCG", flush=True)
    A = ctx.matrix(B.system)
    db = ctx.array(B.b)
    dx, dy = ctx.empty(B.system.neq), ctx.empty(B.system.neq)
    rc, st = noalloc(ctx, A, S.PCG, 1e-8, 10000, db, dx)
    assert rc == S._lib.OK, lib.sfr_last_error(ctx.h)
    ref = ctx.solve_device(A, S.PCG, 1e-8, 10000, db, dy)
    assert np.array_equal(dx.get(), dy.get()) and st["iters"] == ref["iters"]
    assert st["relres"] < 2e-8
    A.destroy(); db.free(); dx.free(); dy.free()


def test_noalloc_refuses_beyond_scratch(ctx):
    """BiCGSTAB keeps six vectors: 112^3 = 1,404,928 rows exceed its
    1,354,304-row limit; the matrix is untouched and a smaller solve on the
    same context still works afterwards."""
    B = systems.poisson3d(112, seed=2, reorder=False)
    A = ctx.matrix(B.system)
    db = ctx.array(B.b)
    dx = ctx.empty(B.system.neq)
    rc, _ = noalloc(ctx, A, S.PBICGSTAB, 1e-8, 100, db, dx)
    assert rc == S._lib.BAD_ARG
    assert b"exceed" in lib.sfr_last_error(ctx.h)
    rc, _ = noalloc(ctx, A, S.GMRES, 1e-8, 100, db, dx)
    assert rc == S._lib.NOT_SUPPORTED
    A.destroy(); db.free(); dx.free()
    S2 = systems.poisson3d(10, seed=3)
    A = ctx.matrix(S2.system)
    db, dx = ctx.array(S2.b), ctx.empty(S2.system.neq)
    rc, st = noalloc(ctx, A, S.PCG, 1e-8, 1000, db, dx)
    xo, _ = oracle.pcg(S2.system, S2.b, 1e-8, 1000)
    assert rc == S._lib.OK and relerr(dx.get(), xo) < 1e-6
    A.destroy(); db.free(); dx.free()
