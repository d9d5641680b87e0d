"""GPU parity on the 2-D lid-driven cavity workload: the pressure-correction
systems of 12 SIMPLE outer iterations on the 128x128 mesh (n = 16,383, one
reference cell pinned), in CSRD and ELLPACK storage, solved by the HIP
Jacobi-PCG and compared with the oracle's PCG, which reproduces the
reference's compiled solverC bit for bit (tests/test_cavity.py).  Also the
SIMPLE loop itself driven by the GPU solve, as SFR3D drives solverC."""
import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems as Y

pytestmark = pytest.mark.gpu

TOL = 1e-8
XTOL = 1e-6


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


@pytest.fixture(scope="module")
def cavity128():
    sysl, _ = Y.cavity2d(n=128, outer=12)
    return sysl


@pytest.mark.parametrize("storage", [S.CSRD, S.ELLPACK])
def test_cavity_pressure_systems_match_oracle(ctx, cavity128, storage):
    for k, rec in enumerate(cavity128):
        B = rec[storage]
        print(f"cavity outer {k} storage {storage}", flush=True)
        x, st = ctx.solve(B.system, S.PCG, TOL, 50000, B.b)
        xo, so = oracle.pcg(B.system, B.b, TOL, 50000)
        assert st["status"] == S._lib.OK
        assert abs(st["iters"] - so["iters"]) <= 2, (k, st["iters"], so["iters"])
        assert relerr(x, xo) < XTOL
        assert abs(st["xKx"] - so["xKx"]) <= 1e-8 * abs(so["xKx"])


def test_cavity_device_resident_value_updates(ctx, cavity128):
    """One device matrix, new values every outer iteration (SFR3D's SIMPLE
    re-assembles on a fixed structure), FP64 and mixed precision."""
    first = cavity128[0][S.CSRD].system
    A = ctx.matrix(first)
    n = first.neq
    db, dx = ctx.empty(n), ctx.empty(n)
    for k, rec in enumerate(cavity128):
        B = rec[S.CSRD]
        A.set_values(B.system.al, B.system.ad)
        db.set(B.b)
        xo, so = oracle.pcg(B.system, B.b, TOL, 50000)
        for prec in (S.PREC_FP64, S.PREC_MIXED):
            st = ctx.solve_device(A, S.PCG, TOL, 50000, db, dx, precision=prec)
            assert st["status"] == S._lib.OK
            assert st["norm_r_m"] < st["conv"]
            assert relerr(dx.get(), xo) < XTOL, (k, prec)
    A.destroy(); db.free(); dx.free()


def test_simple_loop_driven_by_gpu_solve(ctx):
    def gpu_p(sysm, b):
        x, st = ctx.solve(sysm, S.PCG, TOL, 50000, b)
        assert st["status"] == S._lib.OK
        return x

    _, f_gpu = Y.cavity2d(n=32, outer=300, p_solve=gpu_p, record=False)
    _, f_ref = Y.cavity2d(n=32, outer=300, record=False)
    assert f_gpu["mass"][-1] < 1e-6 * f_gpu["mass"][0]
    # same fixed point: the tol-1e-8 pressure solves leave only a small drift
    assert np.abs(f_gpu["u"] - f_ref["u"]).max() < 1e-6
    assert np.abs(f_gpu["p"] - f_ref["p"]).max() < 1e-5 * np.abs(f_ref["p"]).max()
