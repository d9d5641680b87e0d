"""GPU tests of the solverC drop-in (sfr_solverC, the prototype-identical
twin of the reference's solverC, include/Solv.h:59-70) for the two named
solvers, and of storage CSR through sfr_solve.

- The SIMPLE call pattern (src/Simple.c:2983-3136): momentum (BiCGSTAB on
  CSRD unsym and CSRC unsym) and pressure / energy (PCG on CSRC sym,
  ELLPACK, CSRD sym with a pinned cell) in turn, new values every round on
  the same ia/ja.  Each x against the oracle to 1e-6 at tol 1e-10; each log
  line parsed field by field against src/Cg.c:201-208 and
  src/Bicgstab.c:235-242 (tol as printed, iterations within 2 of the
  oracle's, xKx and ||x|| to 1e-8); the pattern cache against an LRU model
  (value-only uploads while a pattern stays cached).
- The maxit exits: PCG's message and EXIT_FAILURE (src/Cg.c:195-200),
  BiCGSTAB's and EXIT_SOLVER = -2 (src/Bicgstab.c:225-231).
- Storage CSR (diagonal inside al) through sfr_solve, symmetric and not,
  against the oracle's solve of the same operator in CSRD storage.
Each drop-in case runs tests/emu/run_dropin.py in a child process, since
the drop-in ends the process on failure as the reference does."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems
from sfr3d_amd._lib import LIB_PATH

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    from tests import dropin_cases as D
    path = str(tmp_path_factory.mktemp("dropin_gpu") / "dropin.npz")
    D.write(path, n=16)
    return path


def _run(cases, mode):
    print(f"drop-in child: {mode}", flush=True)
    return subprocess.run([sys.executable, os.path.join(HERE, "emu", "run_dropin.py"),
                           LIB_PATH, cases, mode], capture_output=True, text=True,
                          timeout=110)


def test_dropin_alternating_patterns(cases):
    res = _run(cases, "alternate")
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr[-3000:]
    assert "DROPIN OK" in res.stdout and "FAIL" not in res.stdout
    assert res.stdout.count("PASS") == 26


@pytest.mark.parametrize("mode,code,out", [
    ("maxit_pcg", 1, "name : sistEqTest\nPCG: *** WARNING: no convergende reached !\n"
                     "MAXIT = 3 \n"),
    ("maxit_bicg", 254, "name : sistEqTest\n PBICGSTAB *** WARNING: no convergende "
                        "reached !\nMAXIT = 3 \n"),
])
def test_dropin_maxit_exit(cases, mode, code, out):
    res = _run(cases, mode)
    assert res.returncode == code, res.stdout + res.stderr[-2000:]
    assert res.stdout == out


def relerr(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("unsym", [False, True])
def test_storage_csr_matches_oracle_in_csrd(ctx, unsym):
    """The operator a CSR system stores (diagonal among the row's entries),
    solved on the device, against the oracle's solve of the same matrix
    assembled in CSRD storage (same cells, same numbering)."""
    n = 14
    if unsym:
        Bc, Bd = systems.momentum3d(n, storage=S.CSR), systems.momentum3d(n, storage=S.CSRD)
        solver, orc = S.PBICGSTAB, oracle.pbicgstab
    else:
        Bc = systems.poisson3d(n, seed=8, storage=S.CSR)
        Bd = systems.poisson3d(n, seed=8, storage=S.CSRD)
        solver, orc = S.PCG, oracle.pcg
    assert np.array_equal(Bc.b, Bd.b)
    print(f"storage CSR unsym={unsym} n={n}", flush=True)
    x, st = ctx.solve(Bc.system, solver, 1e-10, 10000, Bc.b)
    xo, so = orc(Bd.system, Bd.b, 1e-10, 10000)
    assert st["status"] == S._lib.OK
    assert abs(st["iters"] - so["iters"]) <= 2
    assert relerr(x, xo) < 1e-6
    assert abs(st["xKx"] - so["xKx"]) <= 1e-8 * abs(so["xKx"])
    # the stored operator, not the reference's aliased diagonal: the true
    # residual of the CSRD system is at the tolerance
    r = Bd.b - oracle.matvec(Bd.system, x)
    assert np.linalg.norm(r) / np.linalg.norm(Bd.b) < 1e-8
