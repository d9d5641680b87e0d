"""Shared-pattern batch (SFR3D's u/v/w momentum solves, reference
src/Simple.c:2980-3036) in one set of batched launches, against sequential
oracle pbicgstab solves of each component."""
import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems
from sfr3d_amd.solver import System

pytestmark = pytest.mark.gpu


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


@pytest.mark.parametrize("storage", [S.CSRD, S.CSRC, S.ELLPACK])
def test_uvw_batch_matches_sequential(ctx, storage):
    M = systems.momentum3d(14, storage=storage)
    bs = list(M.extra["b_all"])
    print(f"uvw storage={storage}", flush=True)
    xs, sts = ctx.solve_shared(M.system, S.PBICGSTAB, 1e-10, 5000, bs)
    for c in range(3):
        xo, so = oracle.pbicgstab(M.system, bs[c], 1e-10, 5000)
        assert sts[c]["status"] == 0
        assert abs(sts[c]["iters"] - so["iters"]) <= 2
        assert relerr(xs[c], xo) < 1e-6


@pytest.mark.parametrize("nsys", [2, 3, 4])
def test_distinct_diagonals(ctx, nsys):
    M = systems.momentum3d(10, storage=S.CSRD)
    rng = np.random.default_rng(nsys)
    ads = [M.system.ad * (1.0 + 0.3 * c) for c in range(nsys)]
    bs = [rng.uniform(-1, 1, M.system.neq) for _ in range(nsys)]
    xs, sts = ctx.solve_shared(M.system, S.PBICGSTAB, 1e-10, 5000, bs, ads=ads)
    for c in range(nsys):
        sc = System(M.system.neq, M.system.nad, M.system.ia, M.system.ja,
                    M.system.al, ads[c], M.system.storage, M.system.unsym)
        xo, so = oracle.pbicgstab(sc, bs[c], 1e-10, 5000)
        assert abs(sts[c]["iters"] - so["iters"]) <= 2, (c, sts[c], so)
        assert relerr(xs[c], xo) < 1e-6


def test_components_converge_independently(ctx):
    # one component with b = 0, one easy, one harder (smaller diagonal)
    M = systems.momentum3d(10, storage=S.CSRD)
    n = M.system.neq
    ads = [M.system.ad, M.system.ad, M.system.ad * 0.999]
    bs = [np.zeros(n), M.extra["b_all"][0], M.extra["b_all"][1] * 1e3]
    xs, sts = ctx.solve_shared(M.system, S.PBICGSTAB, 1e-10, 5000, bs, ads=ads)
    assert sts[0]["iters"] == 0 and not xs[0].any()
    for c in (1, 2):
        sc = System(n, M.system.nad, M.system.ia, M.system.ja, M.system.al,
                    ads[c], M.system.storage, True)
        xo, so = oracle.pbicgstab(sc, bs[c], 1e-10, 5000)
        assert abs(sts[c]["iters"] - so["iters"]) <= 2
        assert relerr(xs[c], xo) < 1e-6
