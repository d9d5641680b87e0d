"""The benchmark's full-size configurations on the GPU, checked against the
oracle where it finishes in seconds and otherwise through properties that do
not depend on size:
  - 128^3 variable-coefficient Poisson, RCM, device ELL (2,097,152 rows):
    FP64 PCG against the oracle's OpenMP PCG (the reference's recurrences
    over full rows): the same iteration count to within 2, solutions to
    1e-9; mixed precision: the reference's stop test met on the true FP64
    residual, its error against a tol-1e-13 solution below a sanity bound
    (it is about 7x the FP64 solve's), bitwise repeatable.
  - the same system in the device CSR format (several lanes per row):
    FP64 PCG against the oracle the same way, bitwise repeatable.
  - 96^3 Taylor-Green momentum, shared CSRD pattern (884,736 rows x 3):
    the batched u/v/w BiCGSTAB bitwise equal to three single solves and
    within 1e-6 of the oracle's serial BiCGSTAB, every stop test met.
  - the same u/v/w batch with each component's own diagonal.
  - the uneven batch of 256 all-Neumann Poisson systems (36.2 M rows),
    tuned, mixed precision: every system meets its own FP64 stop test, 16 of
    them (the 8 smallest and 8 spread up to 80,000 rows) agree with the
    oracle's PCG to 1e-6.
"""
import copy

import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems

pytestmark = pytest.mark.gpu

TOL = 1e-8


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


@pytest.fixture(scope="module")
def poisson128():
    return systems.poisson3d(128, seed=1, kmin=1e-2, reorder=True, storage=S.ELLPACK)


def test_poisson_128cube_fp64_and_mixed(ctx, poisson128):
    B = poisson128
    n = B.system.neq
    print(f"128^3 Poisson: {n} rows", flush=True)
    A = ctx.matrix(B.system)
    db, dx = ctx.array(B.b), ctx.empty(n)
    st64 = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx)
    x64 = dx.get()
    xo, so = oracle.pcg_omp(oracle.full_rows(B.system), B.system.ad[:n], B.b, TOL, 100000)
    print(f"  fp64 iters {st64['iters']} oracle {so['iters']}", flush=True)
    assert st64["status"] == S._lib.OK and so["status"] == 0
    assert abs(int(st64["iters"]) - int(so["iters"])) <= 2
    assert relerr(x64, xo) < 1e-9
    assert st64["norm_r_m"] < st64["conv"]
    stm = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx, precision=S.PREC_MIXED)
    xm = dx.get()
    print(f"  mixed iters {stm['iters']} folds {stm['outer']} "
          f"relerr vs fp64 {relerr(xm, x64):.2e}", flush=True)
    assert stm["status"] == S._lib.OK
    assert stm["norm_r_m"] < stm["conv"]             # the reference's own test, FP64
    # The contract is the reference's residual test, met above on the true
    # FP64 residual.  The error is not fixed by it: with a white-noise right
    # side the test weighs error in the lowest modes about 1/sqrt(n) as much
    # as the rest, and the FP32 iterations leave more of it there (measured:
    # 7e-7 to 9e-7 against FP64 PCG's 1.0e-7, both against a tol-1e-13
    # solve).
    # The bound below is a sanity bound on that error (over 10x above it),
    # not a claim that mixed precision matches FP64 in error.
    st_ex = ctx.solve_device(A, S.PCG, 1e-13, 100000, db, dx)
    xex = dx.get()
    e64, em = relerr(x64, xex), relerr(xm, xex)
    print(f"  error vs tol-1e-13 solution: fp64 {e64:.2e} mixed {em:.2e} "
          f"({st_ex['iters']} iterations)", flush=True)
    assert st_ex["status"] == S._lib.OK
    assert e64 < 1e-6 and em < 1e-5
    stm2 = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx, precision=S.PREC_MIXED)
    assert stm2["iters"] == stm["iters"] and np.array_equal(dx.get(), xm)
    A.destroy(); db.free(); dx.free()


def test_poisson_128cube_fp64_csr_format(ctx, poisson128):
    """The same system with the device CSR format forced (several lanes per
    row) through the FP64 PCG loop's launch shape from 2^21 rows (1 row per
    thread, the SpMV on at most 1,024 workgroups): the oracle's iteration
    count to within 2 and its solution to 1e-9, and bitwise repeatable."""
    B = poisson128
    n = B.system.neq
    A = ctx.matrix(B.system, S.FMT_CSR)
    assert A.info()["format"] == S.FMT_CSR
    db, dx = ctx.array(B.b), ctx.empty(n)
    st = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx)
    x = dx.get()
    xo, so = oracle.pcg_omp(oracle.full_rows(B.system), B.system.ad[:n], B.b, TOL, 100000)
    print(f"128^3 CSR format: iters {st['iters']} oracle {so['iters']} "
          f"rows per thread {A.info()['rows_per_thread_pcg']}", flush=True)
    assert st["status"] == S._lib.OK
    assert abs(int(st["iters"]) - int(so["iters"])) <= 2
    assert relerr(x, xo) < 1e-9
    st2 = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx)
    assert st2["iters"] == st["iters"] and np.array_equal(dx.get(), x)
    A.destroy(); db.free(); dx.free()


def test_momentum_96cube_uvw_batch(ctx):
    M = systems.momentum3d(96)
    n = M.system.neq
    bs = M.extra["b_all"]
    print(f"96^3 momentum: {n} rows x 3", flush=True)
    xs, sts = ctx.solve_shared(M.system, S.PBICGSTAB, TOL, 5000, list(bs))
    for c in range(3):
        x1, s1 = ctx.solve(M.system, S.PBICGSTAB, TOL, 5000, bs[c])
        xo, so = oracle.pbicgstab(M.system, bs[c], TOL, 5000)
        print(f"  comp {c}: batched iters {sts[c]['iters']} single {s1['iters']} "
              f"oracle {so['iters']}", flush=True)
        assert sts[c]["status"] == S._lib.OK
        assert sts[c]["iters"] == s1["iters"] and np.array_equal(xs[c], x1)
        assert abs(int(sts[c]["iters"]) - int(so["iters"])) <= 2
        assert relerr(xs[c], xo) < 1e-6
        assert sts[c]["norm_r"] < sts[c]["conv"] * 1.0001 or sts[c]["relres"] < 1e-7


def test_momentum_96cube_uvw_own_diagonals(ctx):
    """u, v, w with their own diagonals, as SIMPLE builds them (ad + c*n,
    src/Simple.c:2748-2750; here ad (1 + 0.05 c)): the batch is bitwise three
    single solves of the three systems and within 1e-6 of the oracle's."""
    M = systems.momentum3d(96)
    n = M.system.neq
    bs = M.extra["b_all"]
    ad = M.system.ad[:n]
    ads = [np.ascontiguousarray(ad * (1.0 + 0.05 * c)) for c in range(3)]
    print(f"96^3 momentum, own diagonals: {n} rows x 3", flush=True)
    xs, sts = ctx.solve_shared(M.system, S.PBICGSTAB, TOL, 5000, list(bs), ads=ads)
    for c in range(3):
        sc = copy.copy(M.system)
        sc.ad = ads[c]
        x1, s1 = ctx.solve(sc, S.PBICGSTAB, TOL, 5000, bs[c])
        xo, so = oracle.pbicgstab(sc, bs[c], TOL, 5000)
        print(f"  comp {c}: batched iters {sts[c]['iters']} single {s1['iters']} "
              f"oracle {so['iters']}", flush=True)
        assert sts[c]["status"] == S._lib.OK
        assert sts[c]["iters"] == s1["iters"] and np.array_equal(xs[c], x1)
        assert abs(int(sts[c]["iters"]) - int(so["iters"])) <= 2
        assert relerr(xs[c], xo) < 1e-6


def test_uneven_batch_256_systems(ctx):
    """The benchmark's uneven batch: 256 all-Neumann Poisson systems with
    sides log-uniform in [16, 96] (36.2 M rows), tuned, mixed precision.
    Every system stops on the reference PCG's test applied to its true FP64
    residual; the smallest ones are checked against the oracle."""
    built = systems.uneven_batch(256, seed=1)
    print(f"uneven batch: {sum(b.system.neq for b in built)} rows", flush=True)
    Bt = ctx.batch([b.system for b in built], tune=True)
    hb = np.zeros(Bt.rows)
    for k, b in enumerate(built):
        r0 = Bt.info(k)["row0"]
        hb[r0:r0 + b.system.neq] = b.b
    db, dx = ctx.array(hb), ctx.empty(Bt.rows)
    sts = Bt.solve_device(db, dx, TOL, 100000)
    x = dx.get()
    assert all(st["status"] == S._lib.OK for st in sts)
    assert all(st["norm_r_m"] < st["conv"] for st in sts)
    # against the oracle: the 8 smallest and 8 more spread over the sizes the
    # serial oracle finishes in about a second (up to 80,000 rows)
    sizes = np.array([b.system.neq for b in built])
    order = np.argsort(sizes)
    spread = [k for k in order[8::16] if sizes[k] <= 80000][:8]
    for k in list(order[:8]) + spread:
        b = built[k]
        r0 = Bt.info(int(k))["row0"]
        xo, so = oracle.pcg(b.system, b.b, TOL, 100000)
        assert relerr(x[r0:r0 + b.system.neq], xo) < 1e-6, int(k)
    print(f"  oracle checks: {8 + len(spread)} systems, up to "
          f"{max(sizes[k] for k in spread) if spread else 0} rows", flush=True)
    print(f"  iterations {min(s['iters'] for s in sts)}..{max(s['iters'] for s in sts)}",
          flush=True)
    Bt.destroy(); db.free(); dx.free()
