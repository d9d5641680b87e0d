"""GPU parity of the uneven batch (independent systems of different sizes,
patterns and storages solved together by mixed-precision PCG) against the
oracle's FP64 PCG (the reference's pcg restated, src/Cg.c:47-209), system by
system.

Bar: every system stops on the reference PCG's test applied to its true FP64
residual, sqrt|r.D^-1 r| < tol sqrt|b.D^-1 b|, and its solution agrees with
the oracle's to a relative 1e-6 at tol 1e-8.
"""
import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems

pytestmark = pytest.mark.gpu

TOL = 1e-8
XTOL = 1e-6


def relerr(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


def check_against_oracle(built, xs, sts, maxit=20000):
    for k, (B, x, st) in enumerate(zip(built, xs, sts)):
        xo, so = oracle.pcg(B.system, B.b, TOL, maxit)
        assert st["status"] == S._lib.OK, (k, st)
        assert st["norm_r_m"] < st["conv"], (k, st)
        assert relerr(x, xo) < XTOL, (k, relerr(x, xo))
        r = B.b - oracle.matvec(B.system, x)
        assert np.isclose(st["relres"], np.linalg.norm(r) / np.linalg.norm(B.b),
                          rtol=1e-6, atol=1e-15)


def test_uneven_batch_tuned_matches_oracle(ctx):
    built = systems.uneven_batch(14, seed=3, smin=4, smax=26)
    print("uneven tuned sides", [b.extra["side"] for b in built], flush=True)
    Bt = ctx.batch([b.system for b in built], tune=True)
    xs, sts = Bt.solve([b.b for b in built], TOL, 20000)
    check_against_oracle(built, xs, sts)
    # per-system completion times on the device clock
    total = max(st["time_s"] for st in sts)
    assert all(0.0 < st["time_s"] <= total for st in sts)
    ch = Bt.buckets()
    assert sum(c["nsys"] for c in ch) == len(built)
    for c in ch:
        assert c["format"] in (S.FMT_ELL, S.FMT_CSR)
        assert c["rpt"] in (1, 2, 4, 8, 16)
        assert c["ns_per_row"] > 0.0
    for s, b in enumerate(built):
        info = Bt.info(s)
        assert info["neq"] == b.system.neq and info["row0"] % 64 == 0
    Bt.destroy()


def test_uneven_batch_mixed_storages(ctx):
    """Different storages, dimensions, boundary conditions and sizes (down
    to a handful of rows) in one batch, heuristic launch shapes."""
    built = [
        systems.poisson3d(9, seed=21, storage=S.CSRD, reorder=True),
        systems.poisson3d(7, seed=22, storage=S.ELLPACK, pinned_cell=5, reorder=False),
        systems.poisson2d(30, 17, seed=23, storage=S.CSRD, pinned_cell=0),
        systems.poisson3d(11, seed=24, storage=S.CSRC),
        systems.poisson3d(2, seed=25, storage=S.CSRD),
        systems.poisson2d(24, 24, seed=26, storage=S.ELLPACK, pinned_cell=3),
        systems.poisson3d(13, seed=27, storage=S.CSRDCOO, kmin=1e-3),
    ]
    print("uneven mixed storages", [b.system.neq for b in built], flush=True)
    Bt = ctx.batch([b.system for b in built], tune=False)
    xs, sts = Bt.solve([b.b for b in built], TOL, 20000)
    check_against_oracle(built, xs, sts)
    Bt.destroy()


def test_uneven_zero_rhs_warm_start_and_new_values(ctx):
    built = systems.uneven_batch(5, seed=7, smin=5, smax=12)
    sysl = [b.system for b in built]
    bs = [b.b.copy() for b in built]
    bs[2][:] = 0.0                                  # b = 0: x = 0 at once
    Bt = ctx.batch(sysl, tune=False)
    xs, sts = Bt.solve(bs, TOL, 20000)
    assert sts[2]["status"] == S._lib.OK and sts[2]["iters"] == 0
    assert not np.any(xs[2])
    for k in (0, 1, 3, 4):
        xo, _ = oracle.pcg(sysl[k], bs[k], TOL, 20000)
        assert relerr(xs[k], xo) < XTOL
    # warm start from the solution: the first FP64 test (all but) passes
    xw, stw = Bt.solve(bs, TOL, 20000, xs=xs, newX=False)
    for k in range(5):
        assert stw[k]["status"] == S._lib.OK
        assert stw[k]["iters"] <= 5, (k, stw[k])
        assert relerr(xw[k], xs[k]) < 1e-7 or not np.any(xs[k])
    # new values for system 1 (matrix doubled): its solution halves
    s1 = sysl[1]
    Bt.set_values(1, 2.0 * s1.al, 2.0 * s1.ad)
    xn, stn = Bt.solve(bs, TOL, 20000)
    assert stn[1]["status"] == S._lib.OK
    assert relerr(xn[1], 0.5 * xs[1]) < XTOL
    assert relerr(xn[0], xs[0]) < XTOL
    Bt.destroy()


def test_uneven_maxit_and_repeatability(ctx):
    built = systems.uneven_batch(6, seed=9, smin=6, smax=16)
    Bt = ctx.batch([b.system for b in built], tune=False)
    bs = [b.b for b in built]
    xs, sts = Bt.solve(bs, TOL, 3)
    for st in sts:
        assert st["status"] == S._lib.MAXIT
        assert st["iters"] <= 3
    x1, s1 = Bt.solve(bs, TOL, 20000)
    x2, s2 = Bt.solve(bs, TOL, 20000)
    for a, b, p, q in zip(x1, x2, s1, s2):
        assert np.array_equal(a, b)                 # fixed summation order
        assert p["iters"] == q["iters"] and p["outer"] == q["outer"]
    Bt.destroy()


def test_uneven_device_form(ctx):
    built = systems.uneven_batch(4, seed=12, smin=6, smax=14)
    Bt = ctx.batch([b.system for b in built], tune=False)
    hb = np.zeros(Bt.rows)
    for s, b in enumerate(built):
        r0 = Bt.info(s)["row0"]
        hb[r0:r0 + b.system.neq] = b.b
    db, dx = ctx.array(hb), ctx.empty(Bt.rows)
    sts = Bt.solve_device(db, dx, TOL, 20000)
    x = dx.get()
    for s, b in enumerate(built):
        r0 = Bt.info(s)["row0"]
        xo, _ = oracle.pcg(b.system, b.b, TOL, 20000)
        assert sts[s]["status"] == S._lib.OK
        assert relerr(x[r0:r0 + b.system.neq], xo) < XTOL
    Bt.destroy(); db.free(); dx.free()


def test_uneven_batch_escaped_far_columns(ctx):
    """A batch member whose device ELL keeps 16-bit column offsets with
    escaped far entries: SFR3D's ELLPACK points the slot facing the pinned
    cell at column 0 (src/EllPack.c:54-60), and with the reference's
    assembly (ell_quirk) that slot carries the face coefficient.  The batch
    must solve the reference's operator, far column-0 couplings included
    (the coupling is tiny next to the 5 % heavier diagonal, so CG still
    converges; the FP64 BiCGSTAB oracle solves the same system)."""
    from sfr3d_amd import structure as ST
    from sfr3d_amd.solver import System
    B = systems.poisson3d(34, seed=6, storage=S.ELLPACK, reorder=False,
                          pinned_cell=34 ** 3 - 500)
    al, ad, b = ST.assemble(B.pattern, B.id, B.nelcon, B.nviz, B.extra["lA"],
                            B.extra["lB"], ell_quirk=True)
    s0 = B.system
    sq = System(s0.neq, s0.nad, s0.ia, s0.ja, al, 1.05 * ad, S.ELLPACK, True)
    bq = b[:s0.neq].copy()
    small = systems.uneven_batch(2, seed=4, smin=6, smax=10)
    sysl = [sq] + [x.system for x in small]
    bs = [bq] + [x.b for x in small]
    Bt = ctx.batch(sysl, tune=False)
    xs, sts = Bt.solve(bs, TOL, 20000)
    xo, so = oracle.pbicgstab(sq, bq, 1e-10, 20000)
    assert sts[0]["status"] == S._lib.OK
    assert relerr(xs[0], xo) < XTOL
    for k in (1, 2):
        xk, _ = oracle.pcg(sysl[k], bs[k], TOL, 20000)
        assert relerr(xs[k], xk) < XTOL
    Bt.destroy()
