#!/usr/bin/env python3
"""Benchmark of the linear-solve path on one MI355X.

Workload: 3-D variable-coefficient pressure Poisson on a 128^3 hex mesh
(2,097,152 unknowns, 7-point FV stencil, ELLPACK width 7 = 6 slots + the
diagonal), face conductances exp(U[ln 1e-2, 0]), Dirichlet on z = 0,
homogeneous Neumann elsewhere, RHS U[-1, 1], cells RCM-reordered (GENRCM,
bit-exact with the reference's).  One step = one complete Jacobi-PCG solve to
the reference's stopping test at tol 1e-8 (the true FP64 relative residual is
reported too), matrix and vectors resident on the GPU.

value = mixed-precision solves/s (FP32 inner PCG, FP64 residual/update/test);
the plain FP64 GPU PCG is timed the same way and reported beside it.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
TOL = 1e-8
# HBM bytes per launch of k_mx_a on the 128^3 workload (2 rows per thread,
# 1,024 workgroups), from two rocprofv3 counter runs of
# `bench.py --profile-kernels` (FETCH_SIZE x 2.000 and WRITE_SIZE x 1.000,
# the factors a 256 MiB copy gives): 105.5 MB read + 25.2 MB written
K_MX_A_TRAFFIC_128 = 130.7e6


def finite(v):
    """v, or None when it is not finite (the JSON line stays valid JSON)."""
    return v if np.isfinite(v) else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (u/v/w batch, uneven "
                         "batch, adjoint) and the copy ceiling")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the two SpMV kernels and a plain copy "
                         "of known size (for rocprofv3 counter runs)")
    ap.add_argument("--rpt", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="with --profile-kernels: rows per thread of the "
                         "kernels (0: the library's default)")
    ap.add_argument("--nt-loads", type=int, default=-1, choices=[-1, 0, 1],
                    help="non-temporal matrix loads in the PCG loops' SpMV "
                         "kernels: -1 by size (the library's rule), 0 off, 1 on")
    ap.add_argument("--only", choices=["solver_ids", "species", "uvw", "uneven",
                                       "adjoint", "cells", "poisson160", "poisson200"],
                    help="run one of the other workloads alone and print its "
                         "JSON (for rocprofv3 runs of that workload)")
    ap.add_argument("--solvers", default="PBICGSTAB,PBICGSTABL2,GMRES,PCG,MINRES",
                    help="solver ids of the solver_ids workload (comma list)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is "
              "not supported (a multi-GPU solve would split the rows into "
              "blocks with a halo exchange per SpMV and all-reduced dots)",
              file=sys.stderr)
        sys.exit(2)

    import sfr3d_amd as S
    from sfr3d_amd import systems
    from sfr3d_amd._lib import lib

    if args.only:
        ctx = S.Context(0)
        ctx.set_option("nt_loads", args.nt_loads)
        leg = {"solver_ids": lambda: bench_solvers(ctx, S, systems, cpu=not args.no_cpu,
                                                   names=args.solvers.split(",")),
               "species": lambda: bench_species(ctx, S, systems),
               "cells": lambda: bench_cells(ctx, S, systems),
               "uvw": lambda: bench_uvw(ctx, S, systems, lib),
               "uneven": lambda: bench_uneven(ctx, S, systems),
               "adjoint": lambda: bench_adjoint(ctx, S, systems),
               "poisson160": lambda: bench_out_of_mall(ctx, S, systems, lib, 160),
               "poisson200": lambda: bench_out_of_mall(ctx, S, systems, lib, 200)}[args.only]
        print(json.dumps({args.only: leg()}), flush=True)
        ctx.close()
        return

    n = args.n
    t0 = time.time()
    B = systems.poisson3d(n, seed=1, kmin=1e-2, reorder=True, storage=S.ELLPACK)
    sysm = B.system
    neq = sysm.neq
    log(f"[bench] built {n}^3 system: neq={neq} nad={sysm.nad} "
        f"({time.time() - t0:.1f}s)")

    ctx = S.Context(0)
    ctx.set_option("nt_loads", args.nt_loads)
    A = ctx.matrix(sysm)
    info = A.info()
    db = ctx.array(B.b)
    dx = ctx.empty(neq)

    if args.profile_kernels:
        if args.rpt:
            A.tune(args.rpt)
        profile_kernels(ctx, A, lib)
        return

    def run(precision, steps, warmup):
        st = None
        for _ in range(warmup):
            st = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx,
                                  precision=precision)
        lib.sfr_synchronize(ctx.h)
        times = []
        t_start = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            st = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx,
                                  precision=precision)
            times.append(time.perf_counter() - t1)
        lib.sfr_synchronize(ctx.h)
        wall = time.perf_counter() - t_start
        return st, wall / steps, times

    st64, t64, times64 = run(S.PREC_FP64, args.steps, args.warmup)
    x64 = dx.get()
    log(f"[bench] fp64  : {1.0 / t64:8.3f} solves/s  iters={st64['iters']} "
        f"relres={st64['relres']:.3e} status={st64['status']}")
    stmx, tmx, timesmx = run(S.PREC_MIXED, args.steps, args.warmup)
    xmx = dx.get()
    log(f"[bench] mixed : {1.0 / tmx:8.3f} solves/s  inner iters={stmx['iters']} "
        f"outer={stmx['outer']} relres={stmx['relres']:.3e} status={stmx['status']}")

    # dominant kernels: SpMV (+p.q) of each loop, timed alone on the stream
    ms = C.c_double()
    reps = 50
    lib.sfr_kernel_time(ctx.h, A.h, 1, reps, C.byref(ms))
    t_k32 = ms.value * 1e-3
    lib.sfr_kernel_time(ctx.h, A.h, 0, reps, C.byref(ms))
    t_k64 = ms.value * 1e-3
    s = info["slots"]
    ib = info["index_bytes"]
    # value planes the FP64 SpMV streams: half of them when it reads a
    # symmetric ELL through its face pairs (the other half from cache); the
    # FP32 kernel reads every slot
    pairs = A.info()["pair_planes"]
    sv = s // 2 if pairs else s
    # algorithmic bytes per row of one launch (each stream counted once; ib =
    # bytes per stored column index, 2 for 16-bit row offsets):
    #   fused FP32 kernel: s*4 val + s*ib + r_i, p_old_i (4+4) read,
    #                      p_new_i, q_i (4+4) written, y_i (4+4) read+written
    #   FP64 SpMV        : sv*8 val + s*ib + 8 (ad_i) + 8 (p_i) + 8 (q_i)
    b32 = neq * (s * 4 + s * ib + 24)
    b64 = neq * (sv * 8 + s * ib + 24)
    gbs32 = b32 / t_k32 / 1e9
    gbs64 = b64 / t_k64 / 1e9
    log(f"[bench] fused FP32 kernel {t_k32 * 1e6:.1f} us -> {gbs32:.0f} GB/s; "
        f"FP64 SpMV {t_k64 * 1e6:.1f} us -> {gbs64:.0f} GB/s")

    it64 = st64["iters"]
    itmx, outer = stmx["iters"], stmx["outer"]
    solve_bytes64, solve_bytesmx = solve_bytes(neq, s, ib, st64, stmx, sv)
    canon64, canonmx = canonical_bytes(neq, st64, stmx)

    relerr = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))
    # error of both against an FP64 solve to tol 1e-13 (untimed): the stop
    # test fixes the residual, not the error, and the FP32 iterations leave
    # more error in the lowest modes, which that test barely weighs.  Left
    # out of --no-extra runs, the command the kernel-trace profile uses.
    err64 = errmx = float("nan")
    if not args.no_extra:
        ctx.solve_device(A, S.PCG, 1e-13, 100000, db, dx)
        xex = dx.get()
        nex = np.linalg.norm(xex)
        err64 = float(np.linalg.norm(x64 - xex) / nex)
        errmx = float(np.linalg.norm(xmx - xex) / nex)
        log(f"[bench] error vs tol-1e-13 solve: fp64 {err64:.2e} mixed {errmx:.2e}")

    # the FP64 PCG loop again after the library's measured tuning
    # (sfr_matrix_tune with rpt = 0 picks that loop's rows per thread)
    tuned = None
    if not args.no_extra:
        t1 = time.perf_counter()
        A.tune(0)
        tune_s = time.perf_counter() - t1
        st64t, t64t, _ = run(S.PREC_FP64, args.steps, args.warmup)
        x64t = dx.get()
        log(f"[bench] fp64 tuned: {1.0 / t64t:8.3f} solves/s  iters={st64t['iters']} "
            f"relres={st64t['relres']:.3e} (tuning {tune_s:.2f}s)")
        tuned = {"solves_per_s": round(1.0 / t64t, 4), "iters": st64t["iters"],
                 "relres_true": st64t["relres"],
                 "relerr_vs_untuned": float(np.linalg.norm(x64t - x64) /
                                            np.linalg.norm(x64)),
                 "tuning_s": round(tune_s, 3),
                 "rows_per_thread": A.info()["rows_per_thread_pcg"],
                 "rows_per_thread_default": info["rows_per_thread"],
                 "speedup_mixed_vs_fp64_tuned": round(t64t / tmx, 3)}

    # the FP64 solve with every slot's value read (face pairs off): the same
    # bits, the time of the full read
    # (FP64 PCG, the loop that reads them; at the launch shape in force)
    pairs_off = None
    if not args.no_extra and pairs:
        st1, t1_64, _ = run(S.PREC_FP64, 3, 1)
        x_on = dx.get()
        ctx.set_option("ell_pairs", 0)
        st0, t0_64, _ = run(S.PREC_FP64, 3, 1)
        same64 = bool(np.array_equal(dx.get(), x_on) and st0["iters"] == st1["iters"])
        ctx.set_option("ell_pairs", 1)
        pairs_off = {"fp64_solves_per_s_pairs": round(1.0 / t1_64, 4),
                     "fp64_solves_per_s_full_read": round(1.0 / t0_64, 4),
                     "same_bits": same64,
                     "speedup_pairs": round(t0_64 / t1_64, 3)}
        log(f"[bench] FP64 PCG with face pairs {1.0 / t1_64:.3f}, full read "
            f"{1.0 / t0_64:.3f} solves/s (same bits: {same64})")

    bicg_long = None if args.no_extra else bench_bicg_long(ctx, S, A, db, dx, x64)

    copy_gbs = None if args.no_extra else hbm_copy_ceiling(ctx)
    assembly = None if args.no_extra else bench_assembly(ctx, S, A, B, db)

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        Bc = systems.poisson3d(n, seed=1, kmin=1e-2, reorder=True, storage=S.CSRC)
        assert np.array_equal(Bc.b, B.b)
        log(f"[bench] built the same system in CSRC storage ({time.time() - t0:.1f}s)")
        cpu = cpu_baseline(Bc.system, Bc.b, args.cpu_threads, x64)
        del Bc

    extra = {}
    if not args.no_extra:
        db.free(); dx.free(); A.destroy()
        extra["poisson_160cube"] = bench_out_of_mall(ctx, S, systems, lib, 160)
        extra["poisson_200cube"] = bench_out_of_mall(ctx, S, systems, lib, 200)
        extra["momentum_uvw"] = bench_uvw(ctx, S, systems, lib)
        extra["uneven_batch"] = bench_uneven(ctx, S, systems)
        extra["adjoint"] = bench_adjoint(ctx, S, systems)
        extra["solver_ids"] = bench_solvers(ctx, S, systems, cpu=not args.no_cpu)
        extra["species"] = bench_species(ctx, S, systems)
        extra["cells_pres"] = bench_cells(ctx, S, systems)

    out = {
        "metric": "Krylov solves/s + HBM GB/s, 128³ Poisson, mixed vs FP64 @ relres 1e-8",
        "value": round(1.0 / tmx, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tmx * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-inner/fp64-outer",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 variable-coefficient Poisson, 7-point FV, "
                        "ELLPACK width 7, RCM, Jacobi-PCG to tol 1e-8",
            "neq": neq, "device_format": "ELL" if info["format"] == 1 else "CSR",
            "solver": "PCG", "precision": "mixed",
        },
        "mixed": {"solves_per_s": round(1.0 / tmx, 4),
                  "ms_median": round(float(np.median(timesmx)) * 1e3, 3),
                  "inner_iters": itmx, "outer_steps": outer,
                  "relres_true": stmx["relres"],
                  "ref_criterion": [stmx["norm_r_m"], stmx["conv"]],
                  "relerr_vs_fp64_gpu": relerr,
                  "relerr_vs_tol1e-13": finite(errmx),
                  "solve_GBs_algorithmic": round(solve_bytesmx / tmx / 1e9, 1),
                  "solve_frac": round(solve_bytesmx / tmx / 1e9 / HBM_PEAK_GBS, 4),
                  "solve_GBs_canonical": round(canonmx / tmx / 1e9, 1)},
        "fp64": {"solves_per_s": round(1.0 / t64, 4), "iters": it64,
                 "ms_median": round(float(np.median(times64)) * 1e3, 3),
                 "relres_true": st64["relres"],
                 "ref_criterion": [st64["norm_r_m"], st64["conv"]],
                 "relerr_vs_tol1e-13": finite(err64),
                 "solve_GBs_algorithmic": round(solve_bytes64 / t64 / 1e9, 1),
                 "solve_frac": round(solve_bytes64 / t64 / 1e9 / HBM_PEAK_GBS, 4),
                 "solve_GBs_canonical": round(canon64 / t64 / 1e9, 1)},
        "fp64_tuned": tuned,
        "bicgstab_long": bicg_long,
        "ell_pair_planes": pairs,
        "face_pairs_off": pairs_off,
        "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
        "roofline": {"bound": "hbm",
                     "kernel": "k_mx_a (fused FP32 iteration kernel: ELL SpMV "
                               "of r + beta p, direction and correction "
                               "updates, p.q)",
                     "achieved": round(gbs32, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs32 / HBM_PEAK_GBS, 4),
                     "traffic": K_MX_A_TRAFFIC_128 if n == 128 else None,
                     "traffic_unit": "bytes per launch (rocprofv3 PMC)",
                     "launch_us": round(t_k32 * 1e6, 2),
                     "bytes_per_launch": b32, "index_bytes": ib,
                     "fp64_spmv": {"achieved": round(gbs64, 1),
                                   "frac": round(gbs64 / HBM_PEAK_GBS, 4),
                                   "launch_us": round(t_k64 * 1e6, 2),
                                   "bytes_per_launch": b64}},
        "cpu_baseline": cpu,
        "hbm_copy_GBs": copy_gbs[0] if copy_gbs else None,
        "hbm_copy_GBs_hipmemcpy": copy_gbs[1] if copy_gbs else None,
        "device_assembly": assembly,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    if args.no_extra:
        A.destroy()
    ctx.close()


def solve_bytes(neq, s, ib, st64, stmx, sv=None):
    """Algorithmic HBM bytes of one whole PCG solve, FP64 and mixed, on an
    ELL matrix of s slots with ib-byte column offsets, of which the FP64
    SpMV reads sv value planes (s/2 with face-pair reads of a symmetric
    matrix).  Per row: FP64 iteration = SpMV (8 sv + s ib + 24) + residual
    update 32 + direction 48; mixed iteration = fused FP32 kernel
    (4 s + s ib + 24) + 12 (alpha, r -= alpha q, r.r); a fold = 8 + 12
    (update) + FP64 residual (full rows) + 16."""
    sv = s if sv is None else sv
    m64 = 8 * sv + s * ib + 24
    m32 = 4 * s + s * ib + 24
    mfold = s * (8 + ib) + 24
    b64 = neq * (st64["iters"] * (m64 + 80.0) + 16 + 2 * mfold + 32)
    bmx = neq * (stmx["iters"] * (m32 + 12.0)
                 + (stmx["outer"] + 1) * (20.0 + mfold + 16) + 16 + mfold)
    return b64, bmx


def canonical_bytes(neq, st64, stmx):
    """Bytes of one whole PCG solve priced by the reference's own dataflow
    (src/Cg.c:103-152 as three phases, int32 ELL indices, 13 vector words a
    row): 184 B/row per FP64 iteration, 104 per FP32 iteration,
    124 per FP64 refinement step, set-up and final pass one SpMV (80) and
    8 + 6 vector words each.  The kernels here move fewer bytes than that
    (fused phases, 16-bit offsets), so the rate in this accounting is an
    equivalent rate, above the bytes actually streamed."""
    ends = 2 * 80 + 14 * 8
    b64 = neq * (st64["iters"] * 184.0 + ends)
    bmx = neq * (stmx["iters"] * 104.0 + (stmx["outer"] + 1) * 124.0 + ends)
    return b64, bmx


def bench_out_of_mall(ctx, S, systems, lib, n=160, reps=2):
    """The same Poisson workload at n^3 rows beyond the 256 MB last-level
    cache (MALL): at 160^3 (4,096,000 rows) an SpMV launch streams 246 MB
    (FP32, and FP64 with face pairs), at the cache's edge; at 200^3
    (8,000,000 rows) 480 MB, so every iteration streams from HBM.  Whole solves FP64 and mixed, solve-level algorithmic GB/s and
    their fraction of the 8 TB/s peak, and the two SpMV kernels timed alone."""
    t0 = time.time()
    B = systems.poisson3d(n, seed=1, kmin=1e-2, reorder=True, storage=S.ELLPACK)
    neq = B.system.neq
    log(f"[bench] built {n}^3 system ({time.time() - t0:.1f}s)")
    A = ctx.matrix(B.system)
    info = A.info()
    s, ib = info["slots"], info["index_bytes"]
    db, dx = ctx.array(B.b), ctx.empty(neq)
    out = {"workload": f"{n}^3 variable-coefficient Poisson (the 128^3 "
                       "workload's construction), RCM, ELL, Jacobi-PCG to tol 1e-8",
           "neq": neq}
    sts, ts, xs = {}, {}, {}
    for name, prec in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
        ctx.solve_device(A, S.PCG, TOL, 100000, db, dx, precision=prec)
        lib.sfr_synchronize(ctx.h)
        t1 = time.perf_counter()
        for _ in range(reps):
            sts[name] = ctx.solve_device(A, S.PCG, TOL, 100000, db, dx,
                                         precision=prec)
        lib.sfr_synchronize(ctx.h)
        ts[name] = (time.perf_counter() - t1) / reps
        xs[name] = dx.get()
    sv = s // 2 if A.info()["pair_planes"] else s
    b64, bmx = solve_bytes(neq, s, ib, sts["fp64"], sts["mixed"], sv)
    # bytes one SpMV launch streams (FP64 with face pairs when it read them)
    out["spmv_stream_MB"] = {"fp64": round(neq * (8 * sv + s * ib + 24) / 1e6),
                             "fp32": round(neq * (4 * s + s * ib + 24) / 1e6)}
    out["ell_pair_planes"] = A.info()["pair_planes"]
    ms = C.c_double()
    lib.sfr_kernel_time(ctx.h, A.h, 1, 20, C.byref(ms))
    t32 = ms.value * 1e-3
    lib.sfr_kernel_time(ctx.h, A.h, 0, 20, C.byref(ms))
    t64k = ms.value * 1e-3
    k32 = neq * (4 * s + s * ib + 24) / t32 / 1e9
    k64 = neq * (8 * sv + s * ib + 24) / t64k / 1e9
    for name, nb in (("fp64", b64), ("mixed", bmx)):
        st = sts[name]
        gbs = nb / ts[name] / 1e9
        out[name] = {"solves_per_s": round(1.0 / ts[name], 4),
                     "ms": round(ts[name] * 1e3, 2), "iters": st["iters"],
                     "folds": st["outer"], "status": st["status"],
                     "relres_true": finite(st["relres"]),
                     "ref_criterion_met": bool(st["norm_r_m"] < st["conv"]),
                     "solve_GBs_algorithmic": round(gbs, 1),
                     "solve_frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["relerr_mixed_vs_fp64"] = finite(float(
        np.linalg.norm(xs["mixed"] - xs["fp64"]) / np.linalg.norm(xs["fp64"])))
    out["speedup_mixed_vs_fp64"] = round(ts["fp64"] / ts["mixed"], 3)
    out["k_mx_a"] = {"launch_us": round(t32 * 1e6, 2), "GBs": round(k32, 1),
                     "frac": round(k32 / HBM_PEAK_GBS, 4)}
    out["fp64_spmv"] = {"launch_us": round(t64k * 1e6, 2), "GBs": round(k64, 1),
                        "frac": round(k64 / HBM_PEAK_GBS, 4)}
    log(f"[bench] {n}^3: fp64 {out['fp64']['solve_GBs_algorithmic']} GB/s, "
        f"mixed {out['mixed']['solve_GBs_algorithmic']} GB/s, k_mx_a {k32:.0f} GB/s")
    A.destroy(); db.free(); dx.free()
    return out


def bench_bicg_long(ctx, S, A, db, dx, x64, reps=2):
    """BiCGSTAB on the same 128^3 system, FP64 and mixed (FP32 iterations,
    FP64 reliable updates): a long solve, where the FP32 iterations rather
    than the folds set the time."""
    from sfr3d_amd._lib import lib
    out = {"workload": "the 128^3 Poisson system above, Jacobi-BiCGSTAB to "
                       "tol 1e-8 (reference stop test ||r|| < tol ||b||)"}
    for name, prec in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
        st = ctx.solve_device(A, S.PBICGSTAB, TOL, 100000, db, dx, precision=prec)
        lib.sfr_synchronize(ctx.h)
        t1 = time.perf_counter()
        for _ in range(reps):
            st = ctx.solve_device(A, S.PBICGSTAB, TOL, 100000, db, dx,
                                  precision=prec)
        lib.sfr_synchronize(ctx.h)
        ms = (time.perf_counter() - t1) / reps * 1e3
        x = dx.get()
        err = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
        out[name] = {"ms": round(ms, 2), "iters": st["iters"],
                     "folds": st["outer"], "status": st["status"],
                     "relres_true": finite(st["relres"]),
                     "relerr_vs_fp64_pcg": finite(err)}
        log(f"[bench] BiCGSTAB {name}: {ms:.1f} ms, {st['iters']} iterations")
    out["speedup_mixed"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    return out


def profile_kernels(ctx, A, lib):
    """A few launches of each SpMV kernel (FP32 inner, FP64) and one
    device-to-device copy of 256 MiB, so byte counters can be scaled against
    a kernel whose traffic is known (256 MiB read + 256 MiB written)."""
    ms = C.c_double()
    lib.sfr_kernel_time(ctx.h, A.h, 1, 5, C.byref(ms))
    log(f"[profile] k_mx_a {ms.value * 1e3:.1f} us")
    lib.sfr_kernel_time(ctx.h, A.h, 0, 5, C.byref(ms))
    log(f"[profile] k_pcg_spmv {ms.value * 1e3:.1f} us")
    nbytes = 256 << 20
    src = ctx.empty(nbytes // 4, np.float32)
    dst = ctx.empty(nbytes // 4, np.float32)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = hip.hipMemcpy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, 3)
    lib.sfr_synchronize(ctx.h)
    log(f"[profile] copy 256 MiB done (rc={rc})")
    src.free()
    dst.free()


def _hip():
    hip = C.CDLL("libamdhip64.so")
    vp = C.c_void_p
    hip.hipMemcpyAsync.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
    hip.hipEventCreate.argtypes = [C.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    return hip


def hbm_copy_ceiling(ctx, gib=2, reps=5):
    """Device-to-device copy of `gib` GiB (far above the 256 MB last-level
    cache), timed with HIP events, counted as read + write bytes: the
    library's streaming copy kernel (sfr_copy_time), the HBM rate a pure
    stream attains, and hipMemcpyAsync's device copy beside it."""
    from sfr3d_amd._lib import lib
    hip = _hip()
    n = gib << 30
    src = ctx.empty(n // 8)
    dst = ctx.empty(n // 8)
    ms_k = C.c_double()
    rc = lib.sfr_copy_time(ctx.h, src.ptr, dst.ptr, n // 8, reps, C.byref(ms_k))
    assert rc == 0, rc
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemcpyAsync(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), n, 3, None)
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), n, 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    src.free()
    dst.free()
    gbs = 2.0 * n / (ms_k.value * 1e-3) / 1e9
    gbs_mc = 2.0 * n * reps / (ms.value * 1e-3) / 1e9
    log(f"[bench] HBM copy ceiling ({gib} GiB D2D): streaming kernel {gbs:.0f} GB/s, "
        f"hipMemcpyAsync {gbs_mc:.0f} GB/s")
    return round(gbs, 1), round(gbs_mc, 1)


def bench_assembly(ctx, S, A, B, db, reps=5):
    """Re-assembly of the 128^3 matrix and right-hand side from per-cell
    coefficients (the same values it holds): on the host (sfr_assemble) plus
    the upload, against the device scatter from coefficients already on the
    GPU."""
    from sfr3d_amd import structure as ST
    lA = np.ascontiguousarray(B.extra["lA"])
    lB = np.ascontiguousarray(B.extra["lB"])
    th, td = [], []
    for _ in range(reps):
        t = time.perf_counter()
        al, ad, b = ST.assemble(B.pattern, B.id, B.nelcon, B.nviz, lA, lB)
        A.set_values(al, ad)
        db.set(b[:B.system.neq])
        th.append(time.perf_counter() - t)
    plan = ctx.assembly(A, B.pattern, B.id, B.nelcon, B.nviz)
    d_lA, d_lB = ctx.array(lA), ctx.array(lB)
    for _ in range(reps + 1):
        t = time.perf_counter()
        plan.assemble(d_lA, d_lB, db)
        td.append(time.perf_counter() - t)
    plan.destroy(); d_lA.free(); d_lB.free()
    out = {"host_assemble_and_upload_ms": round(float(np.median(th)) * 1e3, 2),
           "device_assemble_ms": round(float(np.median(td[1:])) * 1e3, 3)}
    log(f"[bench] assembly 128^3: host + upload {out['host_assemble_and_upload_ms']} ms, "
        f"device {out['device_assemble_ms']} ms")
    return out


def bench_uvw(ctx, S, systems, lib, reps=3):
    """u/v/w momentum (96^3 Taylor-Green upwind, shared CSRD pattern, the
    components' own diagonals ad (1 + 0.05 c) as the reference's ad + c*n):
    one batched BiCGSTAB launch stream in FP64 and in mixed precision
    (FP32 iterations, FP64 reliable updates) vs three sequential FP64
    device solves."""
    from sfr3d_amd._lib import SfrStats
    from sfr3d_amd.solver import System
    M = systems.momentum3d(96)
    s0 = M.system
    n = s0.neq
    ads = [s0.ad[:n] * (1.0 + 0.05 * c) for c in range(3)]
    mats = [ctx.matrix(System(n, s0.nad, s0.ia, s0.ja, s0.al, ads[c], s0.storage, s0.unsym))
            for c in range(3)]
    bs = M.extra["b_all"]
    db = ctx.array(np.ascontiguousarray(bs.reshape(-1)))
    dad = ctx.array(np.ascontiguousarray(np.concatenate(ads)))
    dx = ctx.empty(3 * n)
    db1 = [ctx.array(bs[c]) for c in range(3)]
    dx1 = ctx.empty(n)
    st = (SfrStats * 3)()

    def batched(prec):
        ctx.set_option("precision", prec)
        rc = lib.sfr_solve_shared_device(ctx.h, mats[0].h, 3, dad.ptr, db.ptr, dx.ptr,
                                         S.PBICGSTAB, TOL, 5000, 1, st)
        ctx.set_option("precision", S.PREC_FP64)
        assert rc == 0, rc

    def sequential():
        out = []
        for c in range(3):
            s1 = ctx.solve_device(mats[c], S.PBICGSTAB, TOL, 5000, db1[c], dx1)
            out.append((s1, dx1.get()))
        return out

    def timed(f, *a):
        f(*a)
        lib.sfr_synchronize(ctx.h)
        t = time.perf_counter()
        for _ in range(reps):
            r = f(*a)
        lib.sfr_synchronize(ctx.h)
        return (time.perf_counter() - t) / reps, r

    tb, _ = timed(batched, S.PREC_FP64)
    xb = dx.get().reshape(3, n)
    its = [st[c].iters for c in range(3)]
    relres = [st[c].relres for c in range(3)]
    tm, _ = timed(batched, S.PREC_MIXED)
    xm = dx.get().reshape(3, n)
    its_m = [st[c].iters for c in range(3)]
    outer_m = [st[c].outer for c in range(3)]
    relres_m = [st[c].relres for c in range(3)]
    ts, seq = timed(sequential)
    rel = max(float(np.linalg.norm(xb[c] - seq[c][1]) / np.linalg.norm(seq[c][1]))
              for c in range(3))
    rel_m = max(float(np.linalg.norm(xm[c] - xb[c]) / np.linalg.norm(xb[c]))
                for c in range(3))
    log(f"[bench] u/v/w 96^3 BiCGSTAB: batched {tb * 1e3:.2f} ms, 3 sequential "
        f"{ts * 1e3:.2f} ms ({ts / tb:.2f}x), iters {its}, max relerr {rel:.1e}; "
        f"mixed batched {tm * 1e3:.2f} ms, iters {its_m} folds {outer_m}, "
        f"relerr vs FP64 {rel_m:.1e}")
    for m in mats:
        m.destroy()
    db.free(); dad.free(); dx.free(); dx1.free()
    for d in db1:
        d.free()
    return {"workload": "96^3 Taylor-Green upwind momentum, u/v/w, shared "
                        "CSRD pattern, own diagonals ad (1 + 0.05 c), "
                        "Jacobi-BiCGSTAB to tol 1e-8",
            "neq": n, "batched_ms": round(tb * 1e3, 3),
            "sequential_ms": round(ts * 1e3, 3),
            "speedup": round(ts / tb, 3), "iters": its,
            "relres": relres,
            "max_relerr_batched_vs_sequential": rel,
            "mixed": {"batched_ms": round(tm * 1e3, 3), "iters": its_m,
                      "folds": outer_m, "relres": relres_m,
                      "max_relerr_vs_fp64_batched": rel_m,
                      "speedup_vs_fp64_batched": round(tb / tm, 3)}}


def bench_uneven(ctx, S, systems, reps=3):
    """256 independent all-Neumann Poisson systems (sides log-uniform in
    [16, 96], conductances exp(U[ln 1e-3, 0]), one pinned cell each), mixed
    PCG to tol 1e-8 in one stream of launches, tuned per size bucket."""
    t0 = time.time()
    built = systems.uneven_batch(256, seed=1)
    t_build = time.time() - t0
    t0 = time.time()
    Bt = ctx.batch([b.system for b in built], tune=True)
    t_tune = time.time() - t0
    hb = np.zeros(Bt.rows)
    for k, b in enumerate(built):
        r0 = Bt.info(k)["row0"]
        hb[r0:r0 + b.system.neq] = b.b
    db, dx = ctx.array(hb), ctx.empty(Bt.rows)
    sts = Bt.solve_device(db, dx, TOL, 100000)
    walls = []
    for _ in range(reps):
        t = time.perf_counter()
        sts = Bt.solve_device(db, dx, TOL, 100000)
        walls.append(time.perf_counter() - t)
    wall = float(np.median(walls))
    n = np.array([b.system.neq for b in built], float)
    it = np.array([s["iters"] for s in sts], float)
    outer = np.array([s["outer"] for s in sts], float)
    ts = np.array([s["time_s"] for s in sts])
    # algorithmic bytes: 88 per row per FP32 iteration, 168 per fold, set-up
    # and final passes 104
    byt = float((n * (it * 88.0 + (outer + 1) * 168.0 + 104.0)).sum())
    ok = sum(1 for s in sts if s["status"] == 0)
    log(f"[bench] uneven 256: {256 / wall:.1f} solves/s ({wall * 1e3:.1f} ms), "
        f"p50 {np.percentile(ts, 50) * 1e3:.1f} ms p99 {np.percentile(ts, 99) * 1e3:.1f} ms, "
        f"{byt / wall / 1e9:.0f} GB/s algorithmic, ok {ok}/256")
    out = {"workload": "256 all-Neumann 3-D Poisson systems, sides log-uniform "
                       "[16, 96], k = exp(U[ln 1e-3, 0]), one pinned cell each, "
                       "mixed PCG to tol 1e-8, one stream of launches",
           "rows": int(n.sum()), "solves_per_s": round(256 / wall, 2),
           "ms_per_batch": round(wall * 1e3, 2),
           "p50_ms": round(float(np.percentile(ts, 50)) * 1e3, 2),
           "p99_ms": round(float(np.percentile(ts, 99)) * 1e3, 2),
           "converged": ok, "max_relres": max(s["relres"] for s in sts),
           "iters_min_med_max": [int(it.min()), int(np.median(it)), int(it.max())],
           "solve_GBs_algorithmic": round(byt / wall / 1e9, 1),
           "buckets": [{k: c[k] for k in ("bucket", "nsys", "format", "rpt")}
                       for c in Bt.buckets()],
           "build_s": round(t_build, 2), "create_and_tune_s": round(t_tune, 2)}
    Bt.destroy(); db.free(); dx.free()
    return out


def bench_adjoint(ctx, S, systems, reps=5):
    """64^3 variable-coefficient diffusion: forward FP64 PCG plus the adjoint
    solve A^T lambda = dL/dx with the value gradients, device times."""
    B = systems.poisson3d(64, seed=11)
    n = B.system.neq
    A = ctx.matrix(B.system)
    db, dx, dg, dl = ctx.array(B.b), ctx.empty(n), ctx.empty(n), ctx.empty(n)
    dgal, dgad = ctx.empty(len(B.system.al)), ctx.empty(n)
    rng = np.random.default_rng(5)
    probes = rng.choice(n, 512, replace=False)
    st = ctx.solve_device(A, S.PCG, TOL, 20000, db, dx)
    g = np.zeros(n)
    g[probes] = 2.0 * dx.get()[probes]
    dg.set(g)
    ctx.solve_adjoint(A, S.PCG, TOL, 20000, dx, dg, dl, dgal, dgad)
    tf, ta = [], []
    for _ in range(reps):
        t = time.perf_counter()
        st = ctx.solve_device(A, S.PCG, TOL, 20000, db, dx)
        tf.append(time.perf_counter() - t)
        t = time.perf_counter()
        sa = ctx.solve_adjoint(A, S.PCG, TOL, 20000, dx, dg, dl, dgal, dgad)
        ta.append(time.perf_counter() - t)
    out = {"workload": "64^3 variable-coefficient diffusion, loss over 512 probe "
                       "cells; forward FP64 PCG + adjoint solve with value "
                       "gradients, tol 1e-8",
           "neq": n, "forward_ms": round(float(np.median(tf)) * 1e3, 3),
           "adjoint_ms": round(float(np.median(ta)) * 1e3, 3),
           "forward_iters": st["iters"], "adjoint_iters": sa["iters"]}
    log(f"[bench] adjoint 64^3: forward {out['forward_ms']} ms, adjoint "
        f"{out['adjoint_ms']} ms")
    A.destroy()
    for d in (db, dx, dg, dl, dgal, dgad):
        d.free()
    return out


def bench_cells(ctx, S, systems, n=128, reps=20):
    """The SIMPLE pressure-correction cell loop on the device at n^3 hex
    cells (warped, graded mesh; every pressure-side load type): one call
    turns the fields into lA / lB / rCell.  Algorithmic bytes per cell, each
    stream once: six face slots of 9 doubles + 2 ints (80 B each), the
    cell's density, pressure, gradient, velocity and d field (88 B) and its
    face count (1 B), lA (7 doubles), lB and rCell (72 B): 641 B."""
    t0 = time.time()
    m = systems.hex_mesh(n, n, n, ratio=(1.01, 0.99, 1.0),
                         shear=(0.05, 0.02, -0.03), warp=0.02)
    case = systems.simple_pres_case(m, seed=1, diffusion=4, boundary="mixed")
    log(f"[bench] cells: {n}^3 mesh and fields built ({time.time() - t0:.1f}s)")
    numel = m["numel"]
    t1 = time.perf_counter()
    cells = ctx.cells(m, case)
    create_s = time.perf_counter() - t1
    d = {k: ctx.array(case[k]) for k in ("pres", "gradPres", "vel", "dField")}
    d_lA, d_lB, d_rc = ctx.empty(numel * 7), ctx.empty(numel), ctx.empty(numel)
    from sfr3d_amd._lib import lib

    def call():
        cells.pres(4, d["pres"], d["gradPres"], d["vel"], d["dField"], d_lA,
                   d_lB, d_rc)
    call()
    lib.sfr_synchronize(ctx.h)
    t1 = time.perf_counter()
    for _ in range(reps):
        call()
    lib.sfr_synchronize(ctx.h)
    ms = (time.perf_counter() - t1) / reps * 1e3
    nbytes = 641 * numel
    out = {"workload": f"{n}^3 warped hex mesh, SIMPLE pressure-correction "
                       "cell loop (systFormSimplePres / cellSimplePres3D), "
                       "over-relaxed diffusion, mixed boundary loads",
           "cells": numel, "ms_per_call": round(ms, 4),
           "bytes_per_cell": 641,
           "GBs_algorithmic": round(nbytes / (ms * 1e-3) / 1e9, 1),
           "create_s": round(create_s, 3)}
    log(f"[bench] cells: {ms:.3f} ms per call, {out['GBs_algorithmic']} GB/s")
    for a in list(d.values()) + [d_lA, d_lB, d_rc]:
        a.free()
    cells.destroy()
    out["momentum"] = bench_cells_vel(ctx, S, systems, m, reps)
    return out


def bench_cells_vel(ctx, S, systems, m, reps):
    """The SIMPLE momentum cell loop (systFormSimpleVel / cellSimpleVel3D) on
    the same mesh: SoUp advection, over-relaxed diffusion, Rhie-Chow,
    second-order face pressure, explicit viscous terms, eddy viscosity,
    BACKWARD time term, lid and walls.  Algorithmic bytes per cell, each
    stream once: six face slots of 15 doubles + 2 ints (768 B); the cell's
    density, viscosity, volume, face count, pressure, pressure gradient,
    velocity, velocity gradient, d field, eddy viscosity and equation
    number (189 B); lA (7 doubles), the three diagonals and right-hand
    sides, rCell and the new d (152 B): 1,109 B."""
    from sfr3d_amd._lib import lib
    numel = m["numel"]
    case = systems.simple_vel_case(m, seed=1)
    t1 = time.perf_counter()
    cells = ctx.cells(m, dict(case, loadsPres=[]))
    create_s = time.perf_counter() - t1
    neq = int(case["id"].max())
    d = {k: ctx.array(case[k]) for k in ("pres", "gradPres", "vel", "gradVel",
                                         "dField", "eddyVisc")}
    d_id = ctx.array(case["id"])
    d_out, d_lA = ctx.empty(3 * numel), ctx.empty(numel * 7)
    d_ad, d_b, d_rc = ctx.empty(3 * neq), ctx.empty(3 * neq), ctx.empty(3 * numel)

    def call():
        cells.vel(case["opts"], d_id, neq, d["pres"], d["gradPres"], d["vel"],
                  d["gradVel"], d["eddyVisc"], None, d["dField"], d_out, d_lA,
                  d_ad, d_b, d_rc)
    call()
    lib.sfr_synchronize(ctx.h)
    t1 = time.perf_counter()
    for _ in range(reps):
        call()
    lib.sfr_synchronize(ctx.h)
    ms = (time.perf_counter() - t1) / reps * 1e3
    out = {"workload": "the same mesh, SIMPLE momentum cell loop (u, v, w), "
                       "SoUp, over-relaxed diffusion, Rhie-Chow, BACKWARD",
           "ms_per_call": round(ms, 4), "bytes_per_cell": 1109,
           "GBs_algorithmic": round(1109 * numel / (ms * 1e-3) / 1e9, 1),
           "create_s": round(create_s, 3)}
    log(f"[bench] momentum cells: {ms:.3f} ms per call, {out['GBs_algorithmic']} GB/s")
    for a in list(d.values()) + [d_id, d_out, d_lA, d_ad, d_b, d_rc]:
        a.free()
    cells.destroy()
    return out


def bench_species(ctx, S, systems, reps=3):
    """Ten transported species on one 64^3 pattern (FOUP, own diffusivity,
    source and old value each), BiCGSTAB to tol 1e-8: sfr_solve_batched
    against the reference's loop of one solve per species (host arrays both,
    uploads included)."""
    from sfr3d_amd.solver import System
    sysm, sp = systems.species3d(64, 10, seed=4)
    als, ads, bs = [a for a, _, _ in sp], [d for _, d, _ in sp], [b for _, _, b in sp]
    singles = [System(sysm.neq, sysm.nad, sysm.ia, sysm.ja, als[c], ads[c],
                      sysm.storage, sysm.unsym) for c in range(10)]
    xb, sb = ctx.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
    tb = []
    for _ in range(reps):
        t = time.perf_counter()
        xb, sb = ctx.solve_batched(sysm, S.PBICGSTAB, TOL, 5000, als, ads, bs)
        tb.append(time.perf_counter() - t)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        xs = [ctx.solve(singles[c], S.PBICGSTAB, TOL, 5000, bs[c])[0] for c in range(10)]
        ts.append(time.perf_counter() - t)
    tbm, tsm = float(np.median(tb)), float(np.median(ts))
    rel = max(float(np.linalg.norm(xb[c] - xs[c]) / np.linalg.norm(xs[c])) for c in range(10))
    out = {"workload": "10 species, 64^3 Taylor-Green FOUP transport, shared CSRD "
                       "pattern, own values, Jacobi-BiCGSTAB to tol 1e-8, host arrays",
           "neq": sysm.neq, "batched_ms": round(tbm * 1e3, 3),
           "sequential_ms": round(tsm * 1e3, 3), "speedup": round(tsm / tbm, 3),
           "iters": [s["iters"] for s in sb], "max_relerr_vs_sequential": rel}
    log(f"[bench] species 10 x 64^3: batched {out['batched_ms']} ms, sequential "
        f"{out['sequential_ms']} ms ({out['speedup']}x), max relerr {rel:.1e}")
    return out


def bench_solvers(ctx, S, systems, cpu=True, reps=3, names=None):
    """Every solver id on the system it is meant for, device times:
    BiCGSTAB, BiCGSTAB(2) and GMRES(15) on the 96^3 upwind momentum system
    (u component), PCG and MINRES on the 96^3 variable-coefficient Poisson
    system, FP64, tol 1e-8.  The CPU leg times the oracle's serial solvers
    on one bounded sample each (one GMRES restart cycle, 20 iterations of
    the others)."""
    out = {}
    M = systems.momentum3d(96)
    P = systems.poisson3d(96, seed=2)
    legs = [("PBICGSTAB", S.PBICGSTAB, M, M.extra["b_all"][0]),
            ("PBICGSTABL2", S.PBICGSTABL2, M, M.extra["b_all"][0]),
            ("GMRES", S.GMRES, M, M.extra["b_all"][0]),
            ("PCG", S.PCG, P, P.b),
            ("MINRES", S.MINRES, P, P.b)]
    if names:
        legs = [leg for leg in legs if leg[0] in names]
    mats = {}
    for name, sid, B, b in legs:
        key = id(B)
        if key not in mats:
            mats[key] = ctx.matrix(B.system)
        A = mats[key]
        n = B.system.neq
        db, dx = ctx.array(b), ctx.empty(n)
        st = ctx.solve_device(A, sid, TOL, 20000, db, dx)
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            st = ctx.solve_device(A, sid, TOL, 20000, db, dx)
            ts.append(time.perf_counter() - t)
        tm = float(np.median(ts))
        leg = {"neq": n, "solve_ms": round(tm * 1e3, 3), "iters": st["iters"],
               "status": st["status"], "relres_true": st["relres"]}
        if sid == S.GMRES:
            leg["cycles"] = st["outer"]
            leg["ms_per_cycle"] = round(tm * 1e3 / max(st["outer"], 1), 3)
        else:
            leg["us_per_iter"] = round(tm * 1e6 / max(st["iters"], 1), 2)
        if cpu:
            leg["cpu_baseline"] = cpu_baseline_solver(name, B.system, b)
        out[name] = leg
        log(f"[bench] {name:12s} {n} rows: {leg['solve_ms']} ms, iters {st['iters']}"
            + (f", cycles {st['outer']}" if sid == S.GMRES else ""))
        db.free(); dx.free()
    for A in mats.values():
        A.destroy()
    out["workload"] = ("96^3 Taylor-Green upwind momentum (u; BiCGSTAB, "
                       "BiCGSTAB(2), GMRES(15)) and 96^3 variable-coefficient "
                       "Poisson (PCG, MINRES), FP64 to tol 1e-8")
    return out


def cpu_baseline_solver(name, sysm, b):
    """The oracle's serial restatement of the reference's solver, one host
    core, on a bounded sample of the same solve: one GMRES(15) restart
    cycle, or 20 iterations of the others."""
    import oracle
    f = {"PBICGSTAB": oracle.pbicgstab, "PBICGSTABL2": oracle.pbicgstabl2,
         "GMRES": oracle.gmres, "PCG": oracle.pcg, "MINRES": oracle.minres}[name]
    k = 1 if name == "GMRES" else 20
    t = time.perf_counter()
    f(sysm, b, 1e-30, k)
    dt = time.perf_counter() - t
    if name == "GMRES":
        return {"ms_per_cycle": round(dt * 1e3, 2), "cores": 1, "kind": "port",
                "sample": "one restart cycle"}
    return {"us_per_iter": round(dt * 1e6 / k, 1), "cores": 1, "kind": "port",
            "sample": f"{k} iterations"}


def host_threads(requested=0):
    """Threads for the CPU baseline: the cores this process may run on
    (its affinity set), capped by OMP_NUM_THREADS where that is set (the
    GPU box sets it to the share of host cores that goes with one GPU)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(env) if env.isdigit() and int(env) > 0 else avail
    return requested or min(avail, cap), avail


def cpu_baseline(sysm_csrc, b, threads, x_gpu64):
    """The reference's own OpenMP PCG (pcgOmp, src/Cg.c:490-715, compiled
    from its sources into oracle/_ref with -fopenmp) through its solverC on
    the host cores, on the same 128^3 system in CSRC storage (the storage
    of the reference's pressure deck, solverPres.dat); the oracle's OpenMP
    restatement beside it on the same threads.  A full solve when it fits in
    about 25 s; otherwise the restatement's sampled rate scaled to the
    iteration count (the reference's solverC ends the process at maxIt, so
    it cannot be sampled)."""
    import oracle
    nthreads, avail = host_threads(threads)
    rows = oracle.full_rows(sysm_csrc)
    ad = sysm_csrc.ad[:sysm_csrc.neq]
    _, st = oracle.pcg_omp(rows, ad, b, TOL, 20, nthreads)
    per_it = st["time_s"] / max(st["iters"], 1)
    full_iters_guess = 1200
    base = {"unit": "solves/s", "cores": nthreads, "host_cpus": os.cpu_count(),
            "affinity_cpus": avail}
    if per_it * full_iters_guess >= 25.0:
        k = max(5, int(25.0 / per_it))
        _, st = oracle.pcg_omp(rows, ad, b, 1e-300, k, nthreads)
        per_it = st["time_s"] / st["iters"]
        t_solve = per_it * full_iters_guess
        log(f"[bench] cpu   : {1.0 / t_solve:8.4f} solves/s (port, {nthreads} threads, sampled)")
        return dict(base, value=round(1.0 / t_solve, 5), kind="port",
                    sample=f"oracle OpenMP PCG on the 128^3 system (CSRC): {st['iters']} "
                           f"iterations timed, scaled to {full_iters_guess}")
    xp, sp = oracle.pcg_omp(rows, ad, b, TOL, 100000, nthreads)
    t_port = sp["time_s"]
    if oracle.ref_omp_available():
        t = time.perf_counter()
        xr, sr = oracle.ref_solve_omp(sysm_csrc, b, TOL, 100000, 1, nthreads)
        t_ref = time.perf_counter() - t
        rel = float(np.linalg.norm(xr - x_gpu64) / np.linalg.norm(xr))
        log(f"[bench] cpu   : {1.0 / t_ref:8.4f} solves/s (reference pcgOmp, {nthreads} "
            f"threads, {sr['iters']} iterations; restatement {t_port:.3f} s)")
        return dict(base, value=round(1.0 / t_ref, 5), kind="reference",
                    sample=f"the reference's solverC -> pcgOmp (compiled from its sources, "
                           f"-O3 -fopenmp) on the same 128^3 system in CSRC storage: one full "
                           f"solve, {sr['iters']} iterations, {t_ref:.3f} s wall",
                    iters=sr["iters"], solve_s=round(t_ref, 4),
                    reference_log_s=sr["time_s"],
                    restatement={"solve_s": round(t_port, 4), "iters": sp["iters"],
                                 "time_ratio_vs_reference": round(t_port / t_ref, 3)},
                    relerr_gpu_fp64_vs_cpu=rel)
    rel = float(np.linalg.norm(xp - x_gpu64) / np.linalg.norm(xp))
    log(f"[bench] cpu   : {1.0 / t_port:8.4f} solves/s (port, {nthreads} threads)")
    return dict(base, value=round(1.0 / t_port, 5), kind="port",
                sample=f"oracle OpenMP PCG on the same 128^3 system (CSRC): one full solve "
                       f"({sp['iters']} iterations)", relerr_gpu_fp64_vs_cpu=rel)


if __name__ == "__main__":
    main()
