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
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the two SpMV kernels and a plain copy "
                         "of known size (for rocprofv3 counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is "
              "not supported (see DESIGN.md for how the solve would shard)",
              file=sys.stderr)
        sys.exit(2)

    import sfr3d_amd as S
    from sfr3d_amd import systems
    from sfr3d_amd._lib import lib

    n = args.n
    t0 = time.time()
    B = systems.poisson3d(n, seed=1, kmin=1e-2, reorder=True, storage=S.ELLPACK)
    sysm = B.system
    neq = sysm.neq
    log(f"[bench] built {n}^3 system: neq={neq} nad={sysm.nad} "
        f"({time.time() - t0:.1f}s)")

    ctx = S.Context(0)
    A = ctx.matrix(sysm)
    info = A.info()
    db = ctx.array(B.b)
    dx = ctx.empty(neq)

    if args.profile_kernels:
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

    st64, t64, _ = run(S.PREC_FP64, args.steps, args.warmup)
    x64 = dx.get()
    log(f"[bench] fp64  : {1.0 / t64:8.3f} solves/s  iters={st64['iters']} "
        f"relres={st64['relres']:.3e} status={st64['status']}")
    stmx, tmx, _ = run(S.PREC_MIXED, args.steps, args.warmup)
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
    # algorithmic bytes per row of one launch (each stream counted once):
    #   FP32 inner SpMV: s*(4 val + 4 col) + 4 (read p_i) + 4 (write q_i)
    #   FP64 SpMV      : s*(8 val + 4 col) + 8 (ad_i) + 8 (p_i) + 8 (q_i)
    b32 = neq * (s * 8 + 8)
    b64 = neq * (s * 12 + 24)
    gbs32 = b32 / t_k32 / 1e9
    gbs64 = b64 / t_k64 / 1e9
    log(f"[bench] FP32 SpMV {t_k32 * 1e6:.1f} us -> {gbs32:.0f} GB/s; "
        f"FP64 SpMV {t_k64 * 1e6:.1f} us -> {gbs64:.0f} GB/s")

    # whole-solve algorithmic bytes (per row): FP64 PCG iteration 176 B
    # (spmv 96 + resid 32 + dir 48); MIXED inner iteration 88 B (spmv 56 +
    # resid 12 + dir 20), outer step 8+12 (update) + 96+4+12 (residual)
    it64 = st64["iters"]
    solve_bytes64 = neq * (it64 * 176.0 + 16 + 96 + 32 + 96)
    itmx, outer = stmx["iters"], stmx["outer"]
    solve_bytesmx = neq * (itmx * 88.0 + (outer + 1) * 132.0 + 16 + 96)

    relerr = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sysm, B.b, args.cpu_threads, x64)

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
                  "inner_iters": itmx, "outer_steps": outer,
                  "relres_true": stmx["relres"],
                  "ref_criterion": [stmx["norm_r_m"], stmx["conv"]],
                  "relerr_vs_fp64_gpu": relerr,
                  "solve_GBs_algorithmic": round(solve_bytesmx / tmx / 1e9, 1)},
        "fp64": {"solves_per_s": round(1.0 / t64, 4), "iters": it64,
                 "relres_true": st64["relres"],
                 "ref_criterion": [st64["norm_r_m"], st64["conv"]],
                 "solve_GBs_algorithmic": round(solve_bytes64 / t64 / 1e9, 1)},
        "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
        "roofline": {"bound": "hbm", "kernel": "k_mx_spmv (FP32 ELL SpMV + p.q)",
                     "achieved": round(gbs32, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs32 / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "launch_us": round(t_k32 * 1e6, 2),
                     "bytes_per_launch": b32,
                     "fp64_spmv": {"achieved": round(gbs64, 1),
                                   "frac": round(gbs64 / HBM_PEAK_GBS, 4),
                                   "launch_us": round(t_k64 * 1e6, 2),
                                   "bytes_per_launch": b64}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    A.destroy()
    ctx.close()


def profile_kernels(ctx, A, lib):
    """A few launches of each SpMV kernel (FP32 inner, FP64) and one
    device-to-device copy of 256 MiB, so byte counters can be scaled against
    a kernel whose traffic is known (256 MiB read + 256 MiB written)."""
    ms = C.c_double()
    lib.sfr_kernel_time(ctx.h, A.h, 1, 5, C.byref(ms))
    log(f"[profile] k_mx_spmv {ms.value * 1e3:.1f} us")
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


def cpu_baseline(sysm, b, threads, x_gpu64):
    """The oracle's PCG (reference recurrences, OpenMP over full rows) on the
    host cores, same system.  Bounded: a full solve if it fits in ~25 s, else
    a measured sample of iterations scaled to the full iteration count."""
    import oracle
    nthreads = threads or min(os.cpu_count() or 1, 16)
    rows = oracle.full_rows(sysm)
    ad = sysm.ad[:sysm.neq]
    # probe the per-iteration time
    _, st = oracle.pcg_omp(rows, ad, b, TOL, 20, nthreads)
    per_it = st["time_s"] / max(st["iters"], 1)
    full_iters_guess = 1200
    if per_it * full_iters_guess < 25.0:
        x, st = oracle.pcg_omp(rows, ad, b, TOL, 100000, nthreads)
        sample = f"one full solve ({st['iters']} iterations)"
        t_solve = st["time_s"]
        iters = st["iters"]
        rel = float(np.linalg.norm(x - x_gpu64) / np.linalg.norm(x))
    else:
        k = max(5, int(25.0 / per_it))
        _, st = oracle.pcg_omp(rows, ad, b, 1e-300, k, nthreads)
        per_it = st["time_s"] / st["iters"]
        iters = full_iters_guess
        t_solve = per_it * iters
        sample = f"{st['iters']} iterations timed, scaled to {iters}"
        rel = None
    log(f"[bench] cpu   : {1.0 / t_solve:8.4f} solves/s ({nthreads} threads, {sample})")
    return {"value": round(1.0 / t_solve, 5), "unit": "solves/s",
            "cores": nthreads, "kind": "port",
            "sample": f"oracle PCG (reference recurrences, OpenMP over full "
                      f"rows, FP64) on the same 128^3 system: {sample}",
            "relerr_gpu_fp64_vs_cpu": rel}


if __name__ == "__main__":
    main()
