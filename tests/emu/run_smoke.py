"""TEST INFRASTRUCTURE -- runs the emulated library (argv[1]) on the small
cases of argv[2] (written by tests/test_emulated_kernels.py with the
oracle's solutions) in a process of its own: FP64 and mixed PCG, BiCGSTAB,
BiCGSTAB(2), GMRES, MINRES, ELL face-pair reads, the species batch,
the shared-pattern batch, both adjoint paths, 16-bit column offsets with
escaped far columns, a tuned uneven batch and re-tuned rows per thread.
Prints PASS/FAIL lines and exits non-zero on any failure."""
import ctypes as C, sys, os, numpy as np
lib = C.CDLL(sys.argv[1])
z = np.load(sys.argv[2])
class Sys(C.Structure):
    _fields_ = [("neq", C.c_int32), ("nad", C.c_int32), ("ia", C.c_void_p), ("ja", C.c_void_p),
                ("al", C.c_void_p), ("ad", C.c_void_p), ("au", C.c_void_p), ("storage", C.c_short), ("unsym", C.c_short)]
class St(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_uint), ("outer", C.c_uint)] + [(k, C.c_double) for k in
               ("tol", "norm_b", "conv", "xKx", "norm_x", "norm_r", "norm_r_m", "relres", "time_s")]
vp = C.c_void_p
lib.sfr_last_error.restype = C.c_char_p
lib.sfr_solve_device.argtypes = [vp, vp, C.c_int, C.c_int, C.c_double, C.c_uint, vp, vp, C.c_int, vp]
lib.sfr_solve_adjoint.argtypes = [vp, vp, C.c_int, C.c_double, C.c_uint, vp, vp, vp, vp, vp, vp]
lib.sfr_solve_shared_device.argtypes = [vp, vp, C.c_int, vp, vp, vp, C.c_int, C.c_double, C.c_uint, C.c_int, vp]
lib.sfr_batch_solve.argtypes = [vp, vp, C.c_double, C.c_uint, vp, vp, C.c_int, vp]
ctx = vp(); assert lib.sfr_init(C.byref(ctx), 0) == 0
keep = []
def mk(name):
    a = [np.ascontiguousarray(z[f"{name}_{k}"]) for k in ("ia", "ja", "al", "ad")]
    keep.append(a)
    neq, nad, st_, un, maxit = [int(v) for v in z[f"{name}_meta"]]
    return Sys(neq, nad, *[x.ctypes.data for x in a], None, st_, un), maxit
def rel(x, y): return np.linalg.norm(x - y) / np.linalg.norm(y)
fails = 0
def check(tag, ok):
    global fails
    print(("PASS " if ok else "FAIL ") + tag, flush=True); fails += (not ok)
# Poisson: FP64, mixed, adjoint
s, maxit = mk("pois"); A = vp(); assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A)) == 0
b = np.ascontiguousarray(z["pois_b"]); xo = z["pois_xo"]
for prec in (0, 1):
    x = np.zeros(s.neq); st = St()
    rc = lib.sfr_solve_device(ctx, A, 1, prec, 1e-8, maxit, b.ctypes.data, x.ctypes.data, 1, C.byref(st))
    check(f"pcg prec {prec} rc {rc} it {st.iters} rel {rel(x, xo):.1e}", rc == 0 and rel(x, xo) < 1e-6)
lam = np.zeros(s.neq); gal = np.zeros(len(z["pois_al"])); gad = np.zeros(s.neq); st = St()
rc = lib.sfr_solve_adjoint(ctx, A, 1, 1e-8, maxit, x.ctypes.data, b.ctypes.data, lam.ctypes.data, gal.ctypes.data, gad.ctypes.data, C.byref(st))
check(f"adjoint pcg rc {rc} rel {rel(lam, xo):.1e}", rc == 0 and rel(lam, xo) < 1e-6)
x = np.zeros(s.neq); st = St(); cyc = z["pois_gm_cyc"]
rc = lib.sfr_solve_device(ctx, A, 4, 0, 1e-8, maxit, b.ctypes.data, x.ctypes.data, 1, C.byref(st))
check(f"gmres pois rc {rc} cycles {st.outer}/{cyc[0]} it {st.iters} rel {rel(x, z['pois_gm_xo']):.1e}",
      rc == 0 and abs(int(st.outer) - int(cyc[0])) <= 1 and rel(x, z["pois_gm_xo"]) < 1e-6)
x = np.zeros(s.neq); st = St(); itm = int(z["pois_mr_it"][0])
rc = lib.sfr_solve_device(ctx, A, 5, 0, 1e-8, maxit, b.ctypes.data, x.ctypes.data, 1, C.byref(st))
check(f"minres pois rc {rc} it {st.iters}/{itm} rel {rel(x, z['pois_mr_xo']):.1e}",
      rc == 0 and abs(int(st.iters) - itm) <= 2 and rel(x, z["pois_mr_xo"]) < 1e-6)
lib.sfr_matrix_destroy(ctx, A)
# ELLPACK in SFR3D's face order (hex, quad rows): the FP64 PCG SpMV reads
# the face pairs once the values are seen to mirror, with the same bits as
# with the pairs off (the mixed loop keeps the full read: the option leaves
# it unchanged), and values that stop mirroring turn them off
lib.sfr_set_option.argtypes = [vp, C.c_char_p, C.c_double]
lib.sfr_matrix_set_values.argtypes = [vp, vp, vp, vp, vp]
lib.sfr_matrix_pair_planes.argtypes = [vp]
for name, kind in (("pE3", 1), ("pE2", 2)):
    s, maxit = mk(name); A = vp(); assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A)) == 0
    b = np.ascontiguousarray(z[f"{name}_b"]); xo = z[f"{name}_xo"]
    for prec in (0, 1):
        res = []
        for pairs in (1, 0):
            assert lib.sfr_set_option(ctx, b"ell_pairs", pairs) == 0
            x = np.zeros(s.neq); st = St()
            rc = lib.sfr_solve_device(ctx, A, 1, prec, 1e-8, maxit, b.ctypes.data, x.ctypes.data, 1, C.byref(st))
            res.append((rc, st.iters, x, lib.sfr_matrix_pair_planes(A)))
        (r1, i1, x1, p1), (r0, i0, x0, p0) = res
        check(f"{name} pairs prec {prec} rc {r1},{r0} it {i1},{i0} planes {p1},{p0} rel {rel(x1, xo):.1e}",
              r1 == 0 and r0 == 0 and (prec == 1 or (p1 == kind and p0 == 0)) and i1 == i0
              and np.array_equal(x1, x0) and rel(x1, xo) < 1e-6)
    assert lib.sfr_set_option(ctx, b"ell_pairs", 1) == 0
    al = np.array(z[f"{name}_al"]); ad = np.ascontiguousarray(z[f"{name}_ad"])
    k = int(np.flatnonzero(al)[len(np.flatnonzero(al)) // 2])
    al[k] *= 1.0 + 2.0 ** -40                 # one value no longer mirrored
    planes = []
    for vals in (al, np.ascontiguousarray(z[f"{name}_al"])):
        assert lib.sfr_matrix_set_values(ctx, A, vals.ctypes.data, ad.ctypes.data, None) == 0
        x = np.zeros(s.neq); st = St()
        rc = lib.sfr_solve_device(ctx, A, 1, 0, 1e-8, maxit, b.ctypes.data, x.ctypes.data, 1, C.byref(st))
        planes.append((rc, lib.sfr_matrix_pair_planes(A), rel(x, xo)))
    check(f"{name} pairs after value changes {planes}",
          planes[0][1] == 0 and planes[1][1] == kind and planes[1][0] == 0 and planes[1][2] < 1e-6)
    lib.sfr_matrix_destroy(ctx, A)
# momentum: BiCGSTAB, shared batch, adjoint (transposed operator)
s, maxit = mk("mom"); A = vp(); assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A)) == 0
bb = np.ascontiguousarray(z["mom_b"]); xo = z["mom_xo"]
x = np.zeros(s.neq); st = St()
rc = lib.sfr_solve_device(ctx, A, 2, 0, 1e-8, maxit, bb[0].ctypes.data, x.ctypes.data, 1, C.byref(st))
check(f"bicg rc {rc} it {st.iters} rel {rel(x, xo):.1e}", rc == 0 and rel(x, xo) < 1e-6)
xb = np.zeros(s.neq); st = St(); itb = int(z["mom_b2_it"][0])
rc = lib.sfr_solve_device(ctx, A, 3, 0, 1e-10, maxit, bb[1].ctypes.data, xb.ctypes.data, 1, C.byref(st))
check(f"bicgl2 mom rc {rc} it {st.iters}/{itb} rel {rel(xb, z['mom_b2_xo']):.1e}",
      rc == 0 and abs(int(st.iters) - itb) <= 1 and rel(xb, z["mom_b2_xo"]) < 1e-8)
xg = np.zeros(s.neq); st = St(); cyc = z["mom_gm_cyc"]
rc = lib.sfr_solve_device(ctx, A, 4, 0, 1e-10, maxit, bb[0].ctypes.data, xg.ctypes.data, 1, C.byref(st))
check(f"gmres mom rc {rc} cycles {st.outer}/{cyc[0]} it {st.iters}/{cyc[1]} rel {rel(xg, z['mom_gm_xo']):.1e}",
      rc == 0 and st.outer == cyc[0] and st.iters == cyc[1] and rel(xg, z["mom_gm_xo"]) < 1e-8)
xs = np.zeros(3 * s.neq); sts = (St * 3)(); flat = np.ascontiguousarray(bb.reshape(-1))
rc = lib.sfr_solve_shared_device(ctx, A, 3, None, flat.ctypes.data, xs.ctypes.data, 2, 1e-8, maxit, 1, C.cast(sts, vp))
check(f"shared rc {rc} rel {rel(xs[:s.neq], xo):.1e}", rc == 0 and rel(xs[:s.neq], x) < 1e-12)
lib.sfr_solve_batched.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_double, C.c_uint, C.c_int, vp]
P3 = vp * 3
spal = [np.ascontiguousarray(z[f"sp{c}_al"]) for c in range(3)]
spad = [np.ascontiguousarray(z[f"sp{c}_ad"]) for c in range(3)]
spb = [np.ascontiguousarray(bb[c]) for c in range(3)]
spx = [np.zeros(s.neq) for c in range(3)]
sts = (St * 3)()
rc = lib.sfr_solve_batched(ctx, C.byref(s), 3, C.cast(P3(*[a.ctypes.data for a in spal]), vp),
                           C.cast(P3(*[a.ctypes.data for a in spad]), vp), None,
                           C.cast(P3(*[a.ctypes.data for a in spb]), vp),
                           C.cast(P3(*[a.ctypes.data for a in spx]), vp), 2, 1e-10, maxit, 1, C.cast(sts, vp))
rs = max(rel(spx[c], z[f"sp{c}_xo"]) for c in range(3))
check(f"species rc {rc} it {[sts[c].iters for c in range(3)]} rel {rs:.1e}", rc == 0 and rs < 1e-8)
lam = np.zeros(s.neq); gal = np.zeros(len(z["mom_al"])); gad = np.zeros(s.neq); st = St()
rc = lib.sfr_solve_adjoint(ctx, A, 2, 1e-8, maxit, x.ctypes.data, bb[0].ctypes.data, lam.ctypes.data, gal.ctypes.data, gad.ctypes.data, C.byref(st))
check(f"adjoint bicg rc {rc} it {st.iters}", rc == 0 and st.iters > 0)
# mixed-precision BiCGSTAB (FP32 inner, FP64 reliable updates): one system,
# the u/v/w batch and the species batch, against the oracle's FP64 solves
lib.sfr_set_option.argtypes = [vp, C.c_char_p, C.c_double]
oracle_uvw = z["mom_xo_uvw"]
x = np.zeros(s.neq); st = St()
rc = lib.sfr_solve_device(ctx, A, 2, 1, 1e-8, maxit, bb[0].ctypes.data, x.ctypes.data, 1, C.byref(st))
check(f"mixed bicg rc {rc} it {st.iters} outer {st.outer} rel {rel(x, xo):.1e}",
      rc == 0 and rel(x, xo) < 1e-6 and st.relres < 1e-8)
assert lib.sfr_set_option(ctx, b"precision", 1.0) == 0
xs = np.zeros(3 * s.neq); sts = (St * 3)()
rc = lib.sfr_solve_shared_device(ctx, A, 3, None, flat.ctypes.data, xs.ctypes.data, 2, 1e-8, maxit, 1, C.cast(sts, vp))
rs = max(rel(xs[c * s.neq:(c + 1) * s.neq], oracle_uvw[c]) for c in range(3))
check(f"mixed shared rc {rc} it {[sts[c].iters for c in range(3)]} rel {rs:.1e}",
      rc == 0 and rs < 1e-6 and max(sts[c].relres for c in range(3)) < 1e-8)
spx = [np.zeros(s.neq) for c in range(3)]
sts = (St * 3)()
rc = lib.sfr_solve_batched(ctx, C.byref(s), 3, C.cast(P3(*[a.ctypes.data for a in spal]), vp),
                           C.cast(P3(*[a.ctypes.data for a in spad]), vp), None,
                           C.cast(P3(*[a.ctypes.data for a in spb]), vp),
                           C.cast(P3(*[a.ctypes.data for a in spx]), vp), 2, 1e-10, maxit, 1, C.cast(sts, vp))
rs = max(rel(spx[c], z[f"sp{c}_xo"]) for c in range(3))
check(f"mixed species rc {rc} it {[sts[c].iters for c in range(3)]} rel {rs:.1e}", rc == 0 and rs < 1e-7)
assert lib.sfr_set_option(ctx, b"precision", 0.0) == 0
lib.sfr_matrix_destroy(ctx, A)
# far columns (escaped 16-bit offsets): 3 BiCGSTAB iterations vs the oracle's 3
s, maxit = mk("far"); A = vp(); assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A)) == 0
lib.sfr_matrix_index_bytes.argtypes = [vp]
b = np.ascontiguousarray(z["far_b"]); xo = z["far_xo"]; x = np.zeros(s.neq); st = St()
rc = lib.sfr_solve_device(ctx, A, 2, 0, 1e-300, maxit, b.ctypes.data, x.ctypes.data, 1, C.byref(st))
check(f"far rc {rc} ib {lib.sfr_matrix_index_bytes(A)} it {st.iters} rel {rel(x, xo):.1e}",
      lib.sfr_matrix_index_bytes(A) == 2 and rel(x, xo) < 1e-12)
lib.sfr_matrix_destroy(ctx, A)
# uneven batch (tuned): two copies of the Poisson system
s1, _ = mk("pois")
arr = (Sys * 2)(s1, s1)
Bt = vp(); rc = lib.sfr_batch_create(ctx, 2, arr, 1, C.byref(Bt))
b = np.ascontiguousarray(z["pois_b"]); P = vp * 2
xs = [np.zeros(s1.neq), np.zeros(s1.neq)]; sts = (St * 2)()
rc2 = lib.sfr_batch_solve(ctx, Bt, 1e-8, 1000, C.cast(P(b.ctypes.data, b.ctypes.data), vp),
                          C.cast(P(xs[0].ctypes.data, xs[1].ctypes.data), vp), 1, C.cast(sts, vp))
check(f"batch rc {rc}/{rc2} rel {rel(xs[1], z['pois_xo']):.1e}", rc == 0 and rc2 == 0 and rel(xs[1], z['pois_xo']) < 1e-6)
lib.sfr_batch_destroy(ctx, Bt)
# rows per thread re-tuned between solves (graphs captured again): every
# solve still matches the oracle
s, maxit = mk("pois"); A = vp(); assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A)) == 0
lib.sfr_matrix_tune.argtypes = [vp, vp, C.c_int, C.c_int]
b = np.ascontiguousarray(z["pois_b"]); xo = z["pois_xo"]
ok, its = True, []
for rpt in (8, 2, 1, 8):
    ok &= lib.sfr_matrix_tune(ctx, A, rpt, 4) == 0
    for solver, prec in ((1, 0), (1, 1)):
        x = np.zeros(s.neq); st = St()
        rc = lib.sfr_solve_device(ctx, A, solver, prec, 1e-8, maxit, b.ctypes.data,
                                  x.ctypes.data, 1, C.byref(st))
        ok &= rc == 0 and rel(x, xo) < 1e-6
        its.append(st.iters)
check(f"retune it {its}", ok)
lib.sfr_matrix_destroy(ctx, A)
# the benchmark's solver-id sequence in one context: BiCGSTAB, BiCGSTAB(2),
# GMRES (its workspace grows and the cached loop graphs are dropped) on the
# momentum matrix, then PCG and MINRES on the Poisson matrix
ok, seq = True, []
sm, mm = mk("mom"); Am = vp(); assert lib.sfr_matrix_create(ctx, C.byref(sm), 0, C.byref(Am)) == 0
bm = np.ascontiguousarray(z["mom_b"])
for solver, tol, xo_ in ((2, 1e-8, z["mom_xo"]), (3, 1e-10, z["mom_b2_xo"]),
                         (4, 1e-10, z["mom_gm_xo"])):
    x = np.zeros(sm.neq); st = St()
    rhs = bm[1] if solver == 3 else bm[0]
    rc = lib.sfr_solve_device(ctx, Am, solver, 0, tol, mm, rhs.ctypes.data, x.ctypes.data,
                              1, C.byref(st))
    ok &= rc == 0 and rel(x, xo_) < 1e-6
    seq.append(int(st.iters))
sp_, mp_ = mk("pois"); Ap = vp(); assert lib.sfr_matrix_create(ctx, C.byref(sp_), 0, C.byref(Ap)) == 0
bp = np.ascontiguousarray(z["pois_b"])
for solver, xo_ in ((1, z["pois_xo"]), (5, z["pois_mr_xo"])):
    x = np.zeros(sp_.neq); st = St()
    rc = lib.sfr_solve_device(ctx, Ap, solver, 0, 1e-8, mp_, bp.ctypes.data, x.ctypes.data,
                              1, C.byref(st))
    ok &= rc == 0 and rel(x, xo_) < 1e-6
    seq.append(int(st.iters))
check(f"solver-id sequence it {seq}", ok)
lib.sfr_matrix_destroy(ctx, Am)
lib.sfr_matrix_destroy(ctx, Ap)
# device assembly vs host assembly + set_values: bitwise the same solve
class Pat(C.Structure):
    _fields_ = [("neq", C.c_int32), ("nad", C.c_int32), ("nia", C.c_int32),
                ("nja", C.c_int64), ("nval", C.c_int64), ("storage", C.c_short),
                ("unsym", C.c_short), ("maxViz", C.c_short),
                ("ia", vp), ("ja", vp)]
s, maxit = mk("asm")
pm = [int(v) for v in z["asm_pat"]]
pia, pja = np.ascontiguousarray(z["asm_pia"]), np.ascontiguousarray(z["asm_pja"])
pat = Pat(pm[0], pm[1], pm[2], pm[3], pm[4], pm[5], pm[6], pm[7], pia.ctypes.data, pja.ctypes.data)
aid = np.ascontiguousarray(z["asm_id"]); anel = np.ascontiguousarray(z["asm_nelcon"])
anv = np.ascontiguousarray(z["asm_nviz"])
lA = np.ascontiguousarray(z["asm_lA"]); lB = np.ascontiguousarray(z["asm_lB"])
A1, A2 = vp(), vp()
assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A1)) == 0
assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A2)) == 0
al2 = np.ascontiguousarray(z["asm_al2"]); ad2 = np.ascontiguousarray(z["asm_ad2"])
lib.sfr_matrix_set_values.argtypes = [vp, vp, vp, vp, vp]
assert lib.sfr_matrix_set_values(ctx, A1, al2.ctypes.data, ad2.ctypes.data, None) == 0
plan = vp()
lib.sfr_assembly_create.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, vp]
rc = lib.sfr_assembly_create(ctx, A2, C.byref(pat), aid.ctypes.data, anel.ctypes.data,
                             anv.ctypes.data, anel.shape[0], 0, C.byref(plan))
db = np.zeros(s.neq)
lib.sfr_assemble_device.argtypes = [vp, vp, vp, vp, vp]
rc2 = lib.sfr_assemble_device(ctx, plan, lA.ctypes.data, lB.ctypes.data, db.ctypes.data)
b2 = np.ascontiguousarray(z["asm_b2"])
x1, x2 = np.zeros(s.neq), np.zeros(s.neq)
st1, st2 = St(), St()
lib.sfr_solve_device(ctx, A1, 1, 0, 1e-8, 1000, b2.ctypes.data, x1.ctypes.data, 1, C.byref(st1))
lib.sfr_solve_device(ctx, A2, 1, 0, 1e-8, 1000, db.ctypes.data, x2.ctypes.data, 1, C.byref(st2))
check(f"assembly rc {rc}/{rc2} b {np.array_equal(db, b2)} it {st1.iters}/{st2.iters}",
      rc == 0 and rc2 == 0 and np.array_equal(db, b2) and np.array_equal(x1, x2))
lib.sfr_assembly_destroy(ctx, plan)
lib.sfr_matrix_destroy(ctx, A1); lib.sfr_matrix_destroy(ctx, A2)
# allocation-free solve (static scratch, no graphs): bitwise the graph path,
# twice in a row, the second on a smaller system
lib.sfr_solve_device_noalloc.argtypes = [vp, vp, C.c_int, C.c_double, C.c_uint, vp, vp, C.c_int, vp]
for name, solver in (("mom", 2), ("pois", 1)):
    s, maxit = mk(name); A = vp(); assert lib.sfr_matrix_create(ctx, C.byref(s), 0, C.byref(A)) == 0
    b = np.ascontiguousarray(z[f"{name}_b"][0] if name == "mom" else z[f"{name}_b"])
    x1, x2 = np.zeros(s.neq), np.zeros(s.neq); st1, st2 = St(), St()
    rc1 = lib.sfr_solve_device_noalloc(ctx, A, solver, 1e-8, maxit, b.ctypes.data, x1.ctypes.data, 1, C.byref(st1))
    rc2 = lib.sfr_solve_device(ctx, A, solver, 0, 1e-8, maxit, b.ctypes.data, x2.ctypes.data, 1, C.byref(st2))
    check(f"noalloc {name} rc {rc1}/{rc2} it {st1.iters}/{st2.iters}",
          rc1 == 0 and rc2 == 0 and np.array_equal(x1, x2) and st1.iters == st2.iters
          and st1.xKx == st2.xKx)
    lib.sfr_matrix_destroy(ctx, A)
lib.sfr_finalize(ctx)
print("SMOKE", "FAILED" if fails else "OK", flush=True)
os._exit(1 if fails else 0)
