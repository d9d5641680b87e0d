"""ORACLE -- test infrastructure only.

CPU restatement of the reference's linear-solve path (sfr_oracle.c: pcg,
pbicgstab, pbicgstabl2, gmres, pminres, the serial matvecs, dot, preCondDiag) and of its index builders
(patterns.py: convGraph, GENRCM, numEqV1/V2, csrIa/csrJa, ellPackJa, the csr/
ellPack value scatter), plus, where the reference tree exists, a harness
around the reference's own compiled C (_ref/libsfr_ref.so, built by
oracle/Makefile).

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker / the timed
baseline.  The product library (sfr3d_amd) never loads it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libsfr_ref.so")
REF_OMP_LIB = os.path.join(HERE, "_ref", "libsfr_ref_omp.so")

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_uint), ("xKx", C.c_double),
                ("norm_x", C.c_double), ("norm_r", C.c_double),
                ("norm_r_m", C.c_double), ("norm_b", C.c_double),
                ("conv", C.c_double), ("time_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_orc = None
_ref = None


def orc():
    """liboracle.so (built by `make -C oracle`)."""
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        lib = C.CDLL(ORACLE_LIB)
        for name in ("orc_pcg", "orc_pbicgstab", "orc_minres", "orc_pbicgstabl2"):
            f = getattr(lib, name)
            f.restype = C.c_int
            f.argtypes = [C.c_int, C.c_int, i32p, i32p, f64p, f64p, f64p, f64p,
                          C.c_double, C.c_uint, C.c_int, C.c_int, C.c_int,
                          C.POINTER(OrcStats)]
        lib.orc_gmres.restype = C.c_int
        lib.orc_gmres.argtypes = [C.c_int, C.c_int, i32p, i32p, f64p, f64p,
                                  f64p, f64p, C.c_double, C.c_uint, C.c_int,
                                  C.c_int, C.c_int, C.c_int,
                                  C.POINTER(OrcStats), C.POINTER(C.c_uint)]
        lib.orc_matvec.restype = C.c_int
        lib.orc_matvec.argtypes = [C.c_int, C.c_int, C.c_int, i32p, i32p, f64p,
                                   f64p, f64p, f64p]
        lib.orc_dot.restype = C.c_double
        lib.orc_dot.argtypes = [f64p, f64p, C.c_int]
        lib.orc_full_rows.restype = C.c_int
        lib.orc_full_rows.argtypes = [C.c_int, C.c_int, C.c_int, i32p, i32p,
                                      f64p, i64p, i32p, f64p]
        lib.orc_pcg_omp.restype = C.c_int
        lib.orc_pcg_omp.argtypes = [C.c_int, i64p, i32p, f64p, f64p, f64p,
                                    f64p, C.c_double, C.c_uint, C.c_int,
                                    C.POINTER(OrcStats)]
        lib.orc_max_threads.restype = C.c_int
        _orc = lib
    return _orc


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own compiled C (only where /root/reference exists)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_LIB} missing: run `make -C oracle ref`")
        lib = C.CDLL(REF_LIB)
        lib.ref_init.argtypes = [C.c_long]
        lib.ref_reord.argtypes = [i32p, i32p, i16p, C.c_short, C.c_int, C.c_int]
        lib.ref_genrcm.argtypes = [C.c_int, i32p, i32p, i32p]
        lib.ref_numeq_v1.argtypes = [i32p, i32p, C.c_int]
        lib.ref_numeq_v2.argtypes = [i32p, i32p, i16p, i16p, C.c_int, C.c_short]
        lib.ref_datastruct.argtypes = [C.c_short, C.c_int, i32p, i32p, i32p,
                                       i16p, C.c_int, C.c_short, C.c_int,
                                       np.ctypeslib.ndpointer(dtype=np.int64)]
        lib.ref_datastruct_copy.argtypes = [i32p, i32p, C.c_long, C.c_long]
        lib.ref_assemble.argtypes = [i32p, i32p, i16p, C.c_int, C.c_short,
                                     f64p, f64p, f64p, f64p, f64p, C.c_long,
                                     C.c_int]
        lib.ref_solve.argtypes = [C.c_int, C.c_int, i32p, i32p, f64p, f64p,
                                  C.c_void_p, f64p, f64p, C.c_double, C.c_uint,
                                  C.c_short, C.c_short, C.c_int, C.c_int, f64p,
                                  C.c_char_p, C.c_int]
        lib.ref_matvec.argtypes = [C.c_short, C.c_int, C.c_int, i32p, i32p,
                                   f64p, f64p, f64p, f64p]
        _ref = lib
    return _ref


_ref_omp = None


def ref_omp_available():
    return os.path.exists(REF_OMP_LIB)


def ref_omp():
    """The reference's own C built with -fopenmp (its OpenMP solvers)."""
    global _ref_omp
    if _ref_omp is None:
        if not ref_omp_available():
            raise ImportError(f"{REF_OMP_LIB} missing: run `make -C oracle ref`")
        lib = C.CDLL(REF_OMP_LIB)
        lib.ref_init.argtypes = [C.c_long]
        lib.ref_solve_omp.argtypes = [C.c_int, C.c_int, i32p, i32p, f64p, f64p,
                                      f64p, f64p, C.c_double, C.c_uint,
                                      C.c_short, C.c_short, C.c_int, C.c_int,
                                      f64p, C.c_char_p, C.c_int]
        lib.ref_init(1 << 30)
        _ref_omp = lib
    return _ref_omp


def ref_solve_omp(system, b, tol, maxit, solver, nthreads):
    """The reference's solverC with its OpenMP solvers (pcgOmp for solver 1,
    pbicgstabOmp for 2) on nthreads threads.  Returns (x, stats) with the
    iterations, xKx, ||x|| and time the reference logs."""
    lib = ref_omp()
    x = np.zeros(system.neq)
    out = np.zeros(4)
    buf = C.create_string_buffer(512)
    ad = system.ad if system.ad is not None else system.al
    rc = lib.ref_solve_omp(system.neq, system.nad, system.ia, system.ja,
                           _al_contig(system), np.ascontiguousarray(ad, np.float64),
                           np.ascontiguousarray(b, np.float64), x, tol, maxit,
                           system.storage, solver, 1 if system.unsym else 0,
                           nthreads, out, buf, 512)
    if rc:
        raise RuntimeError(f"ref_solve_omp = {rc}")
    return x, {"iters": int(out[0]), "xKx": out[1], "norm_x": out[2],
               "time_s": out[3], "line": buf.value.decode()}


# ---- convenience wrappers ---------------------------------------------------

def pcg(system, b, tol, maxit, x=None, newX=True):
    """Reference-order serial PCG (oracle).  Returns (x, stats)."""
    return _krylov("orc_pcg", system, b, tol, maxit, x, newX)


def pbicgstab(system, b, tol, maxit, x=None, newX=True):
    return _krylov("orc_pbicgstab", system, b, tol, maxit, x, newX)


def pbicgstabl2(system, b, tol, maxit, x=None, newX=True):
    """Reference-order serial BiCGSTAB(2) (oracle); stats["norm_r_m"] is
    the recurrence residual norm at the end."""
    return _krylov("orc_pbicgstabl2", system, b, tol, maxit, x, newX)


def minres(system, b, tol, maxit, x=None, newX=True):
    """Reference-order serial Jacobi-MINRES (pminres) (oracle)."""
    return _krylov("orc_minres", system, b, tol, maxit, x, newX)


def gmres(system, b, tol, ncycles, x=None, newX=True, nkrylov=15):
    """Reference-order serial GMRES(nkrylov) (oracle).  Returns (x, stats);
    stats["cycles"] is the number of restart cycles run."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(system.neq) if x is None else np.array(x, np.float64)
    st = OrcStats()
    cyc = C.c_uint(0)
    ad = system.ad if system.ad is not None else system.al
    rc = orc().orc_gmres(system.neq, system.nad, system.ia, system.ja,
                         _al_contig(system), np.ascontiguousarray(ad, np.float64),
                         b, x, tol, ncycles, 1 if newX else 0, system.storage,
                         1 if system.unsym else 0, nkrylov, C.byref(st),
                         C.byref(cyc))
    if rc:
        raise RuntimeError(f"orc_gmres = {rc}")
    d = st.as_dict()
    d["cycles"] = cyc.value
    return x, d


def _al_contig(system):
    al = system.al
    if system.au is not None:
        al = np.concatenate([system.al[:system.nad], system.au[:system.nad]])
    return np.ascontiguousarray(al, np.float64)


def _krylov(name, system, b, tol, maxit, x, newX):
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(system.neq) if x is None else np.array(x, np.float64)
    st = OrcStats()
    ad = system.ad if system.ad is not None else system.al
    rc = getattr(orc(), name)(system.neq, system.nad, system.ia, system.ja,
                              _al_contig(system),
                              np.ascontiguousarray(ad, np.float64), b, x, tol,
                              maxit, 1 if newX else 0, system.storage,
                              1 if system.unsym else 0, C.byref(st))
    if rc:
        raise RuntimeError(f"{name} = {rc}")
    return x, st.as_dict()


def matvec(system, x):
    y = np.zeros(system.neq)
    ad = system.ad if system.ad is not None else system.al
    orc().orc_matvec(system.storage, 1 if system.unsym else 0, system.neq,
                     system.ia, system.ja, _al_contig(system),
                     np.ascontiguousarray(ad, np.float64),
                     np.ascontiguousarray(x, np.float64), y)
    return y


def full_rows(system):
    """Full-row CSR (rp, cj, cv) of the off-diagonal part, for the OpenMP
    baseline."""
    n = system.neq
    if system.storage == 4:
        cnt = n * system.ia[0]
    else:
        cnt = int(system.ia[n]) * (1 if system.unsym and system.storage in (1, 2, 5) else 2)
    rp = np.zeros(n + 1, np.int64)
    cj = np.zeros(max(cnt, 1), np.int32)
    cv = np.zeros(max(cnt, 1), np.float64)
    rc = orc().orc_full_rows(system.storage, 1 if system.unsym else 0, n,
                             system.ia, system.ja, _al_contig(system), rp, cj,
                             cv)
    if rc:
        raise RuntimeError("orc_full_rows")
    return rp, cj[:rp[n]].copy(), cv[:rp[n]].copy()


def pcg_omp(rows, ad, b, tol, maxit, nthreads=0):
    """OpenMP PCG over full rows (rows from full_rows); n is the row count of
    rows, and ad and b may be longer (ELLPACK keeps ad[n+1])."""
    rp, cj, cv = rows
    n = len(rp) - 1
    if len(ad) < n or len(b) < n:
        raise ValueError("ad and b need one entry per row")
    x = np.zeros(n)
    st = OrcStats()
    b = np.ascontiguousarray(b[:n], np.float64)
    rc = orc().orc_pcg_omp(n, rp, cj, cv,
                           np.ascontiguousarray(ad[:n], np.float64),
                           np.ascontiguousarray(b, np.float64), x, tol, maxit,
                           nthreads, C.byref(st))
    if rc:
        raise RuntimeError("orc_pcg_omp")
    return x, st.as_dict()


# ---- the SIMPLE pressure-correction cell loop (cells.c) ---------------------

def load_table(loads):
    """Flat arrays (type, np, par[n*8], density) of a list of constant loads,
    each a dict with "type", "par" (at most 8 values) and "density"."""
    n = max(len(loads), 1)
    t = np.zeros(n, np.int16)
    npar = np.zeros(n, np.int16)
    par = np.zeros(n * 8)
    dens = np.zeros(n)
    for i, ld in enumerate(loads):
        t[i] = ld["type"]
        p = list(ld.get("par", []))
        npar[i] = len(p)
        par[i * 8:i * 8 + len(p)] = p
        dens[i] = ld.get("density", 0.0)
    return t, npar, par, dens


def _cells_lib():
    lib = orc()
    if not getattr(lib, "_cells_ready", False):
        f = lib.orc_simple_pres3d
        f.restype = C.c_int
        f.argtypes = ([C.c_int, C.c_short, i16p, i32p, i32p, i32p, f64p, f64p,
                       f64p, f64p, f64p, f64p, f64p, i16p, f64p, i16p, i16p]
                      + [C.c_int, i16p, i16p, f64p, f64p] * 2
                      + [C.c_short, f64p, f64p, f64p, f64p, f64p, f64p, f64p])
        lib._cells_ready = True
    return lib


def simple_pres3d(mesh, case):
    """lA (numel x (maxViz+1), the centre last), lB and rCell of the SIMPLE
    pressure-correction cell loop (systFormSimplePres + cellSimplePres3D) for
    a mesh from sfr3d_amd.systems.hex_mesh and a case (fields, materials,
    boundary markers, loads, diffusion scheme).  Cells without an equation
    keep lA = lB = 0."""
    numel, mv = mesh["numel"], mesh["maxViz"]
    lA = np.zeros(numel * (mv + 1))
    lB = np.zeros(numel)
    rc = np.zeros(numel)
    lv = load_table(case["loadsVel"])
    lp = load_table(case["loadsPres"])
    r = _cells_lib().orc_simple_pres3d(
        numel, mv, mesh["nFace"], mesh["nelcon"], mesh["cellFace"],
        mesh["owner"], mesh["dcca"], mesh["xmcc"], mesh["mksi"], mesh["ksi"],
        mesh["area"], mesh["normal"], mesh["vSkew"], case["mat"],
        case["prop"], case["faceRvel"], case["faceRpres"],
        len(case["loadsVel"]), *lv, len(case["loadsPres"]), *lp,
        case["diffusion"], case["pres"], case["gradPres"], case["vel"],
        case["dField"], lA, lB, rc)
    if r:
        raise ValueError("orc_simple_pres3d: diffusion scheme must be 1 or 4")
    return lA.reshape(numel, mv + 1), lB, rc


def ref_simple_pres(mesh, case, storage=2):
    """The reference's own systFormSimplePres (src/CellLoopInc.c:471) on the
    same mesh and case, assembled by its dataStruct / assbly in `storage`
    (symmetric): (ia, ja, al, ad, b, rCell)."""
    lib = ref()
    if not getattr(lib, "_pres_ready", False):
        f = lib.ref_simple_pres
        f.restype = C.c_int
        f.argtypes = ([C.c_int, C.c_short, i16p, i32p, i32p, i32p, f64p, f64p,
                       f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                       i16p, f64p, C.c_int, i16p, i16p, i32p]
                      + [C.c_int, i16p, i16p, f64p, f64p] * 2
                      + [C.c_short, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                         f64p, C.c_long])
        lib._pres_ready = True
    numel, mv = mesh["numel"], mesh["maxViz"]
    id_ = np.ascontiguousarray(case["id"], np.int32)
    num = np.ascontiguousarray(case["num"], np.int32)
    neq = int(id_.max())
    out = np.zeros(3, np.int64)
    lib.ref_init(1 << 31)          # the arena size the other harness users set
    lib.ref_datastruct(storage, 0, id_, num, mesh["nelcon"], mesh["nFace"],
                       numel, mv, neq, out)
    ia = np.zeros(out[0], np.int32)
    ja = np.zeros(out[1], np.int32)
    lib.ref_datastruct_copy(ia, ja, out[0], out[1])
    al = np.zeros(out[2])
    nrow = neq + 1 if storage == 4 else neq
    ad = np.zeros(nrow)
    b = np.zeros(nrow)
    rc = np.zeros(numel)
    lv = load_table(case["loadsVel"])
    lp = load_table(case["loadsPres"])
    nmat = len(case["prop"]) // 10
    r = lib.ref_simple_pres(
        numel, mv, mesh["nFace"], mesh["nelcon"], mesh["cellFace"],
        mesh["owner"], mesh["volume"], mesh["dcca"], mesh["xmcc"],
        mesh["mksi"], mesh["ksi"], mesh["eta"], mesh["area"], mesh["normal"],
        mesh["xm"], mesh["mvSkew"], mesh["vSkew"], case["mat"], case["prop"],
        nmat, case["faceRvel"], case["faceRpres"], id_,
        len(case["loadsVel"]), *lv, len(case["loadsPres"]), *lp,
        case["diffusion"], case["pres"], case["gradPres"], case["vel"],
        case["dField"], al, ad, b, rc, int(out[2]))
    if r:
        raise RuntimeError("ref_simple_pres")
    return ia, ja, al, ad, b, rc


# ---- the SIMPLE momentum cell loop (cells.c) --------------------------------

VEL_OPTS = ("advection", "diffusion", "turb", "wall_model", "rhie",
            "so_pressure", "viscosity", "residual", "simplec", "time",
            "backward")


def vel_options(case):
    """(opt int32[11], optd f64[3]) of a momentum case."""
    o = case["opts"]
    opt = np.array([int(o[k]) for k in VEL_OPTS], np.int32)
    optd = np.array([o["dt"], o["dt0"], o["underU"]], np.float64)
    return opt, optd


def simple_vel3d(mesh, case, in_place=False):
    """lA (numel x (maxViz+3): faces, then the u, v, w diagonals), lB
    (numel x 3), rCell (3 x numel) and the new d field (numel x 3) of the
    SIMPLE momentum cell loop (systFormSimpleVel + cellSimpleVel3D).
    in_place: cells read the updated d of lower-numbered neighbours, as the
    reference's serial loop does; otherwise every neighbour's input d."""
    lib = orc()
    if not getattr(lib, "_vel_ready", False):
        f = lib.orc_simple_vel3d
        f.restype = C.c_int
        f.argtypes = ([C.c_int, C.c_short, i16p, i32p, i32p, i32p] + [f64p] * 10
                      + [i16p, f64p, i16p, i32p, C.c_int, i16p, i16p, f64p, f64p,
                         i32p, f64p] + [f64p] * 8 + [C.c_int, f64p, f64p, f64p])
        lib._vel_ready = True
    numel, mv = mesh["numel"], mesh["maxViz"]
    lA = np.zeros(numel * (mv + 3))
    lB = np.zeros(numel * 3)
    rc = np.zeros(3 * numel)
    d_in = np.ascontiguousarray(case["dField"], np.float64)
    d_out = d_in.copy() if in_place else np.zeros(3 * numel)
    opt, optd = vel_options(case)
    lv = load_table(case["loadsVel"])
    eddy = np.ascontiguousarray(case.get("eddyVisc", np.zeros(numel)), np.float64)
    wall = np.ascontiguousarray(case.get("wallPar", np.zeros(4 * numel)), np.float64)
    r = lib.orc_simple_vel3d(
        numel, mv, mesh["nFace"], mesh["nelcon"], mesh["cellFace"], mesh["owner"],
        mesh["volume"], mesh["dcca"], mesh["xmcc"], mesh["cc"], mesh["mksi"],
        mesh["ksi"], mesh["area"], mesh["normal"], mesh["xm"], mesh["vSkew"],
        case["mat"], case["prop"], case["faceRvel"],
        np.ascontiguousarray(case["id"], np.int32), len(case["loadsVel"]), *lv,
        opt, optd, case["pres"], case["gradPres"], case["vel"], case["gradVel"],
        eddy, wall, d_in, d_out, 1 if in_place else 0, lA, lB, rc)
    if r:
        raise ValueError(f"orc_simple_vel3d: {r} (advection, diffusion or load type)")
    return lA.reshape(numel, mv + 3), lB.reshape(numel, 3), rc.reshape(3, numel), \
        d_out.reshape(numel, 3)


def ref_simple_vel(mesh, case):
    """The reference's own systFormSimpleVel (src/CellLoopInc.c:98) on the
    same mesh and case, in CSRD unsym: (ia, ja, al, ad[3*neq], b[3*neq],
    rCell[3*numel], dField updated in place)."""
    lib = ref()
    if not getattr(lib, "_vel_ready", False):
        f = lib.ref_simple_vel
        f.restype = C.c_int
        f.argtypes = ([C.c_int, C.c_short, i16p, i32p, i32p, i32p] + [f64p] * 12
                      + [i16p, f64p, C.c_int, i16p, i16p, i32p, C.c_int, i16p,
                         i16p, f64p, f64p, i32p, f64p] + [f64p] * 7
                      + [f64p, f64p, f64p, f64p, C.c_long])
        lib._vel_ready = True
    numel, mv = mesh["numel"], mesh["maxViz"]
    id_ = np.ascontiguousarray(case["id"], np.int32)
    num = np.ascontiguousarray(case["num"], np.int32)
    neq = int(id_.max())
    out = np.zeros(3, np.int64)
    lib.ref_init(1 << 31)
    lib.ref_datastruct(2, 1, id_, num, mesh["nelcon"], mesh["nFace"], numel, mv,
                       neq, out)
    ia = np.zeros(out[0], np.int32)
    ja = np.zeros(out[1], np.int32)
    lib.ref_datastruct_copy(ia, ja, out[0], out[1])
    al = np.zeros(out[2])
    ad = np.zeros(3 * neq)
    b = np.zeros(3 * neq)
    rc = np.zeros(3 * numel)
    d = np.ascontiguousarray(case["dField"], np.float64).copy()
    opt, optd = vel_options(case)
    lv = load_table(case["loadsVel"])
    eddy = np.ascontiguousarray(case.get("eddyVisc", np.zeros(numel)), np.float64)
    wall = np.ascontiguousarray(case.get("wallPar", np.zeros(4 * numel)), np.float64)
    nmat = len(case["prop"]) // 10
    r = lib.ref_simple_vel(
        numel, mv, mesh["nFace"], mesh["nelcon"], mesh["cellFace"], mesh["owner"],
        mesh["volume"], mesh["dcca"], mesh["xmcc"], mesh["cc"], mesh["mksi"],
        mesh["ksi"], mesh["eta"], mesh["area"], mesh["normal"], mesh["xm"],
        mesh["mvSkew"], mesh["vSkew"], case["mat"], case["prop"], nmat,
        case["faceRvel"], case["faceRpres"], id_, len(case["loadsVel"]), *lv,
        opt, optd, case["pres"], case["gradPres"], case["vel"], case["gradVel"],
        eddy, wall, d, al, ad, b, rc, int(out[2]))
    if r:
        raise RuntimeError("ref_simple_vel")
    return ia, ja, al, ad, b, rc, d
