"""Python face of the device solve (C ABI of include/sfr3d.h).

Mirrors the reference's solver boundary, solverC (reference
include/Solv.h:59-70): a system in one of SFR3D's storages, a solver id
(PCG = 1, PBICGSTAB = 2, PBICGSTABL2 = 3, GMRES = 4, MINRES = 5), tol,
maxit, b and x.  Everything here calls the
HIP library; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import lib, SfrSystem, SfrStats, ptr


class SolverError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{lib.sfr_status_string(code).decode()}: {msg}")
        self.code = code


class System:
    """Host arrays of one system in a reference storage (SistEq fields)."""

    def __init__(self, neq, nad, ia, ja, al, ad, storage, unsym, au=None):
        self.neq, self.nad = int(neq), int(nad)
        self.ia = np.ascontiguousarray(ia, np.int32)
        self.ja = np.ascontiguousarray(ja, np.int32)
        self.al = np.ascontiguousarray(al, np.float64)
        self.ad = None if ad is None else np.ascontiguousarray(ad, np.float64)
        self.au = None if au is None else np.ascontiguousarray(au, np.float64)
        self.storage, self.unsym = int(storage), bool(unsym)

    def c(self):
        s = SfrSystem()
        s.neq, s.nad = self.neq, self.nad
        s.ia, s.ja, s.al = ptr(self.ia), ptr(self.ja), ptr(self.al)
        s.ad = ptr(self.ad) if self.ad is not None else ptr(self.al)
        s.au = ptr(self.au)
        s.storage, s.unsym = self.storage, 1 if self.unsym else 0
        return s


class DeviceArray:
    """Plain device allocation owned by a Context."""

    def __init__(self, ctx, nbytes, dtype, shape):
        self.ctx, self.nbytes, self.dtype, self.shape = ctx, nbytes, dtype, shape
        p = C.c_void_p()
        ctx._check(lib.sfr_dev_alloc(ctx.h, nbytes, C.byref(p)))
        self.ptr = p.value

    def set(self, arr):
        arr = np.ascontiguousarray(arr, self.dtype)
        assert arr.nbytes == self.nbytes
        self.ctx._check(lib.sfr_copy_h2d(self.ctx.h, self.ptr, ptr(arr),
                                         arr.nbytes))
        return self

    def get(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(lib.sfr_copy_d2h(self.ctx.h, ptr(out), self.ptr,
                                         out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib.sfr_dev_free(self.ctx.h, self.ptr)
            self.ptr = None


class DeviceMatrix:
    def __init__(self, ctx, system, fmt=L.FMT_AUTO):
        self.ctx = ctx
        self.system = system            # keeps host arrays alive
        h = C.c_void_p()
        s = system.c()
        ctx._check(lib.sfr_matrix_create(ctx.h, C.byref(s), fmt, C.byref(h)))
        self.h = h.value

    def set_values(self, al, ad, au=None):
        al = np.ascontiguousarray(al, np.float64)
        ad = None if ad is None else np.ascontiguousarray(ad, np.float64)
        self._keep = (al, ad, au)
        self.ctx._check(lib.sfr_matrix_set_values(self.ctx.h, self.h, ptr(al),
                                                  ptr(ad), ptr(au)))

    def tune(self, rpt, lanes=4):
        self.ctx._check(lib.sfr_matrix_tune(self.ctx.h, self.h, rpt, lanes))

    def info(self):
        n, f, s, nnz = C.c_int32(), C.c_int(), C.c_int(), C.c_int64()
        lib.sfr_matrix_info(self.h, C.byref(n), C.byref(f), C.byref(s),
                            C.byref(nnz))
        rpt, rpt_pcg = C.c_int(), C.c_int()
        lib.sfr_matrix_rows_per_thread(self.h, C.byref(rpt), C.byref(rpt_pcg))
        return {"neq": n.value, "format": f.value, "slots": s.value,
                "nnz": nnz.value,
                "index_bytes": int(lib.sfr_matrix_index_bytes(self.h)),
                "pair_planes": int(lib.sfr_matrix_pair_planes(self.h)),
                "rows_per_thread": rpt.value, "rows_per_thread_pcg": rpt_pcg.value}

    def destroy(self):
        if self.h:
            lib.sfr_matrix_destroy(self.ctx.h, self.h)
            self.h = None


class Assembly:
    """Device assembly plan of one resident matrix (sfr_assembly_*): per-cell
    coefficients on the device scattered into the matrix's values."""

    def __init__(self, ctx, A, pattern, id_, nelcon, nviz, ell_quirk=False):
        self.ctx, self.A = ctx, A
        nelcon = np.ascontiguousarray(nelcon, np.int32)
        self._keep = (pattern, nelcon)
        h = C.c_void_p()
        ctx._check(lib.sfr_assembly_create(
            ctx.h, A.h, C.byref(pattern._c), np.ascontiguousarray(id_, np.int32),
            nelcon.ravel(), np.ascontiguousarray(nviz, np.int16), nelcon.shape[0],
            1 if ell_quirk else 0, C.byref(h)))
        self.h = h.value

    def assemble(self, d_lA, d_lB=None, d_b=None):
        self.ctx._check(lib.sfr_assemble_device(
            self.ctx.h, self.h, d_lA.ptr, d_lB.ptr if d_lB else None,
            d_b.ptr if d_b else None))

    def destroy(self):
        if self.h:
            lib.sfr_assembly_destroy(self.ctx.h, self.h)
            self.h = None


def load_array(loads):
    """ctypes array of SfrLoad from a list of dicts with "type", "par" (at
    most 8 values) and "density" (SFR3D's constant Loads)."""
    arr = (L.SfrLoad * max(len(loads), 1))()
    for i, ld in enumerate(loads):
        par = list(ld.get("par", []))
        arr[i].type = ld["type"]
        arr[i].np = len(par)
        for k, v in enumerate(par):
            arr[i].par[k] = v
        arr[i].density = ld.get("density", 0.0)
    return arr


class Cells:
    """Cell coefficients on the device (sfr_cells_*): the mesh geometry of
    SFR3D's Mesh (a dict as systems.hex_mesh makes), material properties
    and load tables uploaded once; pres() runs the SIMPLE pressure-correction
    cell loop on device fields into lA / lB / rCell."""

    _MESH = ("nFace", "nelcon", "cellFace", "owner", "volume", "dcca", "xmcc",
             "mksi", "ksi", "area", "normal", "vSkew", "cc", "xm")

    def __init__(self, ctx, mesh, prop, mat, faceRvel, faceRpres, loadsVel,
                 loadsPres):
        self.ctx = ctx
        keep = {k: np.ascontiguousarray(mesh[k]) for k in self._MESH if k in mesh}
        keep["mat"] = np.ascontiguousarray(mat, np.int16)
        keep["faceRvel"] = np.ascontiguousarray(faceRvel, np.int16)
        keep["faceRpres"] = np.ascontiguousarray(faceRpres, np.int16)
        m = L.SfrCellMesh(numel=mesh["numel"], nface=mesh["nface"],
                          maxViz=mesh["maxViz"], ndm=3)
        for k, v in keep.items():
            setattr(m, k, v.ctypes.data)
        prop = np.ascontiguousarray(prop, np.float64)
        lv, lp = load_array(loadsVel), load_array(loadsPres)
        h = C.c_void_p()
        ctx._check(lib.sfr_cells_create(ctx.h, C.byref(m), prop,
                                        len(prop) // 10, lv, len(loadsVel),
                                        lp, len(loadsPres), C.byref(h)))
        self.h = h.value
        self.numel, self.maxViz = mesh["numel"], mesh["maxViz"]

    def pres(self, diffusion, d_pres, d_gradPres, d_vel, d_dField, d_lA, d_lB,
             d_rCell=None):
        self.ctx._check(lib.sfr_cells_pres_device(
            self.ctx.h, self.h, diffusion, d_pres.ptr, d_gradPres.ptr,
            d_vel.ptr, d_dField.ptr, d_lA.ptr, d_lB.ptr,
            d_rCell.ptr if d_rCell else None))

    def vel(self, opts, d_id, neq, d_pres, d_gradPres, d_vel, d_gradVel,
            d_eddyVisc, d_wallPar, d_dField, d_dOut, d_lA, d_ad, d_b, d_rCell):
        """The momentum cell loop: opts a dict with the sfr_vel_opts fields."""
        o = L.SfrVelOpts(**{k: opts[k] for k, _ in L.SfrVelOpts._fields_})
        self.ctx._check(lib.sfr_cells_vel_device(
            self.ctx.h, self.h, C.byref(o), d_id.ptr, neq, d_pres.ptr,
            d_gradPres.ptr, d_vel.ptr, d_gradVel.ptr,
            d_eddyVisc.ptr if d_eddyVisc else None,
            d_wallPar.ptr if d_wallPar else None, d_dField.ptr, d_dOut.ptr,
            d_lA.ptr, d_ad.ptr, d_b.ptr, d_rCell.ptr))

    def destroy(self):
        if self.h:
            lib.sfr_cells_destroy(self.ctx.h, self.h)
            self.h = None


class Batch:
    """Uneven batch of independent systems (sfr_batch_*): mixed-precision
    PCG over all of them in one stream of launches."""

    def __init__(self, ctx, systems, tune=True):
        self.ctx = ctx
        self.systems = list(systems)        # keeps host arrays alive
        n = len(self.systems)
        arr = (SfrSystem * n)(*[s.c() for s in self.systems])
        h = C.c_void_p()
        ctx._check(lib.sfr_batch_create(ctx.h, n, arr, 1 if tune else 0,
                                        C.byref(h)))
        self.h = h.value
        self.nsys = n
        self.rows = int(lib.sfr_batch_rows(self.h))

    def info(self, s):
        r0, n, f, rpt = C.c_int64(), C.c_int32(), C.c_int(), C.c_int()
        self.ctx._check(lib.sfr_batch_info(self.h, s, C.byref(r0), C.byref(n),
                                           C.byref(f), C.byref(rpt)))
        return {"row0": r0.value, "neq": n.value, "format": f.value,
                "rpt": rpt.value}

    def buckets(self):
        out = (L.SfrBucketChoice * 32)()
        k = lib.sfr_batch_buckets(self.h, out, 32)
        return [out[i].as_dict() for i in range(k)]

    def set_values(self, s, al, ad, au=None):
        al = np.ascontiguousarray(al, np.float64)
        ad = None if ad is None else np.ascontiguousarray(ad, np.float64)
        au = None if au is None else np.ascontiguousarray(au, np.float64)
        self.ctx._check(lib.sfr_batch_set_values(self.ctx.h, self.h, s,
                                                 ptr(al), ptr(ad), ptr(au)))

    def solve(self, bs, tol, maxit, xs=None, newX=True):
        """Host arrays: bs[s] (and xs[s]) of system s's neq doubles.
        Returns (list of x, list of stats dicts)."""
        bs = [np.ascontiguousarray(b, np.float64) for b in bs]
        assert len(bs) == self.nsys
        xs = ([np.zeros(s.neq) for s in self.systems] if xs is None
              else [np.array(x, np.float64) for x in xs])
        P = C.c_void_p * self.nsys
        pb = P(*[ptr(b) for b in bs])
        px = P(*[ptr(x) for x in xs])
        st = (SfrStats * self.nsys)()
        rc = lib.sfr_batch_solve(self.ctx.h, self.h, tol, maxit,
                                 C.cast(pb, C.c_void_p), C.cast(px, C.c_void_p),
                                 1 if newX else 0, st)
        self.ctx._check(rc, allow=(L.MAXIT, L.BREAKDOWN))
        return xs, [st[c].as_dict() for c in range(self.nsys)]

    def solve_device(self, d_b, d_x, tol, maxit, newX=True):
        """Device vectors of sfr_batch_rows doubles, system s at info(s)['row0']."""
        st = (SfrStats * self.nsys)()
        rc = lib.sfr_batch_solve_device(self.ctx.h, self.h, tol, maxit,
                                        d_b.ptr, d_x.ptr, 1 if newX else 0, st)
        self.ctx._check(rc, allow=(L.MAXIT, L.BREAKDOWN))
        return [st[c].as_dict() for c in range(self.nsys)]

    def destroy(self):
        if self.h:
            lib.sfr_batch_destroy(self.ctx.h, self.h)
            self.h = None


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib.sfr_init(C.byref(h), device)
        if rc != L.OK:
            raise SolverError(rc, "sfr_init")
        self.h = h.value

    def _check(self, rc, allow=()):
        if rc != L.OK and rc not in allow:
            raise SolverError(rc, lib.sfr_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            lib.sfr_finalize(self.h)
            self.h = None

    def cache_info(self):
        """sfr_cache_info: (matrices held, matrices created, lookups that hit)
        of the host-pointer paths' pattern cache."""
        e, c, h = C.c_int(), C.c_int64(), C.c_int64()
        self._check(lib.sfr_cache_info(self.h, C.byref(e), C.byref(c), C.byref(h)))
        return e.value, c.value, h.value

    def set_option(self, name, value):
        """sfr_set_option: "chunk" (iterations per captured graph chunk),
        "precision", "ru_delta" / "ru_delta_bicg" (mixed-precision
        reliable-update thresholds), "ell_pairs" (0/1) or "nt_loads"
        (-1 by size, 0/1); an unknown name or value raises SolverError."""
        self._check(lib.sfr_set_option(self.h, name.encode(), float(value)))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def array(self, arr):
        arr = np.ascontiguousarray(arr)
        return DeviceArray(self, arr.nbytes, arr.dtype, arr.shape).set(arr)

    def empty(self, shape, dtype=np.float64):
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return DeviceArray(self, nbytes, np.dtype(dtype), shape)

    def matrix(self, system, fmt=L.FMT_AUTO):
        return DeviceMatrix(self, system, fmt)

    def assembly(self, A, pattern, id_, nelcon, nviz, ell_quirk=False):
        return Assembly(self, A, pattern, id_, nelcon, nviz, ell_quirk)

    def batch(self, systems, tune=True):
        return Batch(self, systems, tune)

    def cells_update_value(self, d_id, d_x, d_u, add=False):
        """u[c] = x[id[c]-1] (or +=) on the device, cells with an equation."""
        self._check(lib.sfr_cells_update_value(self.h, d_u.nbytes // 8, d_id.ptr,
                                               d_x.ptr, d_u.ptr, 1 if add else 0))

    def simple_update(self, alpha, d_vel, d_pres, d_presC, d_gradPresC, d_dField):
        """vel -= dField * gradPresC, pres += alpha * presC on the device."""
        self._check(lib.sfr_simple_update(self.h, d_pres.nbytes // 8, alpha,
                                          d_vel.ptr, d_pres.ptr, d_presC.ptr,
                                          d_gradPresC.ptr, d_dField.ptr))

    def cells(self, mesh, case):
        """Cells for a mesh from systems.hex_mesh and a case from
        systems.simple_pres_case (materials, markers and loads)."""
        return Cells(self, mesh, case["prop"], case["mat"], case["faceRvel"],
                     case["faceRpres"], case["loadsVel"], case["loadsPres"])

    def solve(self, system, solver, tol, maxit, b, x=None, newX=True):
        """Host-array solve (sfr_solve).  Returns (x, stats dict)."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(system.neq) if x is None else np.array(x, np.float64)
        st = SfrStats()
        s = system.c()
        rc = lib.sfr_solve(self.h, C.byref(s), solver, tol, maxit, ptr(b),
                           ptr(x), 1 if newX else 0, C.byref(st))
        self._check(rc, allow=(L.MAXIT, L.BREAKDOWN))
        return x, st.as_dict()

    def solve_shared(self, system, solver, tol, maxit, bs, ads=None, xs=None,
                     newX=True):
        """Shared-pattern batch (u/v/w): one system's pattern and
        off-diagonals, per-system diagonals ads[c], right-hand sides bs[c].
        Returns (list of x, list of stats dicts)."""
        nsys = len(bs)
        bs = [np.ascontiguousarray(b, np.float64) for b in bs]
        xs = ([np.zeros(system.neq) for _ in range(nsys)] if xs is None
              else [np.array(x, np.float64) for x in xs])
        if ads is not None:
            ads = [np.ascontiguousarray(a, np.float64) for a in ads]
        P = C.c_void_p * nsys
        pb = P(*[ptr(b) for b in bs])
        px = P(*[ptr(x) for x in xs])
        pad = P(*[ptr(a) for a in ads]) if ads is not None else None
        st = (SfrStats * nsys)()
        s = system.c()
        rc = lib.sfr_solve_shared(self.h, C.byref(s), nsys,
                                  C.cast(pad, C.c_void_p) if pad else None,
                                  C.cast(pb, C.c_void_p), C.cast(px, C.c_void_p),
                                  solver, tol, maxit, 1 if newX else 0, st)
        self._check(rc, allow=(L.MAXIT, L.BREAKDOWN))
        return xs, [st[c].as_dict() for c in range(nsys)]

    def solve_batched(self, system, solver, tol, maxit, als, ads, bs,
                      aus=None, xs=None, newX=True):
        """Species batch (sfr_solve_batched): system's pattern, system c's
        values als[c], ads[c] (aus[c]), right-hand side bs[c].
        Returns (list of x, list of stats dicts)."""
        nsys = len(bs)
        f64 = lambda a: np.ascontiguousarray(a, np.float64)
        als, ads, bs = [f64(a) for a in als], [f64(a) for a in ads], [f64(b) for b in bs]
        aus = [f64(a) for a in aus] if aus is not None else None
        xs = ([np.zeros(system.neq) for _ in range(nsys)] if xs is None
              else [np.array(x, np.float64) for x in xs])
        P = C.c_void_p * nsys
        arr = lambda lst: C.cast(P(*[ptr(a) for a in lst]), C.c_void_p)
        st = (SfrStats * nsys)()
        s = system.c()
        rc = lib.sfr_solve_batched(self.h, C.byref(s), nsys, arr(als), arr(ads),
                                   arr(aus) if aus is not None else None,
                                   arr(bs), arr(xs), solver, tol, maxit,
                                   1 if newX else 0, st)
        self._check(rc, allow=(L.MAXIT, L.BREAKDOWN))
        return xs, [st[c].as_dict() for c in range(nsys)]

    def solve_adjoint(self, A, solver, tol, maxit, d_x, d_g, d_lam,
                      d_grad_al=None, d_grad_ad=None):
        """A^T lambda = g and the value gradients (sfr_solve_adjoint)."""
        st = SfrStats()
        rc = lib.sfr_solve_adjoint(self.h, A.h, solver, tol, maxit, d_x.ptr,
                                   d_g.ptr, d_lam.ptr,
                                   d_grad_al.ptr if d_grad_al else None,
                                   d_grad_ad.ptr if d_grad_ad else None,
                                   C.byref(st))
        self._check(rc, allow=(L.MAXIT, L.BREAKDOWN))
        return st.as_dict()

    def solve_device(self, A, solver, tol, maxit, d_b, d_x, newX=True,
                     precision=L.PREC_FP64):
        st = SfrStats()
        rc = lib.sfr_solve_device(self.h, A.h, solver, precision, tol, maxit,
                                  d_b.ptr, d_x.ptr, 1 if newX else 0,
                                  C.byref(st))
        self._check(rc, allow=(L.MAXIT, L.BREAKDOWN))
        return st.as_dict()
