"""ctypes binding of libsfr3d.so (the C ABI declared in include/sfr3d.h and
include/sfr3d_setup.h).

The library is built in-tree (``__graft_entry__.build()`` or
``make -C sfr3d_amd/csrc``).  Importing this module without it raises: there is
no CPU fallback for any solve.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfr3d.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C sfr3d_amd/csrc` "
        "(the GPU path has no fallback)")

lib = C.CDLL(LIB_PATH)

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

# ---- ids (include/Define.h:392-397, 407-412 of the reference) --------------
PCG, PBICGSTAB, PBICGSTABL2, GMRES, MINRES = 1, 2, 3, 4, 5
CSR, CSRD, CSRC, ELLPACK, CSRDCOO, CSRCCOO = 1, 2, 3, 4, 5, 6
OK, MAXIT, BREAKDOWN, BAD_ARG, HIP_ERROR, NO_DEVICE, NOT_SUPPORTED = range(7)
PREC_FP64, PREC_MIXED = 0, 1
FMT_AUTO, FMT_ELL, FMT_CSR = 0, 1, 2


class SfrSystem(C.Structure):
    _fields_ = [("neq", C.c_int32), ("nad", C.c_int32),
                ("ia", C.c_void_p), ("ja", C.c_void_p),
                ("al", C.c_void_p), ("ad", C.c_void_p), ("au", C.c_void_p),
                ("storage", C.c_short), ("unsym", C.c_short)]


class SfrStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_uint), ("outer", C.c_uint),
                ("tol", C.c_double), ("norm_b", C.c_double),
                ("conv", C.c_double), ("xKx", C.c_double),
                ("norm_x", C.c_double), ("norm_r", C.c_double),
                ("norm_r_m", C.c_double), ("relres", C.c_double),
                ("time_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SfrBucketChoice(C.Structure):
    _fields_ = [("bucket", C.c_int), ("nsys", C.c_int), ("format", C.c_int),
                ("rpt", C.c_int), ("ns_per_row", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SfrPattern(C.Structure):
    _fields_ = [("neq", C.c_int32), ("nad", C.c_int32), ("nia", C.c_int32),
                ("nja", C.c_int64), ("nval", C.c_int64),
                ("storage", C.c_short), ("unsym", C.c_short),
                ("maxViz", C.c_short),
                ("ia", C.POINTER(C.c_int32)), ("ja", C.POINTER(C.c_int32))]


class SfrMtx(C.Structure):
    _fields_ = [("neq", C.c_int32), ("nad", C.c_int32), ("unsym", C.c_short),
                ("ia", C.POINTER(C.c_int32)), ("ja", C.POINTER(C.c_int32)),
                ("al", C.POINTER(C.c_double)), ("ad", C.POINTER(C.c_double))]


class SfrCellMesh(C.Structure):
    _fields_ = [("numel", C.c_int32), ("nface", C.c_int32),
                ("maxViz", C.c_short), ("ndm", C.c_short)] + [
        (k, C.c_void_p) for k in ("nFace", "nelcon", "cellFace", "owner",
                                  "volume", "dcca", "xmcc", "mksi", "ksi",
                                  "area", "normal", "vSkew", "mat",
                                  "faceRvel", "faceRpres", "cc", "xm")]


class SfrVelOpts(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("advection", "diffusion", "turb",
                                       "wall_model", "rhie", "so_pressure",
                                       "viscosity", "residual", "simplec",
                                       "time", "backward")] + [
        ("dt", C.c_double), ("dt0", C.c_double), ("underU", C.c_double)]


class SfrLoad(C.Structure):
    _fields_ = [("type", C.c_short), ("np", C.c_short),
                ("par", C.c_double * 8), ("density", C.c_double)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


vp = C.c_void_p
# context / memory
_sig("sfr_init", C.c_int, C.POINTER(vp), C.c_int)
_sig("sfr_finalize", None, vp)
_sig("sfr_status_string", C.c_char_p, C.c_int)
_sig("sfr_last_error", C.c_char_p, vp)
_sig("sfr_synchronize", C.c_int, vp)
_sig("sfr_dev_alloc", C.c_int, vp, C.c_size_t, C.POINTER(vp))
_sig("sfr_dev_free", C.c_int, vp, vp)
_sig("sfr_copy_h2d", C.c_int, vp, vp, vp, C.c_size_t)
_sig("sfr_copy_d2h", C.c_int, vp, vp, vp, C.c_size_t)
# matrices / solves
_sig("sfr_matrix_create", C.c_int, vp, C.POINTER(SfrSystem), C.c_int,
     C.POINTER(vp))
_sig("sfr_matrix_set_values", C.c_int, vp, vp, vp, vp, vp)
_sig("sfr_matrix_destroy", None, vp, vp)
_sig("sfr_matrix_tune", C.c_int, vp, vp, C.c_int, C.c_int)
_sig("sfr_matrix_info", C.c_int, vp, C.POINTER(C.c_int32),
     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64))
_sig("sfr_matrix_index_bytes", C.c_int, vp)
_sig("sfr_matrix_pair_planes", C.c_int, vp)
_sig("sfr_matrix_rows_per_thread", C.c_int, vp, C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("sfr_solve", C.c_int, vp, C.POINTER(SfrSystem), C.c_int, C.c_double,
     C.c_uint, vp, vp, C.c_int, C.POINTER(SfrStats))
_sig("sfr_solve_device", C.c_int, vp, vp, C.c_int, C.c_int, C.c_double,
     C.c_uint, vp, vp, C.c_int, C.POINTER(SfrStats))
_sig("sfr_solve_device_noalloc", C.c_int, vp, vp, C.c_int, C.c_double, C.c_uint,
     vp, vp, C.c_int, C.POINTER(SfrStats))
_sig("sfr_solve_batched", C.c_int, vp, C.POINTER(SfrSystem), C.c_int, vp, vp, vp,
     vp, vp, C.c_int, C.c_double, C.c_uint, C.c_int, C.POINTER(SfrStats))
_sig("sfr_solve_shared", C.c_int, vp, C.POINTER(SfrSystem), C.c_int, vp, vp,
     vp, C.c_int, C.c_double, C.c_uint, C.c_int, C.POINTER(SfrStats))
_sig("sfr_solve_shared_device", C.c_int, vp, vp, C.c_int, vp, vp, vp, C.c_int,
     C.c_double, C.c_uint, C.c_int, C.POINTER(SfrStats))
_sig("sfr_solve_adjoint", C.c_int, vp, vp, C.c_int, C.c_double, C.c_uint, vp,
     vp, vp, vp, vp, C.POINTER(SfrStats))
_sig("sfr_kernel_time", C.c_int, vp, vp, C.c_int, C.c_int,
     C.POINTER(C.c_double))
_sig("sfr_copy_time", C.c_int, vp, vp, vp, C.c_int64, C.c_int,
     C.POINTER(C.c_double))
_sig("sfr_set_option", C.c_int, vp, C.c_char_p, C.c_double)
# device assembly
_sig("sfr_assembly_create", C.c_int, vp, vp, C.POINTER(SfrPattern), i32p,
     i32p, i16p, C.c_int, C.c_int, C.POINTER(vp))
_sig("sfr_assemble_device", C.c_int, vp, vp, vp, vp, vp)
_sig("sfr_assembly_destroy", None, vp, vp)
# cell coefficients on the device
_sig("sfr_cells_create", C.c_int, vp, C.POINTER(SfrCellMesh), f64p, C.c_int,
     C.POINTER(SfrLoad), C.c_int, C.POINTER(SfrLoad), C.c_int, C.POINTER(vp))
_sig("sfr_cells_pres_device", C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp,
     vp, vp)
_sig("sfr_cells_destroy", None, vp, vp)
_sig("sfr_cells_vel_device", C.c_int, vp, vp, C.POINTER(SfrVelOpts), vp,
     C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("sfr_cells_update_value", C.c_int, vp, C.c_int, vp, vp, vp, C.c_int)
_sig("sfr_simple_update", C.c_int, vp, C.c_int, C.c_double, vp, vp, vp, vp, vp)
# uneven batch
_sig("sfr_batch_create", C.c_int, vp, C.c_int, C.POINTER(SfrSystem), C.c_int,
     C.POINTER(vp))
_sig("sfr_batch_set_values", C.c_int, vp, vp, C.c_int, vp, vp, vp)
_sig("sfr_batch_solve", C.c_int, vp, vp, C.c_double, C.c_uint, vp, vp,
     C.c_int, C.POINTER(SfrStats))
_sig("sfr_batch_solve_device", C.c_int, vp, vp, C.c_double, C.c_uint, vp, vp,
     C.c_int, C.POINTER(SfrStats))
_sig("sfr_batch_rows", C.c_int64, vp)
_sig("sfr_batch_info", C.c_int, vp, C.c_int, C.POINTER(C.c_int64),
     C.POINTER(C.c_int32), C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("sfr_batch_buckets", C.c_int, vp, C.POINTER(SfrBucketChoice), C.c_int)
_sig("sfr_batch_destroy", None, vp, vp)
_sig("sfr_solverC", None, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp,
     vp, vp, vp, vp, vp, vp, C.c_double, C.c_uint, C.c_short, C.c_short, vp,
     C.c_bool, C.c_bool, C.c_bool, C.c_char_p)
_sig("sfr_solverC_context", C.c_int, C.POINTER(vp))
_sig("sfr_cache_info", C.c_int, vp, C.POINTER(C.c_int), C.POINTER(C.c_int64),
     C.POINTER(C.c_int64))
_sig("sfr_matrix_cache_clear", None, vp)
# host set-up
_sig("sfr_mesh_quad_adj", C.c_int, C.c_int, C.c_int, i32p, i16p)
_sig("sfr_mesh_hex_adj", C.c_int, C.c_int, C.c_int, C.c_int, i32p, i16p)
_sig("sfr_mesh_adj_general", C.c_int, C.c_int, C.c_int, C.c_int, i32p, i16p,
     C.c_int, i32p, i16p)
_sig("sfr_reord", C.c_int, i32p, i32p, i16p, C.c_short, C.c_int, C.c_int)
_sig("sfr_genrcm", C.c_int, C.c_int, i64p, i32p, i32p)
_sig("sfr_numeq_v1", C.c_int, i32p, i32p, C.c_int)
_sig("sfr_numeq_v2", C.c_int, i32p, i32p, i16p, i16p, C.c_int, C.c_short)
_sig("sfr_datastruct", C.c_int, C.POINTER(SfrPattern), C.c_short, C.c_int,
     i32p, i32p, i32p, i16p, C.c_int, C.c_short, C.c_int)
_sig("sfr_pattern_free", None, C.POINTER(SfrPattern))
_sig("sfr_assemble", C.c_int, C.POINTER(SfrPattern), i32p, i32p, i16p,
     C.c_int, f64p, f64p, vp, vp, vp, C.c_int)
_sig("sfr_read_solver_config", C.c_int, C.c_char_p, C.POINTER(C.c_short),
     C.POINTER(C.c_double), C.POINTER(C.c_uint), C.POINTER(C.c_short))
_sig("sfr_write_mtx", C.c_int, C.c_char_p, C.POINTER(SfrSystem), C.c_int)
_sig("sfr_write_mtx_vector", C.c_int, C.c_char_p, f64p, C.c_int32)
_sig("sfr_read_mtx", C.c_int, C.c_char_p, C.POINTER(SfrMtx))
_sig("sfr_mtx_free", None, C.POINTER(SfrMtx))
_sig("sfr_solver_id", C.c_short, C.c_char_p)
_sig("sfr_storage_id", C.c_short, C.c_char_p)
_sig("sfr_smachn", C.c_double)


def ptr(a):
    """Address of a numpy array's data (None for None)."""
    if a is None:
        return None
    return a.ctypes.data
