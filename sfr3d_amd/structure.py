"""Host-side set-up of SFR3D systems through the C ABI (include/sfr3d_setup.h):
mesh adjacency in the reference's face order, reordering (reord/GENRCM),
equation numbering, CSR/CSRD/CSRC/ELLPACK index builders and value scatter.

Arrays follow the reference's conventions: nelcon is 1-based with -1 for a
boundary face, num and id are 1-based, ia/ja are 0-based int32.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import lib, SfrPattern, SfrMtx, ptr, CSR, CSRC, CSRCCOO, CSRD, ELLPACK


def _chk(rc, what):
    if rc < 0:
        raise RuntimeError(f"{what} failed (code {rc})")
    return rc


def quad_adjacency(nx, ny):
    """nelcon (numel x 4) and nViz for an nx*ny quad grid, faces S,E,N,W."""
    numel = nx * ny
    nelcon = np.empty(numel * 4, np.int32)
    nviz = np.empty(numel, np.int16)
    _chk(lib.sfr_mesh_quad_adj(nx, ny, nelcon, nviz), "sfr_mesh_quad_adj")
    return nelcon.reshape(numel, 4), nviz


def hex_adjacency(nx, ny, nz):
    """nelcon (numel x 6) and nViz for a hex grid, faces z-,z+,y-,x+,y+,x-."""
    numel = nx * ny * nz
    nelcon = np.empty(numel * 6, np.int32)
    nviz = np.empty(numel, np.int16)
    _chk(lib.sfr_mesh_hex_adj(nx, ny, nz, nelcon, nviz), "sfr_mesh_hex_adj")
    return nelcon.reshape(numel, 6), nviz


def general_adjacency(ndm, el, nen, max_viz):
    """Adjacency of a conforming mesh from 1-based cell-node lists."""
    el = np.ascontiguousarray(el, np.int32)
    numel, max_no = el.shape
    nen = np.ascontiguousarray(nen, np.int16)
    nelcon = np.empty(numel * max_viz, np.int32)
    nviz = np.empty(numel, np.int16)
    _chk(lib.sfr_mesh_adj_general(ndm, numel, max_no, el.ravel(), nen,
                                  max_viz, nelcon, nviz),
         "sfr_mesh_adj_general")
    return nelcon.reshape(numel, max_viz), nviz


def reord(nelcon, nviz, flag=True):
    """Cell ordering num (1-based), reord (reference src/Reord.c:18-112)."""
    nelcon = np.ascontiguousarray(nelcon, np.int32)
    numel, max_viz = nelcon.shape
    num = np.empty(numel, np.int32)
    _chk(lib.sfr_reord(num, nelcon.ravel(), np.ascontiguousarray(nviz, np.int16),
                       max_viz, numel, 1 if flag else 0), "sfr_reord")
    return num


def genrcm(xadj, adj):
    """Reverse Cuthill-McKee permutation (0-based) of a 0-based CSR graph."""
    xadj = np.ascontiguousarray(xadj, np.int64)
    adj = np.ascontiguousarray(adj, np.int32)
    n = len(xadj) - 1
    perm = np.empty(n, np.int32)
    _chk(lib.sfr_genrcm(n, xadj, adj, perm), "sfr_genrcm")
    return perm


def numeq_v1(num):
    id_ = np.empty(len(num), np.int32)
    neq = _chk(lib.sfr_numeq_v1(id_, np.ascontiguousarray(num, np.int32),
                                len(num)), "sfr_numeq_v1")
    return id_, neq


def numeq_v2(num, rt, nface, max_viz):
    id_ = np.empty(len(num), np.int32)
    neq = _chk(lib.sfr_numeq_v2(id_, np.ascontiguousarray(num, np.int32),
                                np.ascontiguousarray(rt, np.int16).ravel(),
                                np.ascontiguousarray(nface, np.int16),
                                len(num), max_viz), "sfr_numeq_v2")
    return id_, neq


@dataclass
class Pattern:
    storage: int
    unsym: bool
    neq: int
    nad: int
    max_viz: int
    ia: np.ndarray
    ja: np.ndarray
    nval: int
    _c: SfrPattern = None

    def value_arrays(self):
        """Zeroed al (storage layout), ad and b of the right sizes."""
        nrow = self.neq + 1 if self.storage == ELLPACK else self.neq
        return (np.zeros(max(self.nval, 1), np.float64),
                np.zeros(nrow, np.float64), np.zeros(nrow, np.float64))


def datastruct(storage, unsym, id_, num, nelcon, nviz, neq):
    """dataStruct (reference src/Datastruct.c:28-613), ndf = 1."""
    nelcon = np.ascontiguousarray(nelcon, np.int32)
    numel, max_viz = nelcon.shape
    p = SfrPattern()
    _chk(lib.sfr_datastruct(C.byref(p), storage, 1 if unsym else 0,
                            np.ascontiguousarray(id_, np.int32),
                            np.ascontiguousarray(num, np.int32),
                            nelcon.ravel(), np.ascontiguousarray(nviz, np.int16),
                            numel, max_viz, neq), "sfr_datastruct")
    ia = np.ctypeslib.as_array(p.ia, shape=(p.nia,)).copy()
    ja = np.ctypeslib.as_array(p.ja, shape=(max(p.nja, 1),))[:p.nja].copy()
    pat = Pattern(storage, bool(unsym), p.neq, p.nad, max_viz, ia, ja, p.nval)
    lib.sfr_pattern_free(C.byref(p))
    # keep a C view on our own copies for sfr_assemble
    c = SfrPattern()
    c.neq, c.nad, c.nia, c.nja, c.nval = p.neq, p.nad, len(ia), len(ja), p.nval
    c.storage, c.unsym, c.maxViz = storage, 1 if unsym else 0, max_viz
    c.ia = ia.ctypes.data_as(C.POINTER(C.c_int32))
    c.ja = ja.ctypes.data_as(C.POINTER(C.c_int32))
    pat._c = c
    return pat


def assemble(pat, id_, nelcon, nviz, lA, lB, ell_quirk=False):
    """Scatter per-cell coefficients (lA: numel x (maxViz+1), diagonal last;
    lB: numel) into (al, ad, b) of the pattern's storage."""
    nelcon = np.ascontiguousarray(nelcon, np.int32)
    numel = nelcon.shape[0]
    lA = np.ascontiguousarray(lA, np.float64)
    lB = np.ascontiguousarray(lB, np.float64)
    al, ad, b = pat.value_arrays()
    _chk(lib.sfr_assemble(C.byref(pat._c), np.ascontiguousarray(id_, np.int32),
                          nelcon.ravel(), np.ascontiguousarray(nviz, np.int16),
                          numel, lA.ravel(), lB, ptr(al),
                          None if pat.storage == CSR else ptr(ad), ptr(b),
                          1 if ell_quirk else 0), "sfr_assemble")
    if pat.storage == CSR:
        ad = al          # the reference aliases ad to al for this storage
    return al, ad, b


def read_solver_config(path):
    s, t, m, st = C.c_short(-1), C.c_double(0), C.c_uint(50000), C.c_short(-1)
    rc = lib.sfr_read_solver_config(path.encode(), C.byref(s), C.byref(t),
                                    C.byref(m), C.byref(st))
    if rc:
        raise RuntimeError(f"sfr_read_solver_config({path}) = {rc}")
    return {"solver": s.value, "tol": t.value, "maxit": m.value,
            "storage": st.value}


# ---------------------------------------------------------------------------
# MatrixMarket (reference src/WriteMtx.c; C side in host/mtx.c)

def write_mtx(path, system, bin=False):
    """The reference's writeCoo of a CSRD system; returns the status."""
    c = system.c()
    return lib.sfr_write_mtx(path.encode(), C.byref(c), 1 if bin else 0)


def write_mtx_vector(path, b):
    b = np.ascontiguousarray(b, np.float64)
    return lib.sfr_write_mtx_vector(path.encode(), b, len(b))


def read_mtx(path):
    """A coordinate real file as a CSRD System (symmetric file: lower rows;
    general: full rows)."""
    from .solver import System
    m = SfrMtx()
    _chk(lib.sfr_read_mtx(path.encode(), C.byref(m)), "sfr_read_mtx")
    try:
        n, nad = m.neq, m.nad
        ia = np.ctypeslib.as_array(m.ia, shape=(n + 1,)).copy()
        ja = (np.ctypeslib.as_array(m.ja, shape=(nad,)).copy() if nad
              else np.zeros(0, np.int32))
        al = (np.ctypeslib.as_array(m.al, shape=(nad,)).copy() if nad
              else np.zeros(0))
        ad = np.ctypeslib.as_array(m.ad, shape=(n,)).copy()
        unsym = bool(m.unsym)
    finally:
        lib.sfr_mtx_free(C.byref(m))
    return System(n, nad, ia, ja, al, ad, CSRD, unsym)
