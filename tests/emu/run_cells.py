"""TEST INFRASTRUCTURE -- drives the device cell loop (sfr_cells_create /
sfr_cells_pres_device) of a library built for the CPU emulation (argv[1]) on
the meshes and cases of argv[2] (written by tests/test_emulated_kernels.py
with the oracle's coefficients), with host memory standing in for device
memory.  Prints PASS/FAIL lines; exits non-zero on any failure."""
import ctypes as C
import sys

import numpy as np

lib = C.CDLL(sys.argv[1])
z = np.load(sys.argv[2])
vp = C.c_void_p
MESH = ("nFace", "nelcon", "cellFace", "owner", "volume", "dcca", "xmcc",
        "mksi", "ksi", "area", "normal", "vSkew", "mat", "faceRvel",
        "faceRpres", "cc", "xm")
OPTS = ("advection", "diffusion", "turb", "wall_model", "rhie", "so_pressure",
        "viscosity", "residual", "simplec", "time", "backward")


class Mesh(C.Structure):
    _fields_ = [("numel", C.c_int32), ("nface", C.c_int32),
                ("maxViz", C.c_short), ("ndm", C.c_short)] + [(k, vp) for k in MESH]


class VelOpts(C.Structure):
    _fields_ = [(k, C.c_int) for k in OPTS] + [("dt", C.c_double), ("dt0", C.c_double),
                                               ("underU", C.c_double)]


class Load(C.Structure):
    _fields_ = [("type", C.c_short), ("np", C.c_short),
                ("par", C.c_double * 8), ("density", C.c_double)]


lib.sfr_init.argtypes = [C.POINTER(vp), C.c_int]
lib.sfr_cells_create.argtypes = [vp, C.POINTER(Mesh), vp, C.c_int, vp, C.c_int,
                                 vp, C.c_int, C.POINTER(vp)]
lib.sfr_cells_pres_device.argtypes = [vp, vp, C.c_int] + [vp] * 7
lib.sfr_cells_destroy.argtypes = [vp, vp]
lib.sfr_cells_vel_device.argtypes = [vp, vp, C.POINTER(VelOpts), vp, C.c_int] + [vp] * 12
lib.sfr_synchronize.argtypes = [vp]
fails = 0


def check(tag, ok):
    global fails
    print(("PASS " if ok else "FAIL ") + tag, flush=True)
    fails += (not ok)


def loads(pfx):
    n = int(z[pfx + "_n"][0])
    arr = (Load * max(n, 1))()
    for i in range(n):
        arr[i].type = int(z[pfx + "_type"][i])
        arr[i].np = int(z[pfx + "_np"][i])
        for k in range(8):
            arr[i].par[k] = float(z[pfx + "_par"][i * 8 + k])
        arr[i].density = float(z[pfx + "_density"][i])
    return arr, n


ctx = vp()
assert lib.sfr_init(C.byref(ctx), 0) == 0
for i in range(int(z["ncases"][0])):
    p = f"c{i}_"
    keep = {k: np.ascontiguousarray(z[p + k]) for k in MESH if p + k in z.files}
    numel, nface, mv, diff = (int(v) for v in z[p + "dims"])
    m = Mesh(numel=numel, nface=nface, maxViz=mv, ndm=3)
    for k, v in keep.items():
        setattr(m, k, v.ctypes.data)
    prop = np.ascontiguousarray(z[p + "prop"])
    lv, nlv = loads(p + "lv")
    lp, nlp = loads(p + "lp")
    h = vp()
    rc = lib.sfr_cells_create(ctx, C.byref(m), prop.ctypes.data, len(prop) // 10,
                              C.cast(lv, vp), nlv, C.cast(lp, vp), nlp, C.byref(h))
    check(f"case {i} create rc {rc}", rc == 0)
    f = {k: np.ascontiguousarray(z[p + k]) for k in ("pres", "gradPres", "vel", "dField")}
    lA = np.zeros(numel * (mv + 1))
    lB = np.zeros(numel)
    rcell = np.full(numel, 7.0)
    rc = lib.sfr_cells_pres_device(ctx, h, diff, f["pres"].ctypes.data,
                                   f["gradPres"].ctypes.data, f["vel"].ctypes.data,
                                   f["dField"].ctypes.data, lA.ctypes.data,
                                   lB.ctypes.data, rcell.ctypes.data)
    lib.sfr_synchronize(ctx)
    check(f"case {i} lA bitwise (diffusion {diff})",
          rc == 0 and np.array_equal(lA, z[p + "lA"].ravel()))
    check(f"case {i} lB, rCell bitwise",
          np.array_equal(lB, z[p + "lB"]) and np.array_equal(rcell, z[p + "rCell"]))
    lib.sfr_cells_destroy(ctx, h)
for i in range(int(z["nvel"][0]) if "nvel" in z else 0):
    p = f"v{i}_"
    keep = {k: np.ascontiguousarray(z[p + k]) for k in MESH if p + k in z.files}
    numel, nface, mv = (int(v) for v in z[p + "dims"])
    m = Mesh(numel=numel, nface=nface, maxViz=mv, ndm=3)
    for k, v in keep.items():
        setattr(m, k, v.ctypes.data)
    prop = np.ascontiguousarray(z[p + "prop"])
    lv, nlv = loads(p + "lv")
    lp, nlp = loads(p + "lp")
    h = vp()
    rc = lib.sfr_cells_create(ctx, C.byref(m), prop.ctypes.data, len(prop) // 10,
                              C.cast(lv, vp), nlv, C.cast(lp, vp), nlp, C.byref(h))
    o = [int(v) for v in z[p + "opt"]]
    od = [float(v) for v in z[p + "optd"]]
    opts = VelOpts(*o, *od)
    f = {k: np.ascontiguousarray(z[p + k]) for k in
         ("pres", "gradPres", "vel", "gradVel", "eddyVisc", "wallPar", "dField", "id")}
    neq = int(f["id"].max())
    dout = np.zeros(3 * numel)
    lA = np.zeros(numel * (mv + 1))
    ad = np.zeros(3 * neq)
    b = np.zeros(3 * neq)
    rcell = np.full(3 * numel, 7.0)
    rc2 = lib.sfr_cells_vel_device(ctx, h, C.byref(opts), f["id"].ctypes.data, neq,
                                   f["pres"].ctypes.data, f["gradPres"].ctypes.data,
                                   f["vel"].ctypes.data, f["gradVel"].ctypes.data,
                                   f["eddyVisc"].ctypes.data, f["wallPar"].ctypes.data,
                                   f["dField"].ctypes.data, dout.ctypes.data,
                                   lA.ctypes.data, ad.ctypes.data, b.ctypes.data,
                                   rcell.ctypes.data)
    lib.sfr_synchronize(ctx)
    check(f"vel case {i} rc {rc}/{rc2} lA faces and u diagonal bitwise",
          rc == 0 and rc2 == 0 and np.array_equal(lA, z[p + "lA1"]))
    check(f"vel case {i} ad, b, rCell, d bitwise",
          np.array_equal(ad, z[p + "ad"]) and np.array_equal(b, z[p + "b"])
          and np.array_equal(rcell, z[p + "rCell"]) and np.array_equal(dout, z[p + "dOut"]))
    lib.sfr_cells_destroy(ctx, h)
sys.exit(1 if fails else 0)
