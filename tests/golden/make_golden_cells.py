"""Writes tests/golden/ref_simple_pres.npz and ref_simple_vel.npz: outputs of
the reference's own compiled SIMPLE cell loops.

ref_simple_pres.npz: outputs of the reference's own
compiled SIMPLE pressure-correction cell loop (systFormSimplePres,
src/CellLoopInc.c:471-679, with cellSimplePres3D, src/CellPres.c:903-1149,
through oracle/_ref) on a small warped hex mesh, assembled by its own
dataStruct / assbly in CSRD.  Run where the reference tree and oracle/_ref
exist: `python tests/golden/make_golden_cells.py`.

The fixture holds, per case, the inputs that vary (fields, markers,
diffusion scheme) and the reference's al, ad, b and rCell, plus the mesh
arrays and the pattern, so the tests rebuild nothing they compare against.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

MESH = dict(nx=5, ny=4, nz=3, ratio=(1.2, 0.9, 1.1), shear=(0.15, 0.05, -0.1),
            warp=0.04)
CASES = [("mixed", 1, 7), ("mixed", 4, 7), ("cavity", 1, None),
         ("cavity", 4, None)]
MESH_KEYS = ("nFace", "nelcon", "cellFace", "owner", "volume", "dcca", "xmcc",
             "mksi", "ksi", "eta", "area", "normal", "xm", "mvSkew", "vSkew")
CASE_KEYS = ("mat", "prop", "pres", "gradPres", "vel", "dField", "faceRvel",
             "faceRpres", "id", "num")


VEL_CASES = [dict(rhie=False),
             dict(rhie=False, advection=1, diffusion=1, time="euler", simplec=True,
                  residual=False, so_pressure=False, viscosity=False, turb=False),
             dict(rhie=False, advection=2, diffusion=2, boundary="channel",
                  wall_model=True, fixed_cell=7, time="steady"),
             dict(rhie=False, diffusion=3, boundary="channel"),
             dict(rhie=True),
             dict(rhie=True, boundary="channel", wall_model=True, fixed_cell=11)]
VEL_KEYS = ("mat", "prop", "vel", "gradVel", "pres", "gradPres", "dField",
            "eddyVisc", "wallPar", "faceRvel", "faceRpres", "id", "num")


def vel_main(m):
    """ref_simple_vel.npz: systFormSimpleVel (src/CellLoopInc.c:98-385,
    cellSimpleVel3D, src/CellSimple.c:106-638) on the same mesh, CSRD unsym,
    over the option sets of VEL_CASES: the reference's al, ad, b (3*neq),
    rCell (3*numel) and its d field after the loop (updated in place)."""
    import oracle
    from sfr3d_amd import systems as Y
    out = {f"mesh_{k}": m[k] for k in MESH_KEYS + ("cc",)}
    out["mesh_dims"] = np.array([m["nx"], m["ny"], m["nz"], m["numel"],
                                 m["nface"], m["maxViz"]], np.int64)
    for i, kw in enumerate(VEL_CASES):
        case = Y.simple_vel_case(m, seed=3 + i, **kw)
        ia, ja, al, ad, b, rc, d = oracle.ref_simple_vel(m, case)
        for k in VEL_KEYS:
            out[f"v{i}_{k}"] = case[k]
        t, npar, par, dens = oracle.load_table(case["loadsVel"])
        out[f"v{i}_lv_type"], out[f"v{i}_lv_np"] = t, npar
        out[f"v{i}_lv_par"], out[f"v{i}_lv_density"] = par, dens
        out[f"v{i}_lv_n"] = np.array([len(case["loadsVel"])])
        out[f"v{i}_opt"], out[f"v{i}_optd"] = oracle.vel_options(case)
        out[f"v{i}_ia"], out[f"v{i}_ja"] = ia, ja
        out[f"v{i}_al"], out[f"v{i}_ad"], out[f"v{i}_b"] = al, ad, b
        out[f"v{i}_rCell"], out[f"v{i}_dOut"] = rc, d
    out["ncases"] = np.array([len(VEL_CASES)])
    path = os.path.join(HERE, "ref_simple_vel.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def main():
    import oracle
    from sfr3d_amd import systems as Y
    m = Y.hex_mesh(MESH["nx"], MESH["ny"], MESH["nz"], MESH["ratio"],
                   MESH["shear"], MESH["warp"])
    out = {f"mesh_{k}": m[k] for k in MESH_KEYS}
    out["mesh_dims"] = np.array([m["nx"], m["ny"], m["nz"], m["numel"],
                                 m["nface"], m["maxViz"]], np.int64)
    for i, (bc, diff, skip) in enumerate(CASES):
        case = Y.simple_pres_case(m, seed=3, diffusion=diff, boundary=bc,
                                  skip_cell=skip)
        ia, ja, al, ad, b, rc = oracle.ref_simple_pres(m, case)
        for k in CASE_KEYS:
            out[f"c{i}_{k}"] = case[k]
        for name, loads in (("lv", case["loadsVel"]), ("lp", case["loadsPres"])):
            t, npar, par, dens = oracle.load_table(loads)
            out[f"c{i}_{name}_type"], out[f"c{i}_{name}_np"] = t, npar
            out[f"c{i}_{name}_par"], out[f"c{i}_{name}_density"] = par, dens
            out[f"c{i}_{name}_n"] = np.array([len(loads)])
        out[f"c{i}_meta"] = np.array([diff, -1 if skip is None else skip,
                                      1 if bc == "mixed" else 0])
        out[f"c{i}_ia"], out[f"c{i}_ja"] = ia, ja
        out[f"c{i}_al"], out[f"c{i}_ad"], out[f"c{i}_b"] = al, ad, b
        out[f"c{i}_rCell"] = rc
    out["ncases"] = np.array([len(CASES)])
    path = os.path.join(HERE, "ref_simple_pres.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")
    vel_main(m)


if __name__ == "__main__":
    main()
