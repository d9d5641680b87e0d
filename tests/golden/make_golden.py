"""Writes the golden fixtures under tests/golden/ (run where the reference
tree and oracle/_ref exist: `python tests/golden/make_golden.py`).

  retangulo_cells.txt        cell-node lists of the reference's 20x10 quad
                             test mesh (test/input/coo.zip, coo/retangulo.dat)
  retangulo_bin.mtx,         the reference's own MatrixMarket pattern goldens
  retangulo_rcm_bin.mtx      (test/output/coo/), reord false / true
  solver_config/*.dat        the reference example's solver config files
  ref_vectors.json           outputs of the reference's compiled C
                             (oracle/_ref): GENRCM orderings (digest + first
                             entries) and solverC results on small systems
  ref_solutions.npz          the solution vectors of those solves
  ref_gmres.json,            solverC with GMRES(15) (cycles, iterations, xKx,
  ref_gmres_solutions.npz    ||x||, digest of x) and the solution vectors
  ref_minres.json,           solverC with MINRES (logged iterations, xKx,
  ref_minres_solutions.npz   ||x||, digest of x) and the solution vectors
  ref_bicgl2.json,           solverC with BiCGSTAB(2) (iterations, xKx, ||x||,
  ref_bicgl2_solutions.npz   digest of x) and the solution vectors
  cavity_ref.json            solverC (PCG, tol 1e-8) on pressure-correction
                             systems of the 128x128 lid-driven cavity (SIMPLE
                             outer iterations 0, 5, 11; CSRD and ELLPACK):
                             iterations, xKx, ||x||, digests of the inputs
                             and of x
"""
import hashlib
import io
import json
import os
import shutil
import sys
import zipfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def mesh_fixture():
    with zipfile.ZipFile(os.path.join(REF, "test/input/coo.zip")) as z:
        txt = z.read("coo/retangulo.dat").decode()
    lines = txt.splitlines()
    i0 = next(i for i, l in enumerate(lines) if l.strip() == "cells") + 1
    i1 = next(i for i, l in enumerate(lines) if l.strip().startswith("endCell"))
    with open(os.path.join(HERE, "retangulo_cells.txt"), "w") as f:
        f.write("# cells of coo/retangulo.dat (reference test/input/coo.zip):"
                " id mat type nen ndm? n1 n2 n3 n4\n")
        for l in lines[i0:i1]:
            f.write(" ".join(l.split()) + "\n")
    for name in ("retangulo_bin.mtx", "retangulo_rcm_bin.mtx"):
        shutil.copy(os.path.join(REF, "test/output/coo", name),
                    os.path.join(HERE, name))
    os.makedirs(os.path.join(HERE, "solver_config"), exist_ok=True)
    for name in ("solverPres.dat", "solverVel.dat"):
        shutil.copy(os.path.join(REF, "exemplos/lid_driven_cavity/config", name),
                    os.path.join(HERE, "solver_config", name))


def ref_vectors():
    import oracle
    from sfr3d_amd import structure as S, systems as Y
    R = oracle.ref()
    R.ref_init(1 << 32)
    out = {"rcm": {}, "solves": []}
    meshes = [("hex", (5, 4, 3)), ("hex", (9, 8, 7)), ("hex", (32, 32, 32)),
              ("hex", (96, 96, 96)), ("hex", (128, 128, 128)),
              ("quad", (128, 128)), ("quad", (20, 10))]
    for kind, dims in meshes:
        nelcon, nviz = (S.hex_adjacency(*dims) if kind == "hex"
                        else S.quad_adjacency(*dims))
        numel, mv = nelcon.shape
        num = np.zeros(numel, np.int32)
        R.ref_reord(num, np.ascontiguousarray(nelcon.ravel()), nviz, mv,
                    numel, 1)
        out["rcm"][f"{kind}_{'x'.join(map(str, dims))}"] = {
            "sha256": sha(num), "head": num[:16].tolist(),
            "tail": num[-16:].tolist()}
        print("rcm", kind, dims, flush=True)
    sols = {}
    cases = [
        ("pcg_poisson3d_12_csrd", lambda: Y.poisson3d(12, seed=3, storage=2), 1, None),
        ("pcg_poisson3d_12_csrc", lambda: Y.poisson3d(12, seed=3, storage=3), 1, None),
        ("pcg_poisson3d_12_ell", lambda: Y.poisson3d(12, seed=3, storage=4), 1, None),
        ("pcg_poisson2d_20x16_pinned", lambda: Y.poisson2d(20, 16, seed=5, storage=2), 1, None),
        ("bicg_momentum_10_csrd", lambda: Y.momentum3d(10), 2, 0),
        ("bicg_momentum_10_csrd_w", lambda: Y.momentum3d(10), 2, 2),
        ("bicg_momentum_10_csrc", lambda: Y.momentum3d(10, storage=3), 2, 1),
        ("bicg_poisson3d_10_ell", lambda: Y.poisson3d(10, seed=4, storage=4), 2, None),
    ]
    for name, make, solver, comp in cases:
        B = make()
        s = B.system
        b = B.b if comp is None else B.extra["b_all"][comp]
        x = np.zeros(s.neq)
        res = np.zeros(4)
        ad = s.ad if s.ad is not None else s.al
        rc = R.ref_solve(s.neq, s.nad, s.ia, s.ja, np.ascontiguousarray(s.al),
                         np.ascontiguousarray(ad), None,
                         np.ascontiguousarray(b), x, 1e-8, 5000, s.storage,
                         solver, 1, int(s.unsym), res, None, 0)
        assert rc == 0
        out["solves"].append({"name": name, "iters": int(res[0]),
                              "xKx": res[1], "norm_x": res[2],
                              "x_sha256": sha(x)})
        sols[name] = x
        print("solve", name, int(res[0]), flush=True)
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_solutions.npz"), **sols)


GMRES_CASES = [
    ("gmres_poisson3d_12_csrd", "poisson3d", dict(n=12, seed=3, storage=2), None),
    ("gmres_poisson3d_12_ell", "poisson3d", dict(n=12, seed=3, storage=4), None),
    ("gmres_momentum_10_csrd", "momentum3d", dict(n=10), 0),
    ("gmres_momentum_10_csrc", "momentum3d", dict(n=10, storage=3), 1),
    ("gmres_momentum_10_ell", "momentum3d", dict(n=10, storage=4), 2),
    ("gmres_poisson2d_20x16_pinned", "poisson2d", dict(nx=20, ny=16, seed=5, storage=2), None),
]


def gmres_vectors():
    """solverC with solver GMRES (4), tol 1e-8 and 1e-11, maxIt 5000."""
    import oracle
    from sfr3d_amd import systems as Y
    R = oracle.ref()
    R.ref_init(1 << 30)
    out, sols = {"solves": []}, {}
    for name, fn, kw, comp in GMRES_CASES:
        B = getattr(Y, fn)(**kw)
        s = B.system
        b = B.b if comp is None else B.extra["b_all"][comp]
        for tol in (1e-8, 1e-11):
            x = np.zeros(s.neq)
            res = np.zeros(5)
            ad = s.ad if s.ad is not None else s.al
            rc = R.ref_solve(s.neq, s.nad, s.ia, s.ja, oracle._al_contig(s),
                             np.ascontiguousarray(ad), None,
                             np.ascontiguousarray(b), x, tol, 5000, s.storage,
                             4, 1, int(s.unsym), res, None, 0)
            assert rc == 0
            key = f"{name}_tol{tol:.0e}"
            out["solves"].append({"name": key, "iters": int(res[0]),
                                  "cycles": int(res[4]), "xKx": res[1],
                                  "norm_x": res[2], "x_sha256": sha(x)})
            sols[key] = x
            print("gmres", key, int(res[4]), int(res[0]), flush=True)
    with open(os.path.join(HERE, "ref_gmres.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_gmres_solutions.npz"), **sols)


BICGL2_CASES = [
    ("bicgl2_momentum_10_csrd", "momentum3d", dict(n=10), 0, 1e-8),
    ("bicgl2_momentum_10_csrc", "momentum3d", dict(n=10, storage=3), 1, 1e-11),
    ("bicgl2_momentum_10_ell", "momentum3d", dict(n=10, storage=4), 2, 1e-8),
    ("bicgl2_poisson3d_12_csrd", "poisson3d", dict(n=12, seed=3, storage=2), None, 1e-8),
    ("bicgl2_poisson2d_20x16_pinned", "poisson2d", dict(nx=20, ny=16, seed=5, storage=2), None, 1e-8),
]


def bicgl2_vectors():
    """solverC with solver PBICGSTABL2 (3), maxIt 5000."""
    import oracle
    from sfr3d_amd import systems as Y
    R = oracle.ref()
    R.ref_init(1 << 30)
    out, sols = {"solves": []}, {}
    for name, fn, kw, comp, tol in BICGL2_CASES:
        B = getattr(Y, fn)(**kw)
        s = B.system
        b = B.b if comp is None else B.extra["b_all"][comp]
        x = np.zeros(s.neq)
        res = np.zeros(5)
        ad = s.ad if s.ad is not None else s.al
        rc = R.ref_solve(s.neq, s.nad, s.ia, s.ja, oracle._al_contig(s),
                         np.ascontiguousarray(ad), None, np.ascontiguousarray(b),
                         x, tol, 5000, s.storage, 3, 1, int(s.unsym), res, None, 0)
        assert rc == 0
        out["solves"].append({"name": name, "tol": tol, "iters": int(res[0]),
                              "xKx": res[1], "norm_x": res[2], "x_sha256": sha(x)})
        sols[name] = x
        print("bicgl2", name, int(res[0]), flush=True)
    with open(os.path.join(HERE, "ref_bicgl2.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_bicgl2_solutions.npz"), **sols)


MINRES_CASES = [
    ("minres_poisson3d_12_csrd", "poisson3d", dict(n=12, seed=3, storage=2), 1e-8, 5000),
    ("minres_poisson3d_12_csrc", "poisson3d", dict(n=12, seed=3, storage=3), 1e-11, 5000),
    ("minres_poisson3d_12_ell", "poisson3d", dict(n=12, seed=3, storage=4), 1e-8, 5000),
    ("minres_poisson2d_20x16_pinned", "poisson2d", dict(nx=20, ny=16, seed=5, storage=2), 1e-8, 5000),
    ("minres_poisson3d_10_maxit", "poisson3d", dict(n=10, seed=7, storage=2), 1e-10, 9),
]


def minres_vectors():
    """solverC with solver MINRES (5).  The reference's log line reports
    maxIt + 1 iterations when the limit is reached (src/Iters.c:540-552)."""
    import oracle
    from sfr3d_amd import systems as Y
    R = oracle.ref()
    R.ref_init(1 << 30)
    out, sols = {"solves": []}, {}
    for name, fn, kw, tol, maxit in MINRES_CASES:
        B = getattr(Y, fn)(**kw)
        s = B.system
        x = np.zeros(s.neq)
        res = np.zeros(5)
        ad = s.ad if s.ad is not None else s.al
        rc = R.ref_solve(s.neq, s.nad, s.ia, s.ja, oracle._al_contig(s),
                         np.ascontiguousarray(ad), None, np.ascontiguousarray(B.b),
                         x, tol, maxit, s.storage, 5, 1, int(s.unsym), res, None, 0)
        assert rc == 0
        out["solves"].append({"name": name, "tol": tol, "maxit": maxit,
                              "log_iters": int(res[0]), "xKx": res[1],
                              "norm_x": res[2], "x_sha256": sha(x)})
        sols[name] = x
        print("minres", name, int(res[0]), flush=True)
    with open(os.path.join(HERE, "ref_minres.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_minres_solutions.npz"), **sols)


CAVITY_N, CAVITY_OUTER, CAVITY_PICK = 128, 12, (0, 5, 11)


def cavity_inputs_digest(B):
    s = B.system
    return sha(np.concatenate([s.ia.view(np.uint8), s.ja.view(np.uint8),
                               s.al.view(np.uint8), s.ad[:s.neq].view(np.uint8),
                               B.b.view(np.uint8)]))


def cavity_vectors():
    import oracle
    from sfr3d_amd import systems as Y
    R = oracle.ref()
    R.ref_init(1 << 31)
    sysl, fields = Y.cavity2d(n=CAVITY_N, outer=CAVITY_OUTER)
    out = {"n": CAVITY_N, "outer": CAVITY_OUTER, "mass": fields["mass"].tolist(),
           "solves": []}
    for k in CAVITY_PICK:
        for st, B in sysl[k].items():
            s = B.system
            x = np.zeros(s.neq)
            res = np.zeros(4)
            rc = R.ref_solve(s.neq, s.nad, s.ia, s.ja, np.ascontiguousarray(s.al),
                             np.ascontiguousarray(s.ad), None,
                             np.ascontiguousarray(B.b), x, 1e-8, 50000, s.storage,
                             1, 1, 0, res, None, 0)
            assert rc == 0
            out["solves"].append({"outer": k, "storage": st, "iters": int(res[0]),
                                  "xKx": res[1], "norm_x": res[2],
                                  "inputs_sha256": cavity_inputs_digest(B),
                                  "x_sha256": sha(x)})
            print("cavity", k, st, int(res[0]), flush=True)
    with open(os.path.join(HERE, "cavity_ref.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    what = sys.argv[1:] or ["mesh", "ref", "gmres", "minres", "bicgl2", "cavity"]
    if "mesh" in what:
        mesh_fixture()
    if "ref" in what:
        ref_vectors()
    if "gmres" in what:
        gmres_vectors()
    if "minres" in what:
        minres_vectors()
    if "bicgl2" in what:
        bicgl2_vectors()
    if "cavity" in what:
        cavity_vectors()
