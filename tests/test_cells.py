"""SIMPLE cell loops -- the pressure correction (systFormSimplePres,
src/CellLoopInc.c:471-679; cellSimplePres3D, src/CellPres.c:903-1149) and the
momentum (systFormSimpleVel, src/CellLoopInc.c:98-385; cellSimpleVel3D,
src/CellSimple.c:106-638): the oracle's restatements (oracle/cells.c) against
  - the reference's own compiled loop, through its golden fixture
    (tests/golden/ref_simple_pres.npz) and, where oracle/_ref is built, live
    on another mesh -- bit for bit, after the bit-exact assembly;
  - closed forms: the coefficient of a uniform orthogonal mesh, the
    pressure-gradient correction vanishing for a linear pressure field;
  - invariants: symmetric face coefficients, zero net interior mass flux.
"""
import os

import numpy as np
import pytest

import oracle
from sfr3d_amd import structure as S
from sfr3d_amd import systems as Y

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_simple_pres.npz")
MESH_KEYS = ("nFace", "nelcon", "cellFace", "owner", "volume", "dcca", "xmcc",
             "mksi", "ksi", "eta", "area", "normal", "xm", "mvSkew", "vSkew")
CASE_KEYS = ("mat", "prop", "pres", "gradPres", "vel", "dField", "faceRvel",
             "faceRpres", "id", "num")


def _loads(z, key):
    n = int(z[f"{key}_n"][0])
    out = []
    for i in range(n):
        k = int(z[f"{key}_np"][i])
        out.append({"type": int(z[f"{key}_type"][i]),
                    "par": list(z[f"{key}_par"][i * 8:i * 8 + k]),
                    "density": float(z[f"{key}_density"][i])})
    return out


def golden_cases():
    """(mesh, case, expected) triples of the fixture."""
    z = np.load(GOLD)
    nx, ny, nz, numel, nface, mv = (int(v) for v in z["mesh_dims"])
    mesh = {k: z[f"mesh_{k}"] for k in MESH_KEYS}
    mesh.update(nx=nx, ny=ny, nz=nz, numel=numel, nface=nface, maxViz=mv)
    out = []
    for i in range(int(z["ncases"][0])):
        case = {k: z[f"c{i}_{k}"] for k in CASE_KEYS}
        case["loadsVel"] = _loads(z, f"c{i}_lv")
        case["loadsPres"] = _loads(z, f"c{i}_lp")
        diff, skip, mixed = (int(v) for v in z[f"c{i}_meta"])
        case["diffusion"] = diff
        exp = {k: z[f"c{i}_{k}"] for k in ("ia", "ja", "al", "ad", "b", "rCell")}
        out.append((mesh, case, exp, (diff, skip, mixed)))
    return out


def assemble(mesh, case, lA, lB):
    nelcon = mesh["nelcon"].reshape(-1, mesh["maxViz"])
    pat = S.datastruct(2, 0, case["id"], case["num"], nelcon, mesh["nFace"],
                       int(case["id"].max()))
    al, ad, b = S.assemble(pat, case["id"], nelcon, mesh["nFace"], lA, lB)
    return pat, al, ad, b


@pytest.mark.parametrize("i", range(4))
def test_oracle_matches_reference_golden(i):
    mesh, case, exp, meta = golden_cases()[i]
    lA, lB, rc = oracle.simple_pres3d(mesh, case)
    pat, al, ad, b = assemble(mesh, case, lA, lB)
    assert np.array_equal(pat.ia, exp["ia"]) and np.array_equal(pat.ja, exp["ja"])
    n = len(exp["al"])
    assert np.array_equal(al[:n], exp["al"])
    assert np.array_equal(ad[:len(exp["ad"])], exp["ad"])
    assert np.array_equal(b[:len(exp["b"])], exp["b"])
    assert np.array_equal(rc, exp["rCell"])
    if meta[1] >= 0:                     # the skipped cell: row untouched
        row = case["id"][meta[1]] - 1
        assert exp["ad"][row] == 0.0 and exp["b"][row] == 0.0
        assert exp["rCell"][meta[1]] == 0.0


@pytest.mark.ref
@pytest.mark.parametrize("bc,diff", [("mixed", 1), ("mixed", 4), ("cavity", 4)])
def test_oracle_matches_compiled_reference(ref, bc, diff):
    m = Y.hex_mesh(7, 5, 4, ratio=(0.85, 1.25, 1.05), shear=(-0.1, 0.2, 0.05),
                   warp=0.06)
    case = Y.simple_pres_case(m, seed=11, diffusion=diff, boundary=bc,
                              skip_cell=20)
    lA, lB, rc = oracle.simple_pres3d(m, case)
    ia, ja, ral, rad, rb, rrc = oracle.ref_simple_pres(m, case)
    pat, al, ad, b = assemble(m, case, lA, lB)
    assert np.array_equal(al[:len(ral)], ral)
    assert np.array_equal(ad[:len(rad)], rad)
    assert np.array_equal(b[:len(rb)], rb)
    assert np.array_equal(rc, rrc)


def test_uniform_mesh_closed_form():
    """Uniform cubes of side h, one density rho, isotropic d: every interior
    coefficient is -rho d h^2 / h = -rho d h (either scheme, ksi = normal),
    the diagonal is their negated sum, and a uniform velocity with a linear
    pressure field (gradient exact) leaves no mass residual inside."""
    n = 6
    h = 1.0 / n
    m = Y.hex_mesh(n, n, n)
    case = Y.simple_pres_case(m, seed=1, boundary="cavity")
    rho, d = 1.3, 2e-3
    case["prop"][1] = case["prop"][11] = rho
    case["dField"] = np.full(3 * m["numel"], d)
    g = np.array([0.3, -0.7, 1.1])
    cc = m["cc"].reshape(-1, 3)
    case["pres"] = cc @ g
    case["gradPres"] = np.tile(g, m["numel"])
    case["vel"] = np.tile([0.4, 0.25, -0.6], m["numel"])
    for diff in (1, 4):
        case["diffusion"] = diff
        lA, lB, rc = oracle.simple_pres3d(m, case)
        nel = m["nelcon"].reshape(-1, 6)
        inner = nel > 0
        assert np.allclose(lA[:, :6][inner], -rho * d * h, rtol=1e-13, atol=0)
        assert np.all(lA[:, :6][~inner] == 0.0)
        assert np.allclose(lA[:, 6], -lA[:, :6].sum(axis=1), rtol=1e-13)
        interior = inner.all(axis=1)
        assert np.abs(lB[interior]).max() < 1e-14


@pytest.mark.parametrize("diff", [1, 4])
def test_symmetry_and_conservation(diff):
    """On a warped, graded mesh with two materials: the coefficient of a face
    is the same seen from both cells (interpolation weights add up to one,
    normal and ksi flip together), and interior mass fluxes cancel, so with
    impermeable walls the residuals sum to zero."""
    m = Y.hex_mesh(6, 5, 4, ratio=(1.3, 0.8, 1.1), shear=(0.1, -0.05, 0.2),
                   warp=0.05)
    case = Y.simple_pres_case(m, seed=7, diffusion=diff, boundary="cavity")
    lA, lB, rc = oracle.simple_pres3d(m, case)
    nel = m["nelcon"].reshape(-1, 6) - 1
    c, f = np.nonzero(nel >= 0)
    opp = Y.HEX_OPPOSITE[f]
    a, b = lA[c, f], lA[nel[c, f], opp]
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    assert abs(rc.sum()) <= 1e-13 * np.abs(rc).sum()
    assert np.array_equal(rc, lB)


def test_bad_scheme_refused():
    m = Y.hex_mesh(2, 2, 2)
    case = Y.simple_pres_case(m, diffusion=2)
    with pytest.raises(ValueError):
        oracle.simple_pres3d(m, case)


# ---- the momentum cell loop (systFormSimpleVel / cellSimpleVel3D) ------------

GOLD_VEL = os.path.join(os.path.dirname(__file__), "golden", "ref_simple_vel.npz")
VEL_KEYS = ("mat", "prop", "vel", "gradVel", "pres", "gradPres", "dField",
            "eddyVisc", "wallPar", "faceRvel", "faceRpres", "id", "num")
OPT_KEYS = ("advection", "diffusion", "turb", "wall_model", "rhie", "so_pressure",
            "viscosity", "residual", "simplec", "time", "backward")


def golden_vel_cases():
    """(mesh, case, expected) triples of the momentum fixture."""
    z = np.load(GOLD_VEL)
    nx, ny, nz, numel, nface, mv = (int(v) for v in z["mesh_dims"])
    mesh = {k: z[f"mesh_{k}"] for k in MESH_KEYS + ("cc",)}
    mesh.update(nx=nx, ny=ny, nz=nz, numel=numel, nface=nface, maxViz=mv)
    out = []
    for i in range(int(z["ncases"][0])):
        case = {k: z[f"v{i}_{k}"] for k in VEL_KEYS}
        case["loadsVel"] = _loads(z, f"v{i}_lv")
        opt, optd = z[f"v{i}_opt"], z[f"v{i}_optd"]
        case["opts"] = dict(zip(OPT_KEYS, (int(v) for v in opt)))
        case["opts"].update(dt=float(optd[0]), dt0=float(optd[1]), underU=float(optd[2]))
        exp = {k: z[f"v{i}_{k}"] for k in ("ia", "ja", "al", "ad", "b", "rCell", "dOut")}
        out.append((mesh, case, exp))
    return out


def vel_assemble(mesh, case, lA, lB):
    """The shared off-diagonals through the bit-exact host assembler, and the
    three diagonals and right-hand sides at c*neq + row (csrSimple's layout)."""
    numel = mesh["numel"]
    nel = mesh["nelcon"].reshape(-1, mesh["maxViz"])
    neq = int(case["id"].max())
    pat = S.datastruct(2, 1, case["id"], case["num"], nel, mesh["nFace"], neq)
    l1 = np.ascontiguousarray(lA[:, :mesh["maxViz"] + 1])
    al, _, _ = S.assemble(pat, case["id"], nel, mesh["nFace"], l1, np.zeros(numel))
    skip = case["faceRvel"].reshape(numel, -1)[np.arange(numel), mesh["nFace"]] > 0
    rows = case["id"] - 1
    ok = (rows >= 0) & ~skip
    ad = np.zeros(3 * neq)
    b = np.zeros(3 * neq)
    for k in range(3):
        ad[k * neq + rows[ok]] = lA[ok, mesh["maxViz"] + k]
        b[k * neq + rows[ok]] = lB[ok, k]
    return pat, al, ad, b


@pytest.mark.parametrize("i", range(6))
def test_vel_oracle_matches_reference_golden(i):
    """Every option set bit for bit, in the reference's own cell order (the d
    field written in place); without Rhie-Chow interpolation the order-free
    form is the same."""
    mesh, case, exp = golden_vel_cases()[i]
    for in_place in (True, False):
        if not in_place and case["opts"]["rhie"]:
            continue
        lA, lB, rc, d = oracle.simple_vel3d(mesh, case, in_place=in_place)
        pat, al, ad, b = vel_assemble(mesh, case, lA, lB)
        assert np.array_equal(pat.ia, exp["ia"]) and np.array_equal(pat.ja, exp["ja"])
        assert np.array_equal(al[:len(exp["al"])], exp["al"])
        assert np.array_equal(ad, exp["ad"]) and np.array_equal(b, exp["b"])
        assert np.array_equal(rc.ravel(), exp["rCell"])
        assert np.array_equal(d.ravel(), exp["dOut"])


@pytest.mark.ref
@pytest.mark.parametrize("kw", [dict(), dict(rhie=False, advection=2, diffusion=4,
                                             boundary="channel", fixed_cell=30),
                                dict(advection=1, diffusion=2, wall_model=True,
                                     simplec=True, time="euler")])
def test_vel_oracle_matches_compiled_reference(ref, kw):
    m = Y.hex_mesh(7, 5, 4, ratio=(0.85, 1.25, 1.05), shear=(-0.1, 0.2, 0.05),
                   warp=0.06)
    case = Y.simple_vel_case(m, seed=17, **kw)
    lA, lB, rc, d = oracle.simple_vel3d(m, case, in_place=True)
    ia, ja, ral, rad, rb, rrc, rd = oracle.ref_simple_vel(m, case)
    pat, al, ad, b = vel_assemble(m, case, lA, lB)
    assert np.array_equal(al[:len(ral)], ral)
    assert np.array_equal(ad, rad) and np.array_equal(b, rb)
    assert np.array_equal(rc.ravel(), rrc) and np.array_equal(d.ravel(), rd)


def test_vel_closed_form_and_symmetry():
    """Fluid at rest on uniform cubes of side h, one material of viscosity mu,
    no turbulence: every interior coefficient is -mu h^2 / h = -mu h (either
    diffusion scheme that has no explicit part), the diagonal is its negated
    sum plus the wall terms, over underU; on a warped mesh with two
    materials the face coefficients are the same seen from both cells (the
    harmonic mean is symmetric in the two cells)."""
    n, h = 5, 1.0 / 5
    m = Y.hex_mesh(n, n, n)
    case = Y.simple_vel_case(m, seed=2, advection=1, diffusion=1, turb=False,
                             rhie=False, time="steady", viscosity=False)
    mu = 2e-2
    case["prop"][0] = case["prop"][10] = mu
    case["vel"][:] = 0.0
    case["gradVel"][:] = 0.0
    lA, lB, rc, d = oracle.simple_vel3d(m, case)
    nel = m["nelcon"].reshape(-1, 6)
    inner = nel > 0
    assert np.allclose(lA[:, :6][inner], -mu * h, rtol=1e-13, atol=0)
    interior = inner.all(axis=1)
    under = case["opts"]["underU"]
    assert np.allclose(lA[interior, 6:9], 6 * mu * h / under, rtol=1e-13)
    # warped mesh, two materials: symmetric face coefficients at rest
    m = Y.hex_mesh(6, 5, 4, ratio=(1.3, 0.8, 1.1), shear=(0.1, -0.05, 0.2), warp=0.05)
    case = Y.simple_vel_case(m, seed=4, advection=3, diffusion=4, rhie=False)
    case["vel"][:] = 0.0
    lA, lB, rc, d = oracle.simple_vel3d(m, case)
    nel = m["nelcon"].reshape(-1, 6) - 1
    c, f = np.nonzero(nel >= 0)
    a, b = lA[c, f], lA[nel[c, f], Y.HEX_OPPOSITE[f]]
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
