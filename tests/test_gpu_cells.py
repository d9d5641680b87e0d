"""Cell coefficients on the device (sfr_cells_create / sfr_cells_pres_device):
SFR3D's SIMPLE pressure-correction cell loop (systFormSimplePres,
src/CellLoopInc.c:471-679; cellSimplePres3D, src/CellPres.c:903-1149) run on
the GPU, against
  - the reference's own compiled loop through its golden fixture
    (tests/golden/ref_simple_pres.npz), after the bit-exact host assembly;
  - the oracle's restatement (oracle/cells.c, itself bit-exact with the
    reference) on larger warped meshes, both diffusion schemes, every load
    type the pressure cell reads and a cell without an equation.
Bar: bitwise (the kernel keeps the reference's operation order with no
fused multiply-adds; double division and square root are correctly
rounded).  Then the whole chain cells -> device assembly -> solve against
the host chain, and an outer loop whose fields, coefficients and matrix
never leave the device."""
import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import structure as ST, systems as Y

from tests.test_cells import golden_cases

pytestmark = pytest.mark.gpu
FIELDS = ("pres", "gradPres", "vel", "dField")


def device_coeffs(ctx, mesh, case, cells=None):
    numel, mv = mesh["numel"], mesh["maxViz"]
    C = cells or ctx.cells(mesh, case)
    d = {k: ctx.array(np.ascontiguousarray(case[k], np.float64)) for k in FIELDS}
    d_lA = ctx.array(np.zeros(numel * (mv + 1)))
    d_lB = ctx.array(np.zeros(numel))
    d_rc = ctx.array(np.full(numel, 7.0))
    C.pres(case["diffusion"], d["pres"], d["gradPres"], d["vel"], d["dField"],
           d_lA, d_lB, d_rc)
    out = d_lA.get().reshape(numel, mv + 1), d_lB.get(), d_rc.get()
    for a in list(d.values()) + [d_lA, d_lB, d_rc]:
        a.free()
    if cells is None:
        C.destroy()
    return out


@pytest.mark.parametrize("i", range(4))
def test_device_matches_reference_golden(ctx, i):
    mesh, case, exp, meta = golden_cases()[i]
    print(f"cells golden case {i} (diffusion {meta[0]}, skip {meta[1]})", flush=True)
    lA, lB, rc = device_coeffs(ctx, mesh, case)
    nelcon = mesh["nelcon"].reshape(-1, mesh["maxViz"])
    pat = ST.datastruct(2, 0, case["id"], case["num"], nelcon, mesh["nFace"],
                        int(case["id"].max()))
    al, ad, b = ST.assemble(pat, case["id"], nelcon, mesh["nFace"], lA, lB)
    assert np.array_equal(al[:len(exp["al"])], exp["al"])
    assert np.array_equal(ad[:len(exp["ad"])], exp["ad"])
    assert np.array_equal(b[:len(exp["b"])], exp["b"])
    assert np.array_equal(rc, exp["rCell"])


@pytest.mark.parametrize("bc,diff", [("mixed", 1), ("mixed", 4), ("cavity", 1)])
def test_device_matches_oracle(ctx, bc, diff):
    m = Y.hex_mesh(24, 20, 18, ratio=(1.02, 0.98, 1.01), shear=(0.1, 0.05, -0.1),
                   warp=0.03)
    case = Y.simple_pres_case(m, seed=5, diffusion=diff, boundary=bc,
                              skip_cell=1234)
    print(f"cells 24x20x18 {bc} diffusion {diff}", flush=True)
    lA, lB, rc = device_coeffs(ctx, m, case)
    olA, olB, orc = oracle.simple_pres3d(m, case)
    assert np.array_equal(lA, olA)
    assert np.array_equal(lB, olB)
    assert np.array_equal(rc, orc)


def test_bad_arguments(ctx):
    m = Y.hex_mesh(3, 3, 3)
    case = Y.simple_pres_case(m, boundary="mixed")
    bad = dict(case, faceRpres=case["faceRpres"].copy())
    bad["faceRpres"][0] = 99                      # no such load
    with pytest.raises(S.SolverError):
        ctx.cells(m, bad)
    C = ctx.cells(m, case)
    d = ctx.array(np.zeros(3 * m["numel"]))
    with pytest.raises(S.SolverError):
        C.pres(2, d, d, d, d, d, d)               # MINIMAL: undefined upstream
    d.free()
    C.destroy()


def _pinned_case(n, seed):
    m = Y.hex_mesh(n, n, n, ratio=(1.03, 1.0, 0.97), warp=0.02)
    case = Y.simple_pres_case(m, seed=seed, diffusion=4, boundary="cavity")
    nel = m["nelcon"].reshape(-1, 6)
    rt = np.zeros((m["numel"], 7), np.int16)
    rt[m["numel"] // 2, 6] = 1                    # the pinned cell
    num = np.arange(1, m["numel"] + 1, dtype=np.int32)
    id_, neq = ST.numeq_v2(num, rt, m["nFace"], 6)
    case["id"], case["num"] = id_, num
    pat = ST.datastruct(S.CSRD, 0, id_, num, nel, m["nFace"], neq)
    return m, case, nel, pat, neq


def test_chain_cells_assembly_solve(ctx):
    """Coefficients, matrix and right-hand side made on the device solve
    bitwise like the host chain (oracle coefficients, host assembly, upload)."""
    m, case, nel, pat, neq = _pinned_case(40, 9)
    print(f"cells chain 40^3, neq {neq}", flush=True)
    olA, olB, _ = oracle.simple_pres3d(m, case)
    al, ad, b = ST.assemble(pat, case["id"], nel, m["nFace"], olA, olB)
    sysm = S.System(neq, pat.nad, pat.ia, pat.ja, al, ad, S.CSRD, False)
    A_host = ctx.matrix(sysm)
    A_dev = ctx.matrix(sysm)
    plan = ctx.assembly(A_dev, pat, case["id"], nel, m["nFace"])
    numel = m["numel"]
    d_lA, d_lB, d_b = ctx.array(np.zeros(numel * 7)), ctx.array(np.zeros(numel)), ctx.empty(neq)
    C = ctx.cells(m, case)
    d = {k: ctx.array(case[k]) for k in FIELDS}
    C.pres(4, d["pres"], d["gradPres"], d["vel"], d["dField"], d_lA, d_lB)
    plan.assemble(d_lA, d_lB, d_b)
    assert np.array_equal(d_b.get(), b[:neq])
    hb = ctx.array(b[:neq].copy())
    x1, x2 = ctx.empty(neq), ctx.empty(neq)
    st1 = ctx.solve_device(A_host, S.PCG, 1e-8, 5000, hb, x1)
    st2 = ctx.solve_device(A_dev, S.PCG, 1e-8, 5000, d_b, x2)
    assert st1["status"] == st2["status"] == S._lib.OK
    assert st1["iters"] == st2["iters"]
    assert np.array_equal(x1.get(), x2.get())
    for a in list(d.values()) + [d_lA, d_lB, d_b, hb, x1, x2]:
        a.free()
    C.destroy(); plan.destroy(); A_host.destroy(); A_dev.destroy()


def test_outer_loop_stays_on_device(ctx):
    """Three SIMPLE outer iterations of the pressure correction with
    everything on the device: cell coefficients, assembly, solve, the
    correction back to the cells (updateCellValue) and the update
    vel -= d grad(p'), p += 0.3 p' (simpleUpdate, with a fixed correction
    gradient: gradient reconstruction stays with the caller).  No value
    crosses the host link between iterations.  The host loop (oracle
    coefficients, host assembly, value upload, numpy updates) gives bitwise
    the same pressure and velocity after every iteration."""
    m, case, nel, pat, neq = _pinned_case(32, 13)
    print(f"cells outer loop 32^3, neq {neq}", flush=True)
    numel = m["numel"]
    id_ = case["id"]
    idx = np.nonzero(id_ > 0)[0]
    gpc = 0.5 * case["gradPres"][::-1].copy()           # the fixed p' gradient
    lA0, lB0, _ = oracle.simple_pres3d(m, case)
    al0, ad0, _ = ST.assemble(pat, id_, nel, m["nFace"], lA0, lB0)
    sysm = S.System(neq, pat.nad, pat.ia, pat.ja, al0, ad0, S.CSRD, False)
    A_dev, A_host = ctx.matrix(sysm), ctx.matrix(sysm)
    plan = ctx.assembly(A_dev, pat, id_, nel, m["nFace"])
    C = ctx.cells(m, case)
    d = {k: ctx.array(case[k]) for k in FIELDS}
    d_id, d_gpc = ctx.array(id_), ctx.array(gpc)
    d_lA, d_lB = ctx.array(np.zeros(numel * 7)), ctx.array(np.zeros(numel))
    d_b, d_x, d_pc = ctx.empty(neq), ctx.empty(neq), ctx.array(np.zeros(numel))
    h = {k: case[k].copy() for k in FIELDS}
    hpc = np.zeros(numel)
    hb, hx = ctx.empty(neq), ctx.empty(neq)
    for it in range(3):
        C.pres(4, d["pres"], d["gradPres"], d["vel"], d["dField"], d_lA, d_lB)
        plan.assemble(d_lA, d_lB, d_b)
        st = ctx.solve_device(A_dev, S.PCG, 1e-8, 5000, d_b, d_x)
        ctx.cells_update_value(d_id, d_x, d_pc)
        ctx.simple_update(0.3, d["vel"], d["pres"], d_pc, d_gpc, d["dField"])
        # host loop
        olA, olB, _ = oracle.simple_pres3d(m, dict(case, **h))
        al, ad, bh = ST.assemble(pat, id_, nel, m["nFace"], olA, olB)
        A_host.set_values(al, ad)
        hb.set(bh[:neq].copy())
        sth = ctx.solve_device(A_host, S.PCG, 1e-8, 5000, hb, hx)
        hpc[idx] = hx.get()[id_[idx] - 1]
        h["vel"] = h["vel"] - h["dField"] * gpc
        h["pres"] = h["pres"] + 0.3 * hpc
        assert st["status"] == sth["status"] == S._lib.OK, it
        assert st["iters"] == sth["iters"], it
        assert np.array_equal(d["pres"].get(), h["pres"]), it
        assert np.array_equal(d["vel"].get(), h["vel"]), it
    for a in list(d.values()) + [d_id, d_gpc, d_lA, d_lB, d_b, d_x, d_pc, hb, hx]:
        a.free()
    C.destroy(); plan.destroy(); A_dev.destroy(); A_host.destroy()
