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
from sfr3d_amd._lib import lib

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


# ---- the momentum cell loop ---------------------------------------------------

from tests.test_cells import golden_vel_cases, vel_assemble  # noqa: E402
import ctypes as C  # noqa: E402

VEL_FIELDS = ("pres", "gradPres", "vel", "gradVel", "dField")


def device_vel(ctx, mesh, case, cells=None):
    numel, mv = mesh["numel"], mesh["maxViz"]
    C_ = cells or ctx.cells(mesh, dict(case, loadsPres=[]))
    neq = int(case["id"].max())
    d = {k: ctx.array(np.ascontiguousarray(case[k], np.float64)) for k in VEL_FIELDS}
    o = case["opts"]
    d_eddy = ctx.array(case["eddyVisc"]) if o["turb"] else None
    d_wall = ctx.array(case["wallPar"]) if o["wall_model"] else None
    d_id = ctx.array(np.ascontiguousarray(case["id"], np.int32))
    d_out = ctx.array(np.zeros(3 * numel))
    d_lA = ctx.array(np.zeros(numel * (mv + 1)))
    d_ad, d_b = ctx.array(np.zeros(3 * neq)), ctx.array(np.zeros(3 * neq))
    d_rc = ctx.array(np.full(3 * numel, 7.0))
    C_.vel(o, d_id, neq, d["pres"], d["gradPres"], d["vel"], d["gradVel"], d_eddy,
           d_wall, d["dField"], d_out, d_lA, d_ad, d_b, d_rc)
    out = (d_lA.get().reshape(numel, mv + 1), d_ad.get(), d_b.get(),
           d_rc.get(), d_out.get())
    for a in list(d.values()) + [d_id, d_out, d_lA, d_ad, d_b, d_rc, d_eddy, d_wall]:
        if a is not None:
            a.free()
    if cells is None:
        C_.destroy()
    return out


def oracle_vel_expect(mesh, case):
    """The oracle's order-free momentum loop in the device's output layout."""
    lA, lB, rc, dn = oracle.simple_vel3d(mesh, case, in_place=False)
    pat, al, ad, b = vel_assemble(mesh, case, lA, lB)
    numel, mv = mesh["numel"], mesh["maxViz"]
    skip = case["faceRvel"].reshape(numel, -1)[np.arange(numel), mesh["nFace"]] > 0
    l1 = np.zeros((numel, mv + 1))
    l1[~skip] = lA[~skip, :mv + 1]
    return l1, ad, b, rc.ravel(), dn.ravel()


@pytest.mark.parametrize("i", range(6))
def test_vel_device_matches_reference_golden(ctx, i):
    """Without Rhie-Chow interpolation the device's coefficients are the
    reference's own (assembled by the bit-exact host assembler); with it, the
    device takes every neighbour's input d and is compared with the oracle's
    order-free form (itself pinned against the reference where the two
    orders agree)."""
    mesh, case, exp = golden_vel_cases()[i]
    print(f"momentum golden case {i} {case['opts']}", flush=True)
    lA, ad, b, rc, dn = device_vel(ctx, mesh, case)
    if not case["opts"]["rhie"]:
        numel = mesh["numel"]
        nel = mesh["nelcon"].reshape(-1, 6)
        neq = int(case["id"].max())
        pat = ST.datastruct(2, 1, case["id"], case["num"], nel, mesh["nFace"], neq)
        al, _, _ = ST.assemble(pat, case["id"], nel, mesh["nFace"], lA, np.zeros(numel))
        assert np.array_equal(al[:len(exp["al"])], exp["al"])
        assert np.array_equal(ad, exp["ad"]) and np.array_equal(b, exp["b"])
        assert np.array_equal(rc, exp["rCell"]) and np.array_equal(dn, exp["dOut"])
    l1, oad, ob, orc, odn = oracle_vel_expect(mesh, case)
    assert np.array_equal(lA, l1)
    assert np.array_equal(ad, oad) and np.array_equal(b, ob)
    assert np.array_equal(rc, orc) and np.array_equal(dn, odn)


@pytest.mark.parametrize("kw", [dict(), dict(advection=1, diffusion=1, time="euler",
                                             simplec=True, residual=False),
                                dict(advection=2, diffusion=2, boundary="channel",
                                     wall_model=True, fixed_cell=777, time="steady"),
                                dict(diffusion=3, so_pressure=False, viscosity=False)])
def test_vel_device_matches_oracle(ctx, kw):
    m = Y.hex_mesh(24, 20, 18, ratio=(1.02, 0.98, 1.01), shear=(0.1, 0.05, -0.1),
                   warp=0.03)
    case = Y.simple_vel_case(m, seed=9, **kw)
    print(f"momentum 24x20x18 {kw}", flush=True)
    got = device_vel(ctx, m, case)
    exp = oracle_vel_expect(m, case)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_vel_chain_shared_solve(ctx):
    """u, v, w coefficients made on the device, the shared off-diagonals
    scattered by the device assembly, the three diagonals and right-hand
    sides written in csrSimple's layout: the u/v/w BiCGSTAB batch solves
    bitwise like the host chain (oracle coefficients, host assembly)."""
    m = Y.hex_mesh(32, 30, 28, ratio=(1.01, 1.0, 0.99), warp=0.02)
    case = Y.simple_vel_case(m, seed=21, boundary="channel", rhie=True)
    numel, neq = m["numel"], int(case["id"].max())
    nel = m["nelcon"].reshape(-1, 6)
    print(f"momentum chain {numel} cells", flush=True)
    l1, ad, b, _, _ = oracle_vel_expect(m, case)
    pat = ST.datastruct(S.CSRD, 1, case["id"], case["num"], nel, m["nFace"], neq)
    al, _, _ = ST.assemble(pat, case["id"], nel, m["nFace"], l1, np.zeros(numel))
    sysm = S.System(neq, pat.nad, pat.ia, pat.ja, al, ad[:neq].copy(), S.CSRD, True)
    A_host, A_dev = ctx.matrix(sysm), ctx.matrix(sysm)
    plan = ctx.assembly(A_dev, pat, case["id"], nel, m["nFace"])
    cells = ctx.cells(m, dict(case, loadsPres=[]))
    d = {k: ctx.array(np.ascontiguousarray(case[k], np.float64)) for k in VEL_FIELDS}
    d_eddy, d_id = ctx.array(case["eddyVisc"]), ctx.array(case["id"])
    d_out, d_lA = ctx.array(np.zeros(3 * numel)), ctx.array(np.zeros(numel * 7))
    d_ad, d_b, d_rc = ctx.empty(3 * neq), ctx.empty(3 * neq), ctx.empty(3 * numel)
    cells.vel(case["opts"], d_id, neq, d["pres"], d["gradPres"], d["vel"], d["gradVel"],
              d_eddy, None, d["dField"], d_out, d_lA, d_ad, d_b, d_rc)
    plan.assemble(d_lA)
    h_ad, h_b = ctx.array(ad), ctx.array(b)
    x1, x2 = ctx.array(np.zeros(3 * neq)), ctx.array(np.zeros(3 * neq))
    st1, st2 = (S._lib.SfrStats * 3)(), (S._lib.SfrStats * 3)()
    rc1 = lib.sfr_solve_shared_device(ctx.h, A_host.h, 3, h_ad.ptr, h_b.ptr, x1.ptr,
                                      S.PBICGSTAB, 1e-8, 2000, 1, st1)
    rc2 = lib.sfr_solve_shared_device(ctx.h, A_dev.h, 3, d_ad.ptr, d_b.ptr, x2.ptr,
                                      S.PBICGSTAB, 1e-8, 2000, 1, st2)
    assert rc1 == rc2 == 0
    assert [st1[c].iters for c in range(3)] == [st2[c].iters for c in range(3)]
    assert np.array_equal(d_ad.get(), ad) and np.array_equal(d_b.get(), b)
    assert np.array_equal(x1.get(), x2.get())
    for a in list(d.values()) + [d_eddy, d_id, d_out, d_lA, d_ad, d_b, d_rc, h_ad,
                                 h_b, x1, x2]:
        a.free()
    cells.destroy(); plan.destroy(); A_host.destroy(); A_dev.destroy()
