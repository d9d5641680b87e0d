"""Device assembly (sfr_assembly_create / sfr_assemble_device): per-cell
coefficients that live on the GPU are scattered into a resident matrix with
the semantics of the reference's assbly / csr / ellPack (src/Assbly.c:43-95,
src/Csr.c:549-975, src/EllPack.c:113-149), which the host assembler
sfr_assemble reproduces bit for bit (tests/test_structure.py).

Bar: bitwise.  A matrix re-assembled on the device from new coefficients
must hold exactly the values sfr_assemble + sfr_matrix_set_values give, so
the same solve returns the same x bit for bit, and b is the same array."""
import numpy as np
import pytest

import sfr3d_amd as S
from sfr3d_amd import structure as ST, systems

pytestmark = pytest.mark.gpu

TOL = 1e-8


def poisson_case(storage, pinned=None, unsym=False):
    B = systems.poisson3d(10, seed=3, storage=storage, reorder=True,
                          pinned_cell=pinned)
    # new coefficients of the same structure: another conductance field
    B2 = systems.poisson3d(10, seed=4, storage=storage, reorder=True,
                           pinned_cell=pinned)
    return B, B2.extra["lA"], B2.extra["lB"], S.PCG


def momentum_case(storage):
    M = systems.momentum3d(8, storage=storage)
    lA2 = M.extra["lA"].copy()
    lA2[:, :6] *= 1.1                        # new off-diagonals, same pattern
    rhs = M.extra["b_all"]
    return M, lA2, None, S.PBICGSTAB


CASES = [
    ("csrd", lambda: poisson_case(S.CSRD)),
    ("csrc", lambda: poisson_case(S.CSRC)),
    ("ell", lambda: poisson_case(S.ELLPACK)),
    ("ell_pinned", lambda: poisson_case(S.ELLPACK, pinned=123)),
    ("csr", lambda: poisson_case(S.CSR)),
    ("csrd_unsym", lambda: momentum_case(S.CSRD)),
    ("csrc_unsym", lambda: momentum_case(S.CSRC)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_device_assembly_matches_host(ctx, name, make):
    B, lA2, lB2, solver = make()
    print(f"device assembly {name}", flush=True)
    s = B.system
    pat = B.pattern
    lB2 = np.random.default_rng(1).uniform(-1, 1, lA2.shape[0]) if lB2 is None else lB2
    # host reference: sfr_assemble of the new coefficients
    al, ad, b = ST.assemble(pat, B.id, B.nelcon, B.nviz, lA2, lB2)
    b = b[:s.neq].copy()
    A_host = ctx.matrix(s)
    A_host.set_values(al, None if s.storage == S.CSR else ad)
    # device: same structure, coefficients uploaded, scattered on the GPU
    A_dev = ctx.matrix(s)
    plan = ctx.assembly(A_dev, pat, B.id, B.nelcon, B.nviz)
    d_lA, d_lB = ctx.array(np.ascontiguousarray(lA2)), ctx.array(lB2)
    d_b = ctx.empty(s.neq)
    plan.assemble(d_lA, d_lB, d_b)
    assert np.array_equal(d_b.get(), b)
    hb = ctx.array(b)
    x1, x2 = ctx.empty(s.neq), ctx.empty(s.neq)
    st1 = ctx.solve_device(A_host, solver, TOL, 5000, hb, x1)
    st2 = ctx.solve_device(A_dev, solver, TOL, 5000, d_b, x2)
    assert st1["status"] == st2["status"] == S._lib.OK
    assert st1["iters"] == st2["iters"]
    assert np.array_equal(x1.get(), x2.get())           # bit for bit
    # and again with the mixed precision path (FP32 copy rebuilt)
    if solver == S.PCG:
        st1 = ctx.solve_device(A_host, solver, TOL, 5000, hb, x1, precision=S.PREC_MIXED)
        st2 = ctx.solve_device(A_dev, solver, TOL, 5000, d_b, x2, precision=S.PREC_MIXED)
        assert st1["iters"] == st2["iters"]
        assert np.array_equal(x1.get(), x2.get())
    plan.destroy(); A_host.destroy(); A_dev.destroy()
    for d in (d_lA, d_lB, d_b, hb, x1, x2):
        d.free()


def test_assembly_plan_rejects_a_foreign_pattern(ctx):
    B = systems.poisson3d(6, seed=1, storage=S.CSRD)
    C2 = systems.poisson3d(7, seed=1, storage=S.CSRD)
    A = ctx.matrix(B.system)
    with pytest.raises(S.SolverError):
        ctx.assembly(A, C2.pattern, C2.id, C2.nelcon, C2.nviz)
    A.destroy()
