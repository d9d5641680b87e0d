"""Pins the oracle (oracle/sfr_oracle.c, the CPU restatement of pcg /
pbicgstab / matvecs) before it is used to check the GPU:
  - bit-exact x and identical iteration counts against the reference's own
    compiled solverC (oracle/_ref), when built, and against the golden vectors
    that code produced (tests/golden/ref_*), which travel without it;
  - the matvecs against dense linear algebra (including the reference's two
    documented storage-CSR quirks);
  - the solutions against a direct sparse solve;
  - closed forms: a diagonal system converges in one iteration; the CG
    residual history is monotone in the A-norm of the error.
A mistake such as a dropped term, a wrong sign or a transposed operand in
the restatement fails one of these.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from sfr3d_amd import systems as Y
from sfr3d_amd.solver import System

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def dense(sysm):
    """Dense matrix the storage represents (mathematical definition)."""
    n = sysm.neq
    A = np.zeros((n, n))
    ia, ja, al = sysm.ia, sysm.ja, sysm.al
    st, uns = sysm.storage, sysm.unsym
    if st == 4:
        s = ia[0]
        for i in range(n):
            A[i, i] += sysm.ad[i]
            for f in range(s):
                A[i, ja[i * s + f]] += al[i * s + f]
        return A
    for i in range(n):
        if st != 1:
            A[i, i] = sysm.ad[i]
        for k in range(ia[i], ia[i + 1]):
            j = ja[k]
            if st == 1 and j == i:
                A[i, i] = al[k]
                continue
            A[i, j] = al[k]
            if st in (3, 6) and uns:
                A[j, i] = al[sysm.nad + k]
            elif not uns:
                A[j, i] = al[k]
    return A


CASES = [
    ("pcg_poisson3d_12_csrd", lambda: Y.poisson3d(12, seed=3, storage=2), "pcg", None),
    ("pcg_poisson3d_12_csrc", lambda: Y.poisson3d(12, seed=3, storage=3), "pcg", None),
    ("pcg_poisson3d_12_ell", lambda: Y.poisson3d(12, seed=3, storage=4), "pcg", None),
    ("pcg_poisson2d_20x16_pinned", lambda: Y.poisson2d(20, 16, seed=5, storage=2), "pcg", None),
    ("bicg_momentum_10_csrd", lambda: Y.momentum3d(10), "bicg", 0),
    ("bicg_momentum_10_csrd_w", lambda: Y.momentum3d(10), "bicg", 2),
    ("bicg_momentum_10_csrc", lambda: Y.momentum3d(10, storage=3), "bicg", 1),
    ("bicg_poisson3d_10_ell", lambda: Y.poisson3d(10, seed=4, storage=4), "bicg", None),
]


@pytest.mark.parametrize("name,make,kind,comp", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_reference_golden(name, make, kind, comp):
    gold = {g["name"]: g for g in json.load(open(os.path.join(GOLD, "ref_vectors.json")))["solves"]}[name]
    xs = np.load(os.path.join(GOLD, "ref_solutions.npz"))[name]
    B = make()
    b = B.b if comp is None else B.extra["b_all"][comp]
    f = oracle.pcg if kind == "pcg" else oracle.pbicgstab
    x, st = f(B.system, b, 1e-8, 5000)
    assert st["iters"] == gold["iters"]
    assert np.array_equal(x, xs)                   # bit for bit
    assert sha(x) == gold["x_sha256"]
    assert abs(st["xKx"] - gold["xKx"]) <= 1e-11 * abs(gold["xKx"])  # log line has 12 digits


@pytest.mark.ref
@pytest.mark.parametrize("storage", [2, 3, 4])
def test_oracle_matches_reference_code(ref, storage):
    import ctypes
    for seed in (1, 2):
        B = Y.poisson3d(9, seed=seed, storage=storage)
        s = B.system
        x = np.zeros(s.neq)
        out = np.zeros(4)
        buf = ctypes.create_string_buffer(512)
        assert ref.ref_solve(s.neq, s.nad, s.ia, s.ja, s.al, s.ad, None, B.b, x,
                             1e-10, 5000, storage, 1, 1, 0, out, buf, 512) == 0
        xo, so = oracle.pcg(s, B.b, 1e-10, 5000)
        assert so["iters"] == int(out[0]) and np.array_equal(x, xo)
        line = buf.value.decode()
        assert line.startswith("PCG: tol") and f"iteration {so['iters']} " in line


@pytest.mark.parametrize("storage,unsym", [(2, 0), (2, 1), (3, 0), (3, 1), (4, 0), (4, 1)])
def test_matvec_against_dense(storage, unsym):
    B = (Y.momentum3d(5, storage=storage) if unsym
         else Y.poisson3d(5, seed=2, storage=storage))
    A = dense(B.system)
    x = np.random.default_rng(0).standard_normal(B.system.neq)
    y = oracle.matvec(B.system, x)
    assert np.allclose(y, A @ x, rtol=1e-13, atol=1e-13)
    if not unsym:
        assert np.allclose(A, A.T)


def test_matvec_reference_quirks():
    # storage CSR, symmetric: matVecCsrSym zeroes y[i] after the scatter of
    # earlier rows, so only diag + upper survive (src/HccaBlas.c:3315-3325)
    B = Y.poisson3d(4, seed=1, storage=1)
    A = dense(B.system)
    x = np.random.default_rng(1).standard_normal(B.system.neq)
    assert np.allclose(oracle.matvec(B.system, x), np.triu(A) @ x)
    # ELLPACK widths other than 6/4/3 leave y untouched
    n, s = 4, 5
    sysm = System(n, 0, np.array([s, (n + 1) * s], np.int32),
                  np.zeros((n + 1) * s, np.int32), np.ones((n + 1) * s),
                  np.ones(n + 1), 4, True)
    assert not oracle.matvec(sysm, np.ones(n)).any()


@pytest.mark.parametrize("storage", [2, 3, 4])
def test_pcg_against_direct_solve(storage):
    B = Y.poisson3d(10, seed=7, storage=storage)
    A = sp.csr_matrix(dense(B.system))
    xd = spla.spsolve(A.tocsc(), B.b)
    x, st = oracle.pcg(B.system, B.b, 1e-12, 10000)
    assert st["status"] == 0
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_bicgstab_against_direct_solve():
    B = Y.momentum3d(8)
    A = sp.csr_matrix(dense(B.system))
    for comp in range(3):
        b = B.extra["b_all"][comp]
        xd = spla.spsolve(A.tocsc(), b)
        x, st = oracle.pbicgstab(B.system, b, 1e-12, 5000)
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_diagonal_system_one_iteration():
    n = 50
    d = np.linspace(1, 5, n)
    sysm = System(n, 0, np.zeros(n + 1, np.int32), np.zeros(1, np.int32),
                  np.zeros(1), d, 2, False)
    b = np.random.default_rng(2).standard_normal(n)
    x, st = oracle.pcg(sysm, b, 1e-12, 10)
    assert st["iters"] == 1 and np.allclose(x, b / d, rtol=1e-15)
    x, st = oracle.pbicgstab(sysm, b, 1e-12, 10)
    assert st["iters"] == 1 and np.allclose(x, b / d, rtol=1e-15)


def test_pcg_error_energy_norm_decreases():
    """CG minimises the A-norm of the error over growing Krylov spaces: the
    error after k iterations is non-increasing in k."""
    B = Y.poisson3d(6, seed=9, storage=2)
    A = dense(B.system)
    xd = np.linalg.solve(A, B.b)
    errs = []
    for k in range(1, 25):
        x = np.zeros(B.system.neq)
        import ctypes
        st = oracle.OrcStats()
        oracle.orc().orc_pcg(B.system.neq, B.system.nad, B.system.ia, B.system.ja,
                             B.system.al, B.system.ad, B.b, x, 1e-300, k, 1, 2, 0,
                             ctypes.byref(st))
        e = x - xd
        errs.append(e @ A @ e)
    assert all(b <= a * (1 + 1e-10) for a, b in zip(errs, errs[1:]))
