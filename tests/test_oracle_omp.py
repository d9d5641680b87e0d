"""Pins the OpenMP half of the oracle (oracle/sfr_oracle.c: orc_full_rows,
orc_pcg_omp), which checks the 128^3 GPU PCG solve and is the CPU baseline's
restatement leg:
  - orc_full_rows against the dense matrix each storage represents, for all
    nine storage/symmetry combinations of the reference's dataStruct
    (CSRC unsym's [al | au] and storage CSR's in-row diagonal included);
  - orc_pcg_omp at one and at several threads against the serial PCG
    (oracle.pcg, bit-exact to the reference's compiled solverC) and against
    the reference's own golden solutions (tests/golden/ref_solutions.npz):
    identical iteration counts, x to 1e-12 -- on CSRD and CSRC symmetric
    storage, ELLPACK, a pinned cell, and the ELLPACK slot that faces the
    pinned cell (ja = 0 holding the face coefficient, src/EllPack.c:54-60).
A dropped mirror entry, a transposed au, a lost diagonal or a wrong
recurrence fails one of these.
"""
import json
import os

import numpy as np
import pytest

import oracle
from sfr3d_amd import structure as ST
from sfr3d_amd import systems as Y
from sfr3d_amd.solver import System
from tests.test_oracle import dense

GOLD = os.path.join(os.path.dirname(__file__), "golden")
STORAGES = [(1, 0), (1, 1), (2, 0), (2, 1), (3, 0), (3, 1), (4, 0), (5, 0),
            (6, 1)]


def _random_system(storage, unsym, dims, pinned, seed):
    """The reference's structure for `storage` on a quad or hex mesh with one
    pinned cell, filled with random per-cell coefficients (nonsymmetric
    values; a symmetric storage keeps the lower or upper ones)."""
    nelcon, nviz = (ST.quad_adjacency(*dims) if len(dims) == 2
                    else ST.hex_adjacency(*dims))
    numel, mv = nelcon.shape
    num = ST.reord(nelcon, nviz, 1)
    rt = np.zeros((numel, mv + 1), np.int16)
    rt[pinned, mv] = 1
    id_, neq = ST.numeq_v2(num, rt, nviz, mv)
    pat = ST.datastruct(storage, unsym, id_, num, nelcon, nviz, neq)
    rng = np.random.default_rng(seed)
    lA = rng.standard_normal((numel, mv + 1))
    lB = rng.standard_normal(numel)
    al, ad, _ = ST.assemble(pat, id_, nelcon, nviz, lA, lB, ell_quirk=True)
    return System(neq, pat.nad, pat.ia, pat.ja, al, ad, storage, unsym)


@pytest.mark.parametrize("storage,unsym", STORAGES)
def test_full_rows_against_dense(storage, unsym):
    for dims, pinned, seed in (((7, 5), 9, 1), ((4, 3, 5), 17, 2)):
        s = _random_system(storage, unsym, dims, pinned, seed)
        n = s.neq
        rp, cj, cv = oracle.full_rows(s)
        assert rp[0] == 0 and (np.diff(rp) >= 0).all() and rp[n] == len(cj)
        R = np.zeros((n, n))
        for i in range(n):
            for k in range(rp[i], rp[i + 1]):
                R[i, cj[k]] += cv[k]
        A = dense(s)
        if storage == 1:
            # storage CSR keeps the diagonal inside al; full rows leave it out
            D = np.zeros(n)
            for i in range(n):
                for k in range(s.ia[i], s.ia[i + 1]):
                    if s.ja[k] == i:
                        D[i] = s.al[k]
            assert not np.diag(R).any()
            R += np.diag(D)
        else:
            R += np.diag(s.ad[:n])
        assert np.allclose(R, A, rtol=0, atol=1e-14)
        if not unsym and storage != 4:
            assert np.array_equal(R, R.T)


def _pcg_omp(sysm, b, tol, maxit, threads):
    return oracle.pcg_omp(oracle.full_rows(sysm), sysm.ad, b, tol, maxit, threads)


def _ell_quirk_system():
    """ELLPACK with a pinned cell whose neighbours' slots keep ja = 0 and the
    face coefficient (the reference's values, ell_quirk), diagonally
    strengthened so the quirky operator stays SPD-like for CG."""
    B = Y.poisson3d(9, seed=6, storage=4, reorder=False, pinned_cell=9 ** 3 - 40)
    al, ad, b = ST.assemble(B.pattern, B.id, B.nelcon, B.nviz, B.extra["lA"],
                            B.extra["lB"], ell_quirk=True)
    s = System(B.system.neq, B.system.nad, B.system.ia, B.system.ja, al,
               1.05 * ad, 4, False)
    return s, b[:s.neq].copy()


CASES = {
    "csrd_sym": lambda: (lambda B: (B.system, B.b))(Y.poisson3d(12, seed=3, storage=2)),
    "csrc_sym": lambda: (lambda B: (B.system, B.b))(Y.poisson3d(12, seed=3, storage=3)),
    "ell": lambda: (lambda B: (B.system, B.b))(Y.poisson3d(12, seed=3, storage=4)),
    "csrd_pinned_2d": lambda: (lambda B: (B.system, B.b))(
        Y.poisson2d(20, 16, seed=5, storage=2)),
    "ell_pinned_ja0": _ell_quirk_system,
}


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("case", sorted(CASES))
def test_pcg_omp_matches_serial_oracle(case, threads):
    sysm, b = CASES[case]()
    if case == "ell_pinned_ja0":
        assert (sysm.ja.reshape(-1, 6)[:sysm.neq] == 0).sum() > 6
    xs, ss = oracle.pcg(sysm, b, 1e-8, 5000)
    xo, so = _pcg_omp(sysm, b, 1e-8, 5000, threads)
    assert so["status"] == 0 and ss["status"] == 0
    assert so["iters"] == ss["iters"]
    assert np.linalg.norm(xo - xs) / np.linalg.norm(xs) <= 1e-12
    r = b - oracle.matvec(sysm, xo)
    assert np.isclose(np.linalg.norm(r), so["norm_r"], rtol=1e-6)
    if threads == 1 and sysm.storage == 4:
        # one thread, ELLPACK: the same operation order as the serial
        # pcg and matVecEllPack, so bit for bit
        assert np.array_equal(xo, xs)


GOLDEN = {"pcg_poisson3d_12_csrd": lambda: Y.poisson3d(12, seed=3, storage=2),
          "pcg_poisson3d_12_csrc": lambda: Y.poisson3d(12, seed=3, storage=3),
          "pcg_poisson3d_12_ell": lambda: Y.poisson3d(12, seed=3, storage=4),
          "pcg_poisson2d_20x16_pinned": lambda: Y.poisson2d(20, 16, seed=5, storage=2)}


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_pcg_omp_matches_reference_golden(name):
    gold = {g["name"]: g for g in json.load(open(os.path.join(GOLD, "ref_vectors.json")))["solves"]}[name]
    xg = np.load(os.path.join(GOLD, "ref_solutions.npz"))[name]
    B = GOLDEN[name]()
    for threads in (1, 3, 8):
        x, st = _pcg_omp(B.system, B.b, 1e-8, 5000, threads)
        assert st["iters"] == gold["iters"]
        assert np.linalg.norm(x - xg) / np.linalg.norm(xg) <= 1e-12


def test_pcg_omp_maxit_and_fixed_iterations():
    """At tol 0 the loop runs exactly maxit iterations and reports the
    reference's non-convergence (j == maxIt); its iterate matches the serial
    one after the same number of steps."""
    B = Y.poisson3d(8, seed=2, storage=2)
    for k in (1, 5, 17):
        xo, so = _pcg_omp(B.system, B.b, 0.0, k, 4)
        xs, ss = oracle.pcg(B.system, B.b, 0.0, k)
        assert so["status"] == 1 and ss["status"] == 1
        assert np.linalg.norm(xo - xs) / np.linalg.norm(xs) <= 1e-12


@pytest.mark.ref
@pytest.mark.parametrize("threads", [1, 4])
def test_reference_openmp_solvers_match_serial_oracle(threads):
    """The reference's own OpenMP solvers (pcgOmp, pbicgstabOmp through
    solverC, oracle/_ref/libsfr_ref_omp.so -- the CPU baseline) against the
    serial oracle: identical iteration counts, x to 1e-12, bit for bit on
    one thread (its row partition, scatter buffer and reductions then
    follow the serial order)."""
    if not oracle.ref_omp_available():
        pytest.skip("oracle/_ref/libsfr_ref_omp.so not built (reference tree absent)")
    for B, solver, f in ((Y.poisson3d(14, seed=3, storage=3), 1, oracle.pcg),
                         (Y.poisson3d(14, seed=3, storage=2), 1, oracle.pcg),
                         (Y.momentum3d(12), 2, oracle.pbicgstab)):
        xs, ss = f(B.system, B.b, 1e-10, 5000)
        xr, sr = oracle.ref_solve_omp(B.system, B.b, 1e-10, 5000, solver, threads)
        assert sr["iters"] == ss["iters"]
        assert np.linalg.norm(xr - xs) / np.linalg.norm(xs) <= 1e-12
        assert abs(sr["xKx"] - ss["xKx"]) <= 1e-11 * abs(ss["xKx"])
        if threads == 1:
            assert np.array_equal(xr, xs)
