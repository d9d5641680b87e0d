"""Pins the oracle's BiCGSTAB(2) (oracle/sfr_oracle.c orc_pbicgstabl2, the
reference's pbicgstabl2 of src/Bicgstab.c:1172-1378 as solverC drives it,
src/Solv.c:217-295) before it checks the GPU:
  - bit-exact x, the same iteration count and xKx as the reference's own
    compiled solverC (golden vectors tests/golden/ref_bicgl2*, and
    oracle/_ref live where it is built);
  - the solution against a direct sparse solve on nonsymmetric systems;
  - consistency of the recurrences: after any number of double steps the
    recurrence residual equals b - Ax to rounding, which fails for a dropped
    or mis-signed term in the x or r updates (the two are updated
    separately from the same coefficients).
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from sfr3d_amd import systems as Y

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle import dense, sha  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def bicgl2_cases():
    return json.load(open(os.path.join(GOLD, "ref_bicgl2.json")))["solves"]


def make_case(name):
    table = {
        "bicgl2_momentum_10_csrd": (lambda: Y.momentum3d(10), 0),
        "bicgl2_momentum_10_csrc": (lambda: Y.momentum3d(10, storage=3), 1),
        "bicgl2_momentum_10_ell": (lambda: Y.momentum3d(10, storage=4), 2),
        "bicgl2_poisson3d_12_csrd": (lambda: Y.poisson3d(12, seed=3, storage=2), None),
        "bicgl2_poisson2d_20x16_pinned": (lambda: Y.poisson2d(20, 16, seed=5, storage=2), None),
    }
    mk, comp = table[name]
    B = mk()
    return B, (B.b if comp is None else B.extra["b_all"][comp])


@pytest.mark.parametrize("case", bicgl2_cases(), ids=lambda c: c["name"])
def test_bicgl2_oracle_matches_reference_golden(case):
    xs = np.load(os.path.join(GOLD, "ref_bicgl2_solutions.npz"))[case["name"]]
    B, b = make_case(case["name"])
    x, st = oracle.pbicgstabl2(B.system, b, case["tol"], 5000)
    assert st["status"] == 0 and st["iters"] == case["iters"]
    assert np.array_equal(x, xs)                   # bit for bit
    assert sha(x) == case["x_sha256"]
    assert abs(st["xKx"] - case["xKx"]) <= 1e-11 * abs(case["xKx"])


@pytest.mark.ref
@pytest.mark.parametrize("storage,unsym", [(2, 1), (3, 1), (4, 1), (2, 0)])
def test_bicgl2_oracle_matches_reference_code(ref, storage, unsym):
    for seed in (11, 12):
        B = (Y.momentum3d(7, seed=seed, storage=storage) if unsym
             else Y.poisson3d(8, seed=seed, storage=storage))
        s = B.system
        x = np.zeros(s.neq)
        out = np.zeros(5)
        buf = ctypes.create_string_buffer(512)
        assert ref.ref_solve(s.neq, s.nad, s.ia, s.ja, oracle._al_contig(s),
                             np.ascontiguousarray(s.ad), None, B.b, x, 1e-10,
                             5000, storage, 3, 1, unsym, out, buf, 512) == 0
        xo, so = oracle.pbicgstabl2(s, B.b, 1e-10, 5000)
        assert so["status"] == 0 and so["iters"] == int(out[0])
        assert np.array_equal(x, xo)
        assert buf.value.decode().startswith("PBICGSTAB(2): tol")


def test_bicgl2_against_direct_solve():
    B = Y.momentum3d(8, seed=4)
    A = sp.csr_matrix(dense(B.system))
    for comp in range(3):
        b = B.extra["b_all"][comp]
        xd = spla.spsolve(A.tocsc(), b)
        x, st = oracle.pbicgstabl2(B.system, b, 1e-12, 5000)
        assert st["status"] == 0
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


@pytest.mark.parametrize("unsym", [0, 1])
def test_bicgl2_recurrence_matches_true_residual(unsym):
    B = Y.momentum3d(7, seed=6) if unsym else Y.poisson3d(7, seed=6, storage=2)
    nb = np.linalg.norm(B.b)
    for k in range(1, 7):
        x, st = oracle.pbicgstabl2(B.system, B.b, 1e-30, k)
        assert st["status"] == 1 and st["iters"] == k
        assert abs(st["norm_r"] - st["norm_r_m"]) <= 1e-11 * nb, k
        assert st["norm_r"] < nb
