"""Pins the oracle's GMRES(15) (oracle/sfr_oracle.c orc_gmres, the
reference's gmres of src/Iters.c:587-790 as solverC drives it,
src/Solv.c:302-372) before it checks the GPU:
  - bit-exact x, equal cycle and iteration counts and xKx against the
    reference's own compiled solverC (golden vectors tests/golden/ref_gmres*,
    and oracle/_ref live where it is built);
  - the converged solution against a direct sparse solve;
  - closed forms of the reference's truncations: with a Krylov dimension of
    1 the update sum over j < k-1 is empty, so x never moves; with a
    dimension of 2 each cycle reduces to the preconditioned minimal-residual
    step x += z (z.w)/(w.w), z = D^-1 (b - A x), w = D^-1 A z;
  - the reported iteration count is k * cycles + 1 (Arnoldi steps + 1).
A dropped rotation, a wrong sign in the Givens update, a back substitution
over the wrong range or a missing preconditioner fails one of these.
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


def gmres_cases():
    spec = json.load(open(os.path.join(GOLD, "ref_gmres.json")))["solves"]
    return [g["name"] for g in spec]


def make_case(key):
    base, tol = key.rsplit("_tol", 1)
    table = {
        "gmres_poisson3d_12_csrd": (lambda: Y.poisson3d(12, seed=3, storage=2), None),
        "gmres_poisson3d_12_ell": (lambda: Y.poisson3d(12, seed=3, storage=4), None),
        "gmres_momentum_10_csrd": (lambda: Y.momentum3d(10), 0),
        "gmres_momentum_10_csrc": (lambda: Y.momentum3d(10, storage=3), 1),
        "gmres_momentum_10_ell": (lambda: Y.momentum3d(10, storage=4), 2),
        "gmres_poisson2d_20x16_pinned": (lambda: Y.poisson2d(20, 16, seed=5, storage=2), None),
    }
    mk, comp = table[base]
    B = mk()
    b = B.b if comp is None else B.extra["b_all"][comp]
    return B, b, float(tol)


@pytest.mark.parametrize("key", gmres_cases())
def test_gmres_oracle_matches_reference_golden(key):
    gold = {g["name"]: g for g in json.load(open(os.path.join(GOLD, "ref_gmres.json")))["solves"]}[key]
    xs = np.load(os.path.join(GOLD, "ref_gmres_solutions.npz"))[key]
    B, b, tol = make_case(key)
    x, st = oracle.gmres(B.system, b, tol, 5000)
    assert st["status"] == 0
    assert st["cycles"] == gold["cycles"] and st["iters"] == gold["iters"]
    assert st["iters"] == 15 * st["cycles"] + 1
    assert np.array_equal(x, xs)                   # bit for bit
    assert sha(x) == gold["x_sha256"]
    assert abs(st["xKx"] - gold["xKx"]) <= 1e-11 * abs(gold["xKx"])
    assert st["norm_r_m"] < st["conv"]


@pytest.mark.ref
@pytest.mark.parametrize("storage,unsym", [(2, 0), (4, 0), (2, 1), (3, 1), (4, 1)])
def test_gmres_oracle_matches_reference_code(ref, storage, unsym):
    for seed in (11, 12):
        B = (Y.momentum3d(7, seed=seed, storage=storage) if unsym
             else Y.poisson3d(8, seed=seed, storage=storage))
        s = B.system
        x = np.zeros(s.neq)
        out = np.zeros(5)
        buf = ctypes.create_string_buffer(512)
        ad = s.ad if s.ad is not None else s.al
        assert ref.ref_solve(s.neq, s.nad, s.ia, s.ja, oracle._al_contig(s),
                             np.ascontiguousarray(ad), None, B.b, x, 1e-10,
                             5000, storage, 4, 1, unsym, out, buf, 512) == 0
        xo, so = oracle.gmres(s, B.b, 1e-10, 5000)
        assert so["iters"] == int(out[0]) and so["cycles"] == int(out[4])
        assert np.array_equal(x, xo)
        line = buf.value.decode()
        assert line.startswith("GMRES: tol") and f"Cycles {so['cycles']} " in line


@pytest.mark.parametrize("unsym", [0, 1])
def test_gmres_against_direct_solve(unsym):
    B = Y.momentum3d(8, seed=4) if unsym else Y.poisson3d(8, seed=5, storage=2)
    A = sp.csr_matrix(dense(B.system))
    xd = spla.spsolve(A.tocsc(), B.b)
    x, st = oracle.gmres(B.system, B.b, 1e-12, 20000)
    assert st["status"] == 0
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    r = B.b - A @ x
    assert np.isclose(st["norm_r"], np.linalg.norm(r), rtol=1e-8)


def test_gmres_krylov_one_never_moves_x():
    B = Y.momentum3d(6, seed=2)
    x0 = np.random.default_rng(3).standard_normal(B.system.neq)
    x, st = oracle.gmres(B.system, B.b, 1e-8, 4, x=x0, newX=False, nkrylov=1)
    assert st["status"] == 1 and st["cycles"] == 4
    assert np.array_equal(x, x0)


@pytest.mark.parametrize("unsym", [0, 1])
def test_gmres_krylov_two_is_minimal_residual_step(unsym):
    B = Y.momentum3d(6, seed=8) if unsym else Y.poisson3d(6, seed=9, storage=2)
    A = dense(B.system)
    d = np.diag(A).copy()
    b = B.b
    x = np.zeros(B.system.neq)
    for cycles in range(1, 6):
        z = (b - A @ x) / d
        w = (A @ z) / d
        x = x + z * (z @ w) / (w @ w)
        xo, st = oracle.gmres(B.system, b, 1e-30, cycles, nkrylov=2)
        assert st["status"] == 1 and st["iters"] == 2 * cycles + 1
        assert np.linalg.norm(xo - x) <= 1e-12 * np.linalg.norm(x)


def test_gmres_maxit_status():
    B = Y.poisson3d(10, seed=2, storage=2)
    x, st = oracle.gmres(B.system, B.b, 1e-10, 3)
    assert st["status"] == 1 and st["cycles"] == 3 and st["iters"] == 46
    assert st["norm_r_m"] >= st["conv"]
