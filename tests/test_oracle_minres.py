"""Pins the oracle's MINRES (oracle/sfr_oracle.c orc_minres, the reference's
pminres of src/Iters.c:309-555 with symOrtho, src/Iters.c:1171-1208, as
solverC drives it, src/Solv.c:380-445) before it checks the GPU:
  - bit-exact x, the same iteration count and xKx as the reference's own
    compiled solverC (golden vectors tests/golden/ref_minres*, and
    oracle/_ref live where it is built);
  - the solution against a direct sparse solve, on an SPD system and on a
    symmetric indefinite one (MINRES's reason to exist next to CG);
  - the defining property, by brute force: after k steps the iterate
    minimises (r.D^-1 r)^(1/2) over x0 + K_k(D^-1 A, D^-1 r0), computed by
    a dense least-squares solve over an orthonormal basis of that space.
A wrong sign in the three-term recurrence, a dropped rotation or a
preconditioner applied on the wrong side fails one of these.
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
from sfr3d_amd.solver import System

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle import dense, sha  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def minres_cases():
    return json.load(open(os.path.join(GOLD, "ref_minres.json")))["solves"]


def make_case(name):
    table = {
        "minres_poisson3d_12_csrd": lambda: Y.poisson3d(12, seed=3, storage=2),
        "minres_poisson3d_12_csrc": lambda: Y.poisson3d(12, seed=3, storage=3),
        "minres_poisson3d_12_ell": lambda: Y.poisson3d(12, seed=3, storage=4),
        "minres_poisson2d_20x16_pinned": lambda: Y.poisson2d(20, 16, seed=5, storage=2),
        "minres_poisson3d_10_maxit": lambda: Y.poisson3d(10, seed=7, storage=2),
    }
    return table[name]()


@pytest.mark.parametrize("case", minres_cases(), ids=lambda c: c["name"])
def test_minres_oracle_matches_reference_golden(case):
    xs = np.load(os.path.join(GOLD, "ref_minres_solutions.npz"))[case["name"]]
    B = make_case(case["name"])
    x, st = oracle.minres(B.system, B.b, case["tol"], case["maxit"])
    if st["status"] == 1:                # the reference logs maxIt + 1
        assert st["iters"] == case["maxit"] and case["log_iters"] == case["maxit"] + 1
    else:
        assert st["iters"] == case["log_iters"]
    assert np.array_equal(x, xs)                   # bit for bit
    assert sha(x) == case["x_sha256"]
    assert abs(st["xKx"] - case["xKx"]) <= 1e-11 * abs(case["xKx"])


@pytest.mark.ref
@pytest.mark.parametrize("storage", [2, 3, 4])
def test_minres_oracle_matches_reference_code(ref, storage):
    for seed in (11, 12):
        B = Y.poisson3d(8, seed=seed, storage=storage)
        s = B.system
        x = np.zeros(s.neq)
        out = np.zeros(5)
        buf = ctypes.create_string_buffer(512)
        assert ref.ref_solve(s.neq, s.nad, s.ia, s.ja, oracle._al_contig(s),
                             np.ascontiguousarray(s.ad), None, B.b, x, 1e-10,
                             5000, storage, 5, 1, 0, out, buf, 512) == 0
        xo, so = oracle.minres(s, B.b, 1e-10, 5000)
        assert so["status"] == 0 and so["iters"] == int(out[0])
        assert np.array_equal(x, xo)
        assert buf.value.decode().startswith("PMINRES: tol")


def shifted(B, frac):
    """B's matrix minus sigma I with sigma a fraction of its smallest
    diagonal entry: symmetric, positive diagonal, and indefinite once sigma
    passes the smallest eigenvalue."""
    s = B.system
    sigma = frac * float(np.min(s.ad))
    return System(s.neq, s.nad, s.ia, s.ja, s.al.copy(), s.ad - sigma,
                  s.storage, s.unsym), sigma


@pytest.mark.parametrize("frac", [0.0, 0.3])
def test_minres_against_direct_solve(frac):
    B = Y.poisson3d(8, seed=5, storage=2)
    sysm, sigma = shifted(B, frac)
    A = dense(sysm)
    ev = np.linalg.eigvalsh(A)
    if frac > 0:
        assert ev[0] < 0 < ev[-1]                 # indefinite
    xd = np.linalg.solve(A, B.b)
    x, st = oracle.minres(sysm, B.b, 1e-12, 20000)
    assert st["status"] == 0
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    r = B.b - A @ x
    assert np.isclose(st["norm_r"], np.linalg.norm(r), rtol=1e-8)


@pytest.mark.parametrize("frac", [0.0, 0.3])
def test_minres_iterate_minimises_preconditioned_residual(frac):
    B = Y.poisson3d(5, seed=9, storage=2)
    sysm, _ = shifted(B, frac)
    A = dense(sysm)
    d = np.diag(A).copy()
    rng = np.random.default_rng(4)
    x0 = rng.standard_normal(sysm.neq)
    r0 = B.b - A @ x0
    Bop = A / d[:, None]
    for k in range(1, 7):
        cols = [r0 / d]
        for _ in range(k - 1):
            cols.append(Bop @ cols[-1])
        Q, _ = np.linalg.qr(np.array(cols).T)
        W = (A @ Q) / np.sqrt(d)[:, None]
        y, *_ = np.linalg.lstsq(W, r0 / np.sqrt(d), rcond=None)
        xk = x0 + Q @ y
        xo, st = oracle.minres(sysm, B.b, 1e-300, k, x=x0, newX=False)
        assert st["status"] == 1 and st["iters"] == k
        assert np.linalg.norm(xo - xk) <= 1e-9 * np.linalg.norm(xk), k
