"""The 2-D lid-driven cavity workload: SIMPLE (collocated, first-order
upwind, Rhie-Chow) on a 128x128 quad mesh; every outer iteration's
pressure-correction system is assembled by the set-up layer in SFR3D's CSRD
(symmetric, lower) and ELLPACK (width 4 + diagonal) storages with the
reference cell pinned (numEqV2).

Pinned here:
  - the pattern sizes the reference's builders give (n = 16,383, 32,510
    lower entries);
  - both storages hold the same operator;
  - the oracle's PCG reproduces, bit for bit, the reference's own compiled
    solverC on identical systems (tests/golden/cavity_ref.json, written by
    tests/golden/make_golden.py from oracle/_ref);
  - the generator itself: SIMPLE drives the mass imbalance down by eight
    orders of magnitude and the centreline velocity agrees with Ghia, Ghia &
    Shin (J. Comput. Phys. 48, 1982), Re = 100, to first-order accuracy.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import systems as Y

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def inputs_digest(B):
    s = B.system
    return sha(np.concatenate([s.ia.view(np.uint8), s.ja.view(np.uint8),
                               s.al.view(np.uint8), s.ad[:s.neq].view(np.uint8),
                               B.b.view(np.uint8)]))


@pytest.fixture(scope="module")
def cavity128():
    gold = json.load(open(os.path.join(GOLD, "cavity_ref.json")))
    sysl, fields = Y.cavity2d(n=gold["n"], outer=gold["outer"])
    return gold, sysl, fields


def test_cavity_pattern_sizes(cavity128):
    _, sysl, _ = cavity128
    csrd, ell = sysl[0][S.CSRD].system, sysl[0][S.ELLPACK].system
    assert csrd.neq == ell.neq == 128 * 128 - 1
    assert csrd.nad == 2 * 128 * 127 - 2          # the pinned corner's 2 couplings go
    assert ell.ia[0] == 4 and len(ell.ja) == 4 * (csrd.neq + 1)


def test_cavity_storages_hold_the_same_operator(cavity128):
    _, sysl, _ = cavity128
    rng = np.random.default_rng(0)
    for rec in (sysl[0], sysl[-1]):
        x = rng.standard_normal(rec[S.CSRD].system.neq)
        y1 = oracle.matvec(rec[S.CSRD].system, x)
        y2 = oracle.matvec(rec[S.ELLPACK].system, x)
        assert np.allclose(y1, y2, rtol=1e-14, atol=1e-16)
        assert np.array_equal(rec[S.CSRD].b, rec[S.ELLPACK].b)


def test_cavity_oracle_matches_reference_golden(cavity128):
    gold, sysl, fields = cavity128
    assert np.allclose(fields["mass"], gold["mass"], rtol=1e-12, atol=0)
    for g in gold["solves"]:
        B = sysl[g["outer"]][g["storage"]]
        assert inputs_digest(B) == g["inputs_sha256"]   # identical assembled system
        x, st = oracle.pcg(B.system, B.b, 1e-8, 50000)
        assert st["iters"] == g["iters"]
        assert sha(x) == g["x_sha256"]                  # bit for bit
        assert abs(st["xKx"] - g["xKx"]) <= 1e-11 * abs(g["xKx"])


@pytest.mark.ref
def test_cavity_oracle_matches_reference_code(ref, cavity128):
    _, sysl, _ = cavity128
    for k in (1, 7):
        for st_id, B in sysl[k].items():
            s = B.system
            x = np.zeros(s.neq)
            out = np.zeros(4)
            assert ref.ref_solve(s.neq, s.nad, s.ia, s.ja, s.al, s.ad, None, B.b,
                                 x, 1e-8, 50000, st_id, 1, 1, 0, out, None, 0) == 0
            xo, so = oracle.pcg(s, B.b, 1e-8, 50000)
            assert so["iters"] == int(out[0]) and np.array_equal(x, xo)


GHIA_Y = np.array([0.0547, 0.0625, 0.0703, 0.1016, 0.1719, 0.2813, 0.4531, 0.5,
                   0.6172, 0.7344, 0.8516, 0.9531, 0.9609, 0.9688, 0.9766])
GHIA_U = np.array([-0.03717, -0.04192, -0.04775, -0.06434, -0.10150, -0.15662,
                   -0.21090, -0.20581, -0.13641, 0.00332, 0.23151, 0.68717,
                   0.73722, 0.78871, 0.84123])


def centreline_u(u):
    n = u.shape[0]
    y = (np.arange(n) + 0.5) / n
    return np.interp(GHIA_Y, y, 0.5 * (u[:, n // 2 - 1] + u[:, n // 2]))


def test_simple_generator_converges_to_ghia():
    _, f = Y.cavity2d(n=32, outer=400, record=False)
    assert f["mass"][-1] < 1e-7 * f["mass"][0]
    # first-order upwind on 32x32 is diffusive: within 0.03 of Ghia's values
    assert np.abs(centreline_u(f["u"]) - GHIA_U).max() < 0.03
