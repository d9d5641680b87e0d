"""MatrixMarket output and input (reference src/WriteMtx.c, through the
Matrix Market I/O library's mm_write_mtx_crd).

Pinned by the reference's own output: the pattern files it wrote for the
20x10 quad test mesh (test/output/coo/retangulo_bin.mtx, reord false and
true) are reproduced byte for byte.  Values round-trip to the 15 significant
digits the %.14e format keeps; a system read back from its file solves to
the same solution on the GPU."""
import os

import numpy as np
import pytest

import oracle
import sfr3d_amd as S
from sfr3d_amd import structure as ST, systems
from sfr3d_amd.solver import System

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def retangulo_system(flag):
    rows = [l.split() for l in open(os.path.join(GOLD, "retangulo_cells.txt"))
            if not l.startswith("#")]
    el = np.array([[int(t) for t in r[5:9]] for r in rows], np.int32)
    nelcon, nviz = ST.general_adjacency(2, el, np.full(len(el), 4, np.int16), 4)
    num = ST.reord(nelcon, nviz, flag)
    id_, neq = ST.numeq_v1(num)
    pat = ST.datastruct(S.CSRD, False, id_, num, nelcon, nviz, neq)
    return System(neq, pat.nad, pat.ia, pat.ja, np.zeros(pat.nad), np.zeros(neq),
                  S.CSRD, False)


@pytest.mark.parametrize("flag,name", [(0, "retangulo_bin.mtx"),
                                       (1, "retangulo_rcm_bin.mtx")])
def test_writer_reproduces_reference_files(tmp_path, flag, name):
    out = str(tmp_path / name)
    assert ST.write_mtx(out, retangulo_system(flag), bin=True) == 0
    assert open(out, "rb").read() == open(os.path.join(GOLD, name), "rb").read()


@pytest.mark.parametrize("kind", ["sym", "unsym"])
def test_values_round_trip(tmp_path, kind):
    B = (systems.poisson3d(7, seed=2, storage=S.CSRD) if kind == "sym"
         else systems.momentum3d(6, storage=S.CSRD))
    s = B.system
    path = str(tmp_path / "a.mtx")
    assert ST.write_mtx(path, s) == 0
    r = ST.read_mtx(path)
    assert r.neq == s.neq and r.nad == s.nad and r.unsym == s.unsym
    assert np.array_equal(r.ia, s.ia) and np.array_equal(r.ja, s.ja)
    assert np.allclose(r.al, s.al, rtol=1e-14, atol=0)
    assert np.allclose(r.ad, s.ad[:s.neq], rtol=1e-14, atol=0)
    # the operator itself: same products
    x = np.random.default_rng(0).standard_normal(s.neq)
    assert np.allclose(oracle.matvec(r, x), oracle.matvec(s, x), rtol=1e-12, atol=1e-12)
    vb = str(tmp_path / "b.mtx")
    assert ST.write_mtx_vector(vb, B.b) == 0
    lines = open(vb).read().splitlines()
    assert lines[0] == "%%MatrixMarket matrix array real general"
    assert lines[1] == f"{s.neq} 1"
    assert np.allclose(np.array(lines[2:], float), B.b, rtol=1e-14, atol=0)


def test_reader_symmetric_upper_entries_and_comments(tmp_path):
    path = str(tmp_path / "u.mtx")
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% comment\n"
                "3 3 5\n1 1 4.0\n1 2 -1.0\n2 2 4.0\n3 2 -2.0\n3 3 4.0\n")
    r = ST.read_mtx(path)
    assert list(r.ia) == [0, 0, 1, 2] and list(r.ja) == [0, 1]
    assert list(r.al) == [-1.0, -2.0] and list(r.ad) == [4.0, 4.0, 4.0]


def test_unsupported_storage(tmp_path):
    B = systems.poisson3d(4, seed=1, storage=S.ELLPACK)
    assert ST.write_mtx(str(tmp_path / "e.mtx"), B.system) == S._lib.NOT_SUPPORTED


@pytest.mark.gpu
def test_system_read_back_solves_on_gpu(ctx, tmp_path):
    B = systems.poisson3d(12, seed=5, storage=S.CSRD)
    path = str(tmp_path / "p.mtx")
    assert ST.write_mtx(path, B.system) == 0
    r = ST.read_mtx(path)
    x, st = ctx.solve(r, S.PCG, 1e-8, 10000, B.b)
    xo, _ = oracle.pcg(B.system, B.b, 1e-8, 10000)
    assert st["status"] == S._lib.OK
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6
