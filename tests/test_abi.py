"""The C ABI: every function include/*.h declares is exported by the built
library, the library loads without a GPU, the context refuses cleanly when no
device is present, and the host-side keyword/config parsing follows the
reference (src/Solv.c:507-622, src/Datastruct.c:1707-1752)."""
import ctypes as C
import os
import re

import pytest

import sfr3d_amd as S
from sfr3d_amd import _lib, structure

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    names = set()
    for h in ("sfr3d.h", "sfr3d_setup.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(sfr_[A-Za-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 30
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_internal_symbols_are_hidden():
    lib = C.CDLL(_lib.LIB_PATH)
    for n in ("sfr_k_pcg_iter", "sfr_view", "sfr_workspace", "sfr_set_error"):
        assert not hasattr(lib, n), n


def test_no_device_is_reported():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(S.SolverError) as e:
        S.Context(0)
    assert e.value.code == _lib.NO_DEVICE


def test_solver_and_storage_words():
    lib = _lib.lib
    assert lib.sfr_solver_id(b"cg") == 1 and lib.sfr_solver_id(b"bicgstab") == 2
    assert lib.sfr_solver_id(b"bicgstabl1") == 3 and lib.sfr_solver_id(b"gmres") == 4
    assert lib.sfr_solver_id(b"nope") == -1
    words = {b"csr": 1, b"card": 2, b"csrc": 3, b"ellpack": 4, b"csrdcoo": 5,
             b"csrccoo": 6, b"csrd": -1}
    for w, v in words.items():
        assert lib.sfr_storage_id(w) == v
    eps = lib.sfr_smachn()
    assert eps == 2.0 ** -53 and 1.0 + eps == 1.0 and 1.0 + 2 * eps != 1.0


def test_reference_solver_config_files(tmp_path):
    p = structure.read_solver_config(os.path.join(GOLD, "solver_config", "solverPres.dat"))
    assert p == {"solver": 1, "tol": 1e-11, "maxit": 10000, "storage": 3}
    v = structure.read_solver_config(os.path.join(GOLD, "solver_config", "solverVel.dat"))
    assert v == {"solver": 2, "tol": 1e-11, "maxit": 1000, "storage": 3}
    f = tmp_path / "s.dat"
    f.write_text("// comment line\nname cg\ntol 0\nstorage EllPack\nend\n")
    c = structure.read_solver_config(str(f))
    assert c["tol"] == 2.0 ** -53 and c["storage"] == 4 and c["maxit"] == 50000


def test_batch_entry_points_refuse_bad_arguments():
    lib = _lib.lib
    h = C.c_void_p()
    assert lib.sfr_batch_create(None, 1, None, 0, C.byref(h)) == _lib.BAD_ARG
    assert lib.sfr_batch_rows(None) == 0
    assert lib.sfr_batch_info(None, 0, None, None, None, None) == _lib.BAD_ARG
    assert lib.sfr_batch_buckets(None, None, 0) == 0
    lib.sfr_batch_destroy(None, None)


def test_solve_entry_points_refuse_bad_arguments():
    lib = _lib.lib
    assert lib.sfr_solve_batched(None, None, 3, None, None, None, None, None,
                                 _lib.PBICGSTAB, 1e-8, 100, 1, None) == _lib.BAD_ARG
    assert lib.sfr_solve_shared(None, None, 3, None, None, None, _lib.PBICGSTAB,
                                1e-8, 100, 1, None) == _lib.BAD_ARG
    assert lib.sfr_solve_device(None, None, _lib.GMRES, 0, 1e-8, 10, None, None, 1,
                                None) == _lib.BAD_ARG


def test_matrix_tune_refuses_bad_arguments():
    lib = _lib.lib
    assert lib.sfr_matrix_tune(None, None, 0, 4) == _lib.BAD_ARG
    assert lib.sfr_matrix_rows_per_thread(None, None, None) == _lib.BAD_ARG


def test_bench_refuses_several_gpus():
    """The benchmark runs on one GPU and says so for --gpus > 1, before it
    touches a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120, cwd=root)
    assert res.returncode == 2
    assert "one GPU" in res.stderr and not res.stdout
