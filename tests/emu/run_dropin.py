"""TEST INFRASTRUCTURE -- drives sfr_solverC (the solverC drop-in) of a
library built for the CPU emulation (argv[1]) or for the GPU, in a process
of its own, on the cases of argv[2] (written by
tests/test_emulated_kernels.py or tests/test_gpu_dropin.py with the oracle's
results).  argv[3] selects what runs:
  alternate   the reference's SIMPLE call pattern: velocity (BiCGSTAB),
              pressure (PCG) and a third pattern in turn with new values
              every round, then patterns enough to evict one; checks x and
              every field of each log line against the oracle and the
              pattern cache (created / hits) against an LRU model;
  maxit_pcg, maxit_bicg, csr, badid
              a call that ends the process the way the reference's solverC
              does (src/Cg.c:195-200, src/Bicgstab.c:225-231,
              src/Solv.c:496); the caller checks the exit status and output.
Prints PASS/FAIL lines; exits non-zero on any failure."""
import ctypes as C
import os
import re
import sys
import tempfile

import numpy as np

lib = C.CDLL(sys.argv[1])
z = np.load(sys.argv[2])
mode = sys.argv[3]
vp = C.c_void_p
lib.sfr_solverC.restype = None
lib.sfr_solverC.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp,
                            vp, vp, vp, vp, vp, vp, C.c_double, C.c_uint,
                            C.c_short, C.c_short, vp, C.c_bool, C.c_bool,
                            C.c_bool, C.c_char_p]
lib.sfr_solverC_context.argtypes = [C.POINTER(vp)]
lib.sfr_cache_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64)]
libc = C.CDLL(None)
libc.fopen.restype = vp
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [vp]
fails = 0


def check(tag, ok):
    global fails
    print(("PASS " if ok else "FAIL ") + tag, flush=True)
    fails += (not ok)


_cases = {}


def case(name):
    """One pattern's arrays, loaded once: SFR3D keeps each system's ia/ja for
    the whole run, and the drop-in recognises a pattern by their addresses."""
    if name not in _cases:
        meta = [int(v) for v in z[f"{name}_meta"]]
        arr = {k: np.ascontiguousarray(z[f"{name}_{k}"]) for k in ("ia", "ja", "al", "ad")}
        _cases[name] = (meta, arr)
    return _cases[name]


def solverC(name, al, ad, b, x, tol, maxit, solver, log_path=None, newX=True, tag="eq"):
    (neq, nad, storage, unsym), arr = case(name)
    f = libc.fopen(log_path.encode(), b"w") if log_path else None
    lib.sfr_solverC(None, neq, neq, nad, 0, arr["ia"].ctypes.data, arr["ja"].ctypes.data,
                    al.ctypes.data, ad.ctypes.data, None, b.ctypes.data, x.ctypes.data,
                    None, None, tol, maxit, storage, solver, f, f is not None, newX,
                    bool(unsym), tag.encode())
    if f:
        libc.fclose(f)


def cache():
    ctx = vp()
    assert lib.sfr_solverC_context(C.byref(ctx)) == 0
    e, c, h = C.c_int(), C.c_int64(), C.c_int64()
    assert lib.sfr_cache_info(ctx, C.byref(e), C.byref(c), C.byref(h)) == 0
    return e.value, c.value, h.value


LINE = {1: re.compile(r"^PCG: tol (.{20}) iteration (\d+) xKx (.{20}) "
                      r"norma\(x\*x\) (.{20}) time (.{20})\n$"),
        2: re.compile(r"^PBICGSTAB: tol (.{20}) iteration (\d+) xKx (.{20}) "
                      r"\|\|x\*x\|\| (.{20}) time (.{20})\n$")}


def rel(a, b):
    return abs(a - b) / abs(b)


if mode == "alternate":
    tmp = tempfile.mkdtemp()
    names = [str(s) for s in z["alt_names"]]
    lru, created, hits = [], 0, 0
    seq = [(names[k % 3], k // 3) for k in range(9)] + [(names[3], 0), (names[4], 0),
                                                        (names[1], 3), (names[0], 3)]
    for k, (name, rnd) in enumerate(seq):
        (neq, nad, storage, unsym), arr = case(name)
        solver = int(z[f"{name}_solver"][0])
        al = np.ascontiguousarray(z[f"{name}_al{rnd}"])
        ad = np.ascontiguousarray(z[f"{name}_ad{rnd}"])
        b = np.ascontiguousarray(z[f"{name}_b{rnd}"])
        x = np.zeros(neq)
        log = os.path.join(tmp, f"log{k}.txt")
        solverC(name, al, ad, b, x, 1e-10, 5000, solver, log)
        xo = z[f"{name}_xo{rnd}"]
        it_o, xkx_o, nx_o = [float(v) for v in z[f"{name}_st{rnd}"]]
        line = open(log).read()
        m = LINE[solver].match(line)
        ok = m is not None
        if ok:
            ok = (m.group(1) == "%20.2e" % 1e-10 and abs(int(m.group(2)) - it_o) <= 2
                  and rel(float(m.group(3)), xkx_o) < 1e-8
                  and rel(float(m.group(4)), nx_o) < 1e-8 and float(m.group(5)) >= 0.0)
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        check(f"{name} round {rnd} solver {solver} relerr {err:.1e} line {line.strip()!r}",
              ok and err < 1e-6)
        if name in lru:
            hits += 1
            lru.remove(name)
        else:
            created += 1
            if len(lru) == 4:
                lru.pop(0)
        lru.append(name)
        e, c, h = cache()
        check(f"cache after {name}: entries {e} created {c} hits {h} "
              f"(model {len(lru)}/{created}/{hits})",
              (e, c, h) == (len(lru), created, hits))
elif mode in ("maxit_pcg", "maxit_bicg"):
    name = "alt_p" if mode == "maxit_pcg" else "alt_v"
    (neq, nad, storage, unsym), arr = case(name)
    x = np.zeros(neq)
    b = np.ascontiguousarray(z[f"{name}_b0"])
    solverC(name, arr["al"], arr["ad"], b, x, 1e-30, 3, 1 if mode == "maxit_pcg" else 2,
            tag="sistEqTest")
    print("NOT REACHED", flush=True)
elif mode == "csr":
    name = "csr"
    (neq, nad, storage, unsym), arr = case(name)
    x = np.zeros(neq)
    solverC(name, arr["al"], arr["ad"], np.ones(neq), x, 1e-8, 100, 1)
    print("NOT REACHED", flush=True)
elif mode == "badid":
    name = "alt_p"
    (neq, nad, storage, unsym), arr = case(name)
    x = np.zeros(neq)
    solverC(name, arr["al"], arr["ad"], np.ones(neq), x, 1e-8, 100, 9)
    print("NOT REACHED", flush=True)
print("DROPIN", "FAILED" if fails else "OK", flush=True)
sys.stdout.flush()
os._exit(1 if fails else 0)
