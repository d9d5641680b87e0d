"""TEST INFRASTRUCTURE -- builds the CPU emulation of the library: the host C
files with gcc and the .hip kernel sources with g++ against the kernel
emulation (emu.hpp, include/hip/hip_runtime.h) and the host stand-ins of the
HIP runtime (stub.cpp).  Used by tests/test_emulated_kernels.py only; the
product library is the hipcc build of the same sources."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "sfr3d_amd", "csrc")


def build(out_dir=None, asan=False, tune_rows=2000):
    """Compile and link; returns the path of libsfr3d_emu.so.  tune_rows
    shrinks the batch tuner's trial size to what an emulated run affords."""
    out = out_dir or os.path.join(HERE, "_build" + ("_asan" if asan else ""))
    os.makedirs(out, exist_ok=True)
    flags = (["-fsanitize=address", "-fno-omit-frame-pointer", "-g", "-O1"]
             if asan else ["-O1", "-g0"])
    flags += ["-fPIC", "-D__HIP_PLATFORM_AMD__", f"-DTUNE_ROWS={tune_rows}"]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I/opt/rocm/include"]
    emu_inc = ["-I" + os.path.join(HERE, "include")]
    jobs = []
    for f in sorted(os.listdir(os.path.join(CSRC, "host"))):
        if f.endswith(".c"):
            jobs.append(["gcc", "-std=c11", *flags, *inc, "-c",
                         os.path.join(CSRC, "host", f),
                         "-o", os.path.join(out, "h_" + f[:-2] + ".o")])
    for f in sorted(os.listdir(os.path.join(CSRC, "device"))):
        if f.endswith(".hip"):
            jobs.append(["g++", "-x", "c++", "-std=c++20", *flags, *emu_inc, *inc,
                         "-c", os.path.join(CSRC, "device", f),
                         "-o", os.path.join(out, "d_" + f[:-4] + ".o")])
    jobs.append(["g++", "-std=c++20", *flags, *emu_inc, *inc, "-c",
                 os.path.join(HERE, "stub.cpp"), "-o", os.path.join(out, "stub.o")])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for cmd, res in zip(jobs, ex.map(lambda c: subprocess.run(
                c, capture_output=True, text=True), jobs)):
            if res.returncode:
                raise RuntimeError(f"{' '.join(cmd)}\n{res.stderr[-4000:]}")
    lib = os.path.join(out, "libsfr3d_emu.so")
    objs = [os.path.join(out, f) for f in sorted(os.listdir(out)) if f.endswith(".o")]
    link_san = flags[:4] if asan else []
    res = subprocess.run(["g++", "-shared", *link_san, "-o", lib,
                          *objs, "-Wl,-Bsymbolic", "-lpthread"],
                         capture_output=True, text=True)
    if res.returncode:
        raise RuntimeError(res.stderr[-4000:])
    return lib
