"""Build libnegf_gpu.so in-tree for gfx950 (hipcc), no JIT cache involved.

    python -m negftrans_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnegf_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled in all HIP code (-packed-fp32-ops).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-shared", "-fPIC",
             "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
SOURCES = ["negf_rgf.hip", "negf_mlp.hip"]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    run([HIPCC] + HIP_FLAGS + srcs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")
F_MODULES = ["negf_gpu_c.f90", "global.f90", "tools.f90", "math_kernel.f90", "grid.f90", "negf.f90"]
F_MAIN = "negftrans_gpu.f90"
EXE = os.path.join(HERE, "negftrans_gpu.x")


def build_fortran():
    """The Fortran host (flang): the reference's module interfaces over the C-ABI and the
    NEGFtrans-compatible driver negftrans_gpu.x, linked against libnegf_gpu.so."""
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for f in F_MODULES + [F_MAIN]:
        obj = os.path.join(bdir, f.replace(".f90", ".o"))
        run([FLANG, "-O2", "-module-dir", bdir, "-I", bdir, "-c", os.path.join(FORTRAN, f), "-o", obj])
        objs.append(obj)
    tmp = EXE + ".tmp"
    run([FLANG, "-O2"] + objs + ["-L" + HERE, "-lnegf_gpu", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, EXE)
    return EXE


def main():
    build_hip()
    build_fortran()


if __name__ == "__main__":
    sys.exit(main())
