"""Build recipe for the gfx950 library libtcsem.so and the drop-in driver tcsem_fdtd.

Compiles in-tree with hipcc (no JIT cache): the shared library next to this file and the driver
executable in the package directory.  Used by __graft_entry__.build() and by the tests.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtcsem.so")
DRIVER = os.path.join(HERE, "tcsem_fdtd")
ARCH = "gfx950"
SOURCES = ["tcsem_host.cpp", "tcsem_kernels.hip", "tcsem_upcont.hip"]

# Packed-FP32 instructions are disabled for all device code of this project.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    """Compile every HIP source for gfx950 and link libtcsem.so (always recompiles)."""
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([hipcc()] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the driver is host code over the C ABI: g++ only, linked against the library beside it
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "tcsem_driver.cpp"), "-o", DRIVER,
          "-L" + HERE, "-ltcsem", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
