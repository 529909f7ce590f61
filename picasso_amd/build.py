"""Build recipe for the gfx950 shared library picasso_amd/libpicasso_amd.so.

Compiles every HIP translation unit under picasso_amd/csrc with hipcc for
gfx950 (one object per unit, in parallel, under _build/) and links them into
one shared object next to this file.  Run directly
(``python -m picasso_amd.build``) or through __graft_entry__.build().
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpicasso_amd.so")
OBJDIR = os.path.join(ROOT, "_build", "picasso_amd")
SOURCES = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMPILE_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + os.path.join(HERE, "csrc"),
]
LINK_FLAGS = ["--offload-arch=gfx950", "-shared", "-fPIC"]
# roctx ranges around the entry points (rocprofv3 --marker-trace)
LINK_LIBS = ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    subprocess.run([HIPCC] + COMPILE_FLAGS + ["-c", src, "-o", obj], check=True)
    return obj


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    if verbose:
        print(" ".join([HIPCC] + COMPILE_FLAGS + ["-c", "<src>"]))
    with ThreadPoolExecutor(_jobs()) as ex:
        objs = list(ex.map(_compile, SOURCES))
    subprocess.run([HIPCC] + LINK_FLAGS + objs + LINK_LIBS + ["-o", LIB + ".tmp"], check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
