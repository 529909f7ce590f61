"""Builds the C++ test programs of tests/cpp (GPU programs that drive the
library through its C++ facade and check it against the oracle) next to
their sources.  Called by __graft_entry__.build() after the library and the
oracle are built; the binaries are found at run time relative to their own
location ($ORIGIN rpaths), so the tree can move."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
PROGRAMS = ["facade_step"]


def build():
    for name in PROGRAMS:
        subprocess.run(
            [HIPCC, "-std=c++17", "-O2", "-Wall", "-Werror",
             "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle"),
             os.path.join(HERE, name + ".cpp"), "-o", os.path.join(HERE, name),
             "-L" + os.path.join(ROOT, "picasso_amd"), "-lpicasso_amd",
             "-L" + os.path.join(ROOT, "oracle"), "-lorpic",
             "-Wl,-rpath,$ORIGIN/../../picasso_amd:$ORIGIN/../../oracle"],
            check=True)


if __name__ == "__main__":
    build()
