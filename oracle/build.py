"""Build recipe of the oracle's C++/OpenMP restatement (test infrastructure).

``python -m oracle.build`` compiles ``oracle/cpu_step.cpp`` with g++ into
``oracle/libajcpu.so`` (git-ignored; it travels with the tree).  The flags
target plain x86-64 so that a library built here also runs on other hosts.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_step.cpp")
LIB = os.path.join(HERE, "libajcpu.so")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-march=x86-64-v2",
         "-fno-fast-math"]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = ["g++", *FLAGS, SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"g++ failed for cpu_step.cpp:\n{out.stdout}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
