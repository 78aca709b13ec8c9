"""Build recipe for the oracle (test infrastructure only).

Compiles oracle_sweep.cpp, the CPU restatement of the reference's sweep +
chemistry, into oracle/liboracle_sweep.so with -O2 -ffp-contract=off (Rust
does not contract floating-point expressions). OpenMP is linked so that tests
can spread a sweep's directions over threads (bitwise the same results); it
runs on one thread unless asked.

The reference itself is Rust and cannot be built here (no cargo/rustc, and its
git dependencies need the network), so there is no oracle/_ref build.
"""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_sweep.so")


def build() -> str:
    src = os.path.join(HERE, "oracle_sweep.cpp")
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp", src, "-o", tmp]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
