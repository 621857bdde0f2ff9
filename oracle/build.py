"""Build recipe of the oracle library (test infrastructure).

    python oracle/build.py        -> oracle/liboracle_pi.so

g++ only, no contraction, no fast-math, so FP32 arithmetic is plain IEEE binary32
in source order; OpenMP for the element-wise loops (orc_set_threads, default 1).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle_pi.so")


def build(verbose=False):
    src = os.path.join(HERE, "pi_oracle.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fno-finite-math-only",
           "-fopenmp", "-Wall", "-fPIC", "-shared", "-o", OUT + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
