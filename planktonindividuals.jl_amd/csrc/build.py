"""Build recipe of libpi_hip.so (gfx950).

    python planktonindividuals.jl_amd/csrc/build.py

Flags: no FP contraction (the reference's Julia code never fuses), IEEE
divide/sqrt (HIP's default), denormals kept, packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libpi_hip.so")
SOURCES = ["pi_host.hip", "k_particles.hip", "k_tracers.hip", "k_util.hip"]
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
         "-fno-gpu-flush-denormals-to-zero",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def build(verbose=False, jobs=4):
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, "_obj_" + src.replace(".hip", ".o"))
        cmd = ["hipcc"] + [f for f in FLAGS if f != "-shared"] + ["-c", os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(obj)
    bad = [s for p, s in procs if p.wait() != 0]
    if bad:
        raise RuntimeError(f"hipcc failed on {bad}")
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    for o in objs:
        os.remove(o)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
