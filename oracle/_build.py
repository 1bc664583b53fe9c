"""Build recipe for the oracle (test infrastructure only).

Compiles oracle_mpm.cpp (the plain step) and oracle_mlp.cpp (the learned
correction's network) with g++ into oracle/liboracle_mpm.so.  No FP
contraction (the reference's default x86-64 g++ build emits no FMA), OpenMP
for the timed all-core CPU baseline.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_mpm.so")


def build(verbose=False):
    srcs = [os.path.join(HERE, "oracle_mpm.cpp"), os.path.join(HERE, "oracle_mlp.cpp")]
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
           "-Wall"] + srcs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
