"""Build recipe for the oracle (test infrastructure only).

Compiles ``oracle/cforacle.cpp`` (the CPU restatement of ChannelFlow.jl's
step path) into ``oracle/libcforacle.so`` with g++ and OpenMP.  The product
library never links this file.

The reference itself is Julia (ChannelFlow.jl + the un-vendored
ChebyshevHelmoltzSolvers.jl + FFTW); no Julia toolchain exists in this image,
so no ``oracle/_ref`` build is possible.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cforacle.cpp")
LIB = os.path.join(HERE, "libcforacle.so")

# x86-64-v3 (AVX2/FMA) keeps the shared object runnable on any current x86 host.
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-shared", "-fPIC", "-std=c++17",
         "-Wall", "-Wno-unused-function"]


def build(verbose=False):
    cmd = ["g++", *FLAGS, SRC, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
