"""Build recipe for the oracle (test infrastructure only).

Compiles ``oracle/cforacle.cpp`` (the CPU restatement of ChannelFlow.jl's
step path) with g++ and OpenMP, twice:

  oracle/libcforacle.so     FP64 (the oracle)
  oracle/libcforacle_x.so   -DORC_EXTENDED: the same source in long double,
                            the accuracy reference for FP64 results

The product library never links either.  The reference itself is Julia
(ChannelFlow.jl + the un-vendored ChebyshevHelmoltzSolvers.jl + FFTW); no
Julia toolchain exists in this image, so no ``oracle/_ref`` build is possible.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cforacle.cpp")
LIB = os.path.join(HERE, "libcforacle.so")
LIB_X = os.path.join(HERE, "libcforacle_x.so")

# x86-64-v3 (AVX2/FMA) keeps the shared object runnable on any current x86 host.
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-shared", "-fPIC", "-std=c++17",
         "-Wall", "-Wno-unused-function"]


def _one(lib, extra, verbose):
    cmd = ["g++", *FLAGS, *extra, SRC, "-o", lib + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(lib + ".tmp", lib)
    return lib


def build(verbose=False):
    with ThreadPoolExecutor(2) as ex:
        jobs = [ex.submit(_one, LIB, [], verbose), ex.submit(_one, LIB_X, ["-DORC_EXTENDED"], verbose)]
        for j in jobs:
            j.result()
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
