"""Writes tests/golden/extended_couette32.npz: an extended-precision
trajectory of the README example workload, the accuracy reference for FP64
steps.

    python tests/golden/make_extended_fixture.py

Workload: plane Couette 32x33x32, Lx = 2 pi/1.14, Lz = 2 pi/2.5,
nu = 1/400, dt = 0.01, base flow u = y, zero pressure gradient, rotational
form (README.md:44-60 of the reference).  Initial state: random_state with
amplitude 0.1 from the documented Philox4x32-10 + Box-Muller stream (seed 42,
tests/philox_ref.py), forward-transformed, projected and given its pressure
by the FP64 oracle.  U0 and P0 are stored as they are, so the fixture does not
depend on how a machine rounds the noise.

Stored (complex128, oracle layout (c, n, l, k)):
  U0, P0        initial state (FP64)
  X100_U, X100_P  the state after 100 CNRK2 steps by the oracle built in long
                double (oracle/libcforacle_x.so, every operation and derived
                constant in 64-bit-significand arithmetic, the state kept
                extended between steps), rounded to FP64 at the end
  oracle_err    relative max-abs distance of the FP64 oracle's own 100-step
                result from X100 (U, P), for the record
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

from oracle import oracle as O  # noqa: E402
import philox_ref  # noqa: E402

OUT = os.path.join(HERE, "extended_couette32.npz")
N = (32, 33, 32)
L = (2 * math.pi / 1.14, 2 * math.pi / 2.5)
NU, DT, AMP, SEED, STEPS = 1 / 400, 0.01, 0.1, 42, 100


def noise(g):
    """random_state's Gaussian fill (initialization.jl:37-42) from the
    documented device stream: component c, element index x + Nxp (z + Nzp y)."""
    n = g.Nxp * g.Nzp * g.Ny
    return np.stack([philox_ref.gaussian(n, AMP, SEED, c)[:n].reshape(g.phys_shape) for c in range(3)])


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def main():
    g = O.Grid(*N, *L)
    ref = O.Problem(g, lambda y: y, NU, DT)
    U0, P0 = ref.random_state(noise(g))
    ext = O.Problem(g, lambda y: y, NU, DT, extended=True)
    XU, XP = ext.step(U0, P0, STEPS)
    FU, FP = ref.step(U0, P0, STEPS)
    err = np.array([rel(FU, XU), rel(FP, XP)])
    print(f"FP64 oracle after {STEPS} steps vs extended: U {err[0]:.3e} P {err[1]:.3e}")
    np.savez(OUT, U0=U0, P0=P0, X100_U=XU, X100_P=XP, oracle_err=err,
             grid=np.array([*N], dtype=np.int64), box=np.array([*L, NU, DT]))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
