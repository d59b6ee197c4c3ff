"""Closure oracle (oracle/closure.py) pinned on the CPU: the exact-arithmetic
backward against central finite differences, padded parameters inert, the
bf16 rounding against known bit patterns, and the box filter against its
closed form on single Fourier modes."""
import math

import numpy as np

from oracle import closure as K
from oracle import oracle as O


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    th = K.init_theta(1).astype(np.float64)
    th[:] += 0.05 * rng.standard_normal(th.shape) * (th != 0)
    X = rng.standard_normal((7, K.NIN))
    T = rng.standard_normal((7, K.NOUT))
    loss, g = K.loss_and_grad(th, X, T, bf=False)
    idx = [j for j in rng.choice(K.NPARAM, 60, replace=False)]
    P = K.unpack(np.arange(K.NPARAM))
    pad = set(P["W1"][:, K.NIN:].ravel().astype(int)) | set(P["W4"][K.NOUT:, :].ravel().astype(int)) | \
        set(P["b4"][K.NOUT:].astype(int))
    for j in idx:
        if j in pad:
            assert g[j] == 0.0
            continue
        h = 1e-6 * max(1.0, abs(th[j]))
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        fd = (K.loss_and_grad(tp, X, T, bf=False)[0] - K.loss_and_grad(tm, X, T, bf=False)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, g[j])


def test_bf16_rounding_bits():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even keeps 1; 1 + 3*2^-9 rounds up
    assert K.bf16(np.float32(1 + 2 ** -8)) == 1.0
    assert K.bf16(np.float32(1 + 3 * 2 ** -9)) == 1 + 2 ** -7
    assert K.bf16(np.float32(-3.0)) == -3.0


def test_box_filter_single_modes():
    gf = O.Grid(16, 17, 12, 4.0, 3.0)
    gc = O.Grid(8, 9, 6, 4.0, 3.0)
    dx, dz = gf.Lx / 8, gf.Lz / 6
    U = np.zeros((gf.Ny, gf.Nz, gf.Nxh), complex)
    U[2, 1, 3] = 1.0 + 2.0j            # k = 3, l = 1, n = 2
    U[4, gf.Nz - 2, 1] = 0.5           # l = -2
    U[12, 0, 0] = 7.0                  # degree above the coarse grid: dropped
    U[0, 0, 6] = 1.0                   # k = 6: outside the coarse grid
    V = K.box_filter(U, gf, gc, dx, dz)
    s = lambda w: math.sin(w / 2) / (w / 2)  # noqa: E731
    a, b = 2 * math.pi / 4.0, 2 * math.pi / 3.0
    assert np.isclose(V[2, 1, 3], (1 + 2j) * s(a * 3 * dx) * s(b * dz))
    assert np.isclose(V[4, gc.Nz - 2, 1], 0.5 * s(a * dx) * s(-2 * b * dz))
    assert np.count_nonzero(V) == 2
