"""The reference's CPU-against-GPU agreement test on its benchmark grids
(test/cuda/runtests.jl:152-175): plane Couette on Grid(N, N+1, N, 2 pi, 2 pi),
nu = 1/400, dt = 0.002, N = 32, 64, 128, from the field
(0.1 (1-y^2) exp(cos x + cos z), 0.05 (1-y^2)^2 cos z, 0.2 y (1-y^2) sin z),
projected and with its pressure, advanced over (0, 0.002) through the Flows
adapter.  The initial state is built by the oracle (forward transform,
project!, pressure) and handed to both sides, as the reference adapts its CPU
state to the GPU.  Checked at the reference's own bound (isapprox on whole
arrays: norm(a - b) <= max(atol, rtol max(norm a, norm b)), rtol 1e-8,
atol 1e-10) and, far tighter, at the oracle-parity bound used elsewhere:
max(1e-12, 16 x the oracle's own roundoff envelope)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import _split, assert_parity, roundoff_envelope, say

pytestmark = pytest.mark.gpu

NU, DT = 1 / 400, 0.002


def isapprox(a, b, rtol=1e-8, atol=1e-10):
    """Julia's isapprox for arrays (the reference's `≈` with rtol/atol)."""
    d = np.linalg.norm((a - b).ravel())
    return d <= max(atol, rtol * max(np.linalg.norm(a.ravel()), np.linalg.norm(b.ravel()))), d


@pytest.mark.parametrize("N", [32, 64, 128])
def test_reference_grid_one_step(gpu, N):
    import torch
    cf = gpu
    say(f"reference grid {N}x{N + 1}x{N}")
    go = O.Grid(N, N + 1, N, 2 * math.pi, 2 * math.pi)
    ref = O.Problem(go, lambda y: y, NU, DT)
    u = np.stack([go.sample(lambda x, y, z: 0.1 * (1 - y ** 2) * np.exp(np.cos(x) + np.cos(z))),
                  go.sample(lambda x, y, z: 0.05 * (1 - y ** 2) ** 2 * np.cos(z)),
                  go.sample(lambda x, y, z: 0.2 * y * (1 - y ** 2) * np.sin(z))])
    U0 = ref.project(np.stack([O.forward(go, u[c]) for c in range(3)]))
    P0 = ref.pressure(U0)
    U1, P1 = ref.step(U0, P0, 1)

    g = cf.Grid(N, N + 1, N, 2 * math.pi, 2 * math.pi)
    prob = cf.CouetteFlow(g, NU, DT)
    st = cf.State(torch.as_tensor(U0, device="cuda"), torch.as_tensor(P0, device="cuda"))
    cf.flow(prob)(st, (0.0, DT))
    Uh, Ph = st.velocity.cpu().numpy(), st.stagepressure.cpu().numpy()

    for name, a, b in zip(("u", "v", "w", "P"), _split((Uh, Ph)), _split((U1, P1))):
        ok, d = isapprox(a, b)
        say(f"N={N} {name}: |a-b| = {d:.3e} (reference bound rtol 1e-8, atol 1e-10)")
        assert ok, (N, name, d)
    env = roundoff_envelope(lambda a, b: _split(ref.step(a, b, 1)), [U0, P0], _split((U1, P1)), trials=3)
    assert_parity(f"reference grid N={N}", _split((Uh, Ph)), _split((U1, P1)), env)
