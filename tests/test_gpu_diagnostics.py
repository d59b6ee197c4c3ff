"""Device diagnostics (cf_dot / cf_dissipation / cf_power_input) against the
oracle's exact spectral inner product and the laminar closed forms
(postprocessing.jl:21-128, 147-173)."""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import GRIDS, hermitian_random, make, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", GRIDS)
def test_dot_and_dissipation_match_oracle(gpu, dims):
    import torch
    cf = gpu
    say("dot/dissipation", dims)
    gc, go, prob = make(cf, dims)
    rng = np.random.default_rng(21 + sum(dims))
    U = hermitian_random(go, rng, 3)
    V = hermitian_random(go, rng, 3)
    Ud, Vd = torch.as_tensor(U, device="cuda"), torch.as_tensor(V, device="cuda")
    ref = sum(O.dot(go, U[i], V[i]) for i in range(3))
    got = cf.dot(prob, Ud, Vd)
    say(f"  dot hip {got!r} oracle {ref!r}")
    assert abs(got - ref) <= 1e-13 * max(abs(ref), 1e-300) + 1e-15
    nu = 0.37
    W = O.curl(go, U)
    dref = nu * sum(O.dot(go, W[i], W[i]) for i in range(3))
    dgot = cf.dissipation_rate(prob, Ud, nu)
    say(f"  dissipation hip {dgot!r} oracle {dref!r}")
    assert abs(dgot - dref) <= 1e-12 * abs(dref)


@pytest.mark.parametrize("Ny", [9, 17, 33, 65, 97, 129])
def test_laminar_closed_forms(gpu, Ny):
    import torch
    cf = gpu
    nu = 1 / 400
    g = cf.Grid(6, Ny, 4, 2.0, 3.0)
    y = g.y
    for name, profile, E, D, P, grad in (
            ("couette", y, 1 / 6, nu, nu, (0.0, 0.0)),
            ("poiseuille", 1 - y ** 2, 4 / 15, 4 * nu / 3, 4 * nu / 3, (-2 * nu, 0.0))):
        prob = cf.ChannelFlowProblem(g, lambda yy: 0 * yy, nu, 0.01)
        U = g.zeros_spec(3)
        U[0, :, 0, 0] = torch.as_tensor(O.chebcoeffs(profile), device="cuda")
        e = cf.kinetic_energy(prob, U)
        d = cf.dissipation_rate(prob, U, nu)
        p = cf.power_input(prob, U, nu, pressuregradient=grad)
        say(f"  {name} Ny={Ny}: E {e!r} D {d!r} P {p!r}")
        assert abs(e - E) < 1e-14 and abs(d - D) < 1e-14 * max(1, D / nu) and abs(p - P) < 1e-14
