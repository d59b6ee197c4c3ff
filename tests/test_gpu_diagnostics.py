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


def _physical_statistics(go, U, base_values):
    """Reference algorithm on the padded physical grid (statistics.jl:26-43):
    total velocity, x-z plane means and covariances."""
    V = U.copy()
    V[0, :, 0, 0] += O.chebcoeffs(base_values)
    u = [O.inverse(go, V[c]) for c in range(3)]  # (Ny, Nzp, Nxp)
    mean = np.stack([f.mean(axis=(1, 2)) for f in u], axis=1)
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    cov = np.stack([((u[i] - mean[:, i, None, None]) * (u[j] - mean[:, j, None, None])).mean(axis=(1, 2))
                    for i, j in pairs], axis=1)
    return mean, cov


@pytest.mark.parametrize("dims", [(8, 17, 8), (7, 9, 9), (16, 33, 12), (10, 97, 12), (6, 129, 4)])
def test_plane_statistics_match_physical_grid(gpu, dims):
    import torch
    cf = gpu
    say("plane statistics", dims)
    gc, go, prob = make(cf, dims, profile=lambda y: 1 - y ** 2)
    rng = np.random.default_rng(31 + sum(dims))
    U = hermitian_random(go, rng, 3, 0.3)
    mean, cov = cf.plane_statistics(prob, torch.as_tensor(U, device="cuda"))
    mref, cref = _physical_statistics(go, U, 1 - go.y ** 2)
    em = np.abs(mean - mref).max() / np.abs(mref).max()
    ec = np.abs(cov - cref).max() / np.abs(cref).max()
    say(f"  mean rel {em:.2e} covariance rel {ec:.2e}")
    assert em < 1e-13 and ec < 1e-12
    # time averaging: two samples of different fields merge to the pooled statistics
    U2 = hermitian_random(go, rng, 3, 0.2)
    st = cf.ProfileStatistics(go.Ny)
    st.sample(prob, torch.as_tensor(U, device="cuda"), 0.0).sample(prob, torch.as_tensor(U2, device="cuda"), 1.0)
    m2, c2 = _physical_statistics(go, U2, 1 - go.y ** 2)
    pm = (mref + m2) / 2
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    pc = np.stack([(cref[:, k] + c2[:, k]) / 2 + ((mref[:, i] - pm[:, i]) * (mref[:, j] - pm[:, j]) +
                                                  (m2[:, i] - pm[:, i]) * (m2[:, j] - pm[:, j])) / 2
                   for k, (i, j) in enumerate(pairs)], axis=1)
    gm, gcv = st.profiles()
    assert np.abs(gm - pm).max() < 1e-13 * np.abs(pm).max()
    assert np.abs(gcv - pc).max() < 1e-12 * np.abs(pc).max()


@pytest.mark.parametrize("dims", [(8, 17, 8), (16, 33, 12), (12, 65, 10), (4, 255, 3), (10, 97, 12), (4, 257, 3)])
def test_velocity_gradient_matches_oracle(gpu, dims):
    """cf_velocity_gradient: dV_i/dx_j of V = U + Ub e_x at the padded points
    against the oracle's ddx1/ddx2/ddx3 followed by its inverse transform;
    for U = 0 the Poiseuille shear dUb/dy = -2y exactly."""
    import torch
    cf = gpu
    say("velocity gradient", dims)
    gc, go, prob = make(cf, dims, profile=lambda y: 1 - y ** 2)
    if (3 * dims[0] + 1) // 2 % 2:
        with pytest.raises(NotImplementedError):
            cf.velocity_gradient(prob, gc.zeros_spec(3))
        return
    rng = np.random.default_rng(41 + sum(dims))
    U = hermitian_random(go, rng, 3, 0.3)
    V = U.copy()
    V[0, :, 0, 0] += O.chebcoeffs(1 - go.y ** 2)
    got = cf.velocity_gradient(prob, torch.as_tensor(U, device="cuda")).cpu().numpy()
    worst = 0.0
    for i in range(3):
        for j, op in enumerate(("ddx1", "ddx2", "ddx3")):
            ref = O.inverse(go, O.operator(go, op, V[i]))
            worst = max(worst, np.abs(got[i, j] - ref).max() / max(np.abs(ref).max(), 1e-300))
    say(f"  worst rel err {worst:.2e}")
    assert worst < 1e-12
    lam = cf.velocity_gradient(prob, gc.zeros_spec(3)).cpu().numpy()
    y = go.y[:, None, None]
    # derivative of the transformed profile: coefficient roundoff amplified by ~n^2
    tol = 2e-16 * go.Ny ** 2
    assert np.abs(lam[0, 1] - (-2 * y)).max() < tol
    assert np.abs(np.delete(lam.reshape(9, -1), 1, axis=0)).max() < tol
