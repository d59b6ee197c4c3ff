"""forcing(tstage, U, F) on the device (cf_set_forcing; cnrk2.jl:149-207,
forcing.jl:3-11) against the oracle's forcing, which the closed-form heat
equation test pins: a time-dependent forcing F0 + sin(omega t) F1 sampled at
t + (0, 1/3, 3/4) dt, from a nonzero start time, through step_ and through
the Flows adapter with a shortened final step."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import _split, assert_parity, hermitian_random, make, roundoff_envelope, say

pytestmark = pytest.mark.gpu

W = 3.7


def _fields(go, seed):
    rng = np.random.default_rng(seed)
    return hermitian_random(go, rng, 3, 0.05), hermitian_random(go, rng, 3, 0.05)


@pytest.mark.parametrize("dims,case", [((8, 17, 8), "couette"), ((16, 33, 12), "couette"),
                                       ((12, 65, 10), "poiseuille"), ((6, 9, 5), "couette")])
def test_forced_step_matches_oracle(gpu, dims, case):
    import torch
    cf = gpu
    say("forced step", dims, case)
    if case == "couette":
        gc, go, prob = make(cf, dims)
        ref = O.Problem(go, lambda y: y, 1 / 400, 0.01)
    else:
        gc, go, prob = make(cf, dims, profile=lambda y: 1 - y ** 2, bulkvelocity=(2 / 3, 0.0))
        ref = O.Problem(go, lambda y: 1 - y ** 2, 1 / 400, 0.01, bulkvelocity=(2 / 3, 0.0))
    F0, F1 = _fields(go, 7 + sum(dims))
    F0d, F1d = torch.as_tensor(F0, device="cuda"), torch.as_tensor(F1, device="cuda")
    t0, n = 0.3, 4
    U0, P0 = ref.random_state(cf.gaussian_noise(gc, 0.1, 3))
    U, P = torch.as_tensor(U0, device="cuda"), torch.as_tensor(P0, device="cuda")
    calls = []

    def forcing(t, Uf, F):
        calls.append(t)
        F.add_(F0d + math.sin(W * t) * F1d)

    t = cf.step_(prob.scheme, prob.nlterm, U, P, t0, forcing=forcing, nsteps=n, **prob.constraint)
    assert abs(t - (t0 + n * 0.01)) < 1e-12
    want_t = [t0 + s * 0.01 + c * 0.01 for s in range(n) for c in (0.0, 1 / 3, 3 / 4)]
    assert np.allclose(calls, want_t, rtol=0, atol=1e-14)

    def ostep(a, b):
        ref.set_forcing(F0, F1, W, t0)
        return ref.step(a, b, n)

    U1, P1 = ostep(U0, P0)
    env = roundoff_envelope(lambda a, b: _split(ostep(a, b)), [U0, P0], _split((U1, P1)))
    assert_parity(f"forced {dims}", _split((U.cpu().numpy(), P.cpu().numpy())), _split((U1, P1)), env)
    # forcing is uninstalled after the call: a plain step equals the oracle's plain step
    ref.set_forcing(None)
    U2, P2 = ref.step(U1, P1, 1)
    cf.step_(prob.scheme, prob.nlterm, U, P, t, **prob.constraint)
    env2 = roundoff_envelope(lambda a, b: _split(ref.step(a, b, 1)), [U1, P1], _split((U2, P2)))
    assert_parity(f"unforced after {dims}", _split((U.cpu().numpy(), P.cpu().numpy())), _split((U2, P2)),
                  [max(e, v) for e, v in zip(env, env2)])


def test_forced_flow_shortened_step(gpu):
    """Flows adapter with the problem's forcing (ext/ChannelFlowFlowsExt.jl:48-49)."""
    import torch
    cf = gpu
    dims, dt, nu = (8, 17, 8), 0.01, 1 / 400
    gc = cf.Grid(*dims, 5.3, 7.1)
    go = O.Grid(*dims, 5.3, 7.1)
    F0, F1 = _fields(go, 11)
    F0d, F1d = torch.as_tensor(F0, device="cuda"), torch.as_tensor(F1, device="cuda")
    prob = cf.CouetteFlow(gc, nu, dt, forcing=lambda t, U, F: F.add_(F0d + math.sin(W * t) * F1d))
    ref = O.Problem(go, lambda y: y, nu, dt)
    U0, P0 = ref.random_state(cf.gaussian_noise(gc, 0.1, 5))
    st = cf.State(torch.as_tensor(U0, device="cuda"), torch.as_tensor(P0, device="cuda"))
    t0, T = 0.1, 0.1 + 3.4 * dt
    cf.flow(prob)(st, (t0, T))

    def oflow(a, b):
        ref.set_forcing(F0, F1, W, t0)
        a, b = ref.step(a, b, 3)
        r = O.Problem(go, lambda y: y, nu, T - t0 - 3 * dt)
        r.set_forcing(F0, F1, W, t0 + 3 * dt)
        return r.step(a, b, 1)

    U1, P1 = oflow(U0, P0)
    env = roundoff_envelope(lambda a, b: _split(oflow(a, b)), [U0, P0], _split((U1, P1)))
    assert_parity("forced flow", _split((st.velocity.cpu().numpy(), st.stagepressure.cpu().numpy())),
                  _split((U1, P1)), env)


def test_forcing_callback_error_propagates(gpu):
    import torch
    cf = gpu
    gc, go, prob = make(cf, (8, 17, 8))
    s = cf.random_state(prob, 0.1, 1)

    def bad(t, U, F):
        raise RuntimeError("forcing failed on purpose")

    with pytest.raises(RuntimeError, match="on purpose"):
        cf.step_(prob.scheme, prob.nlterm, s.velocity, s.stagepressure, 0.0, forcing=bad)
    torch.cuda.synchronize()


def test_forcing_using_a_helper_problem_leaves_the_step_intact(gpu):
    """A forcing that computes gradients (which use a problem's workspaces)
    on a helper problem of the same grid and adds nothing leaves the step
    bit-identical to the unforced one."""
    import torch
    cf = gpu
    gc, go, prob = make(cf, (16, 33, 12))
    helper = cf.ChannelFlowProblem(gc, prob.baseflow, 1 / 400, 0.01)
    s = cf.random_state(prob, 0.1, 4)
    a, b = s.copy(), s.copy()
    cf.step_(prob.scheme, prob.nlterm, a.velocity, a.stagepressure, 0.0, nsteps=2)
    gbuf = torch.empty((3, 3) + gc.phys_shape, dtype=torch.float64, device="cuda")

    def quiet(t, U, F):
        cf.velocity_gradient(helper, U, out=gbuf)

    cf.step_(prob.scheme, prob.nlterm, b.velocity, b.stagepressure, 0.0, nsteps=2, forcing=quiet)
    assert torch.equal(a.velocity, b.velocity) and torch.equal(a.stagepressure, b.stagepressure)


def test_forcing_and_nonlinear_on_a_side_stream(gpu):
    """step_(..., stream=s) with s not torch's current stream: the callbacks'
    device work runs on s (ordered with the stage kernels), so the result is
    bit-identical to the same step on the current stream, for a forcing and
    for a caller's nonlinear term."""
    import torch
    cf = gpu
    gc, go, prob = make(cf, (16, 33, 12), profile=lambda y: 1 - y ** 2, bulkvelocity=(2 / 3, 0.0))
    helper = cf.ChannelFlowProblem(gc, prob.baseflow, 1 / 400, 0.01)
    F0, F1 = _fields(go, 21)
    F0d, F1d = torch.as_tensor(F0, device="cuda"), torch.as_tensor(F1, device="cuda")
    s0 = cf.random_state(prob, 0.1, 6)

    def forcing(t, U, F):
        F.add_(F0d + math.sin(W * t) * F1d)

    def nonlinear(t, U, N):  # the problem's own form evaluated on a helper, plus a spin of device work
        N.copy_(helper.nlterm(t, U))
        for _ in range(20):
            N.mul_(1.0)

    for kw in ({"forcing": forcing}, {"nlterm": nonlinear}):
        nl = kw.get("nlterm", prob.nlterm)
        extra = {k: v for k, v in kw.items() if k != "nlterm"}
        a, b = s0.copy(), s0.copy()
        torch.cuda.synchronize()
        cf.step_(prob.scheme, nl, a.velocity, a.stagepressure, 0.0, nsteps=3, **extra, **prob.constraint)
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        cf.step_(prob.scheme, nl, b.velocity, b.stagepressure, 0.0, nsteps=3, stream=side, **extra,
                 **prob.constraint)
        side.synchronize()
        say(f"side stream {list(kw)}: max |dU| {float((a.velocity - b.velocity).abs().max()):.3e}")
        assert torch.equal(a.velocity, b.velocity) and torch.equal(a.stagepressure, b.stagepressure)
