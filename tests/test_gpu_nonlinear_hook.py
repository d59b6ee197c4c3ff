"""nonlinear(tstage, U, N) given by the caller (cf_set_nonlinear;
cnrk2.jl:121-126,172-173): the step adds the caller's N to Q as the signed
nonlinear acceleration, in place of its own form.  A callable that evaluates
the problem's form on a second problem of the same grid and adds
sin(W t) F1 must reproduce the oracle's step with F1 sin(W t) as forcing
(the oracle's forcing is pinned by the closed-form heat equation test).
The callable is sampled at t + (0, 1/3, 3/4) dt, composes with a forcing,
and an exception raised in it propagates and leaves the problem usable."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import _split, assert_parity, hermitian_random, make, roundoff_envelope, say

pytestmark = pytest.mark.gpu

W = 3.7
DT = 0.01


def _setup(cf, dims, form, seed):
    import torch
    gc, go, prob = make(cf, dims, form=form)
    helper = cf.ChannelFlowProblem(gc, lambda y: y, 1 / 400, DT, form=form)
    ref = O.Problem(go, lambda y: y, 1 / 400, DT, form=form)
    F1 = hermitian_random(go, np.random.default_rng(seed), 3, 0.05)
    U0, P0 = ref.random_state(cf.gaussian_noise(gc, 0.1, seed))
    return gc, go, prob, helper, ref, F1, torch.as_tensor(F1, device="cuda"), U0, P0


@pytest.mark.parametrize("dims,form", [((8, 17, 8), "rotating"), ((16, 33, 12), "rotating"),
                                       ((12, 65, 10), "convective"), ((8, 97, 6), "divergence")])
def test_caller_nonlinear_matches_oracle(gpu, dims, form):
    import torch
    cf = gpu
    say("caller nonlinear", dims, form)
    gc, go, prob, helper, ref, F1, F1d, U0, P0 = _setup(cf, dims, form, 11 + sum(dims))
    U, P = torch.as_tensor(U0, device="cuda"), torch.as_tensor(P0, device="cuda")
    t0, n = 0.2, 3
    calls = []

    def nonlinear(t, Uc, N):
        calls.append(t)
        helper.nlterm(t, Uc, N)  # overwrites N, on the helper's workspaces
        N.add_(math.sin(W * t) * F1d)

    t = cf.step_(prob.scheme, nonlinear, U, P, t0, nsteps=n, **prob.constraint)
    assert abs(t - (t0 + n * DT)) < 1e-12
    want_t = [t0 + s * DT + c * DT for s in range(n) for c in (0.0, 1 / 3, 3 / 4)]
    assert np.allclose(calls, want_t, rtol=0, atol=1e-14)

    Z = np.zeros_like(F1)

    def ostep(a, b):
        ref.set_forcing(Z, F1, W, t0)
        return ref.step(a, b, n)

    U1, P1 = ostep(U0, P0)
    env = roundoff_envelope(lambda a, b: _split(ostep(a, b)), [U0, P0], _split((U1, P1)))
    assert_parity(f"caller nonlinear {dims} {form}", _split((U.cpu().numpy(), P.cpu().numpy())), _split((U1, P1)), env)
    ref.set_forcing(None)


def test_caller_nonlinear_composes_with_forcing(gpu):
    """N from the callable plus F from forcing equals N + F from the callable."""
    import torch
    cf = gpu
    dims = (16, 33, 12)
    say("caller nonlinear + forcing", dims)
    gc, go, prob, helper, ref, F1, F1d, U0, P0 = _setup(cf, dims, "rotating", 5)

    def nl_plain(t, Uc, N):
        helper.nlterm(t, Uc, N)

    def nl_sum(t, Uc, N):
        helper.nlterm(t, Uc, N)
        N.add_(math.cos(W * t) * F1d)

    def forcing(t, Uc, F):
        F.add_(math.cos(W * t) * F1d)

    Ua, Pa = torch.as_tensor(U0, device="cuda"), torch.as_tensor(P0, device="cuda")
    Ub, Pb = Ua.clone(), Pa.clone()
    cf.step_(prob.scheme, nl_plain, Ua, Pa, 0.1, forcing=forcing, nsteps=3, **prob.constraint)
    cf.step_(prob.scheme, nl_sum, Ub, Pb, 0.1, nsteps=3, **prob.constraint)
    scale = float(torch.view_as_real(Ub).abs().max())
    assert float(torch.view_as_real(Ua - Ub).abs().max()) <= 1e-13 * scale
    assert float(torch.view_as_real(Pa - Pb).abs().max()) <= 1e-13 * max(float(torch.view_as_real(Pb).abs().max()), 1e-300)


def test_caller_nonlinear_error_propagates(gpu):
    import torch
    cf = gpu
    dims = (8, 17, 8)
    say("caller nonlinear error", dims)
    gc, go, prob, helper, ref, F1, F1d, U0, P0 = _setup(cf, dims, "rotating", 3)
    U, P = torch.as_tensor(U0, device="cuda"), torch.as_tensor(P0, device="cuda")

    class Boom(Exception):
        pass

    def bad(t, Uc, N):
        raise Boom("nonlinear failed")

    with pytest.raises(Boom):
        cf.step_(prob.scheme, bad, U, P, 0.0, **prob.constraint)
    with pytest.raises(cf.ArgumentError):
        cf.step_(prob.scheme, "not callable", U, P, 0.0, **prob.constraint)
    # the hook is uninstalled: the problem's own form steps as the oracle's
    U, P = torch.as_tensor(U0, device="cuda"), torch.as_tensor(P0, device="cuda")
    cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=2, **prob.constraint)
    U1, P1 = ref.step(U0, P0, 2)
    env = roundoff_envelope(lambda a, b: _split(ref.step(a, b, 2)), [U0, P0], _split((U1, P1)))
    assert_parity("after error", _split((U.cpu().numpy(), P.cpu().numpy())), _split((U1, P1)), env)
