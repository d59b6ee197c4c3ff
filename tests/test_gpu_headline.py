"""Parity at the headline workload: plane Poiseuille 128x97x128, Lx = 4 pi,
Lz = 4 pi / 3, nu = 1/2800, dt = 0.005, base 1.5 (1 - y^2), bulk velocity
(1, 0), rotational form, autotuned tiles, the seeded kappa^-4 perturbation
of amplitude 0.05 that bench.py steps.  The reference checks CPU against GPU
at N = 32, 64, 128 after one step (test/cuda/runtests.jl:152-175); here the
transforms and the nonlinear term (FFT lengths 192 in z and a 96-point
half-length x transform) are held to 1e-13 / 1e-12, and one and five steps
of U, P and the bulk gradient to the oracle within its roundoff envelope
(8 perturbation trials).  At this shape the oracle's own output moves by
~1e-10 in U and ~1e-9 in P when its inputs are perturbed at 1e-16
(lambda/nu = 3.4e6 in the first stage), so a flat 1e-12 is out of reach of
any FP64 implementation.  After one step the device is judged against the
extended-precision oracle (the same algorithm in long double on the same
inputs): on the mean mode and on the modes outside the few smallest
wavenumbers it must be as accurate as the FP64 oracle (up to 2x); overall,
where those smallest wavenumbers amplify each rounding by ~1e5, its distance
must stay within 2x the larger of the oracle's and the roundoff envelope:
one FP64 rounding sequence against another lands anywhere inside that
envelope (the 8-trial envelope is a sample maximum of that spread)."""
import math

import numpy as np
import pytest

from accuracy import class_errors, report
from oracle import oracle as O
from test_gpu_parity import ENVELOPE_FACTOR, _split, assert_parity, rel, roundoff_envelope, say

pytestmark = pytest.mark.gpu

N = (128, 97, 128)
L = (4 * math.pi, 4 * math.pi / 3)
NU, DT = 1 / 2800, 0.005


def base(y):
    return 1.5 * (1 - y ** 2)


@pytest.fixture(scope="module")
def headline(gpu):
    cf = gpu
    g = cf.Grid(*N, *L)
    prob = cf.ChannelFlowProblem(g, base, NU, DT, bulkvelocity=(1.0, 0.0))
    tiles = cf.tune(prob)
    say(f"headline tiles {tiles}")
    st = cf.perturbation_state(prob, 0.05, seed=42)
    go = O.Grid(*N, *L)
    ref = O.Problem(go, base, NU, DT, bulkvelocity=(1.0, 0.0))
    return cf, g, go, prob, ref, st


def test_headline_transforms(headline):
    import torch
    cf, g, go, prob, ref, st = headline
    rng = np.random.default_rng(128)
    u = rng.standard_normal((3,) + go.phys_shape)
    U3 = cf.FFT(prob, torch.as_tensor(u, device="cuda")).cpu().numpy()
    for c in range(3):
        e = rel(U3[c], O.forward(go, u[c]))
        say(f"forward[{c}] rel err {e:.3e}")
        assert e < 1e-13
    V = st.velocity.cpu().numpy()
    v3 = cf.IFFT(prob, st.velocity).cpu().numpy()
    for c in range(3):
        e = rel(v3[c], O.inverse(go, V[c]))
        say(f"inverse[{c}] rel err {e:.3e}")
        assert e < 1e-13


def test_headline_nonlinear(headline):
    cf, g, go, prob, ref, st = headline
    N3 = prob.nlterm(0.0, st.velocity).cpu().numpy()
    Nr = ref.nonlinear(st.velocity.cpu().numpy())
    for c in range(3):
        e = rel(N3[c], Nr[c])
        say(f"nonlinear[{c}] rel err {e:.3e}")
    e = rel(N3, Nr)
    assert e < 1e-12, e


@pytest.mark.parametrize("nsteps", [1, 5])
def test_headline_steps(headline, nsteps):
    import torch
    cf, g, go, prob, ref, st = headline
    U0, P0 = st.velocity.cpu().numpy(), st.stagepressure.cpu().numpy()
    U = torch.as_tensor(U0, device="cuda")
    P = torch.as_tensor(P0, device="cuda")
    t = cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=nsteps, **prob.constraint)
    assert abs(t - nsteps * DT) < 1e-15
    grad = cf.last_gradient(prob.scheme)
    U1, P1 = ref.step(U0, P0, nsteps)
    gref = ref.last_gradient()
    say(f"headline {nsteps} steps: oracle done")
    env = roundoff_envelope(lambda a, b: _split(ref.step(a, b, nsteps)), [U0, P0], _split((U1, P1)), trials=8)
    Uh, Ph = U.cpu().numpy(), P.cpu().numpy()
    assert_parity(f"headline {nsteps} steps", _split((Uh, Ph)), _split((U1, P1)), env)
    if nsteps == 1:
        ext = O.Problem(go, base, NU, DT, bulkvelocity=(1.0, 0.0), extended=True)
        XU, XP = ext.step(U0, P0, 1)
        ehU, eoU = report(say, "headline U after 1", Uh, U1, XU, go)
        ehP, eoP = report(say, "headline P after 1", Ph, P1, XP, go)
        # envelopes relative to the joint maximum of each field
        envU = max(env[i] * np.abs(U1[i]).max() for i in range(3)) / np.abs(U1).max()
        envP = env[3]
        say(f"headline after 1: envelope U {envU:.3e} P {envP:.3e}")
        assert ehU <= 2.0 * max(eoU, envU) and ehP <= 2.0 * max(eoP, envP), (ehU, eoU, envU, ehP, eoP, envP)
        for name, a, o, x in (("U", Uh, U1, XU), ("P", Ph, P1, XP)):
            ch, co = class_errors(a, x, go), class_errors(o, x, go)
            assert ch["rest"] <= 2.0 * co["rest"], (name, ch, co)
            assert ch["mean"] <= 2.0 * co["mean"] + 1e-15, (name, ch, co)
    dg = max(abs(a - b) / abs(b) if b else abs(a) for a, b in zip(grad, gref))
    say(f"headline {nsteps} steps: gradient {grad} oracle {gref} rel {dg:.3e}")
    assert dg <= max(1e-10, ENVELOPE_FACTOR * env[0]), dg
    # the flux is held: bulk mean of the total streamwise velocity stays 1
    from helpers import bulkmean
    bulk = bulkmean(np.real(Uh[0, :, 0, 0]) + O.chebcoeffs(base(go.y)))
    assert abs(bulk - 1.0) < 1e-13, bulk
