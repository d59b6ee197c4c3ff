"""The learned closure registered as a device-side forcing (cf_set_closure)
against the same closure called back from the host at every stage: the
coarse closure workload's grid (Poiseuille 32x33x32, Lx = 4 pi, Lz = 4 pi / 3,
nu = 1/2800, bulk velocity (1, 0)), an untrained network with its initial
weights and a nonzero feature normalisation, both the full force and its
plane-mean part.  Both paths enqueue the same kernels, so the states agree
bit for bit; the device path is run in several calls so that its cached
step graph is replayed.  Also the registration's argument checks."""
import math

import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu

N = (32, 33, 32)
L = (4 * math.pi, 4 * math.pi / 3)
NU, DT = 1 / 2800, 0.005


def base(y):
    return 1.5 * (1 - y ** 2)


@pytest.fixture(scope="module")
def closure_setup(gpu):
    import torch
    from channelflow import closure as CL
    cf = gpu
    prob = cf.ChannelFlowProblem(cf.Grid(*N, *L), base, NU, DT, bulkvelocity=(1.0, 0.0))
    st = cf.perturbation_state(prob, 0.3, seed=5)
    net = CL.SubgridMLP(seed=3)
    rng = np.random.default_rng(7)
    with torch.no_grad():
        net.norm[:10] = torch.as_tensor(0.1 * rng.standard_normal(10), dtype=torch.float32, device="cuda")
        net.norm[10:20] = torch.as_tensor(0.5 + rng.random(10), dtype=torch.float32, device="cuda")
        net.norm[20:] = 0.02  # a force of a few percent of the nonlinear term
    return cf, prob, st, net


def _run(cf, prob, st, forcing, calls):
    import torch
    U, P = st.velocity.clone(), st.stagepressure.clone()
    t = 0.0
    for n in calls:
        t = cf.step_(prob.scheme, prob.nlterm, U, P, t, nsteps=n, forcing=forcing, **prob.constraint)
    torch.cuda.synchronize()
    return U.cpu().numpy(), P.cpu().numpy()


@pytest.mark.parametrize("modes", ["all", "mean"])
def test_device_closure_matches_host_callback(closure_setup, modes):
    cf, prob, st, net = closure_setup
    host = _run(cf, prob, st, net.forcing(prob, modes=modes, device=False), [4])
    dev_force = net.forcing(prob, modes=modes, device=True)
    dev = _run(cf, prob, st, dev_force, [1, 1, 2])  # capture, then replays of the cached graph
    again = _run(cf, prob, st, dev_force, [2, 2])
    plain = _run(cf, prob, st, None, [4])
    d_force = max(np.abs(host[0] - plain[0]).max(), np.abs(host[1] - plain[1]).max())
    say(f"closure ({modes}): change of the state by the force over 4 steps {d_force:.3e}")
    assert d_force > 1e-8  # the force acts
    for name, a, b in (("U", dev[0], host[0]), ("P", dev[1], host[1]), ("U again", again[0], host[0]),
                       ("P again", again[1], host[1])):
        same = np.array_equal(a, b)
        say(f"closure ({modes}) {name}: device == host callback: {same}")
        assert same, np.abs(a - b).max()
    # the registration ends with the call: a plain step afterwards is the plain step
    assert np.array_equal(_run(cf, prob, st, None, [4])[0], plain[0])


def test_device_closure_registration_checks(closure_setup):
    import torch
    from channelflow import _lib
    cf, prob, st, net = closure_setup
    f = net.forcing(prob, device=True)
    L_ = _lib.load()
    F = prob.grid.zeros_spec(3)
    p = cf._ptr
    args = (p(net.theta), p(net.wbf), p(net.norm), p(f.gbuf), p(f.fbuf), p(F[0]), p(F[1]), p(F[2]), 0)
    h = prob.scheme.h
    assert L_.cf_set_closure(h, h, *args) != 0  # the problem cannot be its own helper
    other = cf.ChannelFlowProblem(cf.Grid(16, 33, 16, *L), base, NU, DT, bulkvelocity=(1.0, 0.0))
    assert L_.cf_set_closure(h, other.scheme.h, *args) != 0  # the helper must have the grid
    assert L_.cf_set_closure(h, f.helper.scheme.h, *args) == 0
    cb = _lib.FORCING_FN(lambda *a: 0)
    assert L_.cf_set_forcing(h, cb, None, p(F[0]), p(F[1]), p(F[2])) != 0  # exclusive with a host forcing
    assert L_.cf_set_closure(h, None, None, None, None, None, None, None, None, None, 0) == 0
    assert L_.cf_set_forcing(h, cb, None, p(F[0]), p(F[1]), p(F[2])) == 0
    assert L_.cf_set_closure(h, f.helper.scheme.h, *args) != 0  # and the other way round
    assert L_.cf_set_forcing(h, _lib.FORCING_FN(), None, None, None, None) == 0
    torch.cuda.synchronize()
