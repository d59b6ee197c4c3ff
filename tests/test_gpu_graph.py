"""The plain step replayed from its cached one-step graph (cf_advance) against
the same step launched kernel by kernel (cf_step_timed): bit-identical, for
several field sets (more than the cache holds, so graphs are evicted and
rebuilt), both mean-flow constraints on one field set, after a change of
tiling, on a side stream, and in the mixed-precision step."""
import math

import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu

N = (32, 33, 32)
L = (2 * math.pi, math.pi)
NU, DT = 1 / 2800, 0.005


def base(y):
    return 1.5 * (1 - y ** 2)


def eager(cf, prob, st, nsteps, **constraint):
    import torch
    U, P = st.velocity.clone(), st.stagepressure.clone()
    for s in range(nsteps):
        cf.step_timed(prob.scheme, U, P, s * DT, **constraint)
    torch.cuda.synchronize()
    return U.cpu().numpy(), P.cpu().numpy()


def graphed(cf, prob, st, nsteps, stream=None, **constraint):
    import torch
    U, P = st.velocity.clone(), st.stagepressure.clone()
    cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=nsteps, stream=stream, **constraint)
    torch.cuda.synchronize()
    return U, P


def same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.fixture(scope="module")
def setup(gpu):
    cf = gpu
    prob = cf.ChannelFlowProblem(cf.Grid(*N, *L), base, NU, DT, bulkvelocity=(1.0, 0.0))
    st = cf.perturbation_state(prob, 0.05, seed=3)
    return cf, prob, st


def test_graph_matches_eager(setup):
    cf, prob, st = setup
    ref = eager(cf, prob, st, 4, **prob.constraint)
    keep = []  # distinct storage for every run: six pointer sets through a cache of four
    for i in range(6):
        U, P = graphed(cf, prob, st, 4, **prob.constraint)
        keep.append((U, P))
        ok = same((U.cpu().numpy(), P.cpu().numpy()), ref)
        say(f"graph run {i}: identical {ok}")
        assert ok
    # replays of cached graphs on the first field sets again
    for U, P in keep[:2]:
        U.copy_(st.velocity)
        P.copy_(st.stagepressure)
        cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=4, **prob.constraint)
        assert same((U.cpu().numpy(), P.cpu().numpy()), ref)


def test_graph_constraints(setup):
    import torch
    cf, prob, st = setup
    U, P = st.velocity.clone(), st.stagepressure.clone()
    for c in (dict(bulkvelocity=(1.0, 0.0)), dict(pressuregradient=(-0.01, 0.0)), dict(bulkvelocity=(1.2, 0.1))):
        ref = eager(cf, prob, st, 3, **c)
        U.copy_(st.velocity)
        P.copy_(st.stagepressure)
        cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=3, **c)
        torch.cuda.synchronize()
        ok = same((U.cpu().numpy(), P.cpu().numpy()), ref)
        say(f"constraint {c}: identical {ok}")
        assert ok


def test_graph_after_retiling_and_on_side_stream(setup):
    import torch
    cf, prob, st = setup
    ref = eager(cf, prob, st, 3, **prob.constraint)
    keep = cf.get_tiles(prob)
    cf.set_tiles(prob, 1, keep[2], 1)
    try:
        U, P = graphed(cf, prob, st, 3, **prob.constraint)
        assert same((U.cpu().numpy(), P.cpu().numpy()), ref)
    finally:
        cf.set_tiles(prob, keep[0], keep[2], keep[3])
    s = torch.cuda.Stream()
    U, P = st.velocity.clone(), st.stagepressure.clone()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=3, stream=s, **prob.constraint)
    s.synchronize()
    assert same((U.cpu().numpy(), P.cpu().numpy()), ref)


def test_graph_mixed(gpu):
    cf = gpu
    prob = cf.ChannelFlowProblem(cf.Grid(*N, *L), base, NU, DT, bulkvelocity=(1.0, 0.0), precision="mixed")
    st = cf.perturbation_state(prob, 0.05, seed=3)
    ref = eager(cf, prob, st, 3, **prob.constraint)
    U, P = graphed(cf, prob, st, 3, **prob.constraint)
    assert same((U.cpu().numpy(), P.cpu().numpy()), ref)


def test_step_under_caller_capture(setup):
    """A caller capturing its own graph around step_ records the step's
    kernels (the library replays no graph of its own inside a capture).
    Every replay, with eager copies and synchronisations between them,
    equals the eager steps."""
    import torch
    cf, prob, st = setup
    ref = eager(cf, prob, st, 2, **prob.constraint)
    U, P = st.velocity.clone(), st.stagepressure.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # first use of these buffers outside the capture
        cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=1, **prob.constraint)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, capture_error_mode="relaxed"):
        cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=2, **prob.constraint)
    torch.cuda.synchronize()
    for rep in range(3):
        U.copy_(st.velocity)
        P.copy_(st.stagepressure)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        ok = same((U.cpu().numpy(), P.cpu().numpy()), ref)
        say(f"caller graph replay {rep}: identical {ok}")
        assert ok
