"""GPU-resident random_state (initialization.jl:32-51): the device Gaussian
stream against the reference Philox4x32-10 + Box-Muller (pinned to the
Random123 known answers on the CPU), and the device FFT -> project! ->
pressure chain against the oracle's random_state on the same noise."""
import numpy as np
import pytest

from oracle import oracle as O
from philox_ref import gaussian
from test_gpu_parity import _split, assert_parity, make, roundoff_envelope, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(8, 17, 8), (7, 9, 9), (32, 33, 32)])
def test_device_noise_stream(gpu, dims):
    cf = gpu
    gc, go, prob = make(cf, dims)
    eps, seed = 0.1, 0x123456789ABCDEF
    noise = cf.device_gaussian_noise(prob, eps, seed).cpu().numpy().reshape(3, -1)
    n = noise.shape[1]
    for comp in range(3):
        ref = gaussian(n, eps, seed, comp)[:n]
        err = np.abs(noise[comp] - ref).max()
        say(f"  {dims} component {comp}: max |device - reference| {err:.2e}")
        assert err < 1e-14
    # same seed reproduces, another seed differs
    again = cf.device_gaussian_noise(prob, eps, seed).cpu().numpy().reshape(3, -1)
    other = cf.device_gaussian_noise(prob, eps, seed + 1).cpu().numpy().reshape(3, -1)
    assert np.array_equal(again, noise) and not np.allclose(other, noise)
    if n >= 50000:
        assert abs(noise.mean()) < 5 * eps / np.sqrt(noise.size)
        assert abs(noise.std() / eps - 1) < 5 / np.sqrt(noise.size)


@pytest.mark.parametrize("dims", [(8, 17, 8), (16, 33, 12), (8, 97, 6)])
def test_random_state_matches_oracle_on_device_noise(gpu, dims):
    cf = gpu
    say("random_state", dims)
    gc, go, prob = make(cf, dims, profile=lambda y: 1 - y ** 2, pressuregradient=(-2 / 400, 0.0))
    ref = O.Problem(go, lambda y: 1 - y ** 2, 1 / 400, 0.01, pressuregradient=(-2 / 400, 0.0))
    st = cf.random_state(prob, 0.1, seed=99)
    noise = cf.device_gaussian_noise(prob, 0.1, 99).cpu().numpy()
    Ur, Pr = ref.random_state(noise)
    env = roundoff_envelope(lambda x: _split(ref.random_state(x)), [noise], _split((Ur, Pr)))
    assert_parity(f"random_state {dims}", _split((st.velocity.cpu().numpy(), st.stagepressure.cpu().numpy())),
                  _split((Ur, Pr)), env)
