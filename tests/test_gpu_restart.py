"""Checkpoint/restart through the portable state file (savefield/loadfield,
src/io.jl:21-54; docs/src/restart.md): stepping n steps, saving, reloading
into a fresh problem and stepping m more reproduces n+m uninterrupted steps
bit for bit; a state written from the oracle's arrays steps on the device
like the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import assert_parity, make, roundoff_envelope, say, _split

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["couette", "poiseuille_bulk"])
def test_restart_reproduces_uninterrupted_run(gpu, tmp_path, case):
    import torch
    cf = gpu
    dims = (16, 33, 12)
    say("restart", case, dims)

    def problem():
        g = cf.Grid(*dims, 2 * np.pi / 1.14, 2 * np.pi / 2.5)
        if case == "couette":
            return cf.CouetteFlow(g, 1 / 400, 0.01)
        return cf.PoiseuilleFlow(g, 1 / 400, 0.01, bulkvelocity=(2 / 3, 0.0))

    prob = problem()
    s = cf.random_state(prob, 0.1, 11)
    t = cf.step_(prob.scheme, prob.nlterm, s.velocity, s.stagepressure, 0.0, **prob.constraint, nsteps=10)
    full = s.copy()

    prob = problem()
    s = cf.random_state(prob, 0.1, 11)
    t5 = cf.step_(prob.scheme, prob.nlterm, s.velocity, s.stagepressure, 0.0, **prob.constraint, nsteps=5)
    path = tmp_path / "chk.cf"
    cf.savefield(path, prob.grid, state=s, time=t5, step=5)
    del s, prob
    prob = problem()
    s = cf.loadfield(path, prob.grid)
    meta = cf.io.loadfield(path)
    t2 = cf.step_(prob.scheme, prob.nlterm, s.velocity, s.stagepressure, meta.time, **prob.constraint, nsteps=5)
    assert meta.step == 5 and abs(t2 - t) < 1e-12
    assert torch.equal(s.velocity, full.velocity) and torch.equal(s.stagepressure, full.stagepressure)


def test_oracle_state_file_steps_like_oracle(gpu, tmp_path):
    import torch
    cf = gpu
    dims = (8, 17, 8)
    gc, go, prob = make(cf, dims)
    ref = O.Problem(go, lambda y: y, 1 / 400, 0.01)
    U0, P0 = ref.random_state(cf.gaussian_noise(gc, 0.1, 5))
    path = tmp_path / "oracle.cf"
    cf.io.savefield(path, gc, velocity=U0, pressure=P0, time=0.0)
    s = cf.loadfield(path, gc)
    assert np.array_equal(s.velocity.cpu().numpy(), U0) and np.array_equal(s.stagepressure.cpu().numpy(), P0)
    cf.step_(prob.scheme, prob.nlterm, s.velocity, s.stagepressure, 0.0, nsteps=3)
    U1, P1 = ref.step(U0, P0, 3)
    env = roundoff_envelope(lambda a, b: _split(ref.step(a, b, 3)), [U0, P0], _split((U1, P1)))
    assert_parity("restart from oracle file", _split((s.velocity.cpu().numpy(), s.stagepressure.cpu().numpy())),
                  _split((U1, P1)), env)
    with pytest.raises(cf.ArgumentError):
        cf.loadfield(path, cf.Grid(8, 17, 10, gc.Lx, gc.Lz))

