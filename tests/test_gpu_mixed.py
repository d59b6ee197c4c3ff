"""Mixed-precision step (CF_PRECISION_MIXED): FP32 y-physical fields, x/z
FFTs and nonlinear products; FP64 state, Chebyshev arithmetic and solves.
Compared with the FP64 device step, which the parity tests pin to the
oracle.  The bound follows from FP32 rounding: each stored physical value and
each FFT pass carries a relative error of about 2^-24, so the nonlinear term
is accurate to ~1e-6 of its own size and a step changes U by dt times that."""
import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu

# even padded Nx (the mixed path's requirement); group sizes L = 8, 16, 32, 64
MIXED_GRIDS = [(8, 17, 8), (16, 33, 12), (32, 33, 32), (12, 65, 10), (4, 255, 3), (16, 97, 16), (4, 257, 4)]


def _pair(cf, dims, form=None, **kw):
    g = cf.Grid(*dims, 2 * np.pi / 1.14, 2 * np.pi / 2.5)
    extra = {} if form is None else {"form": form}
    p64 = cf.CouetteFlow(g, 1 / 400, 0.01, **extra, **kw)
    p32 = cf.CouetteFlow(g, 1 / 400, 0.01, precision="mixed", **extra, **kw)
    return g, p64, p32


@pytest.mark.parametrize("dims", MIXED_GRIDS)
@pytest.mark.parametrize("form", ["rotating", "convective"])
def test_mixed_step_close_to_fp64(gpu, dims, form):
    import torch
    cf = gpu
    say("mixed step", dims, form)
    g, p64, p32 = _pair(cf, dims, form)
    s0 = cf.random_state(p64, 0.1, 3)
    a, b = s0.copy(), s0.copy()
    nsteps = 5
    cf.step_(p64.scheme, p64.nlterm, a.velocity, a.stagepressure, 0.0, nsteps=nsteps)
    cf.step_(p32.scheme, p32.nlterm, b.velocity, b.stagepressure, 0.0, nsteps=nsteps)
    du = (b.velocity - a.velocity).abs().max().item()
    change = (a.velocity - s0.velocity).abs().max().item()
    scale = a.velocity.abs().max().item()
    rel = du / change
    say(f"  |dU| {du:.3e}  step change {change:.3e}  |U| {scale:.3e}  rel-to-change {rel:.3e}")
    assert du > 0.0  # the FP32 path ran
    assert rel < 1e-6  # FP32: ~6e-8 per value x FFT depth, on the nonlinear part of the change
    E64, E32 = cf.kinetic_energy(p64, a.velocity), cf.kinetic_energy(p32, b.velocity)
    assert abs(E32 - E64) <= 1e-6 * E64


def test_mixed_problem_other_calls_stay_fp64(gpu):
    """Only the time step is mixed: transforms and diagnostics of a mixed
    problem equal the FP64 problem's bit for bit."""
    import torch
    cf = gpu
    g, p64, p32 = _pair(cf, (16, 33, 12))
    s = cf.random_state(p64, 0.1, 5)
    s32 = cf.random_state(p32, 0.1, 5)
    assert torch.equal(s.velocity, s32.velocity) and torch.equal(s.stagepressure, s32.stagepressure)
    assert torch.equal(cf.IFFT(p64, s.velocity[0]), cf.IFFT(p32, s.velocity[0]))
    assert cf.dissipation_rate(p64, s.velocity) == cf.dissipation_rate(p32, s.velocity)


def test_mixed_bulk_flux_poiseuille(gpu):
    """Bulk-flux constraint under the mixed step: the constraint is imposed
    by the FP64 mean-mode solve, so the flux holds to FP64 roundoff."""
    import torch
    cf = gpu
    g = cf.Grid(16, 33, 16, 4 * np.pi, 4 * np.pi / 3)
    p = cf.PoiseuilleFlow(g, 1 / 2800, 0.005, bulkvelocity=(2 / 3, 0.0), precision="mixed")
    s = cf.random_state(p, 0.05, 9)
    cf.step_(p.scheme, p.nlterm, s.velocity, s.stagepressure, 0.0, nsteps=10, **p.constraint)
    u00 = s.velocity[0, :, 0, 0].real.cpu().numpy()
    from helpers import bulkmean
    bulk = float(bulkmean(u00))  # perturbation bulk mean (base flow carries 2/3)
    say(f"  perturbation bulk after 10 mixed steps {bulk:.3e}")
    assert abs(bulk) < 1e-13
