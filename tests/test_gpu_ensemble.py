"""Ensembles (cf_ensemble_*): uneven-grid, uneven-Re instances advanced by
batched launches in one captured HIP graph evolve exactly as each problem
stepped on its own by step_ (cnrk2.jl:149-207), bit for bit, with or without
the tile autotuner."""
import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu

# (grid, nu, kind): uneven grids (group sizes L = 8 and 16, several in-place
# widths), Couette and bulk-flux Poiseuille, rotating and convective forms
CASES = [((16, 17, 16), 1 / 300, "couette", "rotating"),
         ((24, 33, 24), 1 / 450, "couette", "rotating"),
         ((8, 17, 8), 1 / 400, "couette", "convective"),
         ((16, 33, 12), 1 / 500, "poiseuille", "rotating"),
         ((32, 33, 32), 1 / 350, "couette", "rotating"),
         ((12, 49, 10), 1 / 420, "couette", "rotating")]


def _problem(cf, dims, nu, kind, form):
    g = cf.Grid(*dims, 2 * np.pi / 1.14, 2 * np.pi / 2.5)
    if kind == "couette":
        return cf.CouetteFlow(g, nu, 0.01, form=form)
    return cf.PoiseuilleFlow(g, nu, 0.01, bulkvelocity=(2 / 3, 0.0), form=form)


@pytest.mark.parametrize("tune", [False, True])
def test_ensemble_matches_individual_steps(gpu, tune):
    import torch
    cf = gpu
    say("ensemble", "tune" if tune else "no tune")
    ens_probs, ens_states, ref = [], [], []
    for i, (dims, nu, kind, form) in enumerate(CASES):
        p = _problem(cf, dims, nu, kind, form)
        s = cf.random_state(p, 0.1, 100 + i)
        ens_probs.append(p)
        ens_states.append(s)
        q = _problem(cf, dims, nu, kind, form)
        ref.append((q, s.copy()))
    ens = cf.Ensemble(ens_probs, ens_states, tune=tune)
    launches, tiles = ens.info()
    say(f"  launches per ensemble step {launches} (instances x 13 = {13 * len(CASES)}), tiles {tiles}")
    assert launches < 13 * len(CASES)
    nsteps = 7
    ens.advance(nsteps)
    for q, s in ref:
        cf.step_(q.scheme, q.nlterm, s.velocity, s.stagepressure, 0.0, nsteps=nsteps, **q.constraint)
    torch.cuda.synchronize()
    for i, ((q, s), e) in enumerate(zip(ref, ens_states)):
        same = torch.equal(e.velocity, s.velocity) and torch.equal(e.stagepressure, s.stagepressure)
        if not same:
            say(f"  instance {i} max |dU| {(e.velocity - s.velocity).abs().max().item():.3e}")
        assert same, i
    del ens


def test_ensemble_rejects_bad_inputs(gpu):
    cf = gpu
    p = _problem(cf, (8, 17, 8), 1 / 400, "couette", "rotating")
    s1, s2 = cf.random_state(p, 0.1, 1), cf.random_state(p, 0.1, 2)
    with pytest.raises(cf.ArgumentError):
        cf.Ensemble([p, p], [s1, s2])  # shared workspaces
    with pytest.raises(cf.ArgumentError):
        cf.Ensemble([p, _problem(cf, (8, 17, 8), 1 / 400, "couette", "rotating")], [s1, s1])  # aliased fields
    g = p.grid
    with pytest.raises(NotImplementedError):
        cf.Ensemble([cf.CouetteFlow(g, 1 / 400, 0.01, precision="mixed")], [s1])
    with pytest.raises(NotImplementedError):
        cf.Ensemble([cf.CouetteFlow(g, 1 / 400, 0.01, form="divergence")], [s1])


def test_problem_tune_keeps_results_bit_identical(gpu):
    """cf_problem_tune changes only tile shapes: the step is bit-identical."""
    import torch
    cf = gpu
    dims = (32, 33, 32)
    a_p = _problem(cf, dims, 1 / 400, "couette", "rotating")
    b_p = _problem(cf, dims, 1 / 400, "couette", "rotating")
    tiles = cf.tune(b_p)
    say(f"  tuned tiles (KT, z ppt, x waves) {tiles}")
    s = cf.random_state(a_p, 0.1, 9)
    a, b = s.copy(), s.copy()
    cf.step_(a_p.scheme, a_p.nlterm, a.velocity, a.stagepressure, 0.0, nsteps=4)
    cf.step_(b_p.scheme, b_p.nlterm, b.velocity, b.stagepressure, 0.0, nsteps=4)
    assert torch.equal(a.velocity, b.velocity) and torch.equal(a.stagepressure, b.stagepressure)
