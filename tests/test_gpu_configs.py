"""The benchmark workloads beyond the headline, each exercised at its own
shape against the oracle (the headline 128x97x128 has tests/test_gpu_headline.py):

  256x193x256 plane Poiseuille (Lx = 4 pi, Lz = 4 pi/3, nu = 1/2800,
      dt = 0.0025, bulk velocity (1, 0), base 1.5 (1 - y^2)): the nonlinear
      term and a stage Stokes solve against the oracle, two FP64 steps
      against the oracle within its roundoff envelope, and five FP32-mixed
      steps against five FP64 device steps (FFT lengths 384 in z, a 192-point
      half-length x transform, 64-lane wall-normal groups);
  the 64-instance Couette ensemble as the benchmark draws it (grids from
      16x17x16, 24x33x24, 32x33x32, 48x49x48, Re uniform in [300, 500]),
      autotuned and graph-captured, bit-identical to per-problem step_;
  the learned-closure data path at 128x97x128 -> 32x33x32: box filter,
      subgrid force filter(N(U)) - N(filter(U)), coarse velocity gradients
      and the network's features for a few snapshots of the fine run."""
import math

import numpy as np
import pytest

from oracle import closure as K
from oracle import oracle as O
from test_gpu_parity import ENVELOPE_FACTOR, _split, assert_parity, exact_errors, rel, roundoff_envelope, say

pytestmark = pytest.mark.gpu

L3 = (4 * math.pi, 4 * math.pi / 3)
N3 = (256, 193, 256)
NU3, DT3 = 1 / 2800, 0.0025


def base(y):
    return 1.5 * (1 - y ** 2)


@pytest.fixture(scope="module")
def big(gpu):
    cf = gpu
    g = cf.Grid(*N3, *L3)
    prob = cf.ChannelFlowProblem(g, base, NU3, DT3, bulkvelocity=(1.0, 0.0))
    st = cf.perturbation_state(prob, 0.1, seed=42)
    go = O.Grid(*N3, *L3)
    ref = O.Problem(go, base, NU3, DT3, bulkvelocity=(1.0, 0.0))
    return cf, g, go, prob, ref, st


def test_big_nonlinear_and_stokes(big):
    import torch
    cf, g, go, prob, ref, st = big
    U0 = st.velocity.cpu().numpy()
    N = prob.nlterm(0.0, st.velocity).cpu().numpy()
    e = rel(N, ref.nonlinear(U0))
    say(f"256x193x256 nonlinear rel err {e:.3e}")
    assert e < 1e-12
    # one stage-1 Stokes solve (lambda = 1/(dt/6)) of a smooth random RHS
    lam = 6.0 / DT3
    R = N + 0.1 * U0
    basec = O.chebcoeffs(base(go.y))
    kw = {"bulkvelocity": (1.0, 0.0)}
    U, P, grads = cf.stokes_solve(prob, NU3, lam, torch.as_tensor(R, device="cuda"), **kw)
    U, P = U.cpu().numpy(), P.cpu().numpy()
    Ur, Pr, gr = O.stokes(go, NU3, lam, R, baseflow=basec, **kw)
    env = roundoff_envelope(lambda R_: O.stokes(go, NU3, lam, R_, baseflow=basec, **kw)[:2], [R], [Ur, Pr])
    eg, eo = exact_errors(go, NU3, lam, R, [(U, P), (Ur, Pr)], max_modes=12)
    for i in range(2):
        say(f"256x193x256 stokes[{i}] vs exact: hip {eg[i]:.3e} oracle {eo[i]:.3e}")
        assert eg[i] <= max(1e-12, 4.0 * eo[i])
    env = [max(x, 4.0 * a / ENVELOPE_FACTOR) for x, a in zip(env, eo)]
    assert_parity("256x193x256 stokes", [U, P], [Ur, Pr], env)
    assert np.allclose(grads, gr, rtol=1e-11, atol=1e-14)


def test_big_fp64_steps(big):
    import torch
    cf, g, go, prob, ref, st = big
    U0, P0 = st.velocity.cpu().numpy(), st.stagepressure.cpu().numpy()
    U = torch.as_tensor(U0, device="cuda")
    P = torch.as_tensor(P0, device="cuda")
    cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=2, **prob.constraint)
    U1, P1 = ref.step(U0, P0, 2)
    say("256x193x256: oracle done")
    env = roundoff_envelope(lambda a, b: _split(ref.step(a, b, 2)), [U0, P0], _split((U1, P1)))
    assert_parity("256x193x256 2 steps", _split((U.cpu().numpy(), P.cpu().numpy())), _split((U1, P1)), env)


def test_big_mixed_against_fp64(big):
    """Five FP32-mixed steps against five FP64 device steps at full size: the
    difference is FP32 rounding of the nonlinear part of the change (~1e-7 of
    it), and the kinetic energies agree to 1e-6."""
    cf, g, go, prob, ref, st = big
    p32 = cf.ChannelFlowProblem(g, base, NU3, DT3, bulkvelocity=(1.0, 0.0), precision="mixed")
    a, b = st.copy(), st.copy()
    cf.step_(prob.scheme, prob.nlterm, a.velocity, a.stagepressure, 0.0, nsteps=5, **prob.constraint)
    cf.step_(p32.scheme, p32.nlterm, b.velocity, b.stagepressure, 0.0, nsteps=5, **p32.constraint)
    du = (b.velocity - a.velocity).abs().max().item()
    change = (a.velocity - st.velocity).abs().max().item()
    E64, E32 = cf.kinetic_energy(prob, a.velocity), cf.kinetic_energy(p32, b.velocity)
    say(f"256x193x256 mixed vs FP64 after 5 steps: |dU| {du:.3e}, step change {change:.3e}, "
        f"rel-to-change {du / change:.3e}, energy rel {abs(E32 - E64) / E64:.3e}")
    assert 0.0 < du and du / change < 1e-6
    assert abs(E32 - E64) <= 1e-6 * E64


ENSEMBLE_GRIDS = [(16, 17, 16), (24, 33, 24), (32, 33, 32), (48, 49, 48)]


def test_ensemble_64_instances_bit_identical(gpu):
    """The batched workload as the benchmark draws it (seed 2024): 64 Couette
    instances, Lx = 2 pi/1.14, Lz = 2 pi/2.5, dt = 0.01, amplitude 0.1 (seed
    1000 + i), autotuned tiles, one captured graph per ensemble step; three
    ensemble steps equal three step_ calls of each problem bit for bit."""
    import torch
    cf = gpu
    rng = np.random.default_rng(2024)
    L = (2 * math.pi / 1.14, 2 * math.pi / 2.5)
    probs, states, refs, drawn = [], [], [], []
    for i in range(64):
        dims = ENSEMBLE_GRIDS[int(rng.integers(len(ENSEMBLE_GRIDS)))]
        Re = float(rng.uniform(300, 500))
        p = cf.CouetteFlow(cf.Grid(*dims, *L), 1.0 / Re, 0.01)
        s = cf.random_state(p, 0.1, seed=1000 + i)
        probs.append(p)
        states.append(s)
        refs.append((cf.CouetteFlow(cf.Grid(*dims, *L), 1.0 / Re, 0.01), s.copy()))
        drawn.append(dims)
    assert (48, 49, 48) in drawn
    ens = cf.Ensemble(probs, states, tune=True)
    launches, tiles = ens.info()
    say(f"64-instance ensemble: {launches} launches per step, grids {dict((d, drawn.count(d)) for d in ENSEMBLE_GRIDS)}")
    ens.advance(3)
    for q, s in refs:
        cf.step_(q.scheme, q.nlterm, s.velocity, s.stagepressure, 0.0, nsteps=3, **q.constraint)
    torch.cuda.synchronize()
    bad = [i for i, ((q, s), e) in enumerate(zip(refs, states))
           if not (torch.equal(e.velocity, s.velocity) and torch.equal(e.stagepressure, s.stagepressure))]
    assert not bad, bad
    del ens


def test_closure_data_path_at_workload_shapes(gpu):
    """128x97x128 fine run -> 32x33x32 coarse (Lx = 4 pi, Lz = 4 pi/3,
    nu = 1/2800, bulk flux): for three snapshots, the device box filter, the
    subgrid force and the coarse velocity-gradient features against the
    oracle (nonlinear terms, transforms and oracle/closure.py's filter)."""
    import torch
    cf = gpu
    from channelflow import closure as CL
    L = (4 * math.pi, 4 * math.pi / 3)
    gf, gc = cf.Grid(128, 97, 128, *L), cf.Grid(32, 33, 32, *L)
    fine = cf.ChannelFlowProblem(gf, base, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    coarse = cf.ChannelFlowProblem(gc, base, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    ofine = O.Problem(O.Grid(128, 97, 128, *L), base, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    ocoarse = O.Problem(O.Grid(32, 33, 32, *L), base, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    st = cf.perturbation_state(fine, 0.3, seed=3)
    dx, dz = gc.Lx / gc.Nx, gc.Lz / gc.Nz
    for snap in range(3):
        cf.step_(fine.scheme, fine.nlterm, st.velocity, st.stagepressure, 0.0, nsteps=5, **fine.constraint)
        U = st.velocity.cpu().numpy()
        F, Ubar = CL.subgrid_force(fine, coarse, st.velocity)
        Ubar_ref = K.box_filter(U, ofine.g, ocoarse.g, dx, dz)
        F_ref = K.box_filter(ofine.nonlinear(U), ofine.g, ocoarse.g, dx, dz) - ocoarse.nonlinear(Ubar_ref)
        eb, ef = rel(Ubar.cpu().numpy(), Ubar_ref), rel(F.cpu().numpy(), F_ref)
        # velocity gradient of the total coarse velocity at the padded points
        grad = cf.velocity_gradient(coarse, Ubar).cpu().numpy()
        V = Ubar_ref.copy()
        V[0, :, 0, 0] += O.chebcoeffs(base(ocoarse.g.y))
        go = ocoarse.g
        derivs = [lambda a: O.operator(go, "ddx1", a), lambda a: O.operator(go, "ddx2", a),
                  lambda a: O.operator(go, "ddx3", a)]
        eg = max(rel(grad[i, j], O.inverse(go, derivs[j](V[i]))) for i in range(3) for j in range(3))
        say(f"closure snapshot {snap}: filter {eb:.2e}, subgrid force {ef:.2e}, velocity gradient {eg:.2e}")
        assert eb < 1e-15 and ef < 1e-12 and eg < 1e-12
        X, T = CL.snapshot_samples(fine, coarse, st.velocity)
        assert X.shape == (go.Nxp * go.Nzp * go.Ny, K.NIN) and T.shape[1] == K.NOUT
        Xh = X.cpu().numpy()
        feat = grad.reshape(9, -1).T.astype(np.float32)
        assert np.array_equal(Xh[:, :9], feat)
        assert bool(torch.isfinite(T).all())
