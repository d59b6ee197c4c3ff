"""Learned-closure device pieces against the closure oracle (oracle/closure.py):
the box filter (fine spectral -> coarse spectral)."""
import numpy as np
import pytest

from oracle import closure as K
from oracle import oracle as O
from test_gpu_parity import hermitian_random, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fine,coarse", [((16, 17, 16), (8, 9, 8)), ((12, 33, 10), (6, 17, 5)),
                                         ((32, 65, 24), (16, 33, 12)), ((9, 17, 9), (8, 17, 8))])
def test_box_filter_matches_oracle(gpu, fine, coarse):
    import torch
    cf = gpu
    say("box filter", fine, coarse)
    L = (4 * np.pi, 4 * np.pi / 3)
    gf, gc = cf.Grid(*fine, *L), cf.Grid(*coarse, *L)
    pf, pc = cf.CouetteFlow(gf, 1 / 400, 0.01), cf.CouetteFlow(gc, 1 / 400, 0.01)
    go_f, go_c = O.Grid(*fine, *L), O.Grid(*coarse, *L)
    U = hermitian_random(go_f, np.random.default_rng(sum(fine)), 3)
    got = cf.box_filter(pf, pc, torch.as_tensor(U, device="cuda")).cpu().numpy()
    ref = K.box_filter(U, go_f, go_c, gc.Lx / gc.Nx, gc.Lz / gc.Nz)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(f"  rel err {err:.2e}")
    assert err < 1e-15
    with pytest.raises(cf.DimensionMismatch):
        cf.box_filter(pc, pf, torch.as_tensor(ref, device="cuda"))


def _norm(rng):
    mean = rng.standard_normal(10).astype(np.float32) * 0.1
    scale = (0.5 + rng.random(10)).astype(np.float32)
    out = (0.5 + rng.random(3)).astype(np.float32)
    return np.concatenate([mean, scale, out]).astype(np.float32)


def _theta(seed):
    rng = np.random.default_rng(seed)
    th = K.init_theta(seed).astype(np.float64)
    P = K.unpack(np.arange(K.NPARAM))
    live = np.ones(K.NPARAM, bool)
    live[P["W1"][:, K.NIN:].ravel().astype(int)] = False
    live[P["W4"][K.NOUT:, :].ravel().astype(int)] = False
    live[P["b4"][K.NOUT:].astype(int)] = False
    th[live] += 0.2 * rng.standard_normal(live.sum())  # biases nonzero too
    return th.astype(np.float32)


def _dev_params(cf, theta, norm):
    import torch
    from channelflow import closure as CL
    th = torch.as_tensor(theta, device="cuda")
    return th, CL.bf16_bits(th), torch.as_tensor(norm, device="cuda")


def test_mlp_inference_matches_oracle(gpu):
    import torch
    cf = gpu
    from channelflow import _lib
    rng = np.random.default_rng(5)
    g = cf.Grid(8, 9, 8, 4.0, 2.0)
    grad = rng.standard_normal((9, g.Ny, g.Nzp, g.Nxp))
    theta, norm = _theta(3), _norm(rng)
    th, wb, nm = _dev_params(cf, theta, norm)
    gd = torch.as_tensor(grad, device="cuda")
    out = torch.empty((3, g.Ny, g.Nzp, g.Nxp), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().cf_mlp_infer(cf._ptr(th), cf._ptr(wb), cf._ptr(nm), cf._ptr(gd), g.Nxp, g.Nzp, g.Ny,
                                        cf._ptr(out), cf._stream()))
    y = np.cos(np.pi * np.arange(g.Ny) / (g.Ny - 1))
    feats = np.concatenate([grad.reshape(9, -1).T.astype(np.float32).astype(np.float64),
                            np.broadcast_to((1 - np.abs(y))[:, None, None], (g.Ny, g.Nzp, g.Nxp)).reshape(-1, 1)], 1)
    Xn = (feats - norm[:10].astype(np.float64)) * norm[10:20].astype(np.float64)
    ref = K.forward(theta, Xn)[0] * norm[20:23].astype(np.float64)
    got = out.cpu().numpy().reshape(3, -1).T
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(f"  mlp inference rel err {err:.3e}")
    # FP32 device arithmetic against float64 sums with the same bf16 rounding points;
    # a rare bf16 rounding flip of an activation moves an output by ~1e-4
    assert err < 1e-3


@pytest.mark.parametrize("batch", [100, 64, 1])
def test_mlp_gradient_matches_oracle(gpu, batch):
    import torch
    cf = gpu
    from channelflow import _lib
    rng = np.random.default_rng(7 + batch)
    n = 300
    X = rng.standard_normal((n, 10)).astype(np.float32) * 2 + 0.3
    T = rng.standard_normal((n, 3)).astype(np.float32) * 0.7
    theta, norm = _theta(11), _norm(rng)
    th, wb, nm = _dev_params(cf, theta, norm)
    idx = rng.choice(n, batch, replace=False).astype(np.int32)
    Xd, Td, Id = (torch.as_tensor(a, device="cuda") for a in (X, T, idx))
    work = torch.empty(int(_lib.load().cf_mlp_work_floats(batch)), dtype=torch.float32, device="cuda")
    grad = torch.empty(K.NPARAM + 1, dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().cf_mlp_grad(cf._ptr(th), cf._ptr(wb), cf._ptr(nm), cf._ptr(Xd), cf._ptr(Td), cf._ptr(Id),
                                       batch, cf._ptr(work), cf._ptr(grad), cf._stream()))
    g = grad.cpu().numpy().astype(np.float64)
    Xn = (X[idx].astype(np.float64) - norm[:10]) * norm[10:20]
    Tn = T[idx].astype(np.float64) / norm[20:23]
    loss, gref = K.loss_and_grad(theta, Xn, Tn, bf=True)
    el = abs(g[-1] - loss) / loss
    eg = np.linalg.norm(g[:-1] - gref) / np.linalg.norm(gref)
    say(f"  batch {batch}: loss rel err {el:.3e}, gradient rel err {eg:.3e}")
    assert el < 1e-5 and eg < 1e-5  # FP32 accumulation vs float64, same bf16 rounding points
    P = K.unpack(np.arange(K.NPARAM))
    pad = np.concatenate([P["W1"][:, 10:].ravel(), P["W4"][3:, :].ravel(), P["b4"][3:]]).astype(int)
    assert np.all(g[pad] == 0.0)


def test_mlp_adam_matches_oracle(gpu):
    import torch
    cf = gpu
    from channelflow import _lib
    from channelflow import closure as CL
    rng = np.random.default_rng(13)
    th = rng.standard_normal(K.NPARAM).astype(np.float32)
    m = rng.standard_normal(K.NPARAM).astype(np.float32) * 0.01
    v = rng.random(K.NPARAM).astype(np.float32) * 0.01
    gr = rng.standard_normal(K.NPARAM).astype(np.float32)
    td, md, vd, gd = (torch.as_tensor(a, device="cuda") for a in (th, m, v, gr))
    wb = CL.bf16_bits(td)
    _lib.check(_lib.load().cf_mlp_adam(cf._ptr(td), cf._ptr(wb), cf._ptr(md), cf._ptr(vd), cf._ptr(gd), 1e-3, 3,
                                       cf._stream()))
    rt, rm, rv = K.adam(th, m, v, gr, 3, 1e-3)
    assert np.allclose(md.cpu().numpy(), rm, rtol=1e-6, atol=1e-9)
    assert np.allclose(vd.cpu().numpy(), rv, rtol=1e-6, atol=1e-12)
    assert np.allclose(td.cpu().numpy(), rt, rtol=1e-6, atol=1e-8)
    bits = wb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    assert np.array_equal(bits.view(np.float32).astype(np.float64), K.bf16(td.cpu().numpy()))


def test_closure_training_gains_over_no_model(gpu):
    """A priori: trained on filtered snapshots of a fine run, the network's
    held-out error is below that of the method without it (zero force)."""
    import torch
    cf = gpu
    from channelflow import closure as CL
    L = (4 * np.pi, 4 * np.pi / 3)
    gf, gc = cf.Grid(48, 49, 48, *L), cf.Grid(16, 17, 16, *L)
    prof = lambda y: 1.5 * (1 - y ** 2)  # noqa: E731
    fine = cf.ChannelFlowProblem(gf, prof, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    coarse = cf.ChannelFlowProblem(gc, prof, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    st = cf.perturbation_state(fine, 0.3, seed=3)
    Xs, Ts = [], []
    for k in range(14):
        cf.step_(fine.scheme, fine.nlterm, st.velocity, st.stagepressure, 0.0, nsteps=5, **fine.constraint)
        X, T = CL.snapshot_samples(fine, coarse, st.velocity)
        Xs.append(X)
        Ts.append(T)
    Xtr, Ttr = torch.cat(Xs[:11]), torch.cat(Ts[:11])
    Xte, Tte = torch.cat(Xs[11:]), torch.cat(Ts[11:])
    net = CL.SubgridMLP(seed=1)
    net.fit_normalization(Xtr, Ttr)
    hist = net.train(Xtr, Ttr, epochs=4, batch=2048, lr=3e-3)
    loss, base = net.evaluate(Xte, Tte)
    say(f"  training loss by epoch {[round(h, 4) for h in hist]}; held-out loss {loss:.4f} vs no model {base:.4f}")
    assert all(np.isfinite(hist)) and hist[-1] < hist[0]
    assert loss < 0.9 * base
    # the closure as a forcing: the coarse step runs and the force is applied
    s = cf.box_filter(fine, coarse, st.velocity)
    sp = cf.pressure(s, coarse)
    a, b = s.clone(), s.clone()
    pa, pb = sp.clone(), sp.clone()
    cf.step_(coarse.scheme, coarse.nlterm, a, pa, 0.0, nsteps=3, **coarse.constraint)
    cf.step_(coarse.scheme, coarse.nlterm, b, pb, 0.0, nsteps=3, forcing=net.forcing(coarse), **coarse.constraint)
    d = (a - b).abs().max().item()
    say(f"  closure changes the coarse state by {d:.3e}")
    assert np.isfinite(d) and d > 0
    # the plane-mean variant forces only the (0,0) mode: after one step the
    # other modes differ only through the nonlinear coupling of later stages
    c, pc = s.clone(), sp.clone()
    e, pe = s.clone(), sp.clone()
    cf.step_(coarse.scheme, coarse.nlterm, c, pc, 0.0, **coarse.constraint)
    cf.step_(coarse.scheme, coarse.nlterm, e, pe, 0.0, forcing=net.forcing(coarse, modes="mean"), **coarse.constraint)
    dm = (c - e)[:, :, 0, 0].abs().max().item()
    do = (c - e).abs().clone()
    do[:, :, 0, 0] = 0
    say(f"  mean-mode forcing: (0,0) change {dm:.3e}, other modes {do.max().item():.3e}")
    assert dm > 0 and do.max().item() < 1e-3 * dm
