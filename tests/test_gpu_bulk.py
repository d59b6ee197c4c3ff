"""The device's constant-bulk-flux step (_bulkpressuregradient,
cnrk2.jl:98-107, computed inside the y stage) against the closed forms that
pin the oracle's (tests/test_oracle_bulk.py): laminar Poiseuille keeps U = 0
with G = -3 nu, and the manufactured forced (0,0)-mode solution converges at
second order in velocity while the Richardson extrapolant of the returned
gradient converges to the exact G(t).  The device gradient must also equal
the oracle's on the same forcing."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import say
import test_oracle_bulk as M

pytestmark = pytest.mark.gpu


def device_run(cf, dt, T=1.0, Ny=17):
    import torch
    gc = cf.Grid(4, Ny, 4, 2.0, 2.0)
    go = O.Grid(4, Ny, 4, 2.0, 2.0)
    prob = cf.ChannelFlowProblem(gc, M.base, M.NU, dt, bulkvelocity=(1.0, 0.0))
    F0, F1, F2 = (torch.as_tensor(F, device="cuda") for F in M.manufactured_forcing(go))

    def forcing(t, U, F):
        F.add_(F0 + math.sin(M.W * t) * F1 + math.cos(M.W * t) * F2)

    U0 = M.initial(go)
    P0 = O.Problem(go, M.base, M.NU, dt, bulkvelocity=(1.0, 0.0)).pressure(U0)
    U = torch.as_tensor(U0, device="cuda")
    P = torch.as_tensor(P0, device="cuda")
    n = int(round(T / dt))
    cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, forcing=forcing, nsteps=n, **prob.constraint)
    grad = cf.last_gradient(prob.scheme)
    Uh = U.cpu().numpy()
    ux, uz, gr = M.exact(go, n * dt)
    return (np.abs(Uh[0, :, 0, 0] - ux).max(), np.abs(Uh[2, :, 0, 0] - uz).max(),
            grad[0] - gr[0], grad[1] - gr[1]), Uh, grad


def test_laminar_poiseuille_bulk_gradient_device(gpu):
    import torch
    cf = gpu
    nu = 1 / 2800
    for dims in [(6, 33, 4), (16, 97, 12)]:
        g = cf.Grid(*dims, 4 * np.pi, 4 * np.pi / 3)
        prob = cf.ChannelFlowProblem(g, M.base, nu, 0.005, bulkvelocity=(1.0, 0.0))
        U = g.zeros_spec(3)
        P = cf.pressure(U, prob)
        for _ in range(3):
            cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, **prob.constraint)
            gx, gz = cf.last_gradient(prob.scheme)
            say(f"laminar {dims}: G = ({gx:.17e}, {gz:.3e}), -3 nu = {-3 * nu:.17e}")
            assert abs(gx + 3 * nu) < 1e-15 and abs(gz) < 1e-15, (gx, gz)
        assert float(torch.view_as_real(U).abs().max()) < 1e-15


def test_manufactured_bulk_flux_device(gpu):
    cf = gpu
    res = []
    for dt in M.DTS:
        errs, Uh, grad = device_run(cf, dt)
        go = O.Grid(4, 17, 4, 2.0, 2.0)
        p = O.Problem(go, M.base, M.NU, dt, bulkvelocity=(1.0, 0.0))
        p.set_forcing(*M.manufactured_forcing(go)[:2], M.W, 0.0, F2=M.manufactured_forcing(go)[2])
        U0 = M.initial(go)
        Ur, _ = p.step(U0, p.pressure(U0), int(round(1.0 / dt)))
        gref = p.last_gradient()
        du = np.abs(Uh - Ur).max() / np.abs(Ur).max()
        dg = max(abs(a - b) for a, b in zip(grad, gref))
        say(f"dt {dt}: err u {errs[0]:.3e} w {errs[1]:.3e} Gx {errs[2]:.3e} Gz {errs[3]:.3e};"
            f" vs oracle U {du:.2e} G {dg:.2e}")
        assert du < 1e-13 and dg < 1e-12
        res.append(errs)
    res = np.array(res)
    for i in range(2):
        e = res[:3, i]
        assert 3.3 < e[0] / e[1] < 5.0 and 3.3 < e[1] / e[2] < 5.0, e
    for c in (2, 3):
        rich = M.richardson(list(res[:, c]))
        assert 3.5 < rich[0] / rich[1] < 5.0 and 3.5 < rich[1] / rich[2] < 5.0, rich
        assert abs(rich[-1]) < 1e-3, rich
