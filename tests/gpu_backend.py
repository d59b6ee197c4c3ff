"""Adapter exposing the HIP product with the oracle's numpy signatures, so the
reference's golden validation cases (golden_cases.py) drive the GPU path.

Trajectories, transforms, nonlinear terms and Stokes solves run on the GPU
through the C ABI; only the diagnostic derivative operators (used to score
energies/dissipation against the golden CSVs) come from the oracle, which is
the checker."""
import numpy as np
import torch

import channelflow as cf
from oracle import oracle as O

Grid = O.Grid
operator = O.operator


def _problem_for(g, profile=lambda y: y, nu=1 / 400, dt=0.01, **kw):
    gc = cf.Grid(g.Nx, g.Ny, g.Nz, g.Lx, g.Lz)
    return cf.ChannelFlowProblem(gc, profile, nu, dt, **kw)


_cache = {}


def _aux(g):
    key = (g.Nx, g.Ny, g.Nz, g.Lx, g.Lz)
    if key not in _cache:
        _cache[key] = _problem_for(g)
    return _cache[key]


def forward(g, u):
    return cf.FFT(_aux(g), torch.as_tensor(np.array(u), device="cuda")).cpu().numpy()


def inverse(g, U):
    return cf.IFFT(_aux(g), torch.as_tensor(np.ascontiguousarray(U), device="cuda")).cpu().numpy()


def stokes(g, nu, lam, R3, bulkvelocity=None, pressuregradient=None, baseflow=None):
    U, P, grads = cf.stokes_solve(_aux(g), nu, lam, torch.as_tensor(R3, device="cuda"),
                                  bulkvelocity=bulkvelocity, pressuregradient=pressuregradient)
    return U.cpu().numpy(), P.cpu().numpy(), grads


class Problem:
    def __init__(self, g, profile, nu, dt, form="rotating", pressuregradient=None, bulkvelocity=None):
        self.g = g
        self.p = _problem_for(g, profile, nu, dt, form=form, pressuregradient=pressuregradient,
                              bulkvelocity=bulkvelocity)

    def nonlinear(self, U3):
        return self.p.nlterm(0.0, torch.as_tensor(U3, device="cuda")).cpu().numpy()

    def step(self, U3, P, nsteps=1):
        U = torch.as_tensor(np.ascontiguousarray(U3), device="cuda").clone()
        Pd = torch.as_tensor(np.ascontiguousarray(P), device="cuda").clone()
        cf.step_(self.p.scheme, self.p.nlterm, U, Pd, 0.0, nsteps=nsteps, **self.p.constraint)
        return U.cpu().numpy(), Pd.cpu().numpy()

    def project(self, U3):
        U = torch.as_tensor(np.ascontiguousarray(U3), device="cuda").clone()
        cf.project_(U, self.p)
        return U.cpu().numpy()

    def pressure(self, U3):
        return cf.pressure(torch.as_tensor(np.ascontiguousarray(U3), device="cuda"), self.p).cpu().numpy()
