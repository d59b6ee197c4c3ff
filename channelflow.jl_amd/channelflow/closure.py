"""Learned subgrid closure for coarse channel DNS.

A small network predicts, at every padded point of a coarse grid, the
acceleration that the coarse step misses because its field is a filtered
one: the subgrid force

    F = filter(N_fine(u)) - N_coarse(filter(u))

where N is the solver's nonlinear term (the signed acceleration the step
uses, nonlinear.jl:173-193, base flow included) and filter is the box filter
of cf_box_filter.  The closure enters the coarse step through the forcing
hook (forcing(tstage, U, F), cnrk2.jl:173-175), so the coarse solver itself
is unchanged and is the fallback when no network is given.

Network and kernels: 10 -> 64 -> 64 -> 64 -> 3 with GELU, inputs the local
velocity-gradient tensor dV_i/dx_j and the wall distance 1 - |y|; bfloat16
weights and activations on the matrix cores with FP32 accumulation
(cf_mlp_infer, cf_mlp_grad), Adam on FP32 master weights (cf_mlp_adam).
"""
import ctypes as C
import math

import numpy as np

from . import _lib

NIN, KIN, HID, NOUT, KOUT = 10, 16, 64, 3, 16
SHAPES = [("W1", (HID, KIN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)), ("b3", (HID,)),
          ("W4", (KOUT, HID)), ("b4", (KOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
assert NPARAM == 10448  # CF_MLP_NPARAM


def _cf():
    import channelflow as cf
    return cf


def initial_parameters(seed=0):
    """He-normal hidden weights, small output weights, zero biases; padded
    input columns and output rows zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("b"):
            parts.append(np.zeros(shp))
            continue
        fan_in = NIN if name == "W1" else HID
        w = rng.standard_normal(shp) * math.sqrt(2.0 / fan_in)
        if name == "W1":
            w[:, NIN:] = 0.0
        if name == "W4":
            w[NOUT:, :] = 0.0
            w *= 0.1
        parts.append(w)
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def bf16_bits(theta):
    """bfloat16 bit patterns (round to nearest even) of a float32 device tensor."""
    torch = _cf()._torch()
    u = theta.contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.to(torch.int32).to(torch.int16)


class SubgridMLP:
    """Parameters, Adam state and normalisation of the subgrid-force network
    (all device-resident)."""

    def __init__(self, seed=0):
        torch = _cf()._torch()
        self.theta = torch.as_tensor(initial_parameters(seed), device="cuda")
        self.wbf = bf16_bits(self.theta)
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.norm = torch.zeros(2 * NIN + NOUT, dtype=torch.float32, device="cuda")
        self.norm[NIN:2 * NIN] = 1.0
        self.norm[2 * NIN:] = 1.0
        self.steps = 0
        self._work = None

    # -- data ---------------------------------------------------------------
    def fit_normalization(self, X, T):
        """Feature means and inverse standard deviations; target rms scales."""
        torch = _cf()._torch()
        mean = X.double().mean(0)
        std = X.double().std(0)
        self.norm[:NIN] = mean.float()
        self.norm[NIN:2 * NIN] = (1.0 / torch.clamp(std, min=1e-30)).float()
        rms = torch.sqrt((T.double() ** 2).mean(0))
        self.norm[2 * NIN:] = torch.clamp(rms, min=1e-30).float()

    # -- training -----------------------------------------------------------
    def _workspace(self, batch):
        torch = _cf()._torch()
        n = int(_lib.load().cf_mlp_work_floats(int(batch)))
        if self._work is None or self._work.numel() < n:
            self._work = torch.empty(n, dtype=torch.float32, device="cuda")
            self._grad = torch.empty(NPARAM + 1, dtype=torch.float32, device="cuda")
        return self._work

    def loss_and_grad(self, X, T, idx):
        """(loss, gradient) device tensors of the batch idx (cf_mlp_grad)."""
        cf = _cf()
        work = self._workspace(idx.numel())
        _lib.check(_lib.load().cf_mlp_grad(cf._ptr(self.theta), cf._ptr(self.wbf), cf._ptr(self.norm), cf._ptr(X),
                                           cf._ptr(T), cf._ptr(idx), int(idx.numel()), cf._ptr(work),
                                           cf._ptr(self._grad), cf._stream()))
        return self._grad[NPARAM], self._grad[:NPARAM]

    def adam(self, lr):
        cf = _cf()
        self.steps += 1
        _lib.check(_lib.load().cf_mlp_adam(cf._ptr(self.theta), cf._ptr(self.wbf), cf._ptr(self.m), cf._ptr(self.v),
                                           cf._ptr(self._grad), C.c_float(lr), self.steps, cf._stream()))

    def train(self, X, T, epochs=3, batch=8192, lr=1e-3, seed=0, lr_final=None):
        """Mini-batch Adam over the samples (X, T); returns the mean training
        loss of each epoch.  The learning rate decays geometrically to
        lr_final (default lr / 10) over the run."""
        torch = _cf()._torch()
        n = X.shape[0]
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        nb = max(1, n // batch)
        total = epochs * nb
        lr_final = lr / 10 if lr_final is None else lr_final
        hist = []
        for e in range(epochs):
            perm = torch.randperm(n, device="cuda", generator=gen).to(torch.int32)
            acc = torch.zeros((), dtype=torch.float64, device="cuda")
            for b in range(nb):
                k = e * nb + b
                rate = lr * (lr_final / lr) ** (k / max(total - 1, 1))
                loss, _ = self.loss_and_grad(X, T, perm[b * batch:(b + 1) * batch])
                acc += loss.double()
                self.adam(rate)
            hist.append(float(acc) / nb)
        return hist

    def evaluate(self, X, T, batch=65536):
        """Mean loss of the network and of the zero prediction (no closure) on
        (X, T), both in the network's normalised units."""
        torch = _cf()._torch()
        n = X.shape[0]
        tot, base = 0.0, 0.0
        for s in range(0, n, batch):
            idx = torch.arange(s, min(s + batch, n), dtype=torch.int32, device="cuda")
            loss, _ = self.loss_and_grad(X, T, idx)
            tot += float(loss) * idx.numel()
            tn = T[s:s + batch].double() / self.norm[2 * NIN:].double()
            base += float((tn ** 2).sum())
        return tot / n, base / n

    # -- inference ----------------------------------------------------------
    def predict(self, problem, U, grad=None, out=None):
        """Predicted subgrid force (3, Ny, Nzp, Nxp) on the padded grid."""
        cf = _cf()
        torch = cf._torch()
        g = problem.grid
        grad = cf.velocity_gradient(problem, U, out=grad)
        if out is None:
            out = torch.empty((3,) + g.phys_shape, dtype=torch.float64, device="cuda")
        _lib.check(_lib.load().cf_mlp_infer(cf._ptr(self.theta), cf._ptr(self.wbf), cf._ptr(self.norm), cf._ptr(grad),
                                            g.Nxp, g.Nzp, g.Ny, cf._ptr(out), cf._stream()))
        return out

    def forcing(self, problem, modes="all", device=True):
        """forcing(t, U, F) adding the predicted subgrid force (spectral) to F.
        modes="mean" keeps only its plane-mean part (the (0,0) Fourier mode,
        a function of y), the part that sets the mean profile.  Passed to
        step_ or a Flow, device=True registers it as a device-side forcing
        (cf_set_closure: no host callback per stage, the step stays
        graph-replayed); device=False calls it back from the host at every
        stage.  Both enqueue the same kernels and give the same result."""
        if modes not in ("all", "mean"):
            raise ValueError("modes must be 'all' or 'mean'")
        return ClosureForcing(self, problem, modes, device)

    # -- persistence ----------------------------------------------------------
    def state_dict(self):
        return {"theta": self.theta.cpu().numpy(), "norm": self.norm.cpu().numpy(), "steps": self.steps}

    def load_state_dict(self, d):
        torch = _cf()._torch()
        self.theta.copy_(torch.as_tensor(np.asarray(d["theta"], np.float32)))
        self.wbf.copy_(bf16_bits(self.theta))
        self.norm.copy_(torch.as_tensor(np.asarray(d["norm"], np.float32)))
        self.steps = int(d.get("steps", 0))


class ClosureForcing:
    """The closure's forcing(t, U, F) on one problem's grid.  The stage's
    nonlinear term is live in the stepping problem's workspaces while the
    force is computed, so gradients and transforms run on a helper problem
    (same grid, base flow and form) with workspaces of its own."""

    def __init__(self, net, problem, modes, device):
        cf = _cf()
        torch = cf._torch()
        g = problem.grid
        self.net, self.modes, self.device = net, modes, device
        self.helper = cf.ChannelFlowProblem(g, problem.baseflow, problem.scheme.nu, problem.scheme.dt,
                                            form=problem.form)
        self.gbuf = torch.empty((3, 3) + g.phys_shape, dtype=torch.float64, device="cuda")
        self.fbuf = torch.empty((3,) + g.phys_shape, dtype=torch.float64, device="cuda")
        self.spec = g.zeros_spec(3)

    def __call__(self, t, U, F):
        cf = _cf()
        self.net.predict(self.helper, U, grad=self.gbuf, out=self.fbuf)
        spec = self.spec
        _lib.check(_lib.load().cf_forward3(self.helper.scheme.h, cf._ptr(self.fbuf), cf._ptr(spec[0]),
                                           cf._ptr(spec[1]), cf._ptr(spec[2]), cf._stream()))
        if self.modes == "mean":
            F[:, :, 0, 0] += spec[:, :, 0, 0]
        else:
            F.add_(spec)

    def _install(self, scheme, U):
        cf = _cf()
        return _DeviceClosure(self, scheme) if self.device else cf._Forcing(scheme, U, self)


class _DeviceClosure:
    """Registers a ClosureForcing on a scheme's handle for one call
    (cf_set_closure) and removes it afterwards.  Its buffers, the network's
    tensors and the scheme's forcing fields keep their addresses from call to
    call, so the step's cached graph is replayed."""

    def __init__(self, force, scheme):
        cf = _cf()
        self.force, self.scheme = force, scheme
        self.F = cf._scheme_buffer(scheme, "_forcing_buffer")

    def __enter__(self):
        cf = _cf()
        f, net, F = self.force, self.force.net, self.F
        _lib.check(_lib.load().cf_set_closure(self.scheme.h, f.helper.scheme.h, cf._ptr(net.theta), cf._ptr(net.wbf),
                                              cf._ptr(net.norm), cf._ptr(f.gbuf), cf._ptr(f.fbuf), cf._ptr(F[0]),
                                              cf._ptr(F[1]), cf._ptr(F[2]), 1 if f.modes == "mean" else 0))
        return self

    def __exit__(self, *exc):
        _lib.load().cf_set_closure(self.scheme.h, None, None, None, None, None, None, None, None, None, 0)
        return False


def subgrid_force(fine, coarse, U, widths=None):
    """Spectral subgrid force of a fine state on the coarse grid:
    filter(N_fine(U)) - N_coarse(filter(U)), and the filtered field."""
    cf = _cf()
    Ubar = cf.box_filter(fine, coarse, U, widths=widths)
    Nf = fine.nlterm(0.0, U)
    F = cf.box_filter(fine, coarse, Nf, widths=widths)
    F -= coarse.nlterm(0.0, Ubar)
    return F, Ubar


def snapshot_samples(fine, coarse, U, widths=None):
    """Training samples (X: points x 10 raw features, T: points x 3 raw
    targets) of one fine snapshot, at the coarse padded points."""
    cf = _cf()
    torch = cf._torch()
    g = coarse.grid
    F, Ubar = subgrid_force(fine, coarse, U, widths)
    f = cf.IFFT(coarse, F)
    grad = cf.velocity_gradient(coarse, Ubar)
    npts = g.Nxp * g.Nzp * g.Ny
    X = torch.empty((npts, NIN), dtype=torch.float32, device="cuda")
    T = torch.empty((npts, NOUT), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().cf_mlp_features(cf._ptr(grad), cf._ptr(f.contiguous()), g.Nxp, g.Nzp, g.Ny, cf._ptr(X),
                                           cf._ptr(T), cf._stream()))
    return X, T
