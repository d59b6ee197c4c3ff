"""channelflow — MI355X (gfx950) device backend for ChannelFlow.jl's CNRK2 step.

Python mirror of the reference's public surface for the hot path
(src/grids.jl, src/problem.jl, src/state.jl, src/initialization.jl,
src/timesteppers/cnrk2.jl): ``Grid``, ``ChannelFlowProblem``, ``CouetteFlow``,
``PoiseuilleFlow``, ``State``, ``zero_state``, ``random_state``, ``step_``
(Julia's ``step!``), ``project_``, ``pressure``, ``FFT``/``IFFT`` and the
energy diagnostics.  Every numerical operation runs in libcfhip.so through
its C ABI; torch provides device memory and streams only.

Storage follows the reference exactly: a spectral field is a complex128
device tensor of shape (Ny, Nz, Nxh) (= (k, l, n) column-major), a vector
field stacks three of them, a padded physical field is a float64 tensor of
shape (Ny, Nzp, Nxp) (= (x, z, y) column-major).
"""
import contextlib
import ctypes as C
import math

import numpy as np

from . import _lib
from . import io  # noqa: F401  (portable state files)
from ._lib import ArgumentError, DimensionMismatch, DeviceError  # noqa: F401

__all__ = ["Grid", "ChannelFlowProblem", "CouetteFlow", "PoiseuilleFlow", "State", "zero_state", "random_state",
           "step_", "project_", "pressure", "FFT", "IFFT", "velocity", "stagepressure", "kinetic_energy",
           "require_device", "RotatingForm", "ConvectiveForm", "DivergenceForm", "dot", "norm", "dissipation_rate",
           "power_input", "plane_statistics", "ProfileStatistics", "flow", "device_gaussian_noise",
           "savefield", "loadfield", "Ensemble", "velocity_gradient", "box_filter", "tune", "get_tiles", "set_tiles"]

# nonlinear forms (src/nonlinear.jl:214-224)
RotatingForm = "rotating"
ConvectiveForm = "convective"
DivergenceForm = "divergence"


def _torch():
    import torch
    return torch


def require_device():
    """Load the HIP library and check that a GPU is visible; raise otherwise."""
    _lib.load()
    torch = _torch()
    if not torch.cuda.is_available():
        raise DeviceError("no HIP device visible")
    return True


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


# ---------------------------------------------------------------------------
# Grid (src/grids.jl:22-96)
# ---------------------------------------------------------------------------
class Grid:
    def __init__(self, Nx, Ny, Nz, Lx, Lz):
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.Lx, self.Lz = float(Lx), float(Lz)
        self.y = np.cos(np.pi * np.arange(self.Ny) / (self.Ny - 1))
        self.Nxp = (3 * self.Nx + 1) // 2
        self.Nzp = (3 * self.Nz + 1) // 2
        self.Nxh = self.Nx // 2 + 1

    def __eq__(self, other):
        return isinstance(other, Grid) and (self.Nx, self.Ny, self.Nz, self.Lx, self.Lz) == (
            other.Nx, other.Ny, other.Nz, other.Lx, other.Lz)

    def physicalsize(self, padded=True):
        return (self.Nxp, self.Nzp, self.Ny) if padded else (self.Nx, self.Nz, self.Ny)

    def spectralsize(self):
        return (self.Nxh, self.Nz, self.Ny)

    @property
    def spec_shape(self):
        return (self.Ny, self.Nz, self.Nxh)

    @property
    def phys_shape(self):
        return (self.Ny, self.Nzp, self.Nxp)

    def points(self):
        """Padded coordinates broadcast to (y, z, x) storage order (grids.jl:89-96)."""
        x = self.Lx * np.arange(self.Nxp) / self.Nxp
        z = self.Lz * np.arange(self.Nzp) / self.Nzp
        return self.y[:, None, None], z[None, :, None], x[None, None, :]

    def zeros_spec(self, n=None, device="cuda"):
        torch = _torch()
        shape = self.spec_shape if n is None else (n,) + self.spec_shape
        return torch.zeros(shape, dtype=torch.complex128, device=device)

    def zeros_phys(self, n=None, device="cuda"):
        torch = _torch()
        shape = self.phys_shape if n is None else (n,) + self.phys_shape
        return torch.zeros(shape, dtype=torch.float64, device=device)


# ---------------------------------------------------------------------------
# State (src/state.jl)
# ---------------------------------------------------------------------------
class State:
    def __init__(self, velocity, stagepressure):
        self.velocity = velocity
        self.stagepressure = stagepressure

    def copy(self):
        return State(self.velocity.clone(), self.stagepressure.clone())


def velocity(state):
    return state.velocity


def stagepressure(state):
    return state.stagepressure


def zero_state(grid):
    """zero_state (initialization.jl:16)."""
    return State(grid.zeros_spec(3), grid.zeros_spec())


# ---------------------------------------------------------------------------
# Problem, scheme and nonlinear term (src/problem.jl, cnrk2.jl, nonlinear.jl)
# ---------------------------------------------------------------------------
class CNRK2:
    """Device CNRK2 scheme: owns the libcfhip problem handle (factored
    stage solvers, plans, workspaces)."""

    def __init__(self, grid, baseflow_values, nu, dt, form, constraint, device=0, precision="fp64"):
        require_device()
        if precision not in _lib.PRECISIONS:
            raise ArgumentError("precision must be 'fp64' or 'mixed'")
        L = _lib.load()
        cfg = _lib.cf_config()
        cfg.Nx, cfg.Ny, cfg.Nz = grid.Nx, grid.Ny, grid.Nz
        cfg.Lx, cfg.Lz = grid.Lx, grid.Lz
        cfg.nu, cfg.dt = float(nu), float(dt)
        self._base = np.ascontiguousarray(baseflow_values, dtype=np.float64)
        if self._base.shape != (grid.Ny,):
            raise DimensionMismatch("baseflow must have one value per wall-normal point")
        cfg.baseflow = self._base.ctypes.data_as(C.POINTER(C.c_double))
        cfg.form = _lib.FORMS[form]
        kind, vals = constraint
        cfg.constraint = kind
        cfg.constraint_values[0], cfg.constraint_values[1] = vals
        cfg.device = int(device)
        cfg.precision = _lib.PRECISIONS[precision]
        h = C.c_void_p()
        _lib.check(L.cf_problem_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.grid, self.nu, self.dt = grid, float(nu), float(dt)
        self.precision = precision

    def __del__(self):
        h = getattr(self, "h", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h and lib is not None:  # at interpreter exit the module may be torn down first
            lib.cf_problem_destroy(h)
            self.h = None


class NonLinearTerm:
    """NonLinearTerm of the problem's form (nonlinear.jl:113-193): rotating
    N = FFT(V x curl V), convective N = -FFT((V . grad) V), divergence
    N = -div FFT(V V), with V = U + Ub e_x."""

    def __init__(self, scheme):
        self.scheme = scheme

    def __call__(self, t, U, N=None, add=False):
        g = self.scheme.grid
        out = g.zeros_spec(3) if (N is None or add) else N
        _check_vec(g, U)
        _check_vec(g, out)
        _lib.check(_lib.load().cf_nonlinear(self.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(out[0]),
                                            _ptr(out[1]), _ptr(out[2]), _stream()))
        if N is not None and add:
            N += out
            return N
        return out


def _constraint(pressuregradient, bulkvelocity):
    if pressuregradient is not None and bulkvelocity is not None:
        raise ArgumentError("specify either pressuregradient or bulkvelocity")
    kind = 1 if bulkvelocity is not None else 0
    given = bulkvelocity if kind else (pressuregradient if pressuregradient is not None else (0.0, 0.0))
    try:
        vals = tuple(float(v) for v in given)
    except TypeError:
        raise ArgumentError("mean-flow constraint must be an (x, z) pair") from None
    if len(vals) != 2:
        raise ArgumentError("mean-flow constraint must be an (x, z) pair")
    if not all(math.isfinite(v) for v in vals):
        raise ArgumentError("mean-flow constraint values must be finite")
    return kind, vals


class ChannelFlowProblem:
    """ChannelFlowProblem(grid, profile, nu, dt; form, forcing, pressuregradient|bulkvelocity)
    (src/problem.jl:69-107).  precision="mixed" runs the time step with FP32
    transforms and nonlinear term and FP64 solves (CF_PRECISION_MIXED)."""

    def __init__(self, grid, profile, nu, dt, form=RotatingForm, forcing=None, pressuregradient=None,
                 bulkvelocity=None, device=0, precision="fp64"):
        if forcing is not None and not callable(forcing):
            raise ArgumentError("forcing must be callable as forcing(t, U, F)")
        kind, vals = _constraint(pressuregradient, bulkvelocity)
        self.grid = grid
        self.baseflow = np.asarray(profile(grid.y) if callable(profile) else profile, dtype=np.float64)
        self.baseflow = np.broadcast_to(self.baseflow, (grid.Ny,)).copy()
        self.scheme = CNRK2(grid, self.baseflow, nu, dt, form, (kind, vals), device, precision)
        self.nlterm = NonLinearTerm(self.scheme)
        self.forcing = forcing
        self.constraint = {"bulkvelocity": vals} if kind == 1 else {"pressuregradient": vals}
        self.form = form


def CouetteFlow(grid, nu, dt, **kw):
    """Base velocity Ub(y) = y (problem.jl:122-123)."""
    return ChannelFlowProblem(grid, lambda y: y, nu, dt, **kw)


def PoiseuilleFlow(grid, nu, dt, bulkvelocity=None, pressuregradient="default", **kw):
    """Base velocity Ub(y) = 1 - y^2, default pressure gradient (-2 nu, 0) (problem.jl:137-144)."""
    if isinstance(pressuregradient, str):
        pressuregradient = None if bulkvelocity is not None else (-2.0 * nu, 0.0)
    return ChannelFlowProblem(grid, lambda y: 1.0 - y ** 2, nu, dt, bulkvelocity=bulkvelocity,
                              pressuregradient=pressuregradient, **kw)


# ---------------------------------------------------------------------------
# step! (cnrk2.jl:149-207)
# ---------------------------------------------------------------------------
def _check_spec(g, f):
    torch = _torch()
    if f.dtype != torch.complex128 or tuple(f.shape) != g.spec_shape:
        raise DimensionMismatch("fields must have the resolved spectral size")
    if not f.is_cuda or not f.is_contiguous():
        raise ArgumentError("fields must be contiguous device tensors")


def _check_vec(g, U):
    if len(U) != 3:
        raise DimensionMismatch("vector field must have three components")
    for c in range(3):
        _check_spec(g, U[c])


@contextlib.contextmanager
def _on_stream(handle):
    """Make the library's stream (a hipStream_t passed to a callback) torch's
    current stream, so that device work a callback enqueues is ordered with
    the step's kernels on that stream."""
    torch = _torch()
    cur = torch.cuda.current_stream()
    if handle is None or int(handle) == cur.cuda_stream:
        yield
        return
    with torch.cuda.stream(torch.cuda.ExternalStream(int(handle))):
        yield


def _scheme_buffer(scheme, name):
    """A three-component coefficient buffer owned by the scheme: it lives as
    long as the problem handle, so kernels still queued on any stream never
    read memory that the caching allocator has handed out again."""
    buf = getattr(scheme, name, None)
    if buf is None:
        buf = scheme.grid.zeros_spec(3)
        setattr(scheme, name, buf)
    return buf


class _Forcing:
    """forcing(tstage, U, F) (cnrk2.jl:173-175) installed on the problem
    handle for one call: the library calls back at every stage; the callback
    zeroes F and lets `forcing` add its spectral acceleration into F with
    device operations enqueued on the step's stream (the callback runs while
    the step is being enqueued, with that stream made torch's current one,
    so it must not read U on the host).  The stage's nonlinear term is live
    in the problem's workspaces at that point: a forcing that needs
    transforms or gradients runs them on another problem of the same grid."""

    def __init__(self, scheme, U, forcing):
        self.scheme, self.U, self.forcing = scheme, U, forcing
        self.F = _scheme_buffer(scheme, "_forcing_buffer")
        self.error = None

        def cb(user, tstage, u0, u1, u2, f0, f1, f2, stream):
            try:
                with _on_stream(stream):
                    self.F.zero_()
                    self.forcing(tstage, self.U, self.F)
                return 0
            except BaseException as e:  # reported after the call returns
                self.error = e
                return 1

        self.fn = _lib.FORCING_FN(cb)

    def __enter__(self):
        F = self.F
        _lib.check(_lib.load().cf_set_forcing(self.scheme.h, self.fn, None, _ptr(F[0]), _ptr(F[1]), _ptr(F[2])))
        return self

    def __exit__(self, *exc):
        _lib.load().cf_set_forcing(self.scheme.h, _lib.FORCING_FN(), None, None, None, None)
        if self.error is not None:  # the callback's own exception, not the library's status
            raise self.error
        return False


def _forcing_hook(scheme, U, forcing):
    """Context installing `forcing` on the scheme's handle for one call: the
    forcing's own installer when it has one (the learned closure registers
    itself as a device-side forcing), else the host callback."""
    install = getattr(forcing, "_install", None)
    return install(scheme, U) if install is not None else _Forcing(scheme, U, forcing)


class _Nonlinear:
    """nonlinear(tstage, U, N) (cnrk2.jl:121-126,172) given by the caller in
    place of the problem's NonLinearTerm, installed for one call: at every
    stage the library calls back and `nonlinear` must overwrite N with the
    signed nonlinear acceleration of the total velocity, with device
    operations enqueued on the step's stream (made torch's current stream
    during the callback).  The same workspace rule as for forcing applies."""

    def __init__(self, scheme, U, nonlinear):
        self.scheme, self.U, self.nonlinear = scheme, U, nonlinear
        self.N = _scheme_buffer(scheme, "_nl_buffer")
        self.error = None

        def cb(user, tstage, u0, u1, u2, n0, n1, n2, stream):
            try:
                with _on_stream(stream):
                    self.nonlinear(tstage, self.U, self.N)
                return 0
            except BaseException as e:  # reported after the call returns
                self.error = e
                return 1

        self.fn = _lib.FORCING_FN(cb)

    def __enter__(self):
        N = self.N
        _lib.check(_lib.load().cf_set_nonlinear(self.scheme.h, self.fn, None, _ptr(N[0]), _ptr(N[1]), _ptr(N[2])))
        return self

    def __exit__(self, *exc):
        _lib.load().cf_set_nonlinear(self.scheme.h, _lib.FORCING_FN(), None, None, None, None)
        if self.error is not None:
            raise self.error
        return False


def step_(scheme, nlterm, U, P, t, forcing=None, pressuregradient=None, bulkvelocity=None, nsteps=1, stream=None):
    """Advance (U, P) in place by scheme.dt (nsteps times); return the new time.
    nlterm(tstage, U, N) overwrites N with the signed nonlinear acceleration
    (the problem's NonLinearTerm runs on the device path of its form; any
    other callable is called back at each stage); forcing(tstage, U, F) adds
    a spectral acceleration into F.  Both are sampled at the stage times
    t + (0, 1/3, 3/4) dt (cnrk2.jl:149-207)."""
    if forcing is not None and not callable(forcing):
        raise ArgumentError("forcing must be callable as forcing(t, U, F)")
    if isinstance(nlterm, NonLinearTerm):
        if nlterm.scheme is not scheme:
            raise ArgumentError("nlterm belongs to another problem")
        custom = None
    elif callable(nlterm):
        custom = nlterm
    else:
        raise ArgumentError("nlterm must be callable as nonlinear(t, U, N)")
    kind, vals = _constraint(pressuregradient, bulkvelocity)
    g = scheme.grid
    _check_vec(g, U)
    _check_spec(g, P)
    arr = (C.c_double * 2)(*vals)
    tout = C.c_double()

    def run():
        _lib.check(_lib.load().cf_advance(scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(P), float(t),
                                          int(nsteps), kind, arr, C.byref(tout), _stream(stream)))

    with contextlib.ExitStack() as hooks:
        if custom is not None:
            hooks.enter_context(_Nonlinear(scheme, U, custom))
        if forcing is not None:
            hooks.enter_context(_forcing_hook(scheme, U, forcing))
        run()
    return tout.value


class Flow:
    """Flows.flow(channel) (ext/ChannelFlowFlowsExt.jl:15-52): F(state, (t0, T))
    advances the State in place from t0 to T with the problem's constant step,
    shortening the final step (solvers factored for its actual duration)."""

    def __init__(self, problem):
        self.problem = problem

    def __call__(self, state, span, stream=None):
        t0, T = (float(x) for x in span)
        p = self.problem
        g = p.grid
        U, P = state.velocity, state.stagepressure
        _check_vec(g, U)
        _check_spec(g, P)
        kind, vals = _constraint(p.constraint.get("pressuregradient"), p.constraint.get("bulkvelocity"))
        arr = (C.c_double * 2)(*vals)
        tout = C.c_double()

        def run():
            _lib.check(_lib.load().cf_advance_to(p.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(P), t0, T, kind,
                                                 arr, C.byref(tout), _stream(stream)))

        if p.forcing is None:  # forcing=channel.forcing (ext/ChannelFlowFlowsExt.jl:48-49)
            run()
        else:
            with _forcing_hook(p.scheme, U, p.forcing):
                run()
        return state


def flow(problem):
    return Flow(problem)


def tune(problem, stream=None):
    """Pick the problem's z tile width and x wavefronts per block by timing
    them on the device (cf_problem_tune); returns (KT, z points per thread,
    x wavefronts).  Results of the step are identical for any tiling."""
    _lib.check(_lib.load().cf_problem_tune(problem.scheme.h, _stream(stream)))
    out = (C.c_int * 3)()
    _lib.check(_lib.load().cf_problem_tiles(problem.scheme.h, out))
    return tuple(out)


X_VARIANTS = {"six": 0, "split": 1, "async": 2}  # x pass variants of the rotating form (even padded Nx)


def get_tiles(problem):
    """(z tile width, z points per thread, x pass variant, x wavefronts per
    block, wall-normal coefficients per lane) of the problem's step
    (cf_problem_get_tiles)."""
    out = (C.c_int * 5)()
    _lib.check(_lib.load().cf_problem_get_tiles(problem.scheme.h, out))
    return tuple(out)


def set_tiles(problem, z_width, x_variant, x_waves, y_coefs=None):
    """Set the step's z tile width, x pass variant (0/"six", 1/"split",
    2/"async"), x wavefronts per block and (optionally) the wall-normal
    kernels' coefficients per lane: the default (4; 6 above 256 points) or
    an alternative built for the size, 6 or 8 (cf_problem_set_tiles); an
    unavailable choice raises and leaves the tiling as it was.  The x and z
    choices give identical results; the wall-normal layout agrees to
    roundoff."""
    v = X_VARIANTS.get(x_variant, x_variant)
    yk = get_tiles(problem)[4] if y_coefs is None else int(y_coefs)
    arr = (C.c_int * 5)(int(z_width), 0, int(v), int(x_waves), yk)
    _lib.check(_lib.load().cf_problem_set_tiles(problem.scheme.h, arr))
    return get_tiles(problem)


def last_gradient(scheme):
    out = (C.c_double * 2)()
    _lib.check(_lib.load().cf_last_gradient(scheme.h, out, _stream()))
    return (out[0], out[1])


# ---------------------------------------------------------------------------
# transforms, projection, pressure, initial states
# ---------------------------------------------------------------------------
def FFT(problem, u):
    """Padded physical field(s) -> resolved spectrum (ffts.jl:65-82, 198-207)."""
    g = problem.grid
    torch = _torch()
    single = u.dim() == 3
    uu = u.unsqueeze(0) if single else u
    out = g.zeros_spec(uu.shape[0])
    if uu.dtype != torch.float64 or tuple(uu.shape[1:]) != g.phys_shape:
        raise DimensionMismatch("forward transform requires padded physical input")
    if uu.shape[0] == 3:  # one launch per pass for the vector field
        src = uu.contiguous()
        _lib.check(_lib.load().cf_forward3(problem.scheme.h, _ptr(src), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                           _stream()))
        return out
    for c in range(uu.shape[0]):
        src = uu[c].contiguous()
        _lib.check(_lib.load().cf_forward(problem.scheme.h, _ptr(src), _ptr(out[c]), _stream()))
    return out[0] if single else out


def IFFT(problem, U):
    """Resolved spectrum -> padded physical field(s) (ffts.jl:151-166, 218-227)."""
    g = problem.grid
    single = U.dim() == 3
    UU = U.unsqueeze(0) if single else U
    out = g.zeros_phys(UU.shape[0])
    if UU.shape[0] == 3:  # one launch per pass for the vector field
        srcs = [UU[c].contiguous() for c in range(3)]
        for src in srcs:
            _check_spec(g, src)
        _lib.check(_lib.load().cf_inverse3(problem.scheme.h, _ptr(srcs[0]), _ptr(srcs[1]), _ptr(srcs[2]), _ptr(out),
                                           _stream()))
        return out
    for c in range(UU.shape[0]):
        src = UU[c].contiguous()
        _check_spec(g, src)
        _lib.check(_lib.load().cf_inverse(problem.scheme.h, _ptr(src), _ptr(out[c]), _stream()))
    return out[0] if single else out


def stokes_solve(problem, nu, lam, R, U=None, P=None, pressuregradient=None, bulkvelocity=None):
    """solve!(StokesSolver(grid, nu, lambda), U, P, R; ...) (stokes.jl:81-108).  Returns (U, P, gradients)."""
    g = problem.grid
    U = g.zeros_spec(3) if U is None else U
    P = g.zeros_spec() if P is None else P
    _check_vec(g, R)
    if pressuregradient is None and bulkvelocity is None:
        kind, vals = -1, (0.0, 0.0)
    else:
        kind, vals = _constraint(pressuregradient, bulkvelocity)
    arr = (C.c_double * 2)(*vals)
    grads = (C.c_double * 2)()
    _lib.check(_lib.load().cf_stokes_solve(problem.scheme.h, float(nu), float(lam), _ptr(R[0]), _ptr(R[1]), _ptr(R[2]),
                                           _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(P), kind, arr, grads, _stream()))
    return U, P, (grads[0], grads[1])


def project_(U, problem):
    """project!(U, problem) (pressure.jl:117-135), in place; returns U."""
    _check_vec(problem.grid, U)
    _lib.check(_lib.load().cf_project(problem.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _stream()))
    return U


def velocity_gradient(problem, U, out=None, stream=None):
    """Velocity-gradient tensor of the total velocity U + Ub e_x on the padded
    grid: a (3, 3, Ny, Nzp, Nxp) float64 device tensor, [i, j] = dV_i/dx_j
    with x_j = (x, y, z) (cf_velocity_gradient)."""
    g = problem.grid
    _check_vec(g, U)
    if out is None:
        out = _torch().empty((3, 3) + g.phys_shape, dtype=_torch().float64, device="cuda")
    _lib.check(_lib.load().cf_velocity_gradient(problem.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(out),
                                                _stream(stream)))
    return out


def box_filter(fine, coarse, U, out=None, widths=None, stream=None):
    """Box-filtered coarse spectral field of the fine problem's U
    (cf_box_filter): top-hat widths (dx, dz), default the coarse grid
    spacings Lx/Nx and Lz/Nz."""
    gf, gc = fine.grid, coarse.grid
    _check_vec(gf, U)
    if out is None:
        out = gc.zeros_spec(3)
    _check_vec(gc, out)
    dx, dz = widths if widths is not None else (gc.Lx / gc.Nx, gc.Lz / gc.Nz)
    _lib.check(_lib.load().cf_box_filter(fine.scheme.h, coarse.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]),
                                         _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), float(dx), float(dz),
                                         _stream(stream)))
    return out


def pressure(U, problem, t=0.0):
    """pressure(U, problem) (pressure.jl:176-185)."""
    g = problem.grid
    _check_vec(g, U)
    P = g.zeros_spec()
    _lib.check(_lib.load().cf_pressure(problem.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(P), _stream()))
    return P


def gaussian_noise(grid, epsilon, seed):
    """epsilon * N(0,1) samples on the padded grid for components u, v, w in the
    reference's evaluation order (x fastest, then z, then y; physicalfield.jl:39-43),
    from numpy's PCG64 stream (Julia's randn stream is not reproducible here)."""
    return epsilon * np.random.default_rng(seed).standard_normal((3,) + grid.phys_shape)


def device_gaussian_noise(problem, epsilon, seed):
    """epsilon * N(0,1) on the padded grid for u, v, w generated on the device
    (Philox4x32-10 + Box-Muller, counter = element; cf_gaussian_noise):
    float64 tensor (3, Ny, Nzp, Nxp)."""
    g = problem.grid
    out = g.zeros_phys(3)
    _lib.check(_lib.load().cf_gaussian_noise(problem.scheme.h, float(epsilon), int(seed) & (2 ** 64 - 1), _ptr(out),
                                             _stream()))
    return out


def random_state(problem, epsilon, seed=42):
    """random_state (initialization.jl:32-51) entirely on the device: Gaussian
    padded field -> FFT -> project! -> pressure (cf_random_state)."""
    if epsilon < 0:
        raise ArgumentError("epsilon must be non-negative")
    g = problem.grid
    U, P = g.zeros_spec(3), g.zeros_spec()
    _lib.check(_lib.load().cf_random_state(problem.scheme.h, float(epsilon), int(seed) & (2 ** 64 - 1), _ptr(U[0]),
                                           _ptr(U[1]), _ptr(U[2]), _ptr(P), _stream()))
    return State(U, P)


def perturbation_state(problem, amplitude, seed=42, decay=4.0):
    """Seeded divergence-free perturbation with spectral energy ~ kappa^-decay:
    Gaussian coefficients of amplitude (1+kappa^2)^(-decay/4) (n+1)^-2, made
    real by a physical round trip, projected (project!), scaled so that the
    largest padded-grid velocity component equals `amplitude`, re-projected,
    and completed by the consistent stage pressure (pressure)."""
    torch = _torch()
    g = problem.grid
    rng = np.random.default_rng(seed)
    k = 2 * np.pi / g.Lx * np.arange(g.Nxh)
    l = 2 * np.pi / g.Lz * np.array([iz if iz <= g.Nz // 2 else iz - g.Nz for iz in range(g.Nz)])
    kap2 = k[None, None, :] ** 2 + l[None, :, None] ** 2
    n = np.arange(g.Ny)[:, None, None]
    amp = (1.0 + kap2) ** (-decay / 4.0) * (n + 1.0) ** -2
    coef = (rng.standard_normal((3,) + g.spec_shape) + 1j * rng.standard_normal((3,) + g.spec_shape)) * amp
    U = torch.as_tensor(coef, device="cuda")
    U = FFT(problem, IFFT(problem, U))  # keep the part representing a real field
    project_(U, problem)
    umax = float(IFFT(problem, U).abs().max())
    U *= amplitude / umax
    project_(U, problem)
    return State(U, pressure(U, problem))


def step_timed(scheme, U, P, t, pressuregradient=None, bulkvelocity=None, stream=None):
    """One step with per-kernel-class device times (ms): prologue, z inverse,
    x pass, z forward, y stage."""
    kind, vals = _constraint(pressuregradient, bulkvelocity)
    arr = (C.c_double * 2)(*vals)
    ms = (C.c_double * 5)()
    tout = C.c_double()
    _lib.check(_lib.load().cf_step_timed(scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(P), float(t), kind, arr,
                                         C.byref(tout), ms, _stream(stream)))
    return list(ms)


# ---------------------------------------------------------------------------
# diagnostics (postprocessing.jl:21-95), torch reductions on the device
# ---------------------------------------------------------------------------
def _scalar():
    return C.c_double(0.0)


def dot(problem, U, V):
    """dot(U::VectorField, V::VectorField) for spectral fields (postprocessing.jl:30-60):
    exact volume average <U . V> (device kernel, Chebyshev mass matrix)."""
    g = problem.grid
    _check_vec(g, U)
    _check_vec(g, V)
    out = _scalar()
    _lib.check(_lib.load().cf_dot(problem.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(V[0]), _ptr(V[1]),
                                  _ptr(V[2]), C.byref(out), _stream()))
    return out.value


def norm(problem, U):
    """norm(U::VectorField): sqrt(<|U|^2>) (postprocessing.jl:70-71)."""
    return math.sqrt(dot(problem, U, U))


def dissipation_rate(problem, U, nu=None):
    """dissipation_rate(U, nu) = nu <|curl U|^2> (postprocessing.jl:87-91)."""
    g = problem.grid
    _check_vec(g, U)
    out = _scalar()
    nu = problem.scheme.nu if nu is None else float(nu)
    _lib.check(_lib.load().cf_dissipation(problem.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), nu, C.byref(out),
                                          _stream()))
    return out.value


def power_input(problem, U, nu=None, pressuregradient=(0.0, 0.0)):
    """power_input(U, nu; pressuregradient) (postprocessing.jl:113-128)."""
    g = problem.grid
    _check_vec(g, U)
    out = _scalar()
    nu = problem.scheme.nu if nu is None else float(nu)
    grad = (C.c_double * 2)(*[float(x) for x in pressuregradient])
    _lib.check(_lib.load().cf_power_input(problem.scheme.h, _ptr(U[0]), _ptr(U[2]), nu, grad, C.byref(out), _stream()))
    return out.value


def plane_statistics(problem, U):
    """x-z plane means and covariances of the TOTAL velocity U + Ub e_x at the
    Lobatto points, from the spectral coefficients on the device.  Returns
    (mean (Ny, 3): U V W, cov (Ny, 6): uu vv ww uv uw vw)."""
    g = problem.grid
    _check_vec(g, U)
    out = np.zeros((g.Ny, 9))
    _lib.check(_lib.load().cf_plane_statistics(problem.scheme.h, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]),
                                               out.ctypes.data_as(C.POINTER(C.c_double)), _stream()))
    return out[:, :3].copy(), out[:, 3:].copy()


STRESS_PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


class ProfileStatistics:
    """Plane- and time-averaged profiles (examples/mkm590/statistics.jl:9-43):
    means of U, V, W and the six velocity covariances about the combined
    plane-and-time mean, merged sample by sample (parallel-variance update)."""

    def __init__(self, Ny):
        self.count = 0
        self.first_time = float("nan")
        self.last_time = float("nan")
        self.mean = np.zeros((Ny, 3))
        self.moment = np.zeros((Ny, 6))

    def sample(self, problem, U, t):
        """sample!(stats, velocity, t) for a spectral perturbation velocity U."""
        mu, cov = plane_statistics(problem, U)
        delta = mu - self.mean
        n = self.count + 1
        for k, (i, j) in enumerate(STRESS_PAIRS):
            self.moment[:, k] += cov[:, k] + (self.count / n) * delta[:, i] * delta[:, j]
        self.mean += delta / n
        if self.count == 0:
            self.first_time = float(t)
        self.count = n
        self.last_time = float(t)
        return self

    def profiles(self):
        """(mean (Ny, 3), covariances (Ny, 6)) of the samples so far."""
        return self.mean.copy(), self.moment / max(self.count, 1)


def savefield(path, grid, field=None, *, state=None, time=None, step=None):
    """savefield(path, field | state) (src/io.jl:21-30) in the portable layout
    documented in channelflow.io; device tensors are copied to the host."""
    if state is not None:
        return io.savefield(path, grid, velocity=state.velocity, pressure=state.stagepressure, time=time, step=step)
    return io.savefield(path, grid, field, time=time, step=step)


def loadfield(path, grid=None, device="cuda"):
    """loadfield(path) (src/io.jl:40-54): a State (velocity, stage pressure) or
    a field tensor on `device`.  With `grid`, its resolution and domain must
    match the file's."""
    torch = _torch()
    f = io.loadfield(path)
    if grid is not None and (f.Nx, f.Ny, f.Nz, f.Lx, f.Lz) != (grid.Nx, grid.Ny, grid.Nz, grid.Lx, grid.Lz):
        raise ArgumentError("file grid does not match the problem grid")
    if f.is_state:
        return State(torch.as_tensor(f.velocity(), device=device), torch.as_tensor(f.pressure(), device=device))
    return torch.as_tensor(f.field(), device=device)


class Ensemble:
    """n independent problems advanced together (cf_ensemble_*): each kernel
    class of the step is one launch per group of instances sharing its tile
    parameters, and one ensemble step is a captured HIP graph.  Every
    instance evolves exactly as step_ would advance it.  The states stay
    bound to the ensemble (their device buffers are baked into the graph)."""

    def __init__(self, problems, states, tune=True):
        if len(problems) != len(states) or not problems:
            raise ArgumentError("need one state per problem")
        L = _lib.load()
        for p, s in zip(problems, states):
            g = p.grid
            _check_vec(g, s.velocity)
            _check_spec(g, s.stagepressure)
        n = len(problems)
        hs = (C.c_void_p * n)(*[p.scheme.h.value for p in problems])
        ptrs = []
        for s in states:
            ptrs += [_ptr(s.velocity[0]), _ptr(s.velocity[1]), _ptr(s.velocity[2]), _ptr(s.stagepressure)]
        fs = (C.c_void_p * (4 * n))(*ptrs)
        h = C.c_void_p()
        _torch().cuda.synchronize()
        _lib.check(L.cf_ensemble_create(hs, fs, n, 1 if tune else 0, C.byref(h)))
        self.h = h
        self.problems, self.states = list(problems), list(states)

    def advance(self, nsteps, stream=None):
        _lib.check(_lib.load().cf_ensemble_advance(self.h, int(nsteps), _stream(stream)))

    def info(self):
        """(launches per ensemble step, [(z tile width, x wavefronts per block)] per instance)."""
        n = len(self.problems)
        launches = C.c_int()
        tiles = (C.c_int * (2 * n))()
        _lib.check(_lib.load().cf_ensemble_info(self.h, C.byref(launches), tiles))
        return launches.value, [(tiles[2 * i], tiles[2 * i + 1]) for i in range(n)]

    def __del__(self):
        h = getattr(self, "h", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h and lib is not None:
            lib.cf_ensemble_destroy(h)
            self.h = None


def kinetic_energy(problem, U):
    """kinetic_energy(U) = <|U|^2>/2 (postprocessing.jl:77)."""
    return dot(problem, U, U) / 2
