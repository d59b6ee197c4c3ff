"""ORACLE — test infrastructure only (ctypes wrapper of oracle/libcforacle.so).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module; it is the checker, never the product.

Array conventions match the reference's column-major storage:
  * spectral field (k, l, n) -> numpy complex128 array of shape (Ny, Nz, Nxh)
  * padded physical field (x, z, y) -> numpy float64 array of shape (Ny, Nzp, Nxp)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcforacle.so")
# the same source built with long double arithmetic: the accuracy reference
# for FP64 results (every function below takes extended=True to use it)
LIB_X = os.path.join(HERE, "libcforacle_x.so")

_libs = {}


def lib(extended=False):
    key = bool(extended)
    if key not in _libs:
        path = LIB_X if key else LIB
        if not os.path.exists(path):
            from oracle import build as _b
            _b.build()
        L = C.CDLL(path)
        _declare(L)
        if L.orc_extended() != int(key):
            raise OracleError(f"{path} is not the {'extended' if key else 'FP64'} build")
        _libs[key] = L
    return _libs[key]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_d = C.c_double
_i = C.c_int


def _declare(L):
    L.orc_last_error.restype = C.c_char_p
    L.orc_extended.restype = _i
    L.orc_num_threads.restype = _i
    L.orc_set_threads.argtypes = [_i]
    L.orc_cheb.argtypes = [_i, _i, _i, _cp, _cp]
    L.orc_forward.argtypes = [_i, _i, _i, _d, _d, _dp, _cp]
    L.orc_inverse.argtypes = [_i, _i, _i, _d, _d, _cp, _dp]
    L.orc_normalize_forward.argtypes = [_i, _i, _i, _d, _cp]
    L.orc_operator.argtypes = [_i, _i, _i, _i, _d, _d, _cp, _cp, _i]
    L.orc_curl.argtypes = [_i, _i, _i, _d, _d, _cp, _cp]
    L.orc_chebcoeffs.argtypes = [_i, _dp, _dp]
    L.orc_diff.argtypes = [_i, _dp, _dp]
    L.orc_helmholtz.argtypes = [_i, _d, _d, _i, _dp, _dp, _d, _d]
    L.orc_helmholtz_c.argtypes = [_i, _d, _d, _i, _cp, _cp, _cp, _cp]
    L.orc_influence_mode.argtypes = [_i, _d, _d, _d, _d, _cp, _cp, _cp, _cp, _cp, _cp, _cp, C.c_void_p]
    L.orc_batched_influence.argtypes = [_i, _i, _dp, _dp, _d, _d, _cp, _cp, _cp, _cp, _cp, _cp, _cp]
    L.orc_mean_mode.argtypes = [_i, _d, _d, _cp, _cp, _cp, _cp, _cp, _cp, _cp, _i, _d, _d, _dp, _dp]
    L.orc_stokes.argtypes = [_i, _i, _i, _d, _d, _d, _d, _cp, _cp, _cp, _i, _d, _d, C.c_void_p, _dp]
    L.orc_problem_create.restype = C.c_void_p
    L.orc_problem_create.argtypes = [_i, _i, _i, _d, _d, _d, _d, _dp, _i, _i, _d, _d]
    L.orc_problem_destroy.argtypes = [C.c_void_p]
    L.orc_base_coeffs.argtypes = [C.c_void_p, _dp, _dp]
    L.orc_nonlinear.argtypes = [C.c_void_p, _cp, _cp, _i]
    L.orc_step.argtypes = [C.c_void_p, _cp, _cp, _i]
    L.orc_last_gradient.argtypes = [C.c_void_p, _dp]
    L.orc_set_bulkgrad_mutation.argtypes = [_i]
    L.orc_set_forcing.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double]
    L.orc_project.argtypes = [C.c_void_p, _cp]
    L.orc_pressure.argtypes = [C.c_void_p, _cp, _cp]
    L.orc_pressure_poisson.argtypes = [_i, _i, _i, _d, _d, _d, _cp, _cp, _cp]
    L.orc_dot.restype = _d
    L.orc_dot.argtypes = [_i, _i, _i, _cp, _cp]


class OracleError(RuntimeError):
    pass


def _check(rc, extended=False):
    if rc != 0:
        raise OracleError(lib(extended).orc_last_error().decode())


def set_bulkgrad_mutation(m):
    """Test hook: plant a known mistake in the oracle's bulk gradient
    (0 none, 1 drop shear, 2 drop base, 3 drop B[N], 4 flip shear sign)."""
    lib().orc_set_bulkgrad_mutation(int(m))


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


# ---------------------------------------------------------------------------
# Grid geometry (src/grids.jl)
# ---------------------------------------------------------------------------
class Grid:
    def __init__(self, Nx, Ny, Nz, Lx, Lz):
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.Lx, self.Lz = float(Lx), float(Lz)
        self.Nxp = (3 * self.Nx + 1) // 2
        self.Nzp = (3 * self.Nz + 1) // 2
        self.Nxh = self.Nx // 2 + 1
        self.y = np.cos(np.pi * np.arange(self.Ny) / (self.Ny - 1))

    @property
    def spec_shape(self):
        return (self.Ny, self.Nz, self.Nxh)

    @property
    def phys_shape(self):
        return (self.Ny, self.Nzp, self.Nxp)

    def args(self):
        return (self.Nx, self.Ny, self.Nz, self.Lx, self.Lz)

    def zeros_spec(self, n=None):
        shape = self.spec_shape if n is None else (n,) + self.spec_shape
        return np.zeros(shape, dtype=np.complex128)

    def points(self):
        """Padded physical coordinates broadcast to (y, z, x) storage order."""
        x = self.Lx * np.arange(self.Nxp) / self.Nxp
        z = self.Lz * np.arange(self.Nzp) / self.Nzp
        return self.y[:, None, None], z[None, :, None], x[None, None, :]

    def sample(self, f):
        """Evaluate f(x, y, z) on the padded grid (helpers.jl:42-48)."""
        y, z, x = self.points()
        return np.ascontiguousarray(np.broadcast_to(f(x, y, z), self.phys_shape), dtype=np.float64)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _r(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
# Transforms and operators
# ---------------------------------------------------------------------------
def forward(g, u, extended=False):
    U = g.zeros_spec()
    _check(lib(extended).orc_forward(*g.args(), _r(u), U), extended)
    return U


def inverse(g, U, extended=False):
    u = np.zeros(g.phys_shape)
    _check(lib(extended).orc_inverse(*g.args(), _c(U), u), extended)
    return u


def cheb(a, inverse=False):
    """plan_cheb/plan_icheb mul! on an array whose LAST axis is the Chebyshev index
    stored slowest, i.e. shape (Ny, B) C-order == (B, Ny) column-major."""
    a = _c(a)
    Ny = a.shape[0]
    B = a.size // Ny
    out = np.zeros_like(a)
    _check(lib().orc_cheb(B, Ny, 1 if inverse else 0, a, out))
    return out


def normalize_forward(g, U, norm):
    U = _c(U).copy()
    _check(lib().orc_normalize_forward(g.Nx, g.Ny, g.Nz, float(norm), U))
    return U


def operator(g, name, U, out=None, add=False):
    op = {"ddx1": 1, "ddx2": 2, "ddx3": 3, "laplacian": 4}[name]
    o = g.zeros_spec() if out is None else _c(out).copy()
    _check(lib().orc_operator(op, *g.args(), _c(U), o, 1 if add else 0))
    return o


def curl(g, U3):
    U3 = _c(U3)
    O = np.zeros_like(U3)
    _check(lib().orc_curl(*g.args(), U3, O))
    return O


def chebcoeffs(values):
    values = _r(values)
    out = np.zeros_like(values)
    _check(lib().orc_chebcoeffs(len(values), values, out))
    return out


def diff(a):
    a = _r(a)
    out = np.zeros_like(a)
    _check(lib().orc_diff(len(a), a, out))
    return out


def helmholtz(P, t0, t1, f, gp, gm, neumann=False, extended=False):
    """Chebyshev-tau solve of rows 0..P-2 of (t0 D^2 - t1) a = f with wall data."""
    L = lib(extended)
    if np.iscomplexobj(f) or np.iscomplexobj(gp) or np.iscomplexobj(gm):
        f = _c(f)
        out = np.zeros(P + 1, dtype=np.complex128)
        _check(L.orc_helmholtz_c(P, t0, t1, 1 if neumann else 0, f, out,
                                 np.array([complex(gp)]), np.array([complex(gm)])), extended)
        return out
    f = _r(f)
    out = np.zeros(P + 1)
    _check(L.orc_helmholtz(P, t0, t1, 1 if neumann else 0, f, out, float(gp), float(gm)), extended)
    return out


def influence_mode(Ny, k, l, nu, lam, Rx, Ry, Rz, responses=False, extended=False):
    outs = [np.zeros(Ny, dtype=np.complex128) for _ in range(4)]
    resp = np.zeros((6, Ny)) if responses else None
    _check(lib(extended).orc_influence_mode(Ny, k, l, nu, lam, _c(Rx), _c(Ry), _c(Rz), *outs,
                                            resp.ctypes.data if responses else None), extended)
    return (outs, resp) if responses else outs


def batched_influence(Ny, k, l, nu, lam, Rx, Ry, Rz, extended=False):
    """Matrices given as numpy (Ny, B) C-order == (B, Ny) column-major."""
    k, l = _r(k), _r(l)
    B = len(k)
    outs = [np.zeros((Ny, B), dtype=np.complex128) for _ in range(4)]
    _check(lib(extended).orc_batched_influence(Ny, B, k, l, nu, lam, _c(Rx), _c(Ry), _c(Rz), *outs), extended)
    return outs


def mean_mode(Ny, nu, lam, Rx, Ry, Rz, bulkvelocity=None, pressuregradient=None, extended=False):
    if bulkvelocity is not None and pressuregradient is not None:
        raise ValueError("specify either pressuregradient or bulkvelocity")
    kind, c = (1, bulkvelocity) if bulkvelocity is not None else (0, pressuregradient or (0.0, 0.0))
    outs = [np.zeros(Ny, dtype=np.complex128) for _ in range(4)]
    grads = np.zeros(2)
    resp = np.zeros(Ny)
    _check(lib(extended).orc_mean_mode(Ny, nu, lam, _c(Rx), _c(Ry), _c(Rz), *outs, kind, float(c[0]),
                                       float(c[1]), grads, resp), extended)
    return outs, tuple(grads), resp


def stokes(g, nu, lam, R3, bulkvelocity=None, pressuregradient=None, baseflow=None, extended=False):
    """StokesSolver solve!: returns (U3, P, gradients)."""
    if bulkvelocity is not None and pressuregradient is not None:
        raise ValueError("specify either pressuregradient or bulkvelocity")
    kind, c = (1, bulkvelocity) if bulkvelocity is not None else (0, pressuregradient or (0.0, 0.0))
    U3 = g.zeros_spec(3)
    P = g.zeros_spec()
    grads = np.zeros(2)
    base = None if baseflow is None else _r(baseflow)
    _check(lib(extended).orc_stokes(*g.args(), nu, lam, _c(R3), U3, P, kind, float(c[0]), float(c[1]),
                                    None if base is None else base.ctypes.data, grads), extended)
    return U3, P, tuple(grads)


def pressure_poisson(g, A3, v, nu, extended=False):
    P = g.zeros_spec()
    _check(lib(extended).orc_pressure_poisson(*g.args(), nu, _c(A3), _c(v), P), extended)
    return P


def dot(g, U, V, extended=False):
    return lib(extended).orc_dot(g.Nx, g.Ny, g.Nz, _c(U), _c(V))


# ---------------------------------------------------------------------------
# Problem (ChannelFlowProblem + CNRK2 + NonLinearTerm)
# ---------------------------------------------------------------------------
FORMS = {"rotating": 0, "convective": 1, "divergence": 2}


class Problem:
    """ChannelFlowProblem + CNRK2 + NonLinearTerm of the oracle.  extended=True
    runs the same algorithm in long double (the accuracy reference); its
    fields still cross as complex128 and stay extended between the steps of
    one step() call."""

    def __init__(self, g, profile, nu, dt, form="rotating", pressuregradient=None, bulkvelocity=None,
                 extended=False):
        if pressuregradient is not None and bulkvelocity is not None:
            raise ValueError("specify either pressuregradient or bulkvelocity")
        self.g = g
        self.nu, self.dt = float(nu), float(dt)
        if bulkvelocity is not None:
            self.kind, c = 1, bulkvelocity
        else:
            self.kind, c = 0, (pressuregradient if pressuregradient is not None else (0.0, 0.0))
        self.c = (float(c[0]), float(c[1]))
        vals = profile(g.y) if callable(profile) else np.asarray(profile, dtype=float)
        self.baseflow_values = _r(np.broadcast_to(vals, (g.Ny,)))
        self.form = form
        self.extended = bool(extended)
        self._L = lib(extended)
        h = self._L.orc_problem_create(*g.args(), self.nu, self.dt, self.baseflow_values, FORMS[form], self.kind,
                                     self.c[0], self.c[1])
        if not h:
            raise OracleError(self._L.orc_last_error().decode())
        self.h = h
        self.base = np.zeros(g.Ny)
        self.curv = np.zeros(g.Ny)
        self._L.orc_base_coeffs(self.h, self.base, self.curv)

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(self._L.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_problem_destroy(self.h)
            self.h = None

    def nonlinear(self, U3, N3=None, add=False):
        N3 = self.g.zeros_spec(3) if N3 is None else _c(N3).copy()
        self._chk(self._L.orc_nonlinear(self.h, _c(U3), N3, 1 if add else 0))
        return N3

    def step(self, U3, P, nsteps=1):
        """Advance copies of (U3, P) by nsteps CNRK2 steps; return them."""
        U3 = _c(U3).copy()
        P = _c(P).copy()
        self._chk(self._L.orc_step(self.h, U3, P, int(nsteps)))
        return U3, P

    def step_inplace(self, U3, P, nsteps=1):
        self._chk(self._L.orc_step(self.h, U3, P, int(nsteps)))

    def set_forcing(self, F0=None, F1=None, omega=0.0, t0=0.0, F2=None):
        """forcing(t, U, N) += F0 + sin(omega t) F1 + cos(omega t) F2 at the
        stage times (cnrk2.jl:173-175); F0=None removes it.  t0: time of the
        next step."""
        if F0 is None:
            self._chk(self._L.orc_set_forcing(self.h, None, None, None, 0.0, float(t0)))
            self._forcing = None
            return
        F0 = _c(np.asarray(F0)).copy()
        F1 = None if F1 is None else _c(np.asarray(F1)).copy()
        F2 = None if F2 is None else _c(np.asarray(F2)).copy()
        self._forcing = (F0, F1, F2)  # keep alive
        self._chk(self._L.orc_set_forcing(self.h, F0.ctypes.data, None if F1 is None else F1.ctypes.data,
                                     None if F2 is None else F2.ctypes.data, float(omega), float(t0)))

    def last_gradient(self):
        g2 = np.zeros(2)
        self._L.orc_last_gradient(self.h, g2)
        return tuple(g2)

    def project(self, U3):
        U3 = _c(U3).copy()
        self._chk(self._L.orc_project(self.h, U3))
        return U3

    def pressure(self, U3):
        P = self.g.zeros_spec()
        self._chk(self._L.orc_pressure(self.h, _c(U3), P))
        return P

    def random_state(self, noise3):
        """random_state (initialization.jl:32-51) from supplied Gaussian samples
        already scaled by epsilon: shape (3,) + padded physical shape."""
        U3 = np.stack([forward(self.g, noise3[i], self.extended) for i in range(3)])
        U3 = self.project(U3)
        return U3, self.pressure(U3)
