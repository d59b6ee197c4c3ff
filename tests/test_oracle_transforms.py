"""Oracle transforms and operators against the reference's known answers.

Restates test/test_ffts.jl, test/test_operators.jl and test/test_nonlinear.jl
(analytic, independent of the implementation) plus closed-form cross-checks
(scipy DCT-I == FFTW REDFT00).
"""
import numpy as np
import pytest
import scipy.fft

from helpers import coefficients
from oracle import oracle as O


@pytest.mark.parametrize("Nx,Nz", [(6, 8), (7, 9), (6, 9), (7, 8)])
def test_analytic_transform(Nx, Nz):
    # test_ffts.jl:5-45
    g = O.Grid(Nx, 9, Nz, 5.3, 7.1)
    a, b = 2 * np.pi / 5.3, 2 * np.pi / 7.1
    u = g.sample(lambda x, y, z: 2 + 0.3 * y + (2 * y ** 2 - 1) * np.cos(a * x) * np.sin(b * z))
    U = O.forward(g, u)
    exact = np.zeros(g.spec_shape, complex)
    exact[0, 0, 0], exact[1, 0, 0] = 2, 0.3
    exact[2, 1, 1], exact[2, -1, 1] = -0.25j, 0.25j
    assert np.abs(U - exact).max() < 2e-13
    assert np.abs(O.inverse(g, U) - u).max() < 2e-13
    U3 = O.forward(g, np.full(g.phys_shape, 3.0))
    exact[:] = 0
    exact[0, 0, 0] = 3
    assert np.abs(U3 - exact).max() < 2e-13


def test_nyquist_planes_excluded():
    # test_ffts.jl:47-53
    g = O.Grid(6, 9, 8, 2 * np.pi, 2 * np.pi)
    U = O.forward(g, g.sample(lambda x, y, z: np.cos(3 * x) + np.cos(4 * z)))
    assert np.abs(U).max() < 2e-13


@pytest.mark.parametrize("N", [7, 8])
def test_quadratic_product_dealiased(N):
    # test_ffts.jl:56-73
    g = O.Grid(N, 9, N, 2 * np.pi, 2 * np.pi)
    for f in (lambda x, y, z: np.cos(3 * x) ** 2 + 0 * z, lambda x, y, z: np.cos(3 * z) ** 2 + 0 * x):
        U = O.forward(g, g.sample(f))
        exact = np.zeros(g.spec_shape, complex)
        exact[0, 0, 0] = 0.5
        assert np.abs(U - exact).max() < 2e-13


@pytest.mark.parametrize("N", [7, 8])
def test_highest_chebyshev_degree(N):
    # test_ffts.jl:75-95
    g = O.Grid(N, 35, N, 2 * np.pi, 2 * np.pi)
    u = g.sample(lambda x, y, z: np.cos(34 * np.arccos(np.clip(y, -1, 1))) * np.cos(2 * x) * np.cos(2 * z))
    U = O.forward(g, u)
    expected = np.zeros(g.spec_shape, complex)
    expected[34, 2, 2] = expected[34, -2, 2] = 0.25
    assert np.abs(U - expected).max() < 2e-12
    assert np.abs(O.inverse(g, U) - u).max() < 2e-12


@pytest.mark.parametrize("Ny", [8, 17, 33, 97])
def test_cheb_is_redft00(Ny):
    # DCT-I closed form: scipy dct(type=1) == FFTW REDFT00; inverse doubles endpoints.
    rng = np.random.default_rng(Ny)
    a = rng.standard_normal((Ny, 12)) + 1j * rng.standard_normal((Ny, 12))
    ref = scipy.fft.dct(a.real, type=1, axis=0) + 1j * scipy.fft.dct(a.imag, type=1, axis=0)
    assert np.abs(O.cheb(a) - ref).max() < 1e-12 * np.abs(ref).max()
    b = a.copy()
    b[0] *= 2
    b[-1] *= 2
    ref = scipy.fft.dct(b.real, type=1, axis=0) + 1j * scipy.fft.dct(b.imag, type=1, axis=0)
    assert np.abs(O.cheb(a, inverse=True) - ref).max() < 1e-12 * np.abs(ref).max()


def test_normalize_forward_matches_definition():
    # ffts.jl:254-269 restated directly
    g = O.Grid(8, 17, 8, 2 * np.pi, 2 * np.pi)
    rng = np.random.default_rng(1)
    U = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    out = O.normalize_forward(g, U, 0.031)
    ref = U * 0.031
    ref[0] *= 0.5
    ref[-1] *= 0.5
    ref[:, :, g.Nx // 2] = 0
    ref[:, g.Nz // 2, :] = 0
    assert np.abs(out - ref).max() < 1e-16


def test_chebcoeffs_known_answer():
    # test_fields.jl:21-22: chebcoeffs(1+2y+3y^2) = [2.5, 2, 1.5, 0...]
    y = np.cos(np.pi * np.arange(9) / 8)
    c = O.chebcoeffs(1 + 2 * y + 3 * y ** 2)
    assert np.abs(c - np.array([2.5, 2, 1.5, 0, 0, 0, 0, 0, 0])).max() < 1e-14


@pytest.mark.parametrize("Ny", [9, 17])
@pytest.mark.parametrize("Nx,Nz", [(40, 42), (41, 43)])
def test_analytic_derivatives(Ny, Nx, Nz):
    # test_operators.jl:6-157
    g = O.Grid(Nx, Ny, Nz, 5.3, 7.1)
    a, b = 2 * np.pi / 5.3, 2 * np.pi / 7.1
    p = lambda y: 1 + 2 * y + 3 * y ** 2 + 4 * y ** 3
    dp = lambda y: 2 + 6 * y + 12 * y ** 2
    d2p = lambda y: 6 + 24 * y
    h = lambda x, z: np.exp(np.cos(a * x) + np.cos(b * z) - 2)
    f = lambda x, y, z: p(y) * h(x, z)
    fx = lambda x, y, z: -a * p(y) * np.sin(a * x) * h(x, z)
    fy = lambda x, y, z: dp(y) * h(x, z)
    fz = lambda x, y, z: -b * p(y) * np.sin(b * z) * h(x, z)
    lap = lambda x, y, z: (d2p(y) + p(y) * (a ** 2 * (np.sin(a * x) ** 2 - np.cos(a * x)) +
                                            b ** 2 * (np.sin(b * z) ** 2 - np.cos(b * z)))) * h(x, z)
    U = O.forward(g, g.sample(f))
    for name, exact in (("ddx1", fx), ("ddx2", fy), ("ddx3", fz)):
        out = O.operator(g, name, U)
        assert np.abs(O.inverse(g, out) - g.sample(exact)).max() < 2e-10
        out2 = O.operator(g, name, U, out=out, add=True)
        assert np.abs(O.inverse(g, out2) - 2 * g.sample(exact)).max() < 4e-10
    L = O.operator(g, "laplacian", U)
    assert np.abs(O.inverse(g, L) - g.sample(lap)).max() < 2e-9
    Om = O.curl(g, np.stack([U, U, U]))
    exact_curl = (lambda x, y, z: fy(x, y, z) - fz(x, y, z),
                  lambda x, y, z: fz(x, y, z) - fx(x, y, z),
                  lambda x, y, z: fx(x, y, z) - fy(x, y, z))
    for i in range(3):
        assert np.abs(O.inverse(g, Om[i]) - g.sample(exact_curl[i])).max() < 4e-10


@pytest.mark.parametrize("form", ["convective", "divergence", "rotating"])
def test_analytic_nonlinear_forms(form):
    # test_nonlinear.jl:1-38: U = (y^2, 1, 0), base (y, 0, 0)
    g = O.Grid(5, 9, 5, 2 * np.pi, 2 * np.pi)
    U = np.stack([O.forward(g, g.sample(lambda x, y, z: y ** 2 + 0 * x + 0 * z)),
                  O.forward(g, np.ones(g.phys_shape)), g.zeros_spec()])
    prob = O.Problem(g, lambda y: y, 0.01, 0.01, form=form)
    expected = [lambda x, y, z: -(2 * y + 1) + 0 * x + 0 * z,
                (lambda x, y, z: (y ** 2 + y) * (2 * y + 1) + 0 * x + 0 * z) if form == "rotating"
                else (lambda x, y, z: 0 * x + 0 * y + 0 * z),
                lambda x, y, z: 0 * x + 0 * y + 0 * z]
    N = prob.nonlinear(U)
    for i in range(3):
        assert np.abs(O.inverse(g, N[i]) - g.sample(expected[i])).max() < 2e-11
    N2 = prob.nonlinear(U, N3=N, add=True)
    for i in range(3):
        assert np.abs(O.inverse(g, N2[i]) - 2 * g.sample(expected[i])).max() < 2e-11


def test_chebyshev_derivative_closed_form():
    # spatial.md:72-77: d/dy T_n via the recurrence equals the dense n*U_{n-1} form
    from helpers import diffmatrix
    rng = np.random.default_rng(3)
    for Ny in (5, 9, 33, 97):
        a = rng.standard_normal(Ny)
        assert np.abs(O.diff(a) - diffmatrix(Ny) @ a).max() < 1e-12 * np.abs(a).max() * Ny ** 2
