"""The oracle's constant-bulk-flux step (_bulkpressuregradient,
cnrk2.jl:98-107, inside step! cnrk2.jl:149-207), pinned by closed forms.

Under a bulk constraint the new uniform gradient G_new is the Lagrange
multiplier of the flux condition, so the velocity does not depend on the
explicit gradient G_old: a wrong G_old shows only in the returned gradient,
G_new = (what the flux condition fixes) - G_old.  The tests therefore check
the returned gradient against the exact one.

G_old = nu (u'(+1) - u'(-1)) / 2 + nu B[Ub''] + B[N] has three terms; each
case below makes the ones it tests nonzero:

* laminar Poiseuille: base Ub = 1.5 (1 - y^2) (bulk 1, Ub'' = -3) with
  bulkvelocity (1, 0) and U = 0 stays at U = 0 with G = -3 nu every stage
  (the base term; the shear and N terms vanish);
* a manufactured (0,0)-mode solution u = A(t) phi(y) e_x + C(t) phi(y) e_z,
  phi = -2 + 15 y^2 - 20 y^4 + 7 y^6: zero at the walls, zero bulk mean,
  phi'(+-1) = -+8 (so the shear term is nonzero) and phi''(+-1) = 0, so the
  forcing F = A' phi - nu A phi'' + c(t) vanishes at the walls (a forcing
  that does not shows the usual Runge-Kutta order reduction at the
  intermediate stages).  Its bulk mean 8 nu A + c(t) is nonzero (the B[N]
  term).  The exact gradient is G_x = c(t) - 3 nu and G_z = d(t).  The
  rotational nonlinear term of a y-only horizontal profile is a pure
  wall-normal force, which the pressure absorbs, so the (0,0) momentum
  equations are exactly the forced heat equations above.

The velocity converges at second order.  The returned gradient is the flux
multiplier of the last stage, which balances the explicit terms sampled at
t + 3 dt / 4, so it converges to the exact G(t + dt) at first order only;
its Richardson extrapolant 2 G(dt/2) - G(dt) converges at second order to
the closed form, which fixes the gradient's limit exactly.  A wrong G_old
term of size s moves the returned gradient by s at every dt.
"""
import numpy as np
import pytest

from oracle import oracle as O

NU = 0.5
W = 5.0
A = (0.7, 1.3)    # A(t) = A0 + A1 sin(W t), streamwise amplitude
CX = (0.4, -0.9)  # c(t) = c0 + c1 sin(W t), y-constant streamwise force
B = (-0.5, 0.8)   # C(t) = B0 + B1 sin(W t), spanwise amplitude
DZ = (0.3, 0.6)   # d(t), y-constant spanwise force
DTS = (0.02, 0.01, 0.005, 0.0025)


def phi(y):
    return -2 + 15 * y ** 2 - 20 * y ** 4 + 7 * y ** 6


def phi2(y):  # phi''
    return 30 - 240 * y ** 2 + 210 * y ** 4


def base(y):
    return 1.5 * (1 - y ** 2)


def manufactured_forcing(g, nu=NU):
    """(F0, F1, F2) with F(t) = F0 + sin(W t) F1 + cos(W t) F2 in the (0,0)
    mode of the x and z components."""
    y = g.y
    S = (3, g.Ny, g.Nz, g.Nxh)
    F0, F1, F2 = (np.zeros(S, complex) for _ in range(3))
    for comp, (a0, a1), (c0, c1) in ((0, A, CX), (2, B, DZ)):
        F0[comp, :, 0, 0] = O.chebcoeffs(-nu * a0 * phi2(y) + c0)
        F1[comp, :, 0, 0] = O.chebcoeffs(-nu * a1 * phi2(y) + c1)
        F2[comp, :, 0, 0] = O.chebcoeffs(a1 * W * phi(y))
    return F0, F1, F2


def exact(g, t, nu=NU):
    """(u coefficients, w coefficients, (G_x, G_z)) of the (0,0) mode at t."""
    amp = A[0] + A[1] * np.sin(W * t)
    bmp = B[0] + B[1] * np.sin(W * t)
    prof = O.chebcoeffs(phi(g.y))
    gx = CX[0] + CX[1] * np.sin(W * t) - 3 * nu
    gz = DZ[0] + DZ[1] * np.sin(W * t)
    return amp * prof, bmp * prof, (gx, gz)


def initial(g):
    U = g.zeros_spec(3)
    ux, uz, _ = exact(g, 0.0)
    U[0, :, 0, 0] = ux
    U[2, :, 0, 0] = uz
    return U


def run(dt, T=1.0, Ny=17):
    """Errors of u, w, signed errors of G_x, G_z, and the largest value of
    every other component and mode, after T / dt steps."""
    g = O.Grid(4, Ny, 4, 2.0, 2.0)
    p = O.Problem(g, base, NU, dt, bulkvelocity=(1.0, 0.0))
    F0, F1, F2 = manufactured_forcing(g)
    p.set_forcing(F0, F1, W, 0.0, F2=F2)
    U = initial(g)
    n = int(round(T / dt))
    U, P = p.step(U, p.pressure(U), n)
    ux, uz, gr = exact(g, n * dt)
    grad = p.last_gradient()
    other = max(np.abs(U[1]).max(), np.abs(U[:, :, 1:, :]).max(), np.abs(U[:, :, :, 1:]).max())
    return (np.abs(U[0, :, 0, 0] - ux).max(), np.abs(U[2, :, 0, 0] - uz).max(),
            grad[0] - gr[0], grad[1] - gr[1], other)


def richardson(errs):
    return [2 * b - a for a, b in zip(errs[:-1], errs[1:])]


def test_manufactured_profile_properties():
    y, w = np.polynomial.legendre.leggauss(8)  # exact for these polynomials
    assert phi(1.0) == 0 and phi(-1.0) == 0 and phi2(1.0) == 0 and phi2(-1.0) == 0
    assert abs(w @ phi(y)) < 1e-14  # zero flux
    assert abs(w @ phi2(y) / 2 + 8) < 1e-13  # B[phi''] = (phi'(1) - phi'(-1)) / 2 = -8


def test_laminar_poiseuille_bulk_gradient():
    g = O.Grid(6, 33, 4, 4 * np.pi, 4 * np.pi / 3)
    nu = 1 / 2800
    p = O.Problem(g, base, nu, 0.005, bulkvelocity=(1.0, 0.0))
    U = g.zeros_spec(3)
    P = p.pressure(U)
    for _ in range(3):
        U, P = p.step(U, P, 1)
        gx, gz = p.last_gradient()
        assert abs(gx - (-3 * nu)) < 1e-15 and abs(gz) < 1e-15, (gx, gz)
    assert np.abs(U).max() < 1e-15


def test_laminar_poiseuille_gradient_constraint_is_steady():
    """The same flow held by its exact gradient -3 nu (pressuregradient
    path, no bulk gradient): U = 0 stays 0."""
    g = O.Grid(6, 33, 4, 4 * np.pi, 4 * np.pi / 3)
    nu = 1 / 2800
    p = O.Problem(g, base, nu, 0.005, pressuregradient=(-3 * nu, 0.0))
    U = g.zeros_spec(3)
    U, P = p.step(U, p.pressure(U), 3)
    assert np.abs(U).max() < 1e-15


def test_manufactured_bulk_flux_velocity_second_order():
    e = [run(dt) for dt in DTS[:3]]
    for r in e:
        assert r[4] < 1e-13  # no other component or mode is excited
    for i, name in enumerate(["u", "w"]):
        e1, e2, e3 = (r[i] for r in e)
        assert e1 < 1e-4, (name, e1)
        r1, r2 = e1 / e2, e2 / e3
        assert 3.3 < r1 < 5.0 and 3.3 < r2 < 5.0, (name, e1, e2, e3)


def test_manufactured_bulk_flux_gradient_limit():
    errs = np.array([run(dt)[2:4] for dt in DTS])
    for c in range(2):
        e = errs[:, c]
        assert 1.6 < e[-2] / e[-1] < 2.4, e  # first order
        rich = richardson(list(e))
        r1, r2 = rich[0] / rich[1], rich[1] / rich[2]
        assert 3.5 < r1 < 5.0 and 3.5 < r2 < 5.0, rich  # second order to the closed form
        assert abs(rich[-1]) < 1e-3, rich


@pytest.mark.parametrize("mutation", [1, 2, 3, 4])
def test_closed_forms_catch_a_planted_mistake(mutation):
    """A mistake planted in the oracle's G_old (a dropped shear, base or B[N]
    term, or a flipped shear sign) leaves the velocity as it was (the flux
    multiplier absorbs G_old) but moves the returned gradient by O(1), so its
    Richardson extrapolant misses the closed form, which the tests above
    reject."""
    good = run(0.01)
    try:
        O.set_bulkgrad_mutation(mutation)
        bad = run(0.01)
        errs = np.array([run(dt)[2:4] for dt in DTS[1:]])
        g = O.Grid(6, 33, 4, 4 * np.pi, 4 * np.pi / 3)
        nu = 1 / 2800
        p = O.Problem(g, base, nu, 0.005, bulkvelocity=(1.0, 0.0))
        U = g.zeros_spec(3)
        p.step(U, p.pressure(U), 1)
        lam_gx = p.last_gradient()[0]
    finally:
        O.set_bulkgrad_mutation(0)
    assert abs(bad[0] - good[0]) < 1e-12 and abs(bad[1] - good[1]) < 1e-12  # velocity unchanged
    rich_x = richardson(list(errs[:, 0]))
    rich_z = richardson(list(errs[:, 1]))
    assert abs(rich_x[-1]) > 0.1, rich_x
    if mutation != 2:  # the base term only acts on G_x
        assert abs(rich_z[-1]) > 0.1, rich_z
    if mutation == 2:  # the laminar case sees only the base term
        assert abs(lam_gx + 3 * nu) > nu
