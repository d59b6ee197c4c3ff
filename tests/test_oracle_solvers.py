"""Oracle modal solvers against manufactured solutions and brute force.

Restates test/test_helmoltz.jl, test_batchedinfluence.jl, test_stokes.jl,
test_projection.jl and test_postprocessing.jl.  The un-vendored CHS Helmholtz
semantics are pinned by the dense tau system (helpers.dense_helmholtz) and
the per-mode Stokes solve by a dense brute-force solve of its square system.
"""
import numpy as np
import pytest

from helpers import (bulkmean, coefficients, dense_helmholtz, dense_stokes_mode, diffmatrix, wallderivative,
                     wallvalue)
from oracle import oracle as O


@pytest.mark.parametrize("Ny", [8, 9, 17])
def test_helmholtz_cubic(Ny):
    # test_helmoltz.jl:1-29
    exact = lambda y: 0.7 + 0.3 * y - 0.2 * y ** 2 + 0.1 * y ** 3
    for nu, lam in ((0.1, 0.0), (0.03, 2.5)):
        for amp in (1.0, -0.4):
            rhs = coefficients(lambda y: amp * (nu * (-0.4 + 0.6 * y) - lam * exact(y)), Ny).real
            expected = coefficients(lambda y: amp * exact(y), Ny).real
            out = O.helmholtz(Ny - 1, nu, lam, rhs, amp * exact(1.0), amp * exact(-1.0))
            assert np.abs(out - expected).max() < 2e-12
            assert abs(wallvalue(out, "right") - amp * exact(1.0)) < 2e-12
            assert abs(wallvalue(out, "left") - amp * exact(-1.0)) < 2e-12


@pytest.mark.parametrize("P", [4, 7, 8, 16, 32, 48, 96, 192])
@pytest.mark.parametrize("neumann", [False, True])
def test_helmholtz_matches_dense_tau_system(P, neumann):
    # Uniqueness of the square tau system: any exact solver must agree.
    rng = np.random.default_rng(P + 7 * neumann)
    for t0, t1 in ((1.0, 0.3), (1.0, 37.0), (1 / 2800, 1200.0 + 25 / 2800), (0.03, 2.5), (1.0, 1e-3)):
        if neumann and t1 == 0:
            continue
        f = rng.standard_normal(P + 1) + 1j * rng.standard_normal(P + 1)
        gp, gm = complex(rng.standard_normal(), 1.0), complex(-0.5, rng.standard_normal())
        ref = dense_helmholtz(P, t0, t1, f, gp, gm, neumann)
        out = O.helmholtz(P, t0, t1, f, gp, gm, neumann)
        assert np.abs(out - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def check_stokes(k, l, nu, lam, u, v, w, p, Rx, Ry, Rz, tol_div=2e-10):
    # test_helmoltz.jl:122-151
    P = len(p) - 1
    D = diffmatrix(len(p))
    assert np.abs(1j * k * u + D @ v + 1j * l * w).max() < tol_div
    shift = lam + nu * (k * k + l * l)
    for field, src, gradp in ((u, Rx, 1j * k * p), (v, Ry, D @ p), (w, Rz, 1j * l * p)):
        res = shift * field - nu * D @ (D @ field) + gradp - src
        assert np.abs(res[: P - 1]).max() < 2e-10
        assert abs(wallvalue(field, "left")) < 2e-11
        assert abs(wallvalue(field, "right")) < 2e-11
    assert abs(wallderivative(v, "left")) < 2e-11
    assert abs(wallderivative(v, "right")) < 2e-11


@pytest.mark.parametrize("Ny", [9, 17, 33])
@pytest.mark.parametrize("k,l", [(1.25, 0.0), (0.0, -2.0), (1.25, -1.75), (1.25, 1.75)])
def test_influence_mode_manufactured(Ny, k, l):
    # test_helmoltz.jl:153-216
    nu, lam = 0.03, 2.5
    kappa2 = k * k + l * l
    shift = lam + nu * kappa2
    c = 0.7 + 0.2j
    vf = lambda y: c * (1 - y ** 2) ** 2 * (1 + 0.2 * y)
    dvf = lambda y: c * (-4 * y + 4 * y ** 3 + 0.2 * (1 - 6 * y ** 2 + 5 * y ** 4))
    d2vf = lambda y: c * (-4 + 12 * y ** 2 + 0.2 * (-12 * y + 20 * y ** 3))
    d3vf = lambda y: c * (24 * y + 0.2 * (-12 + 60 * y ** 2))
    qf = lambda y: (0.15 - 0.1j) * (1 - y ** 2) * (1 + 0.3 * y)
    d2qf = lambda y: (0.15 - 0.1j) * (-2 - 1.8 * y)
    uf = lambda y: 1j * k / kappa2 * dvf(y) + l * qf(y)
    wf = lambda y: 1j * l / kappa2 * dvf(y) - k * qf(y)
    pf = lambda y: (0.4 - 0.3j) * (1 + y + 0.2 * y ** 3)
    dpf = lambda y: (0.4 - 0.3j) * (1 + 0.6 * y ** 2)
    Rx = coefficients(lambda y: shift * uf(y) - nu * (1j * k / kappa2 * d3vf(y) + l * d2qf(y)) + 1j * k * pf(y), Ny)
    Ry = coefficients(lambda y: shift * vf(y) - nu * d2vf(y) + dpf(y), Ny)
    Rz = coefficients(lambda y: shift * wf(y) - nu * (1j * l / kappa2 * d3vf(y) - k * d2qf(y)) + 1j * l * pf(y), Ny)
    u, v, w, p = O.influence_mode(Ny, k, l, nu, lam, Rx, Ry, Rz)
    for field, exact in ((u, uf), (v, vf), (w, wf), (p, pf)):
        assert np.abs(field - coefficients(exact, Ny)).max() < 2e-10
    check_stokes(k, l, nu, lam, u, v, w, p, Rx, Ry, Rz)


@pytest.mark.parametrize("Ny,lam", [(5, 0.0), (5, 2.5), (9, 0.0), (9, 2.5), (17, 2.5)])
def test_influence_mode_tau_high_degree(Ny, lam):
    # test_helmoltz.jl:280-297 plus brute-force comparison with the square system
    k, l, nu = 2.0, -2.0, 1.0
    R = [np.array([complex(np.sin(j + n), np.cos(2 * j + n)) / (n + 1) for n in range(Ny)]) for j in (1, 2, 3)]
    u, v, w, p = O.influence_mode(Ny, k, l, nu, lam, *R)
    check_stokes(k, l, nu, lam, u, v, w, p, *R)
    ref = dense_stokes_mode(Ny, k, l, nu, lam, *R)
    for a, b in zip((u, v, w, p), ref):
        assert np.abs(a - b).max() < 1e-10 * max(1, np.abs(b).max())


@pytest.mark.parametrize("Ny", [5, 9, 17, 33])
@pytest.mark.parametrize("lam", [0.0, 3.0])
def test_batched_vs_scalar(Ny, lam):
    # test_batchedinfluence.jl:1-28
    k = np.array([1.0, 0.0, 1.25, 2.0])
    l = np.array([0.0, -2.0, 1.75, -0.5])
    B, nu = 4, 0.03
    rng = np.random.default_rng(Ny)
    R = [rng.standard_normal((Ny, B)) + 1j * rng.standard_normal((Ny, B)) for _ in range(3)]
    outs = O.batched_influence(Ny, k, l, nu, lam, *R)
    for s in range(B):
        ref = O.influence_mode(Ny, k[s], l[s], nu, lam, *(r[:, s] for r in R))
        for o, t in zip(outs, ref):
            assert np.allclose(o[:, s], t, rtol=2e-11, atol=2e-11)
        check_stokes(k[s], l[s], nu, lam, *(o[:, s] for o in outs), *(r[:, s] for r in R))


@pytest.mark.parametrize("Ny", [9, 16, 33])
@pytest.mark.parametrize("lam", [0.0, 2.5])
def test_mean_mode_manufactured(Ny, lam):
    # test_helmoltz.jl:250-305
    nu = 0.03
    grads, means = (-0.7, 0.3), (0.48, -0.2)
    uf = lambda y: (1 - y ** 2) * (0.7 + 0.2 * y + 0.1 * y ** 2)
    wf = lambda y: (1 - y ** 2) * (-0.3 + 0.15 * y)
    pf = lambda y: y + 0.4 * (y ** 2 - 1 / 3) + 0.2 * y ** 3
    Rx = coefficients(lambda y: lam * uf(y) + nu * (1.2 + 1.2 * y + 1.2 * y ** 2) + grads[0], Ny)
    Ry = coefficients(lambda y: 1 + 0.8 * y + 0.6 * y ** 2, Ny)
    Rz = coefficients(lambda y: lam * wf(y) - nu * (0.6 - 0.9 * y) + grads[1], Ny)
    for fixed in (False, True):
        kw = {"bulkvelocity": means} if fixed else {"pressuregradient": grads}
        (u, v, w, p), actual, _ = O.mean_mode(Ny, nu, lam, Rx, Ry, Rz, **kw)
        assert np.allclose(actual, grads, atol=2e-11)
        for field, exact in ((u, uf), (w, wf), (p, pf)):
            assert np.abs(field - coefficients(exact, Ny)).max() < 2e-10
        assert np.all(v == 0)
        assert abs(bulkmean(p)) < 2e-12
        assert abs(bulkmean(u) - means[0]) < 2e-11
        assert abs(bulkmean(w) - means[1]) < 2e-11


def test_mean_mode_exact_tiny_case():
    # test_helmoltz.jl:357-363: Ny=5, bulk 2/3 -> G=-2nu, u=[0.5,0,-0.5,0,0]
    nu = 0.03
    z = np.zeros(5, complex)
    (u, v, w, p), grads, _ = O.mean_mode(5, nu, 0.0, z, z, z, bulkvelocity=(2 / 3, 0))
    assert abs(grads[0] + 2 * nu) < 2e-12 and abs(grads[1]) < 2e-12
    assert np.allclose(u, [0.5, 0, -0.5, 0, 0], atol=1e-14)


def test_mean_mode_complex_forcing():
    # test_helmoltz.jl:307-348
    Ny, nu, lam = 17, 0.03, 2.5
    R = [np.array([complex(np.sin(i * (n + 1)), np.cos((i + 1) * (n + 1))) / (n + 1) ** 2 for n in range(Ny)])
         for i in (1, 2, 3)]
    (u, v, w, p), grads, _ = O.mean_mode(Ny, nu, lam, *R, bulkvelocity=(0.2, -0.1))
    assert abs(bulkmean(u).real - 0.2) < 2e-12
    assert abs(bulkmean(w).real + 0.1) < 2e-12
    assert abs(bulkmean(p)) < 2e-12
    D = diffmatrix(Ny)
    normal = D @ p - R[1]
    assert np.abs(normal[: Ny - 1]).max() < 2e-11
    assert abs(normal[-1] + R[1][-1]) < 2e-12


STOKES_GRIDS = [(6, 6), (5, 5), (6, 5), (5, 6), (1, 5), (5, 1), (1, 1), (2, 2)]


@pytest.mark.parametrize("Nx,Nz", STOKES_GRIDS)
@pytest.mark.parametrize("profile,base_mean", [(lambda y: 1 - y ** 2, 2 / 3), (lambda y: y, 0.0)])
def test_stokes_assembly(Nx, Nz, profile, base_mean):
    # test_stokes.jl:1-131
    Ny, nu, lam = 9, 0.03, 2.5
    Lx, Lz = 5.3, 7.1
    gradients = (-0.4, 0.15)
    g = O.Grid(Nx, Ny, Nz, Lx, Lz)
    R = g.zeros_spec(3)
    exact = g.zeros_spec(3)
    exactP = g.zeros_spec()
    zmodes = list(range(0, (Nz - 1) // 2 + 1)) + list(range(-(Nz // 2), 0))
    for iz, mz in enumerate(zmodes):
        for ix, mx in enumerate(range(0, Nx // 2 + 1)):
            if (Nx % 2 == 0 and mx == Nx // 2) or (Nz % 2 == 0 and mz == -(Nz // 2)):
                R[:, :, iz, ix] = 7 + 3j
                continue
            k, l = 2 * np.pi * mx / Lx, 2 * np.pi * mz / Lz
            mean = mx == 0 and mz == 0
            if mean:
                uf, vf, wf = (lambda y: 0.2 * (1 - y ** 2)), (lambda y: 0 * y), (lambda y: -0.1 * (1 - y ** 2))
                d2uf, d2vf, d2wf = (lambda y: -0.4 + 0 * y), (lambda y: 0 * y), (lambda y: 0.2 + 0 * y)
                pf, dpf = (lambda y: 0.3 * y + 0.2 * (y ** 2 - 1 / 3)), (lambda y: 0.3 + 0.4 * y)
            else:
                k2 = k * k + l * l
                a, b = 0.1 / (1 + mx ** 2 + mz ** 2), 0.05j
                vf = lambda y, a=a: a * (1 - y ** 2) ** 2
                dvf = lambda y, a=a: a * (-4 * y + 4 * y ** 3)
                d2vf = lambda y, a=a: a * (-4 + 12 * y ** 2)
                d3vf = lambda y, a=a: 24 * a * y
                uf = lambda y, k=k, l=l, k2=k2, b=b, dvf=dvf: 1j * k / k2 * dvf(y) + l * b * (1 - y ** 2)
                wf = lambda y, k=k, l=l, k2=k2, b=b, dvf=dvf: 1j * l / k2 * dvf(y) - k * b * (1 - y ** 2)
                d2uf = lambda y, k=k, l=l, k2=k2, b=b, d3vf=d3vf: 1j * k / k2 * d3vf(y) - 2 * l * b
                d2wf = lambda y, k=k, l=l, k2=k2, b=b, d3vf=d3vf: 1j * l / k2 * d3vf(y) + 2 * k * b
                pf, dpf = (lambda y: 0.3 + 0.2 * y + 0.1 * y ** 2), (lambda y: 0.2 + 0.2 * y)
            shift = lam + nu * (k * k + l * l)
            rx = lambda y: shift * uf(y) - nu * d2uf(y) + (gradients[0] if mean else 1j * k * pf(y))
            ry = lambda y: shift * vf(y) - nu * d2vf(y) + dpf(y)
            rz = lambda y: shift * wf(y) - nu * d2wf(y) + (gradients[1] if mean else 1j * l * pf(y))
            for arr, idx, f in ((exact, 0, uf), (exact, 1, vf), (exact, 2, wf), (R, 0, rx), (R, 1, ry), (R, 2, rz)):
                arr[idx, :, iz, ix] = coefficients(f, Ny)
            exactP[:, iz, ix] = coefficients(pf, Ny)
    basec = O.chebcoeffs(profile(g.y))
    for fixed in (False, True):
        if fixed:
            U, P, actual = O.stokes(g, nu, lam, R, bulkvelocity=(base_mean + 0.2 * 2 / 3, -0.1 * 2 / 3), baseflow=basec)
        else:
            U, P, actual = O.stokes(g, nu, lam, R, pressuregradient=gradients)
        assert np.allclose(actual, gradients, atol=2e-11)
        assert np.abs(U - exact).max() < 2e-10
        assert np.abs(P - exactP).max() < 2e-10
        assert abs(bulkmean(U[0, :, 0, 0]).real - 0.2 * 2 / 3) < 2e-11
        assert abs(bulkmean(U[2, :, 0, 0]).real + 0.1 * 2 / 3) < 2e-11
        assert abs(bulkmean(P[:, 0, 0])) < 2e-12


def test_stokes_rejects_bad_grids():
    g = O.Grid(5, 10, 5, 5.3, 7.1)
    with pytest.raises(O.OracleError):
        O.stokes(g, 0.03, 2.5, g.zeros_spec(3))
    with pytest.raises(O.OracleError):
        O.stokes(O.Grid(5, 9, 5, 0.0, 7.1), 0.03, 2.5, O.Grid(5, 9, 5, 1, 1).zeros_spec(3))


def check_constraints(g, U):
    # helpers.jl:77-87
    div = O.operator(g, "ddx1", U[0])
    div = O.operator(g, "ddx2", U[1], out=div, add=True)
    div = O.operator(g, "ddx3", U[2], out=div, add=True)
    assert np.abs(div).max() < 2e-9
    signs = (-1.0) ** np.arange(g.Ny)
    for f in U:
        assert np.abs(f.sum(axis=0)).max() < 2e-10
        assert np.abs((f * signs[:, None, None]).sum(axis=0)).max() < 2e-10


@pytest.mark.parametrize("Nx,Nz", [(40, 42), (41, 43)])
def test_projection_recovers_exact_field(Nx, Nz):
    # test_projection.jl:5-36
    g = O.Grid(Nx, 13, Nz, 5.3, 7.1)
    a, b = 2 * np.pi / 5.3, 2 * np.pi / 7.1
    q = lambda y: (1 - y ** 2) ** 2
    dq = lambda y: -4 * y + 4 * y ** 3
    h = lambda x, z: np.exp(np.cos(a * x) + np.cos(b * z) - 2)
    hx = lambda x, z: -a * np.sin(a * x) * h(x, z)
    hz = lambda x, z: -b * np.sin(b * z) * h(x, z)
    exact = (lambda x, y, z: dq(y) * h(x, z), lambda x, y, z: -q(y) * hx(x, z), lambda x, y, z: 0 * x + 0 * y + 0 * z)
    corr = (lambda x, y, z: q(y) * hx(x, z), lambda x, y, z: dq(y) * h(x, z), lambda x, y, z: q(y) * hz(x, z))
    U = np.stack([O.forward(g, g.sample(lambda x, y, z, i=i: exact[i](x, y, z) + corr[i](x, y, z))) for i in range(3)])
    prob = O.Problem(g, lambda y: 0 * y, 0.01, 0.01, form="convective")
    U = prob.project(U)
    P = prob.pressure(U)
    for i in range(3):
        assert np.abs(O.inverse(g, U[i]) - g.sample(exact[i])).max() < 2e-10
    assert np.all(np.isfinite(P))
    check_constraints(g, U)
    U2 = prob.project(U)
    assert np.abs(U2 - U).max() < 2e-11
    assert np.abs(prob.pressure(U2) - P).max() < 2e-9


def test_projection_bulk_constraint():
    # test_projection.jl:238-249
    g = O.Grid(5, 9, 5, 2 * np.pi, 2 * np.pi)
    prob = O.Problem(g, lambda y: 0 * y, 0.01, 0.01, bulkvelocity=(0.2, -0.1))
    U = prob.project(g.zeros_spec(3))
    assert abs(bulkmean(U[0, :, 0, 0]).real - 0.2) < 2e-12
    assert abs(bulkmean(U[2, :, 0, 0]).real + 0.1) < 2e-12
    check_constraints(g, U)


def test_pressure_laminar_couette():
    # test_projection.jl:258-268: modified pressure y^2/2 - 1/6
    g = O.Grid(5, 9, 5, 2 * np.pi, 2 * np.pi)
    prob = O.Problem(g, lambda y: y, 0.01, 0.01)
    U = prob.project(g.zeros_spec(3))
    P = prob.pressure(U)
    assert np.abs(O.inverse(g, P) - g.sample(lambda x, y, z: y ** 2 / 2 - 1 / 6 + 0 * x + 0 * z)).max() < 2e-10


def test_random_state_real_and_constrained():
    # test_projection.jl:270-296 (with this project's seeded Gaussian stream)
    g = O.Grid(5, 9, 5, 2 * np.pi, 2 * np.pi)
    prob = O.Problem(g, lambda y: 0 * y, 0.01, 0.01)
    noise = 0.1 * np.random.default_rng(17).standard_normal((3,) + g.phys_shape)
    U, P = prob.random_state(noise)
    U2, P2 = prob.random_state(noise)
    assert np.array_equal(U, U2) and np.array_equal(P, P2)
    assert np.abs(U).max() > 0
    check_constraints(g, U)
    for i in range(3):
        assert np.abs(O.forward(g, O.inverse(g, U[i])) - U[i]).max() < 2e-11


def test_energy_dissipation_closed_forms():
    # test_postprocessing.jl:1-30
    g = O.Grid(8, 17, 8, 2 * np.pi, 2 * np.pi)
    nu = 0.01
    Uc = O.forward(g, g.sample(lambda x, y, z: y + 0 * x + 0 * z))
    Up = O.forward(g, g.sample(lambda x, y, z: 1 - y ** 2 + 0 * x + 0 * z))
    Z = g.zeros_spec()
    E = lambda U: sum(O.dot(g, f, f) for f in U) / 2
    Diss = lambda U: nu * sum(O.dot(g, f, f) for f in O.curl(g, U))
    assert abs(E(np.stack([Uc, Z, Z])) - 1 / 6) < 2e-14
    assert abs(Diss(np.stack([Uc, Z, Z])) - nu) < 2e-14
    assert abs(O.dot(g, Uc, Up)) < 2e-14
    assert abs(E(np.stack([Up, Z, Z])) - 4 / 15) < 2e-14
    assert abs(Diss(np.stack([Up, Z, Z])) - 4 * nu / 3) < 2e-14
