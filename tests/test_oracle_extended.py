"""The extended-precision build of the oracle (oracle/libcforacle_x.so: the
same source in long double) is the accuracy reference for FP64 steps.  These
tests show that it computes the same algorithm (the reference's golden
values and the closed forms that pin the FP64 oracle) and that it is far more
accurate than FP64 where FP64 is limited by conditioning (exact long-double
solutions of the tau systems)."""
import os
import types

import numpy as np
import pytest

import golden_cases as GC
from accuracy import class_errors, rel
from helpers import dense_helmholtz, stokes_mode_exact
from oracle import oracle as O

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "extended_couette32.npz")


def test_builds_are_distinct():
    assert O.lib(False).orc_extended() == 0
    assert O.lib(True).orc_extended() == 1


@pytest.mark.parametrize("P,t0,t1", [(96, 1 / 2800, 1200.0), (96, 1 / 400, 600.0), (128, 1.0, 4.0)])
def test_helmholtz_extended_accuracy(P, t0, t1):
    """Dirichlet tau solve against the dense tau system solved in long double
    with iterative refinement: the extended build is >= 1000x closer."""
    rng = np.random.default_rng(P)
    f = rng.standard_normal(P + 1) / (1 + np.arange(P + 1)) ** 2
    exact = _dense_exact(P, t0, t1, f, 0.3, -0.2)
    e64 = np.abs(O.helmholtz(P, t0, t1, f, 0.3, -0.2) - exact).max() / np.abs(exact).max()
    ex = np.abs(O.helmholtz(P, t0, t1, f, 0.3, -0.2, extended=True) - exact).max() / np.abs(exact).max()
    print(f"P={P} t0={t0:.3g} t1={t1:.3g}: FP64 {e64:.3e} extended {ex:.3e}")
    assert ex <= max(1e-3 * e64, 2.3e-16), (ex, e64)


def _dense_exact(P, t0, t1, f, gp, gm):
    from helpers import diffmatrix
    Ny = P + 1
    D = diffmatrix(Ny).astype(np.longdouble)
    A = t0 * (D @ D) - np.longdouble(t1) * np.eye(Ny, dtype=np.longdouble)
    rows = [A[m] for m in range(P - 1)] + [np.ones(Ny, dtype=np.longdouble),
                                          ((-1.0) ** np.arange(Ny)).astype(np.longdouble)]
    A = np.array(rows)
    b = np.concatenate([np.asarray(f[:P - 1], dtype=np.longdouble), [gp, gm]]).astype(np.longdouble)
    A64 = A.astype(float)
    x = np.linalg.solve(A64, b.astype(float)).astype(np.longdouble)
    for _ in range(6):
        x = x + np.linalg.solve(A64, (b - A @ x).astype(float)).astype(np.longdouble)
    return x


def test_dense_exact_matches_dense_fp64():
    """The long-double refined reference above agrees with the FP64 dense
    tau solve of tests/helpers.py at small conditioning (guards the reference
    itself)."""
    rng = np.random.default_rng(3)
    f = rng.standard_normal(17)
    a = _dense_exact(16, 1.0, 2.0, f, 0.5, 0.25).astype(float)
    b = dense_helmholtz(16, 1.0, 2.0, f, 0.5, 0.25)
    assert np.abs(a - b).max() < 1e-13


@pytest.mark.parametrize("k,l", [(0.5, 0.0), (0.0, 1.5), (3.0, 6.0), (60.0, 90.0)])
def test_stokes_mode_extended_accuracy(k, l):
    """Influence-matrix Stokes solve of one mode at the headline stage-1
    shift (Ny = 97, nu = 1/2800, lambda = 1200) against the exact square
    system in long double: FP64 is limited to ~1e-11..1e-9 by conditioning,
    the extended build reaches ~1e-14 or better."""
    Ny, nu, lam = 97, 1 / 2800, 1200.0
    rng = np.random.default_rng(int(10 * (k + l)))
    R = [(rng.standard_normal(Ny) + 1j * rng.standard_normal(Ny)) / (1 + np.arange(Ny)) ** 2 for _ in range(3)]
    exact = stokes_mode_exact(Ny, k, l, nu, lam, *R)
    out64 = O.influence_mode(Ny, k, l, nu, lam, *R)
    outx = O.influence_mode(Ny, k, l, nu, lam, *R, extended=True)
    for name, i in (("u", 0), ("v", 1), ("w", 2), ("p", 3)):
        ex = np.asarray(exact[i], dtype=np.clongdouble)
        sc = float(np.abs(ex).max())
        if sc == 0:
            continue
        e64 = float(np.abs(out64[i] - ex).max()) / sc
        exx = float(np.abs(outx[i] - ex).max()) / sc
        print(f"k={k} l={l} {name}: FP64 {e64:.3e} extended {exx:.3e}")
        assert exx <= max(1e-2 * e64, 1e-15), (name, exx, e64)


def _extended_backend():
    return types.SimpleNamespace(
        Grid=O.Grid,
        forward=lambda g, u: O.forward(g, u, extended=True),
        inverse=lambda g, U: O.inverse(g, U, extended=True),
        Problem=lambda *a, **kw: O.Problem(*a, extended=True, **kw))


@pytest.mark.parametrize("name", ["Poiseuille", "Couette"])
def test_extended_reproduces_decay_csv(name):
    """validation/results/decay.csv (ConvectiveForm, 5x33x8): the extended
    build reproduces the reference's energy ratios as the FP64 oracle does."""
    ref = [r for r in GC.read_csv("decay.csv") if r["flow"] == name and int(r["n"]) == 1]
    rows = GC.decay_case(_extended_backend(), name, 1, 0.25)
    for row in rows:
        r = [x for x in ref if float(x["dt"]) == 0.25 and float(x["t"]) == row["t"]][0]
        assert abs(row["energy_ratio"] - float(r["energy_ratio"])) <= 1e-13 * abs(float(r["energy_ratio"]))


def test_extended_laminar_bulk_gradient():
    """Laminar Poiseuille under constant bulk flux: U stays 0 and the bulk
    gradient is -3 nu for base 1.5 (1 - y^2), to extended precision."""
    nu = 1 / 2800
    g = O.Grid(6, 33, 4, 4 * np.pi, 4 * np.pi / 3)
    p = O.Problem(g, lambda y: 1.5 * (1 - y ** 2), nu, 0.005, bulkvelocity=(1.0, 0.0), extended=True)
    U = g.zeros_spec(3)
    U1, _ = p.step(U, p.pressure(U), 3)
    gx, gz = p.last_gradient()
    assert abs(gx + 3 * nu) < 1e-17 and abs(gz) < 1e-18, (gx, gz)
    assert np.abs(U1).max() < 2e-16


def test_fixture_inputs_and_first_step():
    """The fixture's initial state is the oracle's random_state of the
    documented Philox noise, and one FP64 oracle step lands within its known
    accuracy of the extended step: ~1e-11 in U and ~1e-10 in P, both in the
    top Chebyshev degrees of the smallest-wavenumber modes (the tau correction
    of the near-singular influence systems)."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(FIXTURE)))
    import make_extended_fixture as M
    fx = np.load(FIXTURE)
    g = O.Grid(*M.N, *M.L)
    ref = O.Problem(g, lambda y: y, M.NU, M.DT)
    U0, P0 = ref.random_state(M.noise(g))
    assert rel(U0, fx["U0"]) < 1e-14 and rel(P0, fx["P0"]) < 1e-13
    U1, P1 = ref.step(fx["U0"], fx["P0"], 1)
    X1 = O.Problem(g, lambda y: y, M.NU, M.DT, extended=True).step(fx["U0"], fx["P0"], 1)
    eu, ep = rel(U1, X1[0]), rel(P1, X1[1])
    print(f"one step: FP64 oracle vs extended U {eu:.3e} P {ep:.3e}; by class P {class_errors(P1, X1[1], g)}")
    assert eu < 3e-11 and ep < 3e-10
    # the recorded 100-step FP64 error is the documented magnitude
    assert fx["oracle_err"][0] < 1e-10 and fx["oracle_err"][1] < 1e-7
