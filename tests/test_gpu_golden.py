"""GPU against the reference's golden validation data and analytic rates.

* Waleffe equilibria (RotatingForm, the default hot path): the GPU trajectory
  must reproduce validation/results/waleffe.csv (energy, dissipation, wall
  power to 17 printed digits) — a well-conditioned, reference-produced pin.
* Exact viscous decay and Tollmien-Schlichting growth on the GPU (the
  analytic checks of the step), with the reference's own assertion bounds
  (validation/viscous_decay.jl:223-259, tollmien_schlichting.jl:128-157).
"""
import numpy as np
import pytest

import golden_cases as GC
from helpers import clenshaw_curtis_weights

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _close(a, b, rtol):
    return abs(float(a) - float(b)) <= rtol * abs(float(b))


@pytest.mark.parametrize("branch", ["LB", "UB"])
@pytest.mark.parametrize("Ny,N", [(35, 32), (49, 48)])
def test_waleffe_csv_on_gpu(gpu, branch, Ny, N):
    import gpu_backend as GB
    say("waleffe gpu", branch, Ny, N)
    ref = [r for r in GC.read_csv("waleffe.csv") if r["branch"] == branch and int(r["Ny"]) == Ny]
    rows = GC.waleffe_case(GB, branch, Ny, N)
    for row, r in zip(rows, ref):
        assert float(r["dt"]) == row["dt"] and float(r["t"]) == row["t"]
        for key, tol in (("energy", 1e-12), ("dissipation", 1e-11), ("input", 1e-11), ("defect", 1e-7),
                         ("drift", 1e-6)):
            say(f"  dt={row['dt']} t={row['t']} {key}: gpu {row[key]!r} csv {r[key]} rel "
                f"{abs(row[key] - float(r[key])) / abs(float(r[key])):.2e}")
            assert _close(row[key], r[key], tol), (key, row[key], r[key])


@pytest.mark.parametrize("name", ["Poiseuille", "Couette"])
@pytest.mark.parametrize("n", [1, 3])
def test_viscous_decay_csv_on_gpu(gpu, name, n):
    """validation/results/decay.csv (ConvectiveForm, 5x33x8) reproduced by the
    GPU trajectory with the oracle's own bounds."""
    import gpu_backend as GB
    say("decay csv gpu", name, n)
    ref = [r for r in GC.read_csv("decay.csv") if r["flow"] == name and int(r["n"]) == n]
    errors = []
    for dt in (0.5, 0.25, 0.125):
        rows = GC.decay_case(GB, name, n, dt)
        for row in rows:
            r = [x for x in ref if float(x["dt"]) == dt and float(x["t"]) == row["t"]][0]
            say(f"  dt={dt} t={row['t']} energy ratio gpu {row['energy_ratio']!r} csv {r['energy_ratio']}")
            assert _close(row["mu"], r["mu"], 1e-15)
            assert _close(row["energy_ratio"], r["energy_ratio"], 1e-13), (row["energy_ratio"], r["energy_ratio"])
            assert _close(row["error"], r["error"], 1e-6), (row["error"], r["error"])
        errors.append(rows[-1]["error"])
    assert 3.5 < errors[0] / errors[1] < 4.5 and 3.5 < errors[1] / errors[2] < 4.5


def test_tollmien_schlichting_csv_on_gpu(gpu):
    """validation/results/ts_history.csv and ts_convergence.csv (ConvectiveForm,
    8x81x1) reproduced by the GPU trajectory with the oracle's own bounds."""
    import gpu_backend as GB
    hist = GC.read_csv("ts_history.csv")
    conv = GC.read_csv("ts_convergence.csv")
    errors = []
    for dt in (0.4, 0.2, 0.1):
        say("ts csv gpu", dt)
        rows, c, _ = GC.ts_case(GB, dt)
        for row in rows:
            r = [h for h in hist if float(h["dt"]) == dt and float(h["t"]) == row["t"]][0]
            assert _close(row["growth"], r["growth"], 1e-9)
            assert _close(row["frequency"], r["frequency"], 1e-9)
            assert _close(row["energy_ratio"], r["energy_ratio"], 1e-9)
        rc = [x for x in conv if float(x["dt"]) == dt][0]
        say(f"  error gpu {c['error']!r} csv {rc['error']}")
        assert abs(c["error"] - float(rc["error"])) < 5e-11
        errors.append(c["error"])
    assert errors[0] / errors[1] > 3.5 and errors[1] / errors[2] > 3.5 and errors[-1] < 1e-5


def _decay_rotating(B, name, n, dt, T=4.0):
    Re, A = 100.0, 0.1
    nu = 1 / Re
    g = B.Grid(5, 33, 8, 2 * np.pi, 2 * np.pi)
    Ny, Nzp, Nxp = g.phys_shape
    w = clenshaw_curtis_weights(Ny)
    energy = lambda u: np.sum(w * np.sum(u ** 2, axis=(1, 2))) / (4 * Nxp * Nzp)
    alpha = n * np.pi / 2
    mu = nu * (alpha ** 2 + 1)
    initial = g.sample(lambda x, y, z: A * np.sin(alpha * (y + 1)) * np.cos(z) + 0 * x)
    E0 = energy(initial)
    profile, grad = (lambda y: 1 - y ** 2, -2 * nu) if name == "Poiseuille" else (lambda y: y, 0.0)
    prob = B.Problem(g, profile, nu, dt, pressuregradient=(grad, 0.0))
    U = g.zeros_spec(3)
    U[0] = B.forward(g, initial)
    P = prob.pressure(U)
    out = []
    for t in (T / 2, T):
        U, P = prob.step(U, P, int(round((T / 2) / dt)))
        u = B.inverse(g, U[0])
        expected = initial * np.exp(-mu * t)
        err = np.abs(u - expected).max() / np.abs(expected).max()
        out.append((t, energy(u) / E0, np.exp(-2 * mu * t), err, mu))
    return out


@pytest.mark.parametrize("name", ["Poiseuille", "Couette"])
@pytest.mark.parametrize("n", [1, 3])
def test_viscous_decay_rate_on_gpu(gpu, name, n):
    import gpu_backend as GB
    from oracle import oracle as O
    errors = []
    for dt in (0.5, 0.25, 0.125):
        rows = _decay_rotating(GB, name, n, dt)
        ref = _decay_rotating(O, name, n, dt)
        for (t, ratio, exact, err, mu), (_, rratio, _, rerr, _) in zip(rows, ref):
            rate = -np.log(ratio) / (2 * t)
            say(f"decay {name} n={n} dt={dt} t={t}: energy ratio {ratio:.15f} exact {exact:.15f} rate {rate:.6e} "
                f"mu {mu:.6e} err {err:.3e} | oracle ratio {rratio:.15f}")
            assert err < 1e-3
            assert abs(ratio / exact - 1) < 2e-3
            assert abs(rate / mu - 1) < 2e-3
            assert abs(ratio - rratio) < 1e-13 * abs(rratio)
        errors.append(rows[-1][3])
    assert 3.5 < errors[0] / errors[1] < 4.5
    assert 3.5 < errors[1] / errors[2] < 4.5


@pytest.mark.parametrize("dt", [0.4, 0.2, 0.1])
def test_tollmien_schlichting_rate_on_gpu(gpu, dt):
    import gpu_backend as GB
    s_ref = 0.002664410371 - 0.2470750602j
    s, u, v, p, residual = GC.poiseuille_ts_mode(81)
    assert abs(s - s_ref) < 5e-10 and residual < 1e-10
    g = GB.Grid(8, 81, 1, 2 * np.pi, 2 * np.pi)
    nu, amp = 1 / 8000, 1e-7
    U0 = g.zeros_spec(3)
    U0[0, :, 0, 1] = amp / 2 * u
    U0[1, :, 0, 1] = amp / 2 * v
    prob = GB.Problem(g, lambda y: 1 - y ** 2, nu, dt, pressuregradient=(-2 * nu, 0.0))
    P0 = prob.pressure(U0)
    import scipy.linalg
    off = np.array([k / np.sqrt(4 * k * k - 1) for k in range(1, 81)])
    nodes, vecs = scipy.linalg.eigh_tridiagonal(np.zeros(81), off)
    Cg = np.cos(np.outer(np.arccos(np.clip(nodes, -1, 1)), np.arange(81)))
    wts = 2 * vecs[0, :] ** 2

    def energy(U):
        total = 0.0
        for f in U:
            for ix in range(f.shape[2]):
                mult = 1 if (ix == 0 or ix == f.shape[2] - 1) else 2
                for iz in range(f.shape[1]):
                    total += mult * np.sum(wts * np.abs(Cg @ f[:, iz, ix]) ** 2) / 4
        return total

    E0 = energy(U0)
    U, P = U0, P0
    previous, phase = 1.0 + 0j, 0.0
    v0 = U0[1, :, 0, 1]
    for t in np.arange(10.0, 50.0 + 1e-9, 10.0):
        U, P = prob.step(U, P, int(round(10.0 / dt)))
        modal = np.vdot(v0, U[1, :, 0, 1]) / np.vdot(v0, v0)
        phase += np.angle(modal / previous)
        previous = modal
        growth, freq = np.log(abs(modal)) / t, -phase / t
        eratio = energy(U) / E0
        say(f"TS dt={dt} t={t}: growth {growth:.6e} (ref {s_ref.real:.6e}) frequency {freq:.8f} "
            f"(ref {-s_ref.imag:.8f}) energy ratio {eratio:.8f} (exact {np.exp(2 * s_ref.real * t):.8f})")
        assert abs(growth / s_ref.real - 1) < 0.02
        assert abs(freq / (-s_ref.imag) - 1) < 2e-3
        assert abs(eratio / np.exp(2 * s_ref.real * t) - 1) < 5e-3
        assert np.abs(U[2]).max() < 1e-13
