"""Golden validation cases of the reference, restated as pure functions.

Each function rebuilds one case of the reference's validation/ (which
produced the committed CSVs in tests/golden/validation) on top of a
``backend`` exposing ``Grid``, ``forward``, ``inverse``, ``operator``,
``stokes`` and ``Problem`` with the oracle's signatures.  The same cases
therefore drive the oracle pinning tests (CPU) and the HIP parity tests (GPU).
"""
import csv
import os
import tarfile

import numpy as np
import scipy.linalg

from helpers import clenshaw_curtis_weights

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "validation")


def read_csv(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return list(csv.DictReader(f))


# ---------------------------------------------------------------------------
# Exact viscous decay (validation/viscous_decay.jl:5-87)
# ---------------------------------------------------------------------------
def decay_case(B, name, n, dt, T=4.0):
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
    prob = B.Problem(g, profile, nu, dt, form="convective", pressuregradient=(grad, 0.0))
    U = g.zeros_spec(3)
    U[0] = B.forward(g, initial)
    P = g.zeros_spec()
    rows = []
    for t in (T / 2, T):
        U, P = prob.step(U, P, int(round((T / 2) / dt)))
        u = B.inverse(g, U[0])
        expected = initial * np.exp(-mu * t)
        err = np.abs(u - expected).max() / np.abs(expected).max()
        rows.append(dict(flow=name, n=n, dt=dt, t=t, mu=mu, energy_ratio=energy(u) / E0,
                         exact_energy=np.exp(-2 * mu * t), error=err, U=U, P=P))
    return rows


# ---------------------------------------------------------------------------
# Waleffe R400 Couette equilibria (validation/waleffe_equilibrium.jl)
# ---------------------------------------------------------------------------
def waleffe_state(B, branch, Ny, N):
    name = f"RRC.a1.14.g2.5.R400.{branch}"
    with tarfile.open(os.path.join(GOLDEN, name + ".tar.gz")) as tf:
        def contents(member):
            return tf.extractfile(f"{name}.{member}").read().decode()
        g = B.Grid(N, Ny, N, 2 * np.pi / 1.14, 2 * np.pi / 2.5)
        U = g.zeros_spec(3)
        for i, comp in enumerate(("u", "v", "w")):
            data = np.array(contents(comp + ".asc").split(), dtype=float)
            assert data.size == 32 * 32 * 34
            raw = data.reshape(34, 32, 32)  # Fortran (x, z, n) -> C (n, z, x)
            spec = np.fft.fft(np.fft.rfft(raw, axis=2), axis=1) / 1024  # (n, l, k)
            for l in range(-15, 16):
                for k in range(0, 16):
                    U[i, :34, l % N, k] = spec[:, l % 32, k]
        U[0, 1, 0, 0] -= 1.0  # subtract the Couette profile y = T_1
    return g, U


def equilibrium_diagnostics(B, g, U):
    Ny, Nzp, Nxp = g.phys_shape
    w = clenshaw_curtis_weights(Ny)
    average = lambda a: np.sum(w * np.sum(a, axis=(1, 2))) / (2 * Nxp * Nzp)
    energy = sum(average(B.inverse(g, U[i]) ** 2) for i in range(3)) / 2
    diss, inp = 0.0, 0.0
    for i in range(3):
        for j, op in enumerate(("ddx1", "ddx2", "ddx3")):
            d = B.inverse(g, B.operator(g, op, U[i]))
            if i == 0 and j == 1:
                d = d + 1.0
                inp = (d[0].sum() + d[-1].sum()) / (800 * Nxp * Nzp)
            diss += average(d ** 2) / 400
    return energy, diss, inp


def waleffe_case(B, branch, Ny, N, dts=(0.025, 0.0125)):
    g, U0 = waleffe_state(B, branch, Ny, N)
    E0, D0, I0 = equilibrium_diagnostics(B, g, U0)
    prob0 = B.Problem(g, lambda y: y, 1 / 400, 0.025)
    source = prob0.nonlinear(U0)
    resp, P0, _ = B.stokes(g, 1 / 400, 0.0, source)
    defect = np.sqrt(equilibrium_diagnostics(B, g, resp - U0)[0] / E0)
    rows = []
    for dt in dts:
        prob = B.Problem(g, lambda y: y, 1 / 400, dt)
        U, P = U0.copy(), P0.copy()
        for t in (0.25, 0.5):
            U, P = prob.step(U, P, int(round(0.25 / dt)))
            drift = np.sqrt(equilibrium_diagnostics(B, g, U - U0)[0] / E0)
            e, d, i = equilibrium_diagnostics(B, g, U)
            rows.append(dict(branch=branch, Ny=Ny, N=N, dt=dt, t=t, defect=defect, drift=drift, energy=e,
                             energy_initial=E0, dissipation=d, input=i, U=U, P=P))
    return rows


# ---------------------------------------------------------------------------
# Tollmien-Schlichting wave (validation/tollmien_schlichting.jl)
# ---------------------------------------------------------------------------
def poiseuille_ts_mode(Ny):
    S = np.zeros((Ny, Ny - 4))
    for k in range(Ny - 4):
        S[k, k] = 1
        S[k + 2, k] = -2 * (k + 2) / (k + 3)
        S[k + 4, k] = (k + 1) / (k + 3)
    D = np.zeros((Ny, Ny))
    for n in range(1, Ny):
        for k in range(n - 1, -1, -2):
            D[k, n] = n if k == 0 else 2 * n
    y = np.cos(np.pi * (np.arange(2 * Ny) + 0.5) / (2 * Ny))
    Cm = np.cos(np.outer(np.arccos(y), np.arange(Ny)))
    I = np.eye(Ny)
    V, L = Cm @ S, Cm @ (D @ D - I) @ S
    A = V.T @ (Cm @ np.linalg.matrix_power(D @ D - I, 2) @ S / 8000 - 1j * (1 - y ** 2)[:, None] * L - 2j * V)
    Bm = V.T @ L
    vals, vecs = scipy.linalg.eig(A.astype(complex), Bm.astype(complex))
    target = 0.002664410371 - 0.2470750602j
    idx = np.argmin(np.abs(vals - target))
    s, c = vals[idx], vecs[:, idx]
    residual = np.linalg.norm(A @ c - s * Bm @ c) / (np.linalg.norm(A @ c) + np.linalg.norm(s * Bm @ c))
    v = S @ c
    u = 1j * D @ v
    scale = np.max(np.sqrt(np.abs(Cm @ u) ** 2 + np.abs(Cm @ v) ** 2))
    u, v = u / scale, v / scale
    Y = np.zeros((Ny, Ny))
    Y[1, 0] = 1
    for n in range(1, Ny):
        Y[n - 1, n] = 0.5
        if n + 1 < Ny:
            Y[n + 1, n] = 0.5
    p = (-1j * (I - Y @ Y) @ u + 2 * Y @ v + (D @ D - I) @ u / 8000 - s * u) / 1j
    return s, u, v, p, residual


def ts_case(B, dt, T=50.0):
    s_ref = 0.002664410371 - 0.2470750602j
    s, u, v, p, residual = poiseuille_ts_mode(81)
    g = B.Grid(8, 81, 1, 2 * np.pi, 2 * np.pi)
    nu, amp = 1 / 8000, 1e-7
    U0 = g.zeros_spec(3)
    P0 = g.zeros_spec()
    U0[0, :, 0, 1] = amp / 2 * u
    U0[1, :, 0, 1] = amp / 2 * v
    P0[:, 0, 1] = amp / 2 * p
    off = np.array([n / np.sqrt(4 * n * n - 1) for n in range(1, 81)])
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
    prob = B.Problem(g, lambda y: 1 - y ** 2, nu, dt, form="convective", pressuregradient=(-2 * nu, 0.0))
    U, P = U0.copy(), P0.copy()
    previous, phase = 1.0 + 0j, 0.0
    v0 = U0[1, :, 0, 1]
    rows = []
    for t in np.arange(10.0, T + 1e-9, 10.0):
        U, P = prob.step(U, P, int(round(10.0 / dt)))
        modal = np.vdot(v0, U[1, :, 0, 1]) / np.vdot(v0, v0)
        phase += np.angle(modal / previous)
        previous = modal
        rows.append(dict(dt=dt, t=t, amplitude=abs(modal), phase=phase, growth=np.log(abs(modal)) / t,
                         frequency=-phase / t, energy_ratio=energy(U) / E0, exact_energy=np.exp(2 * s_ref.real * t)))
    exact = U0 * np.exp(s_ref * T)
    err = np.sqrt(np.sum(np.abs(U - exact) ** 2) / np.sum(np.abs(exact) ** 2))
    return rows, dict(dt=dt, error=err, eigenvalue_real=s.real, eigenvalue_imag=s.imag, eigenpair_residual=residual), U
