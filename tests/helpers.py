"""Test helpers independent of both the oracle and the product.

Restates test/helpers.jl: exact Chebyshev coefficients by an independent
DCT-I (scipy, == FFTW REDFT00), wall values, Gauss-Legendre bulk means, dense
Chebyshev coefficient-differentiation matrices and the dense tau systems.
"""
import numpy as np
import scipy.fft
import scipy.linalg


def coefficients(f, Ny):
    """helpers.jl:5-11: DCT-I on Lobatto nodes / P with halved endpoints."""
    P = Ny - 1
    y = np.cos(np.pi * np.arange(P + 1) / P)
    vals = np.asarray(f(y), dtype=complex) * np.ones(P + 1)
    a = (scipy.fft.dct(vals.real, type=1) + 1j * scipy.fft.dct(vals.imag, type=1)) / P
    a[0] /= 2
    a[-1] /= 2
    return a


def wallvalue(a, side):
    n = np.arange(len(a))
    return np.sum(a * (1.0 if side == "right" else (-1.0) ** n))


def wallderivative(a, side):
    n = np.arange(len(a)).astype(float)
    s = np.ones_like(n) if side == "right" else (-1.0) ** (n + 1)
    return np.sum(s * n * n * a)


def bulkmean(a):
    """helpers.jl:23-32: Gauss-Legendre quadrature via the Jacobi matrix."""
    N = len(a)
    off = np.array([n / np.sqrt(4 * n * n - 1) for n in range(1, N)])
    nodes, vecs = scipy.linalg.eigh_tridiagonal(np.zeros(N), off)
    w = vecs[0, :] ** 2
    vals = np.array([np.sum(a * np.cos(np.arange(N) * np.arccos(np.clip(x, -1, 1)))) for x in nodes])
    return np.sum(w * vals)


def diffmatrix(Ny):
    """Dense Chebyshev coefficient differentiation matrix from T_n' = n U_{n-1}
    (validation/tollmien_schlichting.jl:25-28)."""
    D = np.zeros((Ny, Ny))
    for n in range(1, Ny):
        for k in range(n - 1, -1, -2):
            D[k, n] = n if k == 0 else 2 * n
    return D


def dense_helmholtz(P, t0, t1, f, gp, gm, neumann=False):
    """Dense solve of the Chebyshev-tau system (docs/src/spatial.md:448-470)."""
    Ny = P + 1
    D = diffmatrix(Ny)
    A = np.zeros((Ny, Ny))
    A[: P - 1, :] = (t0 * D @ D - t1 * np.eye(Ny))[: P - 1, :]
    n = np.arange(Ny).astype(float)
    if neumann:
        A[P - 1, :] = n * n
        A[P, :] = (-1.0) ** (n + 1) * n * n
    else:
        A[P - 1, :] = 1.0
        A[P, :] = (-1.0) ** n
    b = np.zeros(Ny, dtype=complex)
    b[: P - 1] = np.asarray(f)[: P - 1]
    b[P - 1], b[P] = gp, gm
    return np.linalg.solve(A, b)


def dense_stokes_mode(Ny, k, l, nu, lam, Rx, Ry, Rz):
    """Brute-force solve of one nonzero mode's square Stokes system:
    momentum rows 0..P-2 for u, v, w; continuity at every degree 0..P;
    u, v, w = 0 at both walls (docs/src/spatial.md:528-563)."""
    P = Ny - 1
    D = diffmatrix(Ny)
    I = np.eye(Ny)
    shift = lam + nu * (k * k + l * l)
    H = shift * I - nu * D @ D
    n = np.arange(Ny)
    bp, bm = np.ones(Ny), (-1.0) ** n
    rows, rhs = [], []
    blocks = [(0, 1j * k * I), (1, D), (2, 1j * l * I)]
    srcs = [Rx, Ry, Rz]
    for c, (_, gp) in enumerate(blocks):
        for m in range(P - 1):
            row = [np.zeros(Ny, complex) for _ in range(4)]
            row[c] = H[m].astype(complex)
            row[3] = gp[m].astype(complex)
            rows.append(np.concatenate(row))
            rhs.append(srcs[c][m])
    for m in range(Ny):  # continuity
        rows.append(np.concatenate([1j * k * I[m], D[m].astype(complex), 1j * l * I[m], np.zeros(Ny, complex)]))
        rhs.append(0.0)
    for c in range(3):
        for b in (bp, bm):
            row = [np.zeros(Ny, complex) for _ in range(4)]
            row[c] = b.astype(complex)
            rows.append(np.concatenate(row))
            rhs.append(0.0)
    A = np.array(rows)
    x = np.linalg.lstsq(A, np.array(rhs, dtype=complex), rcond=None)[0]
    return x[:Ny], x[Ny:2 * Ny], x[2 * Ny:3 * Ny], x[3 * Ny:]


def stokes_mode_exact(Ny, k, l, nu, lam, Rx, Ry, Rz, refinements=6):
    """The same square system as dense_stokes_mode, assembled in long double
    and solved by FP64 LU with iterative refinement on long-double residuals:
    accurate to ~cond x 1e-19, i.e. far below FP64 solver roundoff.  Returns
    (u, v, w, p) as clongdouble."""
    P = Ny - 1
    D = diffmatrix(Ny).astype(np.longdouble)
    I = np.eye(Ny, dtype=np.longdouble)
    k, l, lam, nu = (np.longdouble(x) for x in (k, l, lam, nu))
    H = (lam + nu * (k * k + l * l)) * I - nu * (D @ D)
    n = np.arange(Ny)
    walls = [np.ones(Ny, dtype=np.longdouble), ((-1.0) ** n).astype(np.longdouble)]
    z = np.zeros(Ny, dtype=np.clongdouble)
    rows, rhs = [], []
    for c, (gp, src) in enumerate(zip([1j * k * I, D, 1j * l * I], [Rx, Ry, Rz])):
        for m in range(P - 1):
            row = [z, z, z, gp[m]]
            row[c] = H[m]
            rows.append(np.concatenate(row))
            rhs.append(src[m])
    for m in range(Ny):
        rows.append(np.concatenate([1j * k * I[m], D[m], 1j * l * I[m], z]))
        rhs.append(0.0)
    for c in range(3):
        for b in walls:
            row = [z, z, z, z]
            row[c] = b
            rows.append(np.concatenate(row))
            rhs.append(0.0)
    A = np.array(rows, dtype=np.clongdouble)
    b = np.array(rhs, dtype=np.clongdouble)
    A64 = A.astype(complex)
    x = np.linalg.solve(A64, b.astype(complex)).astype(np.clongdouble)
    for _ in range(refinements):
        x = x + np.linalg.solve(A64, (b - A @ x).astype(complex)).astype(np.clongdouble)
    return x[:Ny], x[Ny:2 * Ny], x[2 * Ny:3 * Ny], x[3 * Ny:]


def spectral_derivative(a):
    """Derivative coefficients by the dense matrix (independent of the oracle)."""
    return diffmatrix(len(a)) @ a


def clenshaw_curtis_weights(Ny):
    """Interpolatory weights reproducing integral(T_j, -1, 1) on Lobatto nodes
    (postprocessing.jl:21-29; viscous_decay.jl:195-197)."""
    j = np.arange(Ny)
    C = np.cos(np.pi * np.outer(j, j) / (Ny - 1))
    moments = np.zeros(Ny)
    even = j % 2 == 0
    moments[even] = 2.0 / (1.0 - j[even].astype(float) ** 2)
    return np.linalg.solve(C.T, moments)
