"""Accuracy of FP64 results against the extended-precision oracle
(oracle/libcforacle_x.so): relative distances overall and by mode class.

The step's stage pressure is ill-conditioned at small wavenumbers (the 2x2
influence matrix is near-singular there) and the velocity at large
lambda/nu (tau-system condition number 1e7-1e8), so FP64 implementations
with different summation orders differ from each other by more than 1e-12.
The long-double run of the same algorithm on the same FP64 inputs settles
which of two FP64 results is more accurate: each is judged by its distance
from it."""
import numpy as np


def rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def mode_classes(go, n_low=4):
    """Boolean masks over the (l, k) modes of a spectral field: the (0, 0)
    mean mode, the n_low nonzero wavenumber magnitudes closest to zero
    ("low kappa"), and the rest."""
    Nz, Nxh = go.Nz, go.Nxh
    iz = np.arange(Nz)[:, None]
    ix = np.arange(Nxh)[None, :]
    l = np.where(iz <= Nz // 2, iz, iz - Nz)
    k2 = (2 * np.pi / go.Lx * ix) ** 2 + (2 * np.pi / go.Lz * l) ** 2
    k2 = np.broadcast_to(k2, (Nz, Nxh))
    mean = k2 == 0
    levels = np.unique(k2[~mean])[:n_low]
    low = np.isin(k2, levels) & ~mean
    return {"mean": mean, "low kappa": low, "rest": ~(mean | low)}


def class_errors(a, x, go):
    """max |a - x| over each mode class, relative to max |x| of the field;
    a, x: (..., Ny, Nz, Nxh) complex arrays."""
    den = np.abs(x).max()
    d = np.abs(a - x)
    d = d.reshape((-1,) + d.shape[-3:]).max(axis=0).max(axis=0)  # (Nz, Nxh)
    return {name: float(d[m].max() / den) if m.any() else 0.0 for name, m in mode_classes(go).items()}


def report(say, name, got, oracle, x, go):
    """Print and return (hip error, oracle error) against x, overall and by class."""
    eh, eo = rel(got, x), rel(oracle, x)
    ch, co = class_errors(got, x, go), class_errors(oracle, x, go)
    cls = "  ".join(f"{k}: hip {ch[k]:.2e} oracle {co[k]:.2e}" for k in ch)
    say(f"{name}: vs extended  hip {eh:.3e}  oracle {eo:.3e}  |  {cls}")
    return eh, eo
