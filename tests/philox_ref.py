"""Reference Philox4x32-10 (Salmon et al., SC 2011) and the Box-Muller
Gaussian stream the device noise kernel documents (test infrastructure)."""
import numpy as np

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) for c in ctr)
    k0, k1 = (np.asarray(k, dtype=np.uint64) for k in key)
    for _ in range(10):
        p0, p1 = M0 * c0, M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0, k1 = (k0 + np.uint64(W0)) & np.uint64(MASK), (k1 + np.uint64(W1)) & np.uint64(MASK)
    return c0, c1, c2, c3


def u01(hi, lo):
    b = ((hi.astype(np.uint64) << np.uint64(21)) ^ (lo.astype(np.uint64) >> np.uint64(11))) & np.uint64((1 << 53) - 1)
    return (b.astype(np.float64) + 1.0) / 9007199254740992.0


def gaussian(n, eps, seed, comp, pairs=None):
    """Samples of component `comp` (first 2*len(pairs) elements of the given pairs)."""
    pairs = np.arange((n + 1) // 2, dtype=np.uint64) if pairs is None else np.asarray(pairs, dtype=np.uint64)
    k0, k1 = seed & MASK, (seed >> 32) & MASK
    r = philox4x32_10((pairs & np.uint64(MASK), pairs >> np.uint64(32), np.full_like(pairs, comp), np.zeros_like(pairs)),
                      (np.uint64(k0), np.uint64(k1)))
    u1, u2 = u01(r[0], r[1]), u01(r[2], r[3])
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * len(pairs))
    out[0::2] = eps * rad * np.cos(2 * np.pi * u2)
    out[1::2] = eps * rad * np.sin(2 * np.pi * u2)
    return out
