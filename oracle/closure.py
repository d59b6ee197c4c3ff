"""CPU restatement of the learned-closure pieces (test infrastructure only:
imported by tests/ and never by the product).

The reference has no closure; these definitions are this project's, fixed
here so that the device kernels can be checked against them:

  box filter   fine spectral (k, l, n) -> coarse: each retained coarse mode
               takes the fine coefficient of the same wavenumbers times
               sinc(alpha k dx/2) sinc(beta l dz/2) (top-hat widths dx, dz);
               coarse Nyquist planes zero; Chebyshev degrees >= coarse Ny
               dropped.
  MLP          10 -> 64 -> 64 -> 64 -> 3 (GELU, exact erf form, after the
               three hidden layers); inputs (x - mean) * scale; outputs
               y * out_scale.  Parameters theta (float32 master) in the
               layout [W1 (64x16) | b1 | W2 (64x64) | b2 | W3 | b3 | W4 (16x64) | b4 (16)];
               padded input columns 10..15 of W1 and output rows 3..15 of
               W4 are zero.
  BF16 points  (bf16=True) weights, inputs, hidden activations, the output
               gradient and the pre-activation gradients are rounded to
               bfloat16 (round to nearest even) where the device feeds them
               to the matrix cores; sums are exact (float64) here, float32 on
               the device.
  loss         L = (1/B) sum_p sum_{i<3} (y_pi - t_pi)^2 on normalised
               targets t = target / out_scale.
"""
import math

import numpy as np

NIN, KIN, H, NOUT, KOUT = 10, 16, 64, 3, 16
SHAPES = [("W1", (H, KIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (KOUT, H)), ("b4", (KOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.vectorize(math.erf)(z / math.sqrt(2.0)))


def gelu_grad(z):
    phi = np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)
    return 0.5 * (1.0 + np.vectorize(math.erf)(z / math.sqrt(2.0))) + z * phi


def init_theta(seed=0):
    rng = np.random.default_rng(seed)
    th = []
    for name, shp in SHAPES:
        if name.startswith("b"):
            th.append(np.zeros(shp))
            continue
        fan_in = NIN if name == "W1" else H
        w = rng.standard_normal(shp) * math.sqrt(2.0 / fan_in)
        if name == "W1":
            w[:, NIN:] = 0.0
        if name == "W4":
            w[NOUT:, :] = 0.0
            w *= 0.1
        th.append(w)
    return np.concatenate([t.ravel() for t in th]).astype(np.float32)


def forward(theta, X, bf=True):
    """X: (B, 10) normalised inputs.  Returns y (B, 3) and the cache."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(theta)
    Xp = np.zeros((X.shape[0], KIN))
    Xp[:, :NIN] = X
    A0 = r(Xp)
    Z1 = A0 @ r(P["W1"]).T + P["b1"]
    A1 = r(gelu(Z1))
    Z2 = A1 @ r(P["W2"]).T + P["b2"]
    A2 = r(gelu(Z2))
    Z3 = A2 @ r(P["W3"]).T + P["b3"]
    A3 = r(gelu(Z3))
    Y = A3 @ r(P["W4"]).T + P["b4"]
    return Y[:, :NOUT], (A0, Z1, A1, Z2, A2, Z3, A3, Y)


def loss_and_grad(theta, X, T, bf=True):
    """Mean-square loss and d loss / d theta (float64), T: normalised targets (B, 3)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(theta)
    y, (A0, Z1, A1, Z2, A2, Z3, A3, Y) = forward(theta, X, bf)
    B = X.shape[0]
    d = y - T
    loss = float(np.sum(d * d) / B)
    dY = np.zeros((B, KOUT))
    dY[:, :NOUT] = 2.0 * d / B
    g = {}
    g["b4"] = dY.sum(0)
    dYr = r(dY)
    g["W4"] = dYr.T @ A3
    dZ3 = (dYr @ r(P["W4"])) * gelu_grad(Z3)
    g["b3"] = dZ3.sum(0)
    dZ3r = r(dZ3)
    g["W3"] = dZ3r.T @ A2
    dZ2 = (dZ3r @ r(P["W3"])) * gelu_grad(Z2)
    g["b2"] = dZ2.sum(0)
    dZ2r = r(dZ2)
    g["W2"] = dZ2r.T @ A1
    dZ1 = (dZ2r @ r(P["W2"])) * gelu_grad(Z1)
    g["b1"] = dZ1.sum(0)
    dZ1r = r(dZ1)
    g["W1"] = dZ1r.T @ A0
    g["W1"][:, NIN:] = 0.0
    g["W4"][NOUT:, :] = 0.0
    g["b4"][NOUT:] = 0.0
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def adam(theta, m, v, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (float32 arithmetic order of the device kernel)."""
    theta, m, v, grad = (np.asarray(a, np.float32) for a in (theta, m, v, grad))
    m = np.float32(b1) * m + np.float32(1 - b1) * grad
    v = np.float32(b2) * v + np.float32(1 - b2) * grad * grad
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    theta = theta - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
    return theta, m, v


def box_filter(Uf, gf, gc, dx, dz):
    """Fine spectral field(s) (..., Nyf, Nzf, Nxhf) -> coarse (..., Nyc, Nzc, Nxhc)."""
    Uf = np.asarray(Uf)
    out = np.zeros(Uf.shape[:-3] + (gc.Ny, gc.Nz, gc.Nxh), complex)
    alpha, beta = 2 * np.pi / gc.Lx, 2 * np.pi / gc.Lz
    for iz in range(gc.Nz):
        l = iz if iz <= gc.Nz // 2 else iz - gc.Nz
        if gc.Nz % 2 == 0 and iz == gc.Nz // 2:
            continue
        izf = l if l >= 0 else l + gf.Nz
        for k in range(gc.Nxh):
            if gc.Nx % 2 == 0 and k == gc.Nx // 2:
                continue
            w = 1.0
            for wav in (alpha * k * dx, beta * l * dz):
                w *= 1.0 if wav == 0 else math.sin(wav / 2) / (wav / 2)
            out[..., :, iz, k] = w * Uf[..., :gc.Ny, izf, k]
    return out
