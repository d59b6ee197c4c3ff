"""Every tiling of the FP64 step gives the same bits.

The z tile width, the x pass variant of the rotating form (six LDS lines,
split: three lines with the velocity held in registers, async: six lines
filled by direct-to-LDS loads) and the x wavefronts per block change how the
work is cut, not the arithmetic: each padded line is transformed by the same
butterflies in the same order.  So three steps under every available tiling
must equal three steps under the default tiling bit for bit, and the default
must match the oracle (the parity tests hold it to the roundoff envelope)."""
import math

import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu

GRIDS = [
    ((32, 33, 32), (2 * math.pi, math.pi)),
    ((16, 33, 12), (2 * math.pi, math.pi)),
    ((12, 25, 10), (2 * math.pi, math.pi)),  # padded Nx 18: odd half-length M = 9
    ((18, 25, 10), (2 * math.pi, math.pi)),  # padded Nx 27: the odd-length x pass (one variant)
    ((128, 97, 128), (4 * math.pi, 4 * math.pi / 3)),
]


def _run(cf, prob, st, nsteps=3):
    import torch
    U = st.velocity.clone()
    P = st.stagepressure.clone()
    cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=nsteps, **prob.constraint)
    torch.cuda.synchronize()
    return U.cpu().numpy(), P.cpu().numpy()


def _valid_tilings(cf, prob):
    """(z width, x variant, x waves) choices the problem accepts."""
    keep = cf.get_tiles(prob)
    out = []
    for kt in (1, 2, 3, 4, 8, 16):
        for var in (0, 1, 2):
            for w in (1, 2, 3, 4):
                try:
                    cf.set_tiles(prob, kt, var, w)
                except (ValueError, NotImplementedError, RuntimeError):
                    continue
                out.append((kt, var, w))
    cf.set_tiles(prob, keep[0], keep[2], keep[3])
    return out


@pytest.mark.parametrize("shape,box", GRIDS, ids=lambda v: "x".join(map(str, v)) if len(v) == 3 else None)
def test_tilings_bit_identical(gpu, shape, box):
    cf = gpu
    g = cf.Grid(*shape, *box)
    prob = cf.ChannelFlowProblem(g, lambda y: 1.5 * (1 - y ** 2), 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    st = cf.perturbation_state(prob, 0.05, seed=7)
    default = cf.get_tiles(prob)
    U0, P0 = _run(cf, prob, st)
    tilings = _valid_tilings(cf, prob)
    say(f"{shape}: default tiles {default}, {len(tilings)} tilings")
    assert tilings, "no tiling accepted"
    variants = sorted({v for _, v, _ in tilings})
    say(f"{shape}: x variants available {variants}")
    # every variant appears with its widest and narrowest blocks; z widths with the default x choice
    chosen = set()
    for var in variants:
        ws = [w for kt, v, w in tilings if v == var and kt == default[0]]
        if ws:
            chosen.update({(default[0], var, min(ws)), (default[0], var, max(ws))})
    for kt in sorted({kt for kt, _, _ in tilings}):
        chosen.add((kt, default[2], default[3]))
    for kt, var, w in sorted(chosen):
        got = cf.set_tiles(prob, kt, var, w)
        U, P = _run(cf, prob, st)
        same = np.array_equal(U, U0) and np.array_equal(P, P0)
        say(f"{shape}: tiles {got} identical {same}")
        assert same, (shape, kt, var, w, np.abs(U - U0).max(), np.abs(P - P0).max())
    cf.set_tiles(prob, default[0], default[2], default[3])


def test_set_tiles_rejects(gpu):
    cf = gpu
    g = cf.Grid(32, 33, 32, 2 * math.pi, math.pi)
    prob = cf.ChannelFlowProblem(g, lambda y: y, 1 / 400, 0.01)
    keep = cf.get_tiles(prob)
    for bad in [(0, 0, 1), (64, 0, 1), (keep[0], 3, 1), (keep[0], 0, 0), (keep[0], 0, 9)]:
        with pytest.raises(ValueError):
            cf.set_tiles(prob, *bad)
        assert cf.get_tiles(prob) == keep
    conv = cf.ChannelFlowProblem(g, lambda y: y, 1 / 400, 0.01, form=cf.ConvectiveForm)
    with pytest.raises(ValueError):
        cf.set_tiles(conv, cf.get_tiles(conv)[0], 1, 1)


@pytest.mark.parametrize("shape,coefs,bad", [((16, 97, 12), (4, 8, 6), 5), ((8, 193, 6), (4, 8, 6), 5),
                                             ((24, 65, 16), (4, 8), 6), ((6, 257, 4), (6, 8), 4)])
def test_wall_normal_layouts(gpu, shape, coefs, bad):
    """The step's stage and prologue kernels with their default Chebyshev
    coefficients per lane (4; 6 above 256 points) and with the alternatives
    built for the size (8 on half the lanes per mode, or on the same 64
    lanes above 256 points; 6 on 32 or 64 lanes): three steps of each agree
    with the oracle within its roundoff envelope (the scans sum in a
    different order, so the layouts agree to roundoff, not bit for bit)."""
    import torch
    from oracle import oracle as O
    from test_gpu_parity import _split, assert_parity, roundoff_envelope
    cf = gpu
    box = (2 * math.pi, math.pi)
    prof = lambda y: 1.5 * (1 - y ** 2)  # noqa: E731
    g = cf.Grid(*shape, *box)
    prob = cf.ChannelFlowProblem(g, prof, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    st = cf.perturbation_state(prob, 0.05, seed=11)
    U0, P0 = st.velocity.cpu().numpy(), st.stagepressure.cpu().numpy()
    go = O.Grid(*shape, *box)
    ref = O.Problem(go, prof, 1 / 2800, 0.005, bulkvelocity=(1.0, 0.0))
    U1, P1 = ref.step(U0, P0, 3)
    env = roundoff_envelope(lambda a, b: _split(ref.step(a, b, 3)), [U0, P0], _split((U1, P1)), trials=8)
    keep = cf.get_tiles(prob)
    try:
        assert keep[4] == coefs[0]
        for yk in coefs:
            got = cf.set_tiles(prob, keep[0], keep[2], keep[3], y_coefs=yk)
            assert got[4] == yk
            U = torch.as_tensor(U0, device="cuda")
            P = torch.as_tensor(P0, device="cuda")
            cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, nsteps=3, **prob.constraint)
            say(f"{shape}: {yk} coefficients per lane")
            assert_parity(f"{shape} K={yk}", _split((U.cpu().numpy(), P.cpu().numpy())), _split((U1, P1)), env)
    finally:
        cf.set_tiles(prob, keep[0], keep[2], keep[3], y_coefs=keep[4])
    with pytest.raises(ValueError):
        cf.set_tiles(prob, keep[0], keep[2], keep[3], y_coefs=bad)
