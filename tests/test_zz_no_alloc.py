"""cf_advance_to_noalloc (the advance that allocates no device memory: its
shortened final step's solver banks live in a fixed device array) against
cf_advance_to (ext/ChannelFlowFlowsExt.jl:36-52 semantics), bit for bit,
through the C ABI: two calls in a row, the second on a smaller grid, and the
size limit rejected before anything runs."""
import ctypes as C
import math

import pytest

pytestmark = pytest.mark.gpu

L2 = (2 * math.pi / 1.14, 2 * math.pi / 2.5)


def _call(fn, prob, U, P, t0, T):
    from channelflow import _lib
    import torch
    tout = C.c_double()
    vals = (C.c_double * 2)(0.0, 0.0)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda x: C.c_void_p(x.data_ptr())
    rc = fn(prob.scheme.h, ptr(U[0]), ptr(U[1]), ptr(U[2]), ptr(P), t0, T, -1, vals, C.byref(tout), stream)
    return rc, tout.value


def _pair(cf, dims, nsteps, seed):
    import torch
    from channelflow import _lib
    L = _lib.load()
    g = cf.Grid(*dims, *L2)
    prob = cf.CouetteFlow(g, 1 / 400, 0.01)
    st = cf.random_state(prob, 0.1, seed=seed)
    U1, P1 = st.velocity.clone(), st.stagepressure.clone()
    U2, P2 = st.velocity.clone(), st.stagepressure.clone()
    T = nsteps * 0.01
    print(f"no-alloc advance: grid {dims}, span {T}", flush=True)
    rc1, t1 = _call(L.cf_advance_to, prob, U1, P1, 0.0, T)
    rc2, t2 = _call(L.cf_advance_to_noalloc, prob, U2, P2, 0.0, T)
    torch.cuda.synchronize()
    assert rc1 == 0 and rc2 == 0, (rc1, rc2, L.cf_last_error())
    assert t1 == t2 == T
    assert torch.equal(U1, U2) and torch.equal(P1, P2)
    assert bool(torch.isfinite(torch.view_as_real(U2)).all())


def test_noalloc_matches_advance_to(gpu):
    _pair(gpu, (16, 33, 12), 3.37, 5)
    _pair(gpu, (8, 17, 8), 2.41, 6)   # second call: smaller grid, same device array
    _pair(gpu, (16, 33, 12), 1.5, 7)  # and back up, a different remainder


def test_noalloc_rejects_oversized_grid(gpu):
    """128x255x128: three banks of 8 (B (Ny + 16) + Ny) bytes exceed 32 MiB."""
    import torch
    from channelflow import _lib
    cf = gpu
    L = _lib.load()
    g = cf.Grid(128, 255, 128, *L2)
    prob = cf.CouetteFlow(g, 1 / 400, 0.01)
    U = torch.zeros((3,) + g.spec_shape, dtype=torch.complex128, device="cuda")
    P = torch.zeros(g.spec_shape, dtype=torch.complex128, device="cuda")
    rc, _ = _call(L.cf_advance_to_noalloc, prob, U, P, 0.0, 0.015)
    assert rc == 4, rc  # CF_ERR_UNSUPPORTED
    assert b"scratch" in L.cf_last_error()
    rc, _ = _call(L.cf_advance_to_noalloc, prob, U, P, 0.0, 0.02)  # whole steps need no scratch
    torch.cuda.synchronize()
    assert rc == 0
