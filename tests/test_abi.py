"""CPU checks of the product boundary: the C-ABI library loads, exports every
symbol include/channelflow_hip.h declares, and rejects bad arguments before
touching a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "channelflow_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cf_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from channelflow import _lib
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 13
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SYMBOLS)


def test_create_rejects_bad_configs_without_device():
    from channelflow import _lib
    L = _lib.load()
    base = np.linspace(-1, 1, 9)

    def make(**kw):
        cfg = _lib.cf_config()
        cfg.Nx, cfg.Ny, cfg.Nz = kw.get("Nx", 8), kw.get("Ny", 9), kw.get("Nz", 8)
        cfg.Lx, cfg.Lz = kw.get("Lx", 1.0), kw.get("Lz", 1.0)
        cfg.nu, cfg.dt = kw.get("nu", 0.01), kw.get("dt", 0.01)
        b = np.ascontiguousarray(kw.get("base", base))
        cfg.baseflow = b.ctypes.data_as(C.POINTER(C.c_double))
        cfg.form = kw.get("form", 0)
        cfg.constraint = kw.get("constraint", 0)
        cfg.constraint_values[0], cfg.constraint_values[1] = kw.get("vals", (0.0, 0.0))
        cfg.precision = kw.get("precision", 0)
        h = C.c_void_p()
        rc = L.cf_problem_create(C.byref(cfg), C.byref(h))
        return rc, b

    # validation happens before any allocation, so these run without a GPU
    assert make(Ny=10)[0] == _lib.CF_ERR_ARGUMENT          # even Ny (stokes.jl:35)
    assert make(Nx=0)[0] == _lib.CF_ERR_ARGUMENT           # stokes.jl:36
    assert make(Lx=0.0)[0] == _lib.CF_ERR_ARGUMENT         # stokes.jl:37-38
    assert make(dt=0.0)[0] == _lib.CF_ERR_ARGUMENT         # cnrk2.jl:60
    assert make(dt=float("inf"))[0] == _lib.CF_ERR_ARGUMENT
    assert make(vals=(float("nan"), 0.0))[0] == _lib.CF_ERR_ARGUMENT  # problem.jl:86-87
    assert make(form=3)[0] == _lib.CF_ERR_ARGUMENT          # forms: rotating, convective, divergence
    assert b"form" in L.cf_last_error()
    assert make(Ny=513)[0] == _lib.CF_ERR_UNSUPPORTED      # a mode's column must fit one wavefront (8 x 64)
    assert make(precision=2)[0] == _lib.CF_ERR_ARGUMENT
    assert make(precision=1, form=2)[0] == _lib.CF_ERR_UNSUPPORTED  # mixed: no divergence form
    assert make(precision=1, Nx=7)[0] == _lib.CF_ERR_UNSUPPORTED    # mixed: odd Nxp (11) rejected


def test_python_mirror_validates_constraints():
    import channelflow as cf
    with pytest.raises(cf.ArgumentError):
        cf._constraint((0, 0), (0, 0))
    with pytest.raises(cf.ArgumentError):
        cf._constraint(None, (float("nan"), 0))
    for bad in (2 / 3, (1.0,), (1.0, 0.0, 0.0)):
        with pytest.raises(cf.ArgumentError):
            cf._constraint(None, bad)
    assert cf._constraint(None, None) == (0, (0.0, 0.0))
    assert cf._constraint(None, (1.0, 0.0)) == (1, (1.0, 0.0))
