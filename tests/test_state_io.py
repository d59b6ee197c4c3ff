"""Portable state files (channelflow.io; reference savefield/loadfield,
src/io.jl:21-54): round trips, header contents and rejection of invalid files."""
import importlib.util
import os
import struct

import numpy as np
import pytest

from conftest import PKG

_spec = importlib.util.spec_from_file_location("cfio", os.path.join(PKG, "channelflow", "io.py"))
io = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(io)


class G:
    def __init__(self, Nx, Ny, Nz, Lx=5.0, Lz=3.0):
        self.Nx, self.Ny, self.Nz, self.Lx, self.Lz = Nx, Ny, Nz, Lx, Lz
        self.Nxh, self.Nxp, self.Nzp = Nx // 2 + 1, (3 * Nx + 1) // 2, (3 * Nz + 1) // 2


def _spec_field(g, rng, n=None):
    shape = (g.Ny, g.Nz, g.Nxh) if n is None else (n, g.Ny, g.Nz, g.Nxh)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("dims", [(8, 17, 8), (7, 9, 5), (32, 33, 32)])
def test_state_round_trip(tmp_path, dims):
    g = G(*dims)
    rng = np.random.default_rng(1)
    U, P = _spec_field(g, rng, 3), _spec_field(g, rng)
    path = tmp_path / "state.cf"
    io.savefield(path, g, velocity=U, pressure=P, time=12.5, step=1250)
    f = io.loadfield(path)
    assert f.is_state and (f.Nx, f.Ny, f.Nz, f.Lx, f.Lz) == (*dims, 5.0, 3.0)
    assert f.time == 12.5 and f.step == 1250
    assert np.array_equal(f.velocity(), U) and np.array_equal(f.pressure(), P)
    assert not os.path.exists(str(path) + ".tmp")
    # layout: payload is u then v, w, P, each with k fastest (reference (k,l,n) column-major)
    raw = path.read_bytes()
    first = np.frombuffer(raw, dtype=np.complex128, count=3, offset=128)
    assert np.array_equal(first, U[0, 0, 0, :3])
    assert struct.unpack_from("<8s", raw)[0] == b"CFLOWST\0"


def test_fields_and_physical(tmp_path):
    g = G(6, 9, 4)
    rng = np.random.default_rng(2)
    spec = _spec_field(g, rng)
    phys = rng.standard_normal((3, g.Ny, g.Nzp, g.Nxp))
    io.savefield(tmp_path / "a", g, spec)
    io.savefield(tmp_path / "b", g, phys)
    a, b = io.loadfield(tmp_path / "a"), io.loadfield(tmp_path / "b")
    assert a.kind == io.KIND_SPECTRAL and np.array_equal(a.field(), spec) and a.time is None and a.step is None
    assert b.kind == io.KIND_PHYSICAL_VECTOR and np.array_equal(b.field(), phys)


def test_invalid_files_rejected(tmp_path):
    g = G(8, 17, 8)
    rng = np.random.default_rng(3)
    path = tmp_path / "s"
    io.savefield(path, g, velocity=_spec_field(g, rng, 3), pressure=_spec_field(g, rng))
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "bad"
    for mutate, what in ((lambda r: r.__setitem__(slice(0, 8), b"NOTSTATE"), "magic"),
                         (lambda r: r.__setitem__(200, r[200] ^ 1), "checksum"),
                         (lambda r: r.__setitem__(slice(8, 12), struct.pack("<I", 9)), "version"),
                         (lambda r: r.__setitem__(slice(24, 28), struct.pack("<i", 10)), "grid")):
        r = bytearray(raw)
        mutate(r)
        bad.write_bytes(bytes(r))
        with pytest.raises(io.FileFormatError):
            io.loadfield(bad)
    bad.write_bytes(bytes(raw[:-16]))
    with pytest.raises(io.FileFormatError):
        io.loadfield(bad)
    with pytest.raises(io.FileFormatError):
        io.savefield(tmp_path / "x", g, rng.standard_normal((4, 4)))
