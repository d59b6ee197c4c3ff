"""Portable field/state files: savefield / loadfield (src/io.jl:21-54).

The reference writes Julia Serialization records (format=:ChannelFlow,
version=2, layout=:xzy); this module writes a documented binary layout that
any language can read and that is validated on load instead of deserialised.

File = 128-byte header + payload, little endian:

  offset  type      field
  0       8 bytes   magic  b"CFLOWST\\0"
  8       uint32    version (1)
  12      uint32    kind: 0 spectral field, 1 spectral vector field,
                          2 state (velocity + stage pressure),
                          3 physical field, 4 physical vector field
  16      uint32    layout: 1 = (x, z, y) column-major (x fastest), as the reference
  20      uint32    dtype: 1 = complex128 (spectral), 2 = float64 (physical)
  24      int32 x3  Nx, Ny, Nz (resolved grid)
  36      uint32    ncomp: 1, 3 or 4 arrays in the payload
  40      float64   Lx
  48      float64   Lz
  56      int64 x3  array dims d0, d1, d2 (fastest first): (Nxh, Nz, Ny) or (Nxp, Nzp, Ny)
  80      float64   simulation time (NaN if not recorded)
  88      int64     step counter (-1 if not recorded)
  96      uint32    CRC-32 (zlib) of the payload
  100     28 bytes  zero
  128     payload: ncomp arrays, each d0*d1*d2 elements in (x, z, y) order;
          a state stores u, v, w, then the stage pressure.

The time step, viscosity, base profile, forcing and constraint are problem
parameters, not part of the file (docs/src/restart.md:1-20); time and step may
be stored for restarts.  Writes go to `path + ".tmp"` and are renamed into
place, so an interrupted write never leaves a truncated file at `path`.
"""
import math
import os
import struct
import zlib

import numpy as np

MAGIC = b"CFLOWST\0"
VERSION = 1
HEADER = struct.Struct("<8sIIII3iId d3qdqI28x")
assert HEADER.size == 128

KIND_SPECTRAL, KIND_VECTOR, KIND_STATE, KIND_PHYSICAL, KIND_PHYSICAL_VECTOR = range(5)
LAYOUT_XZY = 1
DTYPE_C128, DTYPE_F64 = 1, 2


class FileFormatError(ValueError):
    """Raised for files that are not valid state files (mirrors ArgumentError)."""


def _host(a):
    if hasattr(a, "detach"):  # torch tensor (device or host)
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a)


def _classify(grid, arrays):
    spec = (grid.Ny, grid.Nz, grid.Nxh)
    phys = (grid.Ny, grid.Nzp, grid.Nxp)
    shape = arrays[0].shape
    if any(a.shape != shape for a in arrays):
        raise FileFormatError("all arrays must have the same shape")
    if shape == spec and all(a.dtype == np.complex128 for a in arrays):
        return DTYPE_C128, (grid.Nxh, grid.Nz, grid.Ny)
    if shape == phys and all(a.dtype == np.float64 for a in arrays):
        return DTYPE_F64, (grid.Nxp, grid.Nzp, grid.Ny)
    raise FileFormatError(f"arrays of shape {shape} do not match the grid's spectral {spec} or padded {phys} shape")


def savefield(path, grid, field=None, *, velocity=None, pressure=None, time=None, step=None):
    """Write a spectral/physical field (array), a vector field (3 arrays or a
    (3, ...) array) or, with velocity= and pressure=, a State.  Returns path."""
    if velocity is not None or pressure is not None:
        if velocity is None or pressure is None or field is not None:
            raise FileFormatError("a state needs velocity and pressure (and no field)")
        arrays = [_host(velocity[i]) for i in range(3)] + [_host(pressure)]
        kind = KIND_STATE
    else:
        a = _host(field)
        if a.ndim == 4:
            arrays, vector = [a[i] for i in range(a.shape[0])], True
            if len(arrays) != 3:
                raise FileFormatError("a vector field has three components")
        else:
            arrays, vector = [a], False
        kind = None
    dtype, dims = _classify(grid, arrays)
    if kind is None:
        kind = (KIND_VECTOR if vector else KIND_SPECTRAL) if dtype == DTYPE_C128 else (
            KIND_PHYSICAL_VECTOR if vector else KIND_PHYSICAL)
    elif dtype != DTYPE_C128:
        raise FileFormatError("a state holds spectral fields")
    payload = b"".join(np.ascontiguousarray(x).tobytes() for x in arrays)
    header = HEADER.pack(MAGIC, VERSION, kind, LAYOUT_XZY, dtype, grid.Nx, grid.Ny, grid.Nz, len(arrays),
                         float(grid.Lx), float(grid.Lz), *dims, float("nan") if time is None else float(time),
                         -1 if step is None else int(step), zlib.crc32(payload) & 0xFFFFFFFF)
    tmp = os.fspath(path) + ".tmp"
    with open(tmp, "wb") as f:
        f.write(header)
        f.write(payload)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)
    return path


class FieldFile:
    """Contents of a state file: header metadata and host arrays."""

    def __init__(self, kind, Nx, Ny, Nz, Lx, Lz, arrays, time, step):
        self.kind, self.Nx, self.Ny, self.Nz, self.Lx, self.Lz = kind, Nx, Ny, Nz, Lx, Lz
        self.arrays, self.time, self.step = arrays, time, step

    @property
    def is_state(self):
        return self.kind == KIND_STATE

    def velocity(self):
        return np.stack(self.arrays[:3])

    def pressure(self):
        return self.arrays[3]

    def field(self):
        return self.arrays[0] if len(self.arrays) == 1 else np.stack(self.arrays)


def loadfield(path):
    """Read and validate a file written by savefield; returns a FieldFile."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER.size:
        raise FileFormatError("file too short for a header")
    (magic, version, kind, layout, dtype, Nx, Ny, Nz, ncomp, Lx, Lz, d0, d1, d2, time, step,
     crc) = HEADER.unpack_from(raw)
    if magic != MAGIC:
        raise FileFormatError("not a ChannelFlow state file")
    if version != VERSION or layout != LAYOUT_XZY:
        raise FileFormatError("unsupported state file version or layout")
    if kind not in range(5) or dtype not in (DTYPE_C128, DTYPE_F64):
        raise FileFormatError("unknown field kind or element type")
    expect_comp = {KIND_SPECTRAL: 1, KIND_VECTOR: 3, KIND_STATE: 4, KIND_PHYSICAL: 1, KIND_PHYSICAL_VECTOR: 3}[kind]
    if ncomp != expect_comp or (dtype == DTYPE_C128) != (kind in (KIND_SPECTRAL, KIND_VECTOR, KIND_STATE)):
        raise FileFormatError("inconsistent header")
    if dtype == DTYPE_C128:
        ok = (d0, d1, d2) == (Nx // 2 + 1, Nz, Ny)
    else:
        ok = (d0, d1, d2) == ((3 * Nx + 1) // 2, (3 * Nz + 1) // 2, Ny)
    if not ok or min(Nx, Ny, Nz) <= 0 or not (Lx > 0 and Lz > 0):
        raise FileFormatError("array dimensions do not match the grid")
    item = 16 if dtype == DTYPE_C128 else 8
    n = d0 * d1 * d2
    payload = raw[HEADER.size:]
    if len(payload) != ncomp * n * item:
        raise FileFormatError("payload size does not match the header")
    if zlib.crc32(payload) & 0xFFFFFFFF != crc:
        raise FileFormatError("payload checksum mismatch")
    np_dtype = np.complex128 if dtype == DTYPE_C128 else np.float64
    arrays = [np.frombuffer(payload, dtype=np_dtype, count=n, offset=i * n * item).reshape(d2, d1, d0).copy()
              for i in range(ncomp)]
    return FieldFile(kind, Nx, Ny, Nz, Lx, Lz, arrays, None if math.isnan(time) else time, None if step < 0 else step)
