"""Image files of the reference (src/MutoFile.cpp:85-245): a 64-byte header
(u8 length 64, u8 version 1, 16 chars method, u32 rays, u16 iterations, u16
bytes per value, u16 nx, ny, nz, f32 x_min, y_min, z_min, dx, dy, dz, 8 zero
bytes) then nx*ny*nz float32 or float64 values, x fastest, little-endian.
Used to chain PoCA -> MLSD runs and to resume from a saved image."""
import struct

import numpy as np

_HDR = struct.Struct("<BB16sIHHHHH6f8x")
assert _HDR.size == 64


def write_image(path, img, grid, method="", n_ray=0, n_iter=0, precision=4):
    img = np.asarray(img, np.float64).ravel()
    assert img.size == grid.nx * grid.ny * grid.nz
    if precision not in (4, 8):
        raise ValueError("precision must be 4 or 8 bytes")
    hdr = _HDR.pack(64, 1, method.encode()[:16].ljust(16, b"\0"), int(n_ray), int(n_iter),
                    int(precision), grid.nx, grid.ny, grid.nz, grid.x_min, grid.y_min, grid.z_min,
                    grid.dx, grid.dy, grid.dz)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write((img.astype("<f4") if precision == 4 else img.astype("<f8")).tobytes())


def read_image(path):
    """Returns (values as float64, header dict with method, n_ray, n_iter,
    precision and the grid fields)."""
    with open(path, "rb") as f:
        raw = f.read()
    (length, version, method, n_ray, n_iter, prec, nx, ny, nz, x0, y0, z0, dx, dy, dz) = \
        _HDR.unpack_from(raw, 0)
    n = nx * ny * nz
    dt = "<f4" if prec == 4 else "<f8"
    vals = np.frombuffer(raw, dtype=dt, count=n, offset=length).astype(np.float64)
    hdr = dict(method=method.split(b"\0", 1)[0].decode(), version=version, n_ray=n_ray,
               n_iter=n_iter, precision=prec, nx=nx, ny=ny, nz=nz, x_min=x0, y_min=y0, z_min=z0,
               dx=dx, dy=dy, dz=dz)
    return vals, hdr
