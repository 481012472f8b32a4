"""VTK ImageData (``.vti``) writer and reader for sampled volumes (mbem's
src/mbem/vti.py): the XML header of a little-endian ``ImageData`` file with
``appended`` / ``raw`` point-data arrays, each a uint64 byte count followed
by its float32 values at the offset its ``DataArray`` declares; point data
x-fastest, i.e. C order over (nz, ny, nx).  The layout is the file format a
viewer (and the reference's reader) opens, so it is kept as the reference
writes it; ``read`` of a ``write`` is bitwise.
"""
from __future__ import annotations

import pathlib
import re
import struct

import numpy as np

_LEN = "<Q"


def write(path, origin, spacing, dims, arrays: dict) -> pathlib.Path:
    """``arrays``: name -> array with leading axes (nz, ny, nx) (a scalar) or
    (nz, ny, nx, c) (c components); ``dims`` = (nx, ny, nz) point counts."""
    path = pathlib.Path(path)
    nx, ny, nz = (int(d) for d in dims)
    blocks, entries, offset = [], [], 0
    for name, a in arrays.items():
        a = np.asarray(a)
        if a.shape[:3] != (nz, ny, nx):
            raise ValueError(f"{name!r} has leading shape {a.shape[:3]}, expected (nz, ny, nx) = "
                             f"{(nz, ny, nx)}")
        comps = 1 if a.ndim == 3 else int(np.prod(a.shape[3:]))
        data = np.ascontiguousarray(a, dtype="<f4").reshape(-1).tobytes()
        entries.append(f'        <DataArray type="Float32" Name="{name}" '
                       f'NumberOfComponents="{comps}" format="appended" offset="{offset}"/>')
        blocks += [struct.pack(_LEN, len(data)), data]
        offset += struct.calcsize(_LEN) + len(data)
    ext = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    head = "\n".join([
        '<?xml version="1.0"?>',
        '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
        f'  <ImageData WholeExtent="{ext}" Origin="{origin[0]!r} {origin[1]!r} {origin[2]!r}" '
        f'Spacing="{spacing[0]!r} {spacing[1]!r} {spacing[2]!r}">',
        f'    <Piece Extent="{ext}">',
        '      <PointData>',
        *entries,
        '      </PointData>',
        '    </Piece>',
        '  </ImageData>',
        '  <AppendedData encoding="raw">',
        '   _'])
    with open(path, "wb") as fh:
        fh.write(head.encode())
        for b in blocks:
            fh.write(b)
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return path


def read(path) -> dict:
    """{origin, spacing, dims, arrays} of a file ``write`` produced."""
    raw = pathlib.Path(path).read_bytes()
    head, _, body = raw.partition(b'<AppendedData encoding="raw">\n   _')
    text = head.decode()
    grab = lambda key: re.search(key + r'="([^"]+)"', text).group(1).split()  # noqa: E731
    ext = [int(v) for v in grab("WholeExtent")]
    dims = (ext[1] + 1, ext[3] + 1, ext[5] + 1)
    nx, ny, nz = dims
    arrays = {}
    for m in re.finditer(r'<DataArray type="Float32" Name="([^"]+)" NumberOfComponents="(\d+)" '
                         r'format="appended" offset="(\d+)"/>', text):
        name, comps, off = m.group(1), int(m.group(2)), int(m.group(3))
        (nbytes,) = struct.unpack_from(_LEN, body, off)
        a = np.frombuffer(body, dtype="<f4", count=nbytes // 4, offset=off + struct.calcsize(_LEN))
        arrays[name] = a.reshape((nz, ny, nx) if comps == 1 else (nz, ny, nx, comps))
    return {"origin": tuple(float(v) for v in grab("Origin")),
            "spacing": tuple(float(v) for v in grab("Spacing")), "dims": dims, "arrays": arrays}
