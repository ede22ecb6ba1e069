"""VTI (VTK ImageData, raw appended binary) writer whose output SDFReader.read_vti reads back.

The reference ships its writer as a script (scripts/vti.py:10-80): XML header, one DataArray per
field with `format="appended"` and byte offsets, then `<AppendedData encoding="raw">`, `_` and per
array a UInt64 byte count followed by the little-endian values in x-fastest order. Its reader
(src/sdf_reader.cpp:28-152, `SDFReader.read_vti`) takes WholeExtent as *point* extents
(N = hi - lo + 1) and reads the first appended array as Float32. The script writes cell data with
extents `0 N`, which that reader takes for N + 1 points per axis (a size mismatch on every file
it writes). This writer therefore stores point data with extents `0 N-1` by default, so that
`save_vti(f, {"sdf": a})` followed by `SDFReader.read_vti(f)` returns `a` with the same origin
and spacing. `data="cell"` writes the script's cell layout with the extents VTK defines for cells
(`0 N`) for viewers; read_vti, like the reference's reader, does not read that layout back.
"""
from __future__ import annotations

import numpy as np

_TYPES = (("Int32", "<i4"), ("Float64", "<f8"), ("Float32", "<f4"))


def _vtk_type(arr):
    if np.issubdtype(arr.dtype, np.integer) or arr.dtype == np.bool_:
        return _TYPES[0]
    if arr.dtype == np.float64:
        return _TYPES[1]
    return _TYPES[2]


def save_vti(filename, fields, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), data="point"):
    """Write `fields` ({name: array}) to `filename`.

    Arrays are (Nz, Ny, Nx) scalars or (Nz, Ny, Nx, 3) vectors, all of one grid shape; integer
    arrays are stored as Int32, float64 as Float64, everything else as Float32. `spacing` and
    `origin` are zyx, as in the reference script and in SDFReader.read_vti's return values.
    For read_vti, put the SDF first (it reads the first array as Float32).
    """
    if data not in ("point", "cell"):
        raise ValueError("data must be 'point' or 'cell'")
    if not fields:
        raise ValueError("fields must not be empty")
    items = [(str(k), np.asarray(v)) for k, v in fields.items()]
    shape = items[0][1].shape[:3]
    if len(shape) != 3:
        raise ValueError(f"field '{items[0][0]}' must be 3-D (Nz, Ny, Nx) or (Nz, Ny, Nx, 3)")
    nz, ny, nx = (int(s) for s in shape)
    hi = (nx, ny, nz) if data == "cell" else (nx - 1, ny - 1, nz - 1)
    dz, dy, dx = (float(s) for s in spacing)
    oz, oy, ox = (float(s) for s in origin)
    ext = f"0 {hi[0]} 0 {hi[1]} 0 {hi[2]}"
    tag = "CellData" if data == "cell" else "PointData"
    head = [
        '<?xml version="1.0"?>\n',
        '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
        'header_type="UInt64">\n',
        f'  <ImageData WholeExtent="{ext}" Origin="{ox!r} {oy!r} {oz!r}" '
        f'Spacing="{dx!r} {dy!r} {dz!r}">\n',
        f'    <Piece Extent="{ext}">\n',
        f"      <{tag}>\n",
    ]
    offset = 0
    payload = []
    for name, arr in items:
        if arr.shape[:3] != shape or not (arr.ndim == 3 or (arr.ndim == 4 and arr.shape[3] == 3)):
            raise ValueError(f"field '{name}' has shape {arr.shape}; expected {shape} or "
                             f"{shape + (3,)}")
        vtk, dt = _vtk_type(arr)
        a = np.ascontiguousarray(arr, dtype=dt)
        ncomp = 3 if arr.ndim == 4 else 1
        head.append(f'        <DataArray type="{vtk}" Name="{name}" NumberOfComponents="{ncomp}" '
                    f'format="appended" offset="{offset}"/>\n')
        payload.append(a)
        offset += 8 + a.nbytes
    head += [f"      </{tag}>\n", "    </Piece>\n", "  </ImageData>\n",
             '  <AppendedData encoding="raw">\n   _']
    with open(filename, "wb") as f:
        f.write("".join(head).encode("ascii"))
        for a in payload:
            f.write(np.uint64(a.nbytes).astype("<u8").tobytes())
            a.tofile(f)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
