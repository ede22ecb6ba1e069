"""SDFReader.read_vti (sdf_reader.cpp:28-152) on files written here in the raw appended VTI
layout: header with WholeExtent point extents, `_`, UInt64 byte count, Float32 data (CPU)."""
from __future__ import annotations

import numpy as np
import pytest


def write_vti(path, sdf, origin_xyz=(0.0, 0.0, 0.0), spacing_xyz=(1.0, 1.0, 1.0),
              extent_minus_one=True, count=None):
    nz, ny, nx = sdf.shape
    ex = (nx - 1, ny - 1, nz - 1) if extent_minus_one else (nx, ny, nz)
    data = np.ascontiguousarray(sdf, dtype="<f4").tobytes()
    with open(path, "wb") as f:
        f.write(b'<?xml version="1.0"?>\n<VTKFile type="ImageData" version="1.0" '
                b'byte_order="LittleEndian" header_type="UInt64">\n')
        f.write(f'  <ImageData WholeExtent="0 {ex[0]} 0 {ex[1]} 0 {ex[2]}" '
                f'Origin="{origin_xyz[0]} {origin_xyz[1]} {origin_xyz[2]}" '
                f'Spacing="{spacing_xyz[0]} {spacing_xyz[1]} {spacing_xyz[2]}">\n'.encode())
        f.write(b'  </ImageData>\n  <AppendedData encoding="raw">\n   _')
        f.write(np.uint64(len(data) if count is None else count).tobytes())
        f.write(data)
        f.write(b'\n  </AppendedData>\n</VTKFile>\n')


def test_roundtrip(tmp_path, capfd):
    import peclet.pnm as pnm

    rng = np.random.default_rng(0)
    sdf = rng.standard_normal((5, 6, 7)).astype(np.float32)
    p = tmp_path / "a.vti"
    write_vti(p, sdf, (1.5, -2.0, 0.25), (0.5, 1.0, 2.0))
    arr, org, spc = pnm.SDFReader.read_vti(str(p))
    assert arr.shape == (5, 6, 7) and arr.dtype == np.float32
    np.testing.assert_array_equal(arr, sdf)
    assert org == [0.25, -2.0, 1.5] and spc == [2.0, 1.0, 0.5]
    assert "Read data length: 840" in capfd.readouterr().out


def test_errors_and_short_data(tmp_path, capfd):
    import peclet.pnm as pnm

    with pytest.raises(RuntimeError, match="Could not open file"):
        pnm.SDFReader.read_vti(str(tmp_path / "missing.vti"))
    bad = tmp_path / "b.vti"
    bad.write_bytes(b"<VTKFile></VTKFile>")
    with pytest.raises(RuntimeError, match="Could not find WholeExtent"):
        pnm.SDFReader.read_vti(str(bad))
    # byte count larger than the grid: read only the grid, warn
    sdf = np.arange(8, dtype=np.float32).reshape(2, 2, 2)
    p = tmp_path / "c.vti"
    write_vti(p, sdf, count=10_000)
    arr, _, _ = pnm.SDFReader.read_vti(str(p))
    np.testing.assert_array_equal(arr, sdf)
    assert "does not match expected size" in capfd.readouterr().err


def _parse_appended(path):
    """Independent reader of a raw appended VTI: (attrs of ImageData, [(name, type, ncomp,
    array bytes)]) following the offsets in the header."""
    import re

    raw = open(path, "rb").read()
    us = raw.index(b'<AppendedData encoding="raw">')
    us = raw.index(b"_", us) + 1
    head = raw[:us].decode()
    img = dict(re.findall(r'(\w+)="([^"]*)"', re.search(r"<ImageData[^>]*>", head).group(0)))
    arrays = []
    for m in re.finditer(r'<DataArray type="(\w+)" Name="([^"]*)" NumberOfComponents="(\d)" '
                         r'format="appended" offset="(\d+)"/>', head):
        off = us + int(m.group(4))
        nbytes = int(np.frombuffer(raw[off:off + 8], "<u8")[0])
        arrays.append((m.group(2), m.group(1), int(m.group(3)), raw[off + 8:off + 8 + nbytes]))
    return img, arrays


def test_save_vti_roundtrip_through_read_vti(tmp_path):
    """The project's writer: SDFReader.read_vti returns the SDF, origin and spacing bit for bit
    (point extents 0..N-1, which the reference's own script gets wrong for its reader)."""
    import peclet.pnm as pnm
    from peclet.pnm.vti import save_vti

    rng = np.random.default_rng(1)
    sdf = rng.standard_normal((4, 5, 6)).astype(np.float32)
    p = tmp_path / "w.vti"
    save_vti(p, {"sdf": sdf, "seg": np.arange(120, dtype=np.int64).reshape(4, 5, 6)},
             spacing=(2.0, 1.0, 0.5), origin=(0.5, 2.0, -1.0))
    arr, org, spc = pnm.SDFReader.read_vti(str(p))
    np.testing.assert_array_equal(arr.view(np.uint32), sdf.view(np.uint32))
    assert org == [0.5, 2.0, -1.0] and spc == [2.0, 1.0, 0.5]
    img, arrays = _parse_appended(p)
    assert img["WholeExtent"] == "0 5 0 4 0 3"
    assert [a[:3] for a in arrays] == [("sdf", "Float32", 1), ("seg", "Int32", 1)]
    np.testing.assert_array_equal(np.frombuffer(arrays[1][3], "<i4"), np.arange(120))


def test_save_vti_vectors_cells_and_errors(tmp_path):
    from peclet.pnm.vti import save_vti

    u = np.random.default_rng(2).standard_normal((3, 4, 5, 3))
    p = tmp_path / "v.vti"
    save_vti(p, {"u": u}, data="cell")
    img, arrays = _parse_appended(p)
    assert img["WholeExtent"] == "0 5 0 4 0 3"  # cell data: N cells between N+1 points
    name, typ, ncomp, data = arrays[0]
    assert (name, typ, ncomp) == ("u", "Float64", 3)
    np.testing.assert_array_equal(np.frombuffer(data, "<f8").reshape(u.shape), u)
    with pytest.raises(ValueError, match="shape"):
        save_vti(p, {"a": np.zeros((2, 2, 2)), "b": np.zeros((2, 2, 3))})
    with pytest.raises(ValueError):
        save_vti(p, {"a": np.zeros((2, 2, 2))}, data="faces")
