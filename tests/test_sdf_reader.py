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
