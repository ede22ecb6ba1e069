"""Geometries whose Jacobi flood needs many sweeps (pore_extraction.hpp:237-269 loops until a
sweep changes nothing, with no cap).

* `plate_sdf`: a 2x2-voxel solid plate (sdf -0.5: below 0, above the marker threshold -1.5 at
  unit spacing, so it is floodable but holds no marker) attached to one deep grain (sdf -4,
  markers). The flood advances one voxel per sweep along the plate, so a plate of length L
  needs L sweeps plus the one that changes nothing.
* `two_grain_plate_sdf`: the same plate between two grains; the voxel in the middle is reached
  from both in the same sweep and takes the smaller label.
* `TRIG`: the reference's trig field (tests/kokkos_mpi/test_pnm_mpi.cpp:47-51, 169-170) with
  its origin and spacing; the band is a thick shell, about 37 sweeps at 256^3 and 74 at 512^3.
"""
from __future__ import annotations

import numpy as np

# origin and spacing of the reference's trig case, zyx (the reference passes them xyz:
# origin (-1, 2, 0.5), spacing (0.5, 1, 2))
TRIG_ORIGIN_ZYX = (0.5, 2.0, -1.0)
TRIG_SPACING_ZYX = (2.0, 1.0, 0.5)


def _pore_field(shape):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*(np.arange(k, dtype=np.float64) for k in shape), indexing="ij")
    v = 2.0 + np.cos(2 * np.pi * x / 64.0) * np.cos(2 * np.pi * y / ny) + 0.25 * np.sin(
        2 * np.pi * z / nz)
    return v


def plate_sdf(nx=256, plate=(12, 230)):
    """(8, 16, nx) volume: grain at x < 12, plate x in [plate[0], plate[1]), y in {7, 8},
    z in {3, 4}. Needs plate[1] - plate[0] sweeps that label something."""
    shape = (8, 16, nx)
    v = _pore_field(shape)
    v[:, :, : plate[0]] = -4.0
    v[3:5, 7:9, plate[0]: plate[1]] = -0.5
    return v.astype(np.float32)


def two_grain_plate_sdf(nx=256, a_end=12, b=(127, 139)):
    """Grains at x < a_end and x in [b[0], b[1]), plate between them (x in [a_end, b[0])); the
    gap has an odd length, so its middle voxel is equidistant from both grains."""
    shape = (8, 16, nx)
    v = _pore_field(shape)
    v[:, :, :a_end] = -4.0
    v[:, :, b[0]: b[1]] = -4.0
    v[3:5, 7:9, a_end: b[0]] = -0.5
    return v.astype(np.float32)
