"""Small geometries and MAC fields for the network-flow tests.

Geometry helpers build pore SDFs (positive in the pore) from chambers and x/y/z-aligned tubes.
Fields follow the MAC layout of pore_extraction.hpp:476-477: u(i,j,k) lives on the -x face of
cell (i,j,k), v on the -y face, w on the -z face; arrays are (Nz, Ny, Nx).
"""
from __future__ import annotations

import numpy as np


def grid(shape):
    nz, ny, nx = shape
    return np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                       np.arange(nx, dtype=np.float64), indexing="ij")


def _mimg(d, L):
    return d - L * np.round(d / L)


def chambers_and_tubes(shape, chambers, tubes):
    """chambers: [(cx, cy, cz, r)], tubes: [(axis, (cx, cy, cz), r, amp, period)] infinite
    periodic tubes along `axis` through the given point whose radius is
    r + amp*cos(2 pi s / period) (s = coordinate along the axis), so that the tube has no
    plateau ridge. SDF = max over shapes of (radius - distance)."""
    z, y, x = grid(shape)
    nz, ny, nx = shape
    L = {0: nx, 1: ny, 2: nz}
    P = {0: x, 1: y, 2: z}
    best = np.full(shape, -1e9)
    for cx, cy, cz, r in chambers:
        d = np.sqrt(_mimg(x - cx, nx) ** 2 + _mimg(y - cy, ny) ** 2 + _mimg(z - cz, nz) ** 2)
        best = np.maximum(best, r - d)
    for axis, c, r, amp, period in tubes:
        q = 0.0
        for a in range(3):
            if a != axis:
                q = q + _mimg(P[a] - c[a], L[a]) ** 2
        rad = r + amp * np.cos(2 * np.pi * (P[axis] - c[axis]) / period)
        best = np.maximum(best, rad - np.sqrt(q))
    return best.astype(np.float32)


def chain_x(n_ch=3, pitch=24, r_ch=7.0, r_tube=3.0, ny=20, nz=20):
    """A periodic chain of chambers along x joined by one straight x-tube."""
    shape = (nz, ny, n_ch * pitch)
    ch = [(pitch * k + pitch / 2, ny / 2, nz / 2, r_ch) for k in range(n_ch)]
    tube = (0, (pitch / 2, ny / 2, nz / 2), r_tube, 0.5, pitch)
    return chambers_and_tubes(shape, ch, [tube])


def two_chambers_two_tubes(nx=64, ny=24, nz=24):
    """Two chambers on a periodic x-line: one tube joins them directly and, through the wrap,
    also the other way round: pore pair (1, 2) has two disjoint interface patches."""
    shape = (nz, ny, nx)
    ch = [(nx * 0.25, ny / 2, nz / 2, 7.0), (nx * 0.75, ny / 2, nz / 2, 7.0)]
    return chambers_and_tubes(shape, ch, [(0, (nx * 0.25, ny / 2, nz / 2), 3.0, 0.5, nx / 2)])


def uniform_x(shape, U=1.0):
    z = np.zeros(shape)
    return {"u": np.full(shape, U), "v": z.copy(), "w": z.copy(), "p": z.copy()}


def divergence_free(shape, spacing_zyx, seed=0):
    """Discretely divergence-free MAC field from a vector potential (curl on the staggered grid):
    with A on cell edges, u = dAz/dy - dAy/dz etc. using forward differences, so that the MAC
    divergence of every cell is zero up to rounding."""
    rng = np.random.default_rng(seed)
    sz, sy, sx = spacing_zyx
    Ax, Ay, Az = (rng.standard_normal(shape) for _ in range(3))

    def dfw(a, axis, h):  # forward difference a[i+1] - a[i] along numpy axis
        return (np.roll(a, -1, axis=axis) - a) / h

    # numpy axes: z=0, y=1, x=2
    u = dfw(Az, 1, sy) - dfw(Ay, 0, sz)
    v = dfw(Ax, 0, sz) - dfw(Az, 2, sx)
    w = dfw(Ay, 2, sx) - dfw(Ax, 1, sy)
    return {"u": u, "v": v, "w": w, "p": np.zeros(shape)}


def mac_divergence(f, spacing_zyx):
    sz, sy, sx = spacing_zyx
    u, v, w = f["u"], f["v"], f["w"]
    return ((np.roll(u, -1, 2) - u) * (sy * sz) + (np.roll(v, -1, 1) - v) * (sx * sz)
            + (np.roll(w, -1, 0) - w) * (sx * sy))
