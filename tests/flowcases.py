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


def two_chambers_side_by_side_tubes(nx=48, ny=32, nz=24, sep=5.0, r_tube=2.5):
    """Two chambers on a periodic x-line joined by two parallel x-tubes side by side in y: the
    interface between the two basins crosses both tubes (two core patches of one pore pair) and
    the solid wall between them (film faces touching both patches)."""
    shape = (nz, ny, nx)
    ch = [(nx * 0.25, ny / 2, nz / 2, 8.0), (nx * 0.75, ny / 2, nz / 2, 8.0)]
    tubes = [(0, (nx * 0.25, ny / 2 + s * sep, nz / 2), r_tube, 0.3, nx / 2) for s in (-1, 1)]
    return chambers_and_tubes(shape, ch, tubes)


def face_tiers(sdf, flow_lab, opn=None):
    """Brute-force patch label of every face f = 3c + d (the +d face of cell c), -1 for faces
    that are not interface faces, following the tier rules of pore_extraction.hpp:655-818 but
    computed as graph components instead of union-find and label propagation:
      core faces: the minimum face id of their component of core faces;
      film faces: per component of film faces, the minimum core-patch label over the core faces
      adjacent to any face of that component, or (no core face adjacent) the component's
      minimum face id.
    Two faces are adjacent when they have the same pore pair, their cells are 26-neighbours
    (periodic) and |2 dc + e_d2 - e_d|^2 <= 8. opn: optional (ox, oy, oz) arrays (Nz, Ny, Nx)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    shape = sdf.shape
    n = sdf.size
    lab = flow_lab.reshape(shape)
    nb_axis = {0: 2, 1: 1, 2: 0}  # face direction d (x, y, z) -> numpy axis
    itf = np.zeros((n, 3), bool)
    core = np.zeros((n, 3), bool)
    key = np.full((n, 3), -1, np.int64)
    for d in range(3):
        nbl = np.roll(lab, -1, axis=nb_axis[d])
        nbs = np.roll(sdf, -1, axis=nb_axis[d])
        o = np.ones(shape) if opn is None else np.roll(opn[d], -1, axis=nb_axis[d])
        it = (lab != nbl) & (lab > 0) & (nbl > 0) & (o > 0)
        itf[:, d] = it.ravel()
        core[:, d] = (it & (sdf > 0) & (nbs > 0)).ravel()
        lo, hi = np.minimum(lab, nbl).astype(np.int64), np.maximum(lab, nbl).astype(np.int64)
        key[:, d] = np.where(it, (lo << 32) | hi, -1).ravel()
    cell = np.arange(n).reshape(shape)
    rows, cols = [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                c2 = np.roll(cell, (-dz, -dy, -dx), axis=(0, 1, 2)).ravel()
                for d in range(3):
                    for d2 in range(3):
                        D = (2 * dx + (d2 == 0) - (d == 0), 2 * dy + (d2 == 1) - (d == 1),
                             2 * dz + (d2 == 2) - (d == 2))
                        if D[0] ** 2 + D[1] ** 2 + D[2] ** 2 > 8:
                            continue
                        ok = itf[:, d] & itf[c2, d2] & (key[:, d] == key[c2, d2])
                        f1 = 3 * np.flatnonzero(ok) + d
                        f2 = 3 * c2[ok] + d2
                        keep = f1 != f2
                        rows.append(f1[keep])
                        cols.append(f2[keep])
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    nf = 3 * n
    itf_f, core_f = itf.ravel(), core.ravel()
    film_f = itf_f & ~core_f
    out = np.full(nf, -1, np.int64)

    def comps(mask):
        sel = mask[rows] & mask[cols]
        g = coo_matrix((np.ones(sel.sum()), (rows[sel], cols[sel])), shape=(nf, nf))
        _, cc = connected_components(g, directed=False)
        ids = np.flatnonzero(mask)
        root = np.full(cc.max() + 1, np.iinfo(np.int64).max, np.int64)
        np.minimum.at(root, cc[ids], ids)
        return cc, root

    cc_core, root_core = comps(core_f)
    out[core_f] = root_core[cc_core[core_f]]
    cc_film, root_film = comps(film_f)
    big = np.iinfo(np.int64).max
    best = np.full(root_film.size, big, np.int64)
    edge = film_f[rows] & core_f[cols]  # film face next to a core face
    np.minimum.at(best, cc_film[rows[edge]], out[cols[edge]])
    fl = np.flatnonzero(film_f)
    att = best[cc_film[fl]]
    out[fl] = np.where(att < big, att, root_film[cc_film[fl]])
    # how many distinct core patches each film component touches (for tests that need > 1)
    pairs = np.unique(np.stack([cc_film[rows[edge]], out[cols[edge]]], axis=1), axis=0)
    touch = np.bincount(pairs[:, 0], minlength=root_film.size) if pairs.size else \
        np.zeros(root_film.size, int)
    return out, int(touch.max(initial=0))
