"""Independent numpy/scipy formulations of each stage, used to pin the oracle.

Each function computes the same quantity as a stage of src/pore_extraction.hpp, but by a
different route than the oracle's restatement, so that a plausible slip in either (wrong wrap,
sign, tie-break direction, transposed axis, dropped term) shows up as a mismatch:

  peaks        vectorised 26-neighbour dominance test over np.roll views
  centroid     float64 weighted mean (tolerance), exact-zero offset on symmetric peaks
  ccl          graph connected components (scipy.sparse.csgraph) + per-component minimum
  flood        BFS distance layers, then label(v) = min over neighbours one layer closer
  roots        closed-neighbourhood lexicographic argmax, iterated min(depth, 512) times
  relabel      np.unique first-appearance ranks
  topology     np.roll face pairs + np.unique(axis=0)

Volumes are numpy (Nz, Ny, Nx) float32 arrays; flat index = z*Ny*Nx + y*Nx + x.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

OFFSETS26 = [d for d in itertools.product((-1, 0, 1), repeat=3) if d != (0, 0, 0)]  # (dz,dy,dx)


def _nb(a, dz, dy, dx):
    """Value of the periodic neighbour at (z+dz, y+dy, x+dx) for every voxel."""
    return np.roll(a, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def flat_index(shape):
    return np.arange(int(np.prod(shape)), dtype=np.int64).reshape(shape)


def peaks(sdf):
    """Boolean mask of extract_pores peaks: not (v <= 0) and no neighbour dominates."""
    idx = flat_index(sdf.shape)
    dominated = np.zeros(sdf.shape, dtype=bool)
    for dz, dy, dx in OFFSETS26:
        nv = _nb(sdf, dz, dy, dx)
        ni = _nb(idx, dz, dy, dx)
        dominated |= (nv > sdf) | ((nv == sdf) & (ni > idx))
    return ~(sdf <= 0) & ~dominated


def centroid64(sdf, z, y, x):
    """Squared-positive-SDF weighted mean offset in float64."""
    nz, ny, nx = sdf.shape
    sw = 0.0
    p = np.zeros(3)
    for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
        v = float(sdf[(z + dz) % nz, (y + dy) % ny, (x + dx) % nx])
        w = max(v, 0.0) ** 2
        sw += w
        p += w * np.array([dx, dy, dz], dtype=np.float64)
    return p / sw if sw > 1e-6 else np.zeros(3)


def markers(sdf, spacing_zyx):
    thr = np.float32(-1.5) * np.float32(min(np.float32(s) for s in spacing_zyx))
    return sdf < thr


def ccl(mask):
    """Labels: min flat index of the periodic 26-connected component, -1 off the mask."""
    shape = mask.shape
    n = mask.size
    idx = flat_index(shape)
    rows, cols = [], []
    for dz, dy, dx in OFFSETS26:
        both = mask & _nb(mask, dz, dy, dx)
        rows.append(idx[both])
        cols.append(_nb(idx, dz, dy, dx)[both])
    r = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    c = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    g = sp.coo_matrix((np.ones(r.size, dtype=np.int8), (r, c)), shape=(n, n)).tocsr()
    _, comp = connected_components(g, directed=False)
    flat_mask = mask.ravel()
    mins = np.full(comp.max() + 1, np.iinfo(np.int64).max, dtype=np.int64)
    np.minimum.at(mins, comp[flat_mask], np.flatnonzero(flat_mask))
    out = np.full(n, -1, dtype=np.int64)
    out[flat_mask] = mins[comp[flat_mask]]
    return out.astype(np.int32)


def flood_layers(sdf, labels):
    """BFS-layer flood: floodable = !(sdf >= 0) and unlabelled; sources = labelled voxels.
    dist(v) = BFS distance to the labelled set through floodable voxels; label(v) = min label
    of neighbours at dist(v) - 1. Unreached voxels keep -1."""
    shape = sdf.shape
    lab = labels.reshape(shape).astype(np.int64).copy()
    floodable = ~(sdf >= 0) & (lab == -1)
    INF = np.iinfo(np.int64).max
    dist = np.where(lab != -1, 0, INF)
    k = 0
    while True:
        k += 1
        prev_layer = dist == (k - 1)
        reach = np.zeros(shape, dtype=bool)
        for dz, dy, dx in OFFSETS26:
            reach |= _nb(prev_layer, dz, dy, dx)
        new = reach & floodable & (dist == INF)
        if not new.any():
            break
        dist[new] = k
        best = np.full(shape, INF, dtype=np.int64)
        for dz, dy, dx in OFFSETS26:
            nl = _nb(lab, dz, dy, dx)
            nd = _nb(dist, dz, dy, dx)
            cand = np.where(nd == k - 1, nl, INF)
            best = np.minimum(best, cand)
        lab[new] = best[new]
    return lab.ravel().astype(np.int32)


def _walk_next(sdf, sign):
    """Lexicographic (sign * value, index) argmax over the closed 27-neighbourhood, ignoring NaN
    neighbours (a NaN centre keeps itself)."""
    idx = flat_index(sdf.shape)
    bv = sign * sdf.astype(np.float64)
    bi = idx.copy()
    for dz, dy, dx in OFFSETS26:
        nv = sign * _nb(sdf, dz, dy, dx).astype(np.float64)
        ni = _nb(idx, dz, dy, dx)
        better = (nv > bv) | ((nv == bv) & (ni > bi))
        better &= ~np.isnan(nv)
        better &= ~np.isnan(bv)
        bv = np.where(better, nv, bv)
        bi = np.where(better, ni, bi)
    return bi


def ascent_next(sdf, all_voxels):
    """Successor of the gradient walk: ascent (largest value, ties to the larger index) from
    sdf > 0 voxels, or from every voxel when all_voxels; descent (smallest value, ties to the
    larger index) from the others."""
    up = _walk_next(sdf, 1.0)
    if all_voxels:
        return up.ravel()
    return np.where(sdf > 0, up, _walk_next(sdf, -1.0)).ravel()


def roots(sdf, all_voxels=False, cap=512):
    nxt = ascent_next(sdf, all_voxels)
    r = np.arange(nxt.size, dtype=np.int64)
    for _ in range(cap):
        r2 = nxt[r]
        if np.array_equal(r2, r):
            break
        r = r2
    return r.astype(np.int32)


def relabel(sdf, rts, labels):
    """Pore ids by first appearance of the basin root; solid ids by first appearance of the
    label, counted only when that first voxel has sdf <= 0."""
    f = sdf.ravel()
    n = f.size
    seg = np.zeros(n, dtype=np.int64)
    pore = f > 0
    if pore.any():
        keys = rts[pore]
        vox = np.flatnonzero(pore)
        uk, first = np.unique(keys, return_index=True)
        order = np.argsort(vox[first])
        ids = np.empty(uk.size, dtype=np.int64)
        ids[order] = np.arange(1, uk.size + 1)
        seg[pore] = ids[np.searchsorted(uk, keys)]
    sol = ~pore & (labels != -1)
    if sol.any():
        keys = labels[sol]
        vox = np.flatnonzero(sol)
        uk, first = np.unique(keys, return_index=True)
        fv = vox[first]
        valid = f[fv] <= 0
        order = np.argsort(fv)
        ranks = np.cumsum(valid[order]) - valid[order]
        ids = np.zeros(uk.size, dtype=np.int64)
        ids[order] = np.where(valid[order], -(ranks + 1), 0)
        seg[sol] = ids[np.searchsorted(uk, keys)]
    return seg.astype(np.int32)


def segment(sdf, spacing_zyx):
    lab = ccl(markers(sdf, spacing_zyx))
    lab = flood_layers(sdf, lab)
    rt = roots(sdf)
    return relabel(sdf, rt, lab)


def topology(seg, shape_zyx):
    s = np.asarray(seg).reshape(shape_zyx).astype(np.int64)
    pairs = []
    for ax in (2, 1, 0):
        nb = np.roll(s, -1, axis=ax)
        d = s != nb
        lo = np.minimum(s, nb)[d]
        hi = np.maximum(s, nb)[d]
        pairs.append(np.stack([lo, hi], axis=1))
    p = np.concatenate(pairs)
    if p.size == 0:
        return np.zeros((0, 2), dtype=np.int32)
    return np.unique(p, axis=0).astype(np.int32)
