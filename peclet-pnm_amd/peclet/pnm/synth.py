"""Seeded synthetic SDF volumes (numpy float32, shape (Nz, Ny, Nx), C-order).

Generators for the benchmark workloads and the tests. The packing, Boolean-model and
chamber-lattice generators run in C (libpnm_synth.so, OpenMP); the small analytic fields are
numpy. None of this is part of the extraction path.

Reference geometries restated here as inputs:
  * ``sphere_lattice_sdf`` — tests/kokkos_mpi/test_pnm_mpi.cpp:29-45 (2x2x2 solid spheres,
    R = 0.22 min(dims), min-image distance, FP32).
  * ``trig_sdf`` — tests/kokkos_mpi/test_pnm_mpi.cpp:47-51.
  * ``ci_lattice_sdf`` — .github/workflows/ci.yml:61-68 (32^3, 8 spheres R=10, non-periodic).
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _synth():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libpnm_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `python peclet-pnm_amd/build.py`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        i3 = ctypes.POINTER(ctypes.c_int32)
        L.pnms_pack.restype = ctypes.c_int64
        L.pnms_pack.argtypes = [i3, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_uint64, ctypes.c_int, dp, dp, dp, dp,
                                ctypes.c_int64, dp]
        L.pnms_rsa.restype = ctypes.c_int64
        L.pnms_rsa.argtypes = [i3, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_uint64, ctypes.c_int, dp, dp, dp, dp,
                               ctypes.c_int64, dp]
        L.pnms_boolean.restype = ctypes.c_int64
        L.pnms_boolean.argtypes = [i3, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_uint64, dp, dp, dp, dp, ctypes.c_int64]
        L.pnms_sdf_spheres.restype = ctypes.c_int
        L.pnms_sdf_spheres.argtypes = [i3, dp, dp, dp, dp, ctypes.c_int64, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_float)]
        L.pnms_sdf_chambers.restype = ctypes.c_int
        L.pnms_sdf_chambers.argtypes = [i3, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_float), dp]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _dims_xyz(shape_zyx):
    return np.array([shape_zyx[2], shape_zyx[1], shape_zyx[0]], dtype=np.int32)


def sdf_from_spheres(shape_zyx, centers_xyz, radii, margin=12.0):
    """SDF = min_i(|x - c_i| - r_i) with periodic min-image distance (negative in the solid)."""
    L = _synth()
    c = np.ascontiguousarray(centers_xyz, dtype=np.float64)
    r = np.ascontiguousarray(radii, dtype=np.float64)
    cx, cy, cz = (np.ascontiguousarray(c[:, k]) for k in range(3))
    out = np.empty(tuple(shape_zyx), dtype=np.float32)
    dims = _dims_xyz(shape_zyx)
    L.pnms_sdf_spheres(_p(dims, ctypes.c_int32), _p(cx), _p(cy), _p(cz), _p(r), len(r),
                       float(margin), _p(out, ctypes.c_float))
    return out


def pack_spheres(shape_zyx, dist="uniform", a=6.0, b=10.0, solid_fraction=0.6, seed=1001,
                 max_sweeps=4000):
    """Dense non-overlapping periodic packing. dist: 'uniform' radii U[a,b] or 'lognormal'
    (median a, sigma b). Returns (centers_xyz [k,3], radii [k], achieved solid fraction)."""
    L = _synth()
    dims = _dims_xyz(shape_zyx)
    V = float(np.prod(dims))
    rmin = a if dist == "uniform" else a * math.exp(-3 * b)
    cap = int(solid_fraction * V / (4.0 / 3.0 * math.pi * rmin ** 3)) + 64
    cx, cy, cz, r = (np.zeros(cap) for _ in range(4))
    solid = ctypes.c_double(0.0)
    k = L.pnms_pack(_p(dims, ctypes.c_int32), 0 if dist == "uniform" else 1, a, b,
                    solid_fraction, seed, max_sweeps, _p(cx), _p(cy), _p(cz), _p(r), cap,
                    ctypes.byref(solid))
    return np.stack([cx[:k], cy[:k], cz[:k]], axis=1), r[:k].copy(), solid.value


def rsa_spheres(shape_zyx, dist="lognormal", a=9.0, b=0.3, solid_fraction=0.45, seed=2001,
                attempts=400):
    """Random sequential addition (largest radii first) of non-overlapping periodic spheres.
    Returns (centers_xyz [k,3], radii [k], achieved solid fraction)."""
    L = _synth()
    dims = _dims_xyz(shape_zyx)
    V = float(np.prod(dims))
    rmin = a if dist == "uniform" else a * math.exp(-4 * b)
    cap = int(1.5 * solid_fraction * V / (4.0 / 3.0 * math.pi * rmin ** 3)) + 64
    cx, cy, cz, r = (np.zeros(cap) for _ in range(4))
    solid = ctypes.c_double(0.0)
    k = L.pnms_rsa(_p(dims, ctypes.c_int32), 0 if dist == "uniform" else 1, a, b,
                   solid_fraction, seed, attempts, _p(cx), _p(cy), _p(cz), _p(r), cap,
                   ctypes.byref(solid))
    return np.stack([cx[:k], cy[:k], cz[:k]], axis=1), r[:k].copy(), solid.value


def grow_to_porosity(shape_zyx, centers_xyz, radii, porosity, coarse=4, iters=30):
    """Uniform grain overgrowth: the factor f such that spheres of radius f*r (now allowed to
    overlap) leave the target porosity, found by bisection on a grid coarsened `coarse` times."""
    cs = tuple(max(1, k // coarse) for k in shape_zyx)
    scale = np.array([cs[2] / shape_zyx[2], cs[1] / shape_zyx[1], cs[0] / shape_zyx[0]])
    c = centers_xyz * scale
    rs = radii * float(scale.mean())
    lo, hi = 0.5, 2.0
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        phi = float((sdf_from_spheres(cs, c, rs * mid, margin=8.0) > 0).mean())
        if phi > porosity:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def polydisperse_sdf(shape_zyx, median=9.0, sigma=0.3, porosity=0.35, seed=2001):
    """Polydisperse grain pack: RSA of non-overlapping lognormal spheres (solid ~0.4), then
    uniform grain overgrowth to the target porosity. Returns (sdf, achieved porosity, info)."""
    c, r, solid0 = rsa_spheres(shape_zyx, "lognormal", median, sigma, 0.45, seed)
    f = grow_to_porosity(shape_zyx, c, r, porosity)
    sdf = sdf_from_spheres(shape_zyx, c, r * f, margin=max(12.0, float((r * f).max())))
    return sdf, float((sdf > 0).mean()), {"spheres": len(r), "rsa_solid": solid0, "growth": f}


def boolean_spheres(shape_zyx, median=9.0, sigma=0.3, porosity=0.5, seed=2002):
    """Boolean model of overlapping spheres with lognormal radii."""
    L = _synth()
    dims = _dims_xyz(shape_zyx)
    V = float(np.prod(dims))
    er3 = median ** 3 * math.exp(4.5 * sigma ** 2)
    mean = -math.log(porosity) * V / (4.0 / 3.0 * math.pi * er3)
    cap = int(mean + 10 * math.sqrt(mean) + 64)
    cx, cy, cz, r = (np.zeros(cap) for _ in range(4))
    k = L.pnms_boolean(_p(dims, ctypes.c_int32), median, sigma, porosity, seed, _p(cx), _p(cy),
                       _p(cz), _p(r), cap)
    return np.stack([cx[:k], cy[:k], cz[:k]], axis=1), r[:k].copy()


def packing_sdf(shape_zyx, dist="uniform", a=6.0, b=10.0, solid_fraction=0.6, seed=1001):
    c, r, solid = pack_spheres(shape_zyx, dist, a, b, solid_fraction, seed)
    return sdf_from_spheres(shape_zyx, c, r, margin=max(12.0, float(r.max()))), solid


def boolean_sdf(shape_zyx, median=9.0, sigma=0.3, porosity=0.5, seed=2002):
    c, r = boolean_spheres(shape_zyx, median, sigma, porosity, seed)
    return sdf_from_spheres(shape_zyx, c, r, margin=max(12.0, float(r.max())))


def chamber_sdf(shape_zyx, m=5, chamber_r=12.0, tube_rmin=3.0, tube_rmax=8.0, seed=3001):
    """Pore SDF (positive in the pore) of m^3 spherical chambers on a periodic lattice joined by
    capsule tubes along +x/+y/+z with radii U[tube_rmin, tube_rmax]."""
    L = _synth()
    out = np.empty(tuple(shape_zyx), dtype=np.float32)
    tr = np.zeros(m ** 3 * 3)
    dims = _dims_xyz(shape_zyx)
    L.pnms_sdf_chambers(_p(dims, ctypes.c_int32), m, chamber_r, tube_rmin, tube_rmax, seed,
                        _p(out, ctypes.c_float), _p(tr))
    return out


def sphere_lattice_sdf(shape_zyx):
    """2x2x2 lattice of solid spheres (test_pnm_mpi.cpp:29-45), FP32 arithmetic."""
    nz, ny, nx = shape_zyx
    gd = (nx, ny, nz)
    R = np.float32(0.22) * np.float32(min(gd))
    z, y, x = np.meshgrid(*(np.arange(k, dtype=np.float32) for k in (nz, ny, nx)),
                          indexing="ij")
    best = np.full((nz, ny, nx), np.float32(1e30), dtype=np.float32)
    for cz in range(2):
        for cy in range(2):
            for cx in range(2):
                ccx = np.float32(gd[0] * (2 * cx + 1)) / np.float32(4.0)
                ccy = np.float32(gd[1] * (2 * cy + 1)) / np.float32(4.0)
                ccz = np.float32(gd[2] * (2 * cz + 1)) / np.float32(4.0)
                dx = np.abs(x - ccx)
                dy = np.abs(y - ccy)
                dz = np.abs(z - ccz)
                dx = np.minimum(dx, np.float32(gd[0]) - dx)
                dy = np.minimum(dy, np.float32(gd[1]) - dy)
                dz = np.minimum(dz, np.float32(gd[2]) - dz)
                d = np.sqrt(dx * dx + dy * dy + dz * dz) - R
                best = np.minimum(best, d.astype(np.float32))
    return best


def trig_sdf(shape_zyx):
    """sin(x)cos(y) + 0.7 sin(2z) - 0.1 on a periodic grid (test_pnm_mpi.cpp:47-51)."""
    nz, ny, nx = shape_zyx
    z, y, x = np.meshgrid(*(np.arange(k, dtype=np.float32) for k in (nz, ny, nx)),
                          indexing="ij")
    tp = np.float32(2.0) * np.float32(math.pi)
    xs = tp * x / np.float32(nx)
    ys = tp * y / np.float32(ny)
    zs = tp * z / np.float32(nz)
    v = np.sin(xs) * np.cos(ys) + np.float32(0.7) * np.sin(np.float32(2.0) * zs) - np.float32(0.1)
    return v.astype(np.float32)


def ci_lattice_sdf(n=32, R=10.0):
    """The CI smoke geometry: 8 solid spheres of radius R centred at (8|24)^3 (ci.yml:61-68)."""
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float32),) * 3, indexing="ij")
    sdf = np.full((n, n, n), np.inf, dtype=np.float32)
    for cz in (8, 24):
        for cy in (8, 24):
            for cx in (8, 24):
                d = np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) - R
                sdf = np.minimum(sdf, d.astype(np.float32))
    return sdf


def lattice_flow_fields(shape_zyx, with_openness=True):
    """Smooth synthetic MAC fields u, v, w, p (float64) and three openness fields in [0, 1]
    zeroed below 0.02: the consistency-only fields `fld` of test_pnm_flow_mpi.cpp:43-56 (not a
    physical flow; the bookkeeping only needs consistency). Openness k in (x, y, z) is fld(4)
    evaluated at gx + 4 + k, as the reference test does."""
    nz, ny, nx = shape_zyx
    z, y, x = np.meshgrid(*(np.arange(k, dtype=np.float64) for k in (nz, ny, nx)),
                          indexing="ij")
    tp = 2.0 * math.pi

    def coords(gx):
        return tp * gx / nx, tp * y / ny, tp * z / nz

    X, Y, Z = coords(x)
    out = {
        "u": np.sin(X) * np.cos(Y) + 0.3 * np.cos(2 * Z),
        "v": np.cos(X) * np.sin(Z) - 0.2 * np.sin(Y),
        "w": np.sin(Y) * np.sin(Z) + 0.1 * np.cos(X),
        "p": np.cos(X + Y) + 0.5 * np.sin(Z - X),
    }
    if with_openness:
        for c, name in zip((4, 5, 6), ("ox", "oy", "oz")):
            Xc, Yc, Zc = coords(x + c)
            s = 0.5 + 0.5 * np.sin(Xc + 2 * Yc - Zc)
            out[name] = np.where(s < 0.02, 0.0, s)
    return out
