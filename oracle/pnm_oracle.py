"""ctypes wrapper of the serial C oracle (oracle/libpnm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker. The product (libpnm_hip.so, peclet.pnm) never imports it.

All functions take numpy (Nz, Ny, Nx) C-order volumes and zyx origin/spacing, like the reference
Python API, and convert to the xyz convention of the C core (pnm_bindings.cpp:116-118: origin
and spacing are cast to float32).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpnm_oracle.so")

PORE_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("radius", "<f4")])


class _Pore(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float),
                ("radius", ctypes.c_float)]


class _Flow(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int64), ("nt", ctypes.c_int64),
                ("pores", ctypes.POINTER(_Pore)),
                ("pore_pressure", ctypes.POINTER(ctypes.c_double)),
                ("pore_residual", ctypes.POINTER(ctypes.c_double)),
                ("throats", ctypes.POINTER(ctypes.c_int32)),
                ("throat_flow", ctypes.POINTER(ctypes.c_double)),
                ("throat_area", ctypes.POINTER(ctypes.c_double)),
                ("throat_dp", ctypes.POINTER(ctypes.c_double)),
                ("flow_lab", ctypes.POINTER(ctypes.c_int32)),
                ("fpar", ctypes.POINTER(ctypes.c_int64)),
                ("peak", ctypes.POINTER(ctypes.c_int32)),
                ("cx", ctypes.POINTER(ctypes.c_double)),
                ("cy", ctypes.POINTER(ctypes.c_double)),
                ("cz", ctypes.POINTER(ctypes.c_double))]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.pnmo_extract_pores.restype = ctypes.c_int64
        L.pnmo_extract_pores.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64]
        L.pnmo_init_markers.argtypes = [vp, vp, vp, vp]
        L.pnmo_ccl.argtypes = [vp, vp]
        L.pnmo_ccl.restype = ctypes.c_int
        L.pnmo_flood.argtypes = [vp, vp, vp]
        L.pnmo_flood.restype = ctypes.c_int
        L.pnmo_gradient_roots.argtypes = [vp, vp, ctypes.c_int, vp]
        L.pnmo_relabel.argtypes = [vp, vp, vp, ctypes.c_int64, vp]
        L.pnmo_segment_volume.argtypes = [vp, vp, vp, vp, vp, vp]
        L.pnmo_segment_volume.restype = ctypes.c_int
        L.pnmo_extract_topology.restype = ctypes.c_int64
        L.pnmo_extract_topology.argtypes = [vp, vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_int32))]
        L.pnmo_extract_network_flow.restype = ctypes.c_int
        L.pnmo_extract_network_flow.argtypes = [vp] * 12 + [ctypes.c_int, ctypes.POINTER(_Flow)]
        L.pnmo_flow_free.argtypes = [ctypes.POINTER(_Flow)]
        L.pnmo_free.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    """Raw pointer of a numpy array; the caller keeps the array alive across the call."""
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _dims(shape_zyx):
    return np.array([shape_zyx[2], shape_zyx[1], shape_zyx[0]], dtype=np.int32)


def _xyz_f32(v):
    return np.array([v[2], v[1], v[0]], dtype=np.float32)


def _sdf(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError("SDF array must be 3D (Nz,Ny,Nx)")
    return a


def extract_pores(sdf, origin_zyx, spacing_zyx, cap=1000000):
    """Pores in ascending peak-voxel index; returns (structured array, total peaks)."""
    s = _sdf(sdf)
    out = np.zeros(max(cap, 1), dtype=PORE_DTYPE)
    d, o, sp = _dims(s.shape), _xyz_f32(origin_zyx), _xyz_f32(spacing_zyx)
    tot = lib().pnmo_extract_pores(_p(s), _p(d), _p(o), _p(sp), _p(out), cap)
    return out[: min(tot, cap)].copy(), int(tot)


def segment_volume(sdf, spacing_zyx, return_stages=False):
    s = _sdf(sdf)
    n = s.size
    seg = np.zeros(n, dtype=np.int32)
    roots = np.zeros(n, dtype=np.int32)
    labels = np.zeros(n, dtype=np.int32)
    d, sp = _dims(s.shape), _xyz_f32(spacing_zyx)
    lib().pnmo_segment_volume(_p(s), _p(d), _p(sp), _p(seg), _p(roots), _p(labels))
    if return_stages:
        return seg, roots, labels
    return seg


def markers_ccl(sdf, spacing_zyx):
    """Labels after init_markers + CCL (before the flood)."""
    s = _sdf(sdf)
    lab = np.zeros(s.size, dtype=np.int32)
    d, sp = _dims(s.shape), _xyz_f32(spacing_zyx)
    lib().pnmo_init_markers(_p(s), _p(d), _p(sp), _p(lab))
    lib().pnmo_ccl(_p(d), _p(lab))
    return lab


def flood(sdf, labels):
    s = _sdf(sdf)
    lab = np.array(labels, dtype=np.int32, copy=True)
    d = _dims(s.shape)
    sweeps = lib().pnmo_flood(_p(s), _p(d), _p(lab))
    return lab, sweeps


def gradient_roots(sdf, all_voxels=False):
    s = _sdf(sdf)
    r = np.zeros(s.size, dtype=np.int32)
    d = _dims(s.shape)
    lib().pnmo_gradient_roots(_p(s), _p(d), int(bool(all_voxels)), _p(r))
    return r


def relabel(sdf, roots, labels):
    s = _sdf(sdf)
    seg = np.zeros(s.size, dtype=np.int32)
    rt = np.ascontiguousarray(roots, dtype=np.int32)
    lb = np.ascontiguousarray(labels, dtype=np.int32)
    lib().pnmo_relabel(_p(s), _p(rt), _p(lb), s.size, _p(seg))
    return seg


def extract_topology(seg, shape_zyx):
    seg = np.ascontiguousarray(seg, dtype=np.int32).ravel()
    if seg.size != int(np.prod(shape_zyx)):
        raise ValueError("segmentation length does not match shape")
    ptr = ctypes.POINTER(ctypes.c_int32)()
    d = _dims(shape_zyx)
    cnt = lib().pnmo_extract_topology(_p(seg), _p(d), ctypes.byref(ptr))
    if cnt == 0:
        if ptr:
            lib().pnmo_free(ptr)
        return np.zeros((0, 2), dtype=np.int32)
    out = np.ctypeslib.as_array(ptr, shape=(2 * cnt,)).copy().reshape(-1, 2)
    lib().pnmo_free(ptr)
    return out


def extract_pore_network(sdf, origin_zyx, spacing_zyx):
    pores, _ = extract_pores(sdf, origin_zyx, spacing_zyx)
    seg = segment_volume(sdf, spacing_zyx)
    conns = extract_topology(seg, np.shape(sdf))
    return pores, seg, conns


def extract_network_flow(sdf, origin_zyx, spacing_zyx, u, v, w, p, ox=None, oy=None, oz=None,
                         grad_p_zyx=(0.0, 0.0, 0.0), debug=False):
    s = _sdf(sdf)
    f = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
         for a in (u, v, w, p, ox, oy, oz)]
    for a in f:
        if a is not None and a.shape != s.shape:
            raise ValueError("field shape must match the SDF")
    gp = np.array([grad_p_zyx[2], grad_p_zyx[1], grad_p_zyx[0]], dtype=np.float64)
    r = _Flow()
    d, o, sp = _dims(s.shape), _xyz_f32(origin_zyx), _xyz_f32(spacing_zyx)
    lib().pnmo_extract_network_flow(_p(s), _p(d), _p(o), _p(sp), *[_p(a) for a in f], _p(gp),
                                    int(bool(debug)), ctypes.byref(r))
    npores, nt, n = r.np, r.nt, s.size

    def arr(ptr, count, dtype):
        if count == 0 or not ptr:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

    out = {
        "pores": (np.ctypeslib.as_array(ctypes.cast(r.pores, ctypes.POINTER(ctypes.c_float)),
                                        shape=(npores * 4,)).copy().view(PORE_DTYPE)
                  if npores else np.zeros(0, dtype=PORE_DTYPE)),
        "pore_pressure": arr(r.pore_pressure, npores, np.float64),
        "pore_residual": arr(r.pore_residual, npores, np.float64),
        "throats": arr(r.throats, 2 * nt, np.int32).reshape(-1, 2),
        "throat_flow": arr(r.throat_flow, nt, np.float64),
        "throat_area": arr(r.throat_area, nt, np.float64),
        "throat_dp": arr(r.throat_dp, nt, np.float64),
    }
    if debug and npores:
        out["flow_lab"] = arr(r.flow_lab, n, np.int32)
        out["fpar"] = arr(r.fpar, 3 * n, np.int64)
        out["peak"] = arr(r.peak, npores, np.int32)
        out["cx"] = arr(r.cx, nt, np.float64)
        out["cy"] = arr(r.cy, nt, np.float64)
        out["cz"] = arr(r.cz, nt, np.float64)
    lib().pnmo_flow_free(ctypes.byref(r))
    return out


# ---- the OpenMP baseline (oracle/pnm_ref_omp.cpp): the reference's extract_pore_network_k
# restated for CPU threads, timed by bench.py's cpu_baseline and checked by tests/ ----
REF_OMP_PATH = os.path.join(HERE, "libpnm_ref_omp.so")
_ref = None


def ref_omp():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_OMP_PATH):
            build()
        L = ctypes.CDLL(REF_OMP_PATH)
        vp = ctypes.c_void_p
        L.pnmr_threads.restype = ctypes.c_int
        L.pnmr_extract_pore_network.restype = ctypes.c_int64
        L.pnmr_extract_pore_network.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64, vp,
                                                ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)),
                                                ctypes.POINTER(ctypes.c_int64), vp]
        L.pnmr_free.argtypes = [vp]
        _ref = L
    return _ref


def ref_omp_extract_pore_network(sdf, origin_zyx, spacing_zyx):
    """(pores in atomic-slot order, seg, connections, stage_ms [upload, pores, segment,
    topology, download]) from the OpenMP baseline; threads from OMP_NUM_THREADS."""
    s = _sdf(sdf)
    d, o, sp = _dims(s.shape), _xyz_f32(origin_zyx), _xyz_f32(spacing_zyx)
    cap = 1000000
    pores = np.zeros(4 * cap, dtype=np.float32)
    seg = np.zeros(s.size, dtype=np.int32)
    ptr = ctypes.POINTER(ctypes.c_int32)()
    npairs = ctypes.c_int64()
    ms = np.zeros(5, dtype=np.float64)
    npores = ref_omp().pnmr_extract_pore_network(_p(s), _p(d), _p(o), _p(sp), _p(pores), cap,
                                                 _p(seg), ctypes.byref(ptr),
                                                 ctypes.byref(npairs), _p(ms))
    conns = (np.ctypeslib.as_array(ptr, shape=(2 * npairs.value,)).reshape(-1, 2).copy()
             if npairs.value else np.zeros((0, 2), dtype=np.int32))
    ref_omp().pnmr_free(ptr)
    p = pores[: 4 * min(npores, cap)].copy().view(PORE_DTYPE)
    return p, seg, conns, ms
