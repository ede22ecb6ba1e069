"""ctypes view of the C-ABI of libpnm_hip.so (declared in include/pnm_capi.h).

Used for device-resident calls (torch tensors already in HBM) by the benchmark and the GPU
parity tests; the `peclet.pnm` functions themselves go through the compiled `_pnm` module.
Every call raises RuntimeError with the library's message on a non-zero status. There is no
fallback: a missing library or GPU is an error.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpnm_hip.so")

PNM_OK, PNM_EINVAL, PNM_ENOMEM, PNM_EHIP, PNM_ECAPACITY, PNM_ENOCONVERGE = range(6)
PNM_MAX_PORES = 1000000

# Every exported symbol include/pnm_capi.h declares.
EXPORTS = (
    "pnm_version", "pnm_execution_space", "pnm_last_error", "pnm_device_count",
    "pnm_ctx_create", "pnm_ctx_destroy", "pnm_ctx_set_stream", "pnm_ctx_stream", "pnm_ctx_trim",
    "pnm_ctx_arena_bytes", "pnm_ctx_profile", "pnm_ctx_profile_only", "pnm_ctx_profile_read",
    "pnm_free", "pnm_extract_pores", "pnm_segment_volume",
    "pnm_extract_topology", "pnm_extract_pore_network", "pnm_extract_network_flow",
    "pnm_flow_result_free", "pnm_extract_pores_d", "pnm_segment_volume_d",
    "pnm_extract_topology_d", "pnm_extract_pore_network_d", "pnm_extract_network_flow_d",
    "pnm_stage_labels_d", "pnm_stage_roots_d", "pnm_segment_batch_d", "pnm_batch_config",
    "pnm_stokes_solve_d", "pnm_sur_param_count", "pnm_sur_init", "pnm_sur_forward_d",
    "pnm_sur_backward_d", "pnm_sur_loss_d", "pnm_sur_adam_d", "pnm_sur_patches_d",
    "pnm_segment_volume_noalloc_d", "pnm_sur_infer_d", "pnm_throat_geometry_d",
)

# Largest tile-padded volume the no-allocation entry point accepts (PNM_NOALLOC_MAX_PADDED).
PNM_NOALLOC_MAX_PADDED = 1 << 21


def noalloc_padded_voxels(dims_xyz):
    """Voxels after rounding X, Y, Z up to the 64 x 8 x 16 tile (PNM_NOALLOC_PADDED_VOXELS)."""
    x, y, z = (int(a) for a in dims_xyz)
    return (x + 63) // 64 * 64 * ((y + 7) // 8 * 8) * ((z + 15) // 16 * 16)


PNM_BATCH_GRAPH = 1
PNM_BATCH_AUTOTUNE = 2


class Pore(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float),
                ("radius", ctypes.c_float)]


PORE_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("radius", "<f4")])


class FlowResult(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int64), ("nt", ctypes.c_int64),
                ("pores", ctypes.POINTER(Pore)),
                ("pore_pressure", ctypes.POINTER(ctypes.c_double)),
                ("pore_residual", ctypes.POINTER(ctypes.c_double)),
                ("throats", ctypes.POINTER(ctypes.c_int32)),
                ("throat_flow", ctypes.POINTER(ctypes.c_double)),
                ("throat_area", ctypes.POINTER(ctypes.c_double)),
                ("throat_dp", ctypes.POINTER(ctypes.c_double)),
                ("throat_center", ctypes.POINTER(ctypes.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `python peclet-pnm_amd/build.py`")
        from ._runtime import pin_hip_runtime

        pin_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
        st = ctypes.c_int
        L.pnm_version.restype = ctypes.c_char_p
        L.pnm_execution_space.restype = ctypes.c_char_p
        L.pnm_last_error.restype = ctypes.c_char_p
        L.pnm_device_count.restype = ctypes.c_int
        L.pnm_ctx_create.restype = st
        L.pnm_ctx_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.pnm_ctx_destroy.argtypes = [vp]
        L.pnm_ctx_destroy.restype = None
        L.pnm_ctx_set_stream.argtypes = [vp, vp]
        L.pnm_ctx_set_stream.restype = st
        L.pnm_ctx_stream.argtypes = [vp]
        L.pnm_ctx_stream.restype = vp
        L.pnm_ctx_trim.argtypes = [vp]
        L.pnm_ctx_trim.restype = st
        L.pnm_ctx_arena_bytes.argtypes = [vp]
        L.pnm_ctx_arena_bytes.restype = ctypes.c_int64
        L.pnm_ctx_profile.argtypes = [vp, ctypes.c_int]
        L.pnm_ctx_profile.restype = st
        L.pnm_ctx_profile_only.argtypes = [vp, ctypes.c_char_p]
        L.pnm_ctx_profile_only.restype = st
        L.pnm_ctx_profile_read.argtypes = [vp, ctypes.c_char_p, ctypes.c_int64,
                                           ctypes.POINTER(ctypes.c_double), i64p, ctypes.c_int64,
                                           i64p]
        L.pnm_ctx_profile_read.restype = st
        L.pnm_free.argtypes = [vp]
        L.pnm_free.restype = None
        L.pnm_extract_pores_d.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int64, i64p, i64p]
        L.pnm_segment_volume_d.argtypes = [vp, vp, vp, vp, vp, vp]
        L.pnm_segment_volume_noalloc_d.argtypes = [vp, vp, vp, vp, vp, vp]
        L.pnm_extract_topology_d.argtypes = [vp, vp, vp, vp, ctypes.c_int64, i64p]
        L.pnm_extract_pore_network_d.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int64, i64p,
                                                 i64p, vp, vp, ctypes.c_int64, i64p]
        L.pnm_extract_network_flow_d.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                                 vp, ctypes.POINTER(FlowResult)]
        L.pnm_extract_network_flow.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                               vp, ctypes.POINTER(FlowResult)]
        L.pnm_throat_geometry_d.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64,
                                            ctypes.POINTER(FlowResult)]
        L.pnm_throat_geometry_d.restype = st
        L.pnm_flow_result_free.argtypes = [ctypes.POINTER(FlowResult)]
        L.pnm_flow_result_free.restype = None
        L.pnm_segment_batch_d.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp, vp, vp,
                                          ctypes.c_int32]
        L.pnm_segment_batch_d.restype = st
        L.pnm_batch_config.argtypes = [vp, vp, vp, vp, vp]
        L.pnm_batch_config.restype = st
        L.pnm_stokes_solve_d.argtypes = [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.pnm_stokes_solve_d.restype = st
        L.pnm_sur_param_count.restype = ctypes.c_int64
        L.pnm_sur_param_count.argtypes = []
        L.pnm_sur_init.argtypes = [ctypes.c_uint64, vp, ctypes.c_int64]
        L.pnm_sur_init.restype = st
        L.pnm_sur_forward_d.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp]
        L.pnm_sur_forward_d.restype = st
        L.pnm_sur_backward_d.argtypes = [vp, ctypes.c_int32, vp, vp, vp]
        L.pnm_sur_backward_d.restype = st
        L.pnm_sur_loss_d.argtypes = [vp, ctypes.c_int32, vp, vp, ctypes.c_float, vp, vp]
        L.pnm_sur_loss_d.restype = st
        L.pnm_sur_adam_d.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.pnm_sur_adam_d.restype = st
        L.pnm_sur_patches_d.argtypes = [vp, vp, vp, vp, ctypes.c_int32, ctypes.c_float, vp]
        L.pnm_sur_patches_d.restype = st
        L.pnm_sur_infer_d.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_float, vp, vp]
        L.pnm_sur_infer_d.restype = st
        L.pnm_stage_labels_d.argtypes = [vp, vp, vp, vp, vp, vp]
        L.pnm_stage_roots_d.argtypes = [vp, vp, vp, ctypes.c_int, vp]
        for name in ("pnm_extract_pores_d", "pnm_segment_volume_d",
                     "pnm_segment_volume_noalloc_d", "pnm_extract_topology_d",
                     "pnm_extract_pore_network_d", "pnm_extract_network_flow_d",
                     "pnm_extract_network_flow", "pnm_stage_labels_d", "pnm_stage_roots_d"):
            getattr(L, name).restype = st
        _lib = L
    return _lib


def _check(status):
    if status != PNM_OK:
        msg = lib().pnm_last_error().decode()
        raise RuntimeError(f"pnm status {status}: {msg}")


def _f3(v):
    return (ctypes.c_float * 3)(*[float(a) for a in v])


def _i3(v):
    return (ctypes.c_int32 * 3)(*[int(a) for a in v])


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Context:
    """A libpnm_hip context (own HIP stream + scratch arena) on one device."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().pnm_ctx_create(ctypes.byref(h), device))
        self.h = h

    def close(self):
        if self.h:
            lib().pnm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        _check(lib().pnm_ctx_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)))

    @property
    def stream(self):
        return lib().pnm_ctx_stream(self.h)

    def arena_bytes(self):
        return lib().pnm_ctx_arena_bytes(self.h)

    def profile(self, enable=True):
        """Bracket every kernel launch with HIP events on the launch stream."""
        _check(lib().pnm_ctx_profile(self.h, int(bool(enable))))

    def profile_only(self, names=None):
        """Time only the launches of these kernel names (None: every launch)."""
        arg = "\n".join(names).encode() if names else None
        _check(lib().pnm_ctx_profile_only(self.h, arg))

    def profile_read(self):
        """{kernel name: (total ms, launches)} since the last read (synchronises)."""
        cap = 128
        names = ctypes.create_string_buffer(1 << 14)
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_int64 * cap)()
        k = ctypes.c_int64()
        _check(lib().pnm_ctx_profile_read(self.h, names, len(names), ms, cnt, cap,
                                          ctypes.byref(k)))
        keys = names.value.decode().split("\n")[: k.value]
        return {name: (ms[i], cnt[i]) for i, name in enumerate(keys)}

    # ---- device-pointer entry points (torch tensors) ----
    def extract_pores_d(self, sdf, dims_xyz, origin_xyz, spacing_xyz, pores_out):
        """pores_out: torch uint8/float tensor with room for cap records (16 B each)."""
        cap = pores_out.numel() * pores_out.element_size() // 16
        cnt, tot = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().pnm_extract_pores_d(self.h, _ptr(sdf), _i3(dims_xyz), _f3(origin_xyz),
                                         _f3(spacing_xyz), _ptr(pores_out), cap,
                                         ctypes.byref(cnt), ctypes.byref(tot)))
        return cnt.value, tot.value

    def segment_volume_d(self, sdf, dims_xyz, spacing_xyz, seg_out, roots_out=None):
        _check(lib().pnm_segment_volume_d(self.h, _ptr(sdf), _i3(dims_xyz), _f3(spacing_xyz),
                                          _ptr(seg_out), _ptr(roots_out)))

    def segment_volume_noalloc_d(self, sdf, dims_xyz, spacing_xyz, seg_out, roots_out=None):
        """segment_volume_d in the library's fixed per-GPU scratch (no device allocation)."""
        _check(lib().pnm_segment_volume_noalloc_d(self.h, _ptr(sdf), _i3(dims_xyz),
                                                  _f3(spacing_xyz), _ptr(seg_out),
                                                  _ptr(roots_out)))

    def extract_topology_d(self, seg, dims_xyz, pairs_out):
        cap = pairs_out.numel() // 2
        cnt = ctypes.c_int64()
        _check(lib().pnm_extract_topology_d(self.h, _ptr(seg), _i3(dims_xyz), _ptr(pairs_out),
                                            cap, ctypes.byref(cnt)))
        return cnt.value

    def extract_pore_network_d(self, sdf, dims_xyz, origin_xyz, spacing_xyz, pores_out, seg_out,
                               pairs_out):
        cap_p = pores_out.numel() * pores_out.element_size() // 16
        cap_q = pairs_out.numel() // 2
        npores, tot, npairs = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().pnm_extract_pore_network_d(
            self.h, _ptr(sdf), _i3(dims_xyz), _f3(origin_xyz), _f3(spacing_xyz), _ptr(pores_out),
            cap_p, ctypes.byref(npores), ctypes.byref(tot), _ptr(seg_out), _ptr(pairs_out), cap_q,
            ctypes.byref(npairs)))
        return npores.value, tot.value, npairs.value

    def segment_batch_d(self, sdfs, segs, dims_xyz, spacing_xyz, n_pores=None, n_solids=None,
                        graph=True, autotune=False):
        """Segment many device volumes in one persistent launch. sdfs/segs: lists of torch
        tensors (float32 / int32, contiguous); dims_xyz: list of (nx, ny, nz)."""
        nvol = len(sdfs)
        ps = (ctypes.c_void_p * max(nvol, 1))(*[t.data_ptr() for t in sdfs])
        qs = (ctypes.c_void_p * max(nvol, 1))(*[t.data_ptr() for t in segs])
        d = np.ascontiguousarray(np.asarray(dims_xyz, dtype=np.int32).reshape(-1))
        flags = (PNM_BATCH_GRAPH if graph else 0) | (PNM_BATCH_AUTOTUNE if autotune else 0)
        _check(lib().pnm_segment_batch_d(self.h, nvol, ps, qs, ctypes.c_void_p(d.ctypes.data),
                                         _f3(spacing_xyz), _ptr(n_pores), _ptr(n_solids),
                                         flags))

    def batch_config(self):
        b, pc, g = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ms = ctypes.c_float()
        _check(lib().pnm_batch_config(self.h, ctypes.byref(b), ctypes.byref(pc), ctypes.byref(g),
                                      ctypes.byref(ms)))
        return {"block": b.value, "per_cu": pc.value, "groups": g.value, "tuned_ms": ms.value}

    def stage_labels_d(self, sdf, dims_xyz, spacing_xyz, labels_out, ccl_out=None):
        _check(lib().pnm_stage_labels_d(self.h, _ptr(sdf), _i3(dims_xyz), _f3(spacing_xyz),
                                        _ptr(labels_out), _ptr(ccl_out)))

    def stage_roots_d(self, sdf, dims_xyz, all_voxels, roots_out):
        _check(lib().pnm_stage_roots_d(self.h, _ptr(sdf), _i3(dims_xyz), int(bool(all_voxels)),
                                       _ptr(roots_out)))

    def extract_network_flow_d(self, sdf, dims_xyz, origin_xyz, spacing_xyz, u, v, w, p,
                               ox=None, oy=None, oz=None, grad_p_xyz=(0.0, 0.0, 0.0)):
        r = FlowResult()
        gp = (ctypes.c_double * 3)(*[float(a) for a in grad_p_xyz])
        _check(lib().pnm_extract_network_flow_d(
            self.h, _ptr(sdf), _i3(dims_xyz), _f3(origin_xyz), _f3(spacing_xyz), _ptr(u), _ptr(v),
            _ptr(w), _ptr(p), _ptr(ox), _ptr(oy), _ptr(oz), gp, ctypes.byref(r)))
        return _flow_to_numpy(r)

    def throat_geometry_d(self, sdf, dims_xyz, origin_xyz, spacing_xyz, seg=None, centers=None,
                          raw=None):
        """Pores, throats, throat_area and throat_center of extract_network_flow under binary
        openness, without fields. seg: optional device int32 segmentation of this SDF.
        centers / raw: optional device float32 [cap, 3] / [cap, 4] receiving the surrogate
        inputs when the throat count fits."""
        r = FlowResult()
        cap = 0
        if centers is not None or raw is not None:
            if centers is None or raw is None:
                raise ValueError("pass both centers and raw, or neither")
            cap = min(centers.numel() // 3, raw.numel() // 4)
        _check(lib().pnm_throat_geometry_d(
            self.h, _ptr(sdf), _ptr(seg), _i3(dims_xyz), _f3(origin_xyz), _f3(spacing_xyz),
            _ptr(centers), _ptr(raw), ctypes.c_int64(cap), ctypes.byref(r)))
        out = _flow_to_numpy(r)
        for k in ("pore_pressure", "pore_residual", "throat_flow", "throat_dp"):
            del out[k]
        return out


    def stokes_solve_d(self, sdf, dims_xyz, spacing_xyz, force_xyz, u, v, w, p, ox=None, oy=None,
                       oz=None, mu=1.0, rtol=1e-10, max_iter=20000):
        """Project MAC Stokes solve into device float64 tensors u, v, w, p (and the binary
        openness into ox, oy, oz when given). Returns (iterations, relative residual)."""
        it = ctypes.c_int32()
        rr = ctypes.c_double()
        f = (ctypes.c_double * 3)(*[float(a) for a in force_xyz])
        _check(lib().pnm_stokes_solve_d(
            self.h, _ptr(sdf), _i3(dims_xyz), _f3(spacing_xyz), f, ctypes.c_double(mu),
            ctypes.c_double(rtol), ctypes.c_int32(max_iter), _ptr(u), _ptr(v), _ptr(w), _ptr(p),
            _ptr(ox), _ptr(oy), _ptr(oz), ctypes.byref(it), ctypes.byref(rr)))
        return it.value, rr.value


    # ---- throat surrogate ----
    def sur_forward_d(self, batch, patches, scalars, params, out):
        _check(lib().pnm_sur_forward_d(self.h, int(batch), _ptr(patches), _ptr(scalars),
                                       _ptr(params), _ptr(out)))

    def sur_backward_d(self, batch, dout, params, grads):
        _check(lib().pnm_sur_backward_d(self.h, int(batch), _ptr(dout), _ptr(params),
                                        _ptr(grads)))

    def sur_loss_d(self, batch, pred, target, dpred, loss=None, delta=0.0):
        _check(lib().pnm_sur_loss_d(self.h, int(batch), _ptr(pred), _ptr(target),
                                    float(delta), _ptr(dpred), _ptr(loss)))

    def sur_adam_d(self, params, grads, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999,
                   eps=1e-8):
        _check(lib().pnm_sur_adam_d(self.h, _ptr(params), _ptr(grads), _ptr(m), _ptr(v),
                                    int(step), lr, beta1, beta2, eps))

    def sur_patches_d(self, sdf, dims_xyz, centers, batch, scale, out):
        _check(lib().pnm_sur_patches_d(self.h, _ptr(sdf), _i3(dims_xyz), _ptr(centers),
                                       int(batch), float(scale), _ptr(out)))

    def sur_infer_d(self, sdf, dims_xyz, centers, scalars, batch, scale, params, out):
        """Fused surrogate inference from the device SDF (no patches, no activations in HBM)."""
        _check(lib().pnm_sur_infer_d(self.h, _ptr(sdf), _i3(dims_xyz), _ptr(centers),
                                     _ptr(scalars), int(batch), float(scale), _ptr(params),
                                     _ptr(out)))


def sur_param_count():
    return int(lib().pnm_sur_param_count())


def sur_init(seed):
    """Initial surrogate parameters (host float32 array)."""
    n = sur_param_count()
    out = np.zeros(n, dtype=np.float32)
    _check(lib().pnm_sur_init(ctypes.c_uint64(seed), out.ctypes.data_as(ctypes.c_void_p), n))
    return out


def _flow_to_numpy(r):
    npores, nt = r.np, r.nt

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
        "throat_center": arr(r.throat_center, 3 * nt, np.float64).reshape(-1, 3),
    }
    lib().pnm_flow_result_free(ctypes.byref(r))
    return out
