"""C-ABI library and Python binding checks that need no GPU: exports, argument validation and
the reference's RuntimeError messages (pnm_bindings.cpp:58-59, 67-69, 173-174)."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from peclet.pnm import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pnm_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pnm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(capi.EXPORTS)


def test_identity_strings():
    L = capi.lib()
    assert L.pnm_execution_space() == b"HIP"
    assert L.pnm_version()


def test_validation_before_device_work():
    L = capi.lib()
    dims = (ctypes.c_int32 * 3)(4, 4, 4)
    f3 = (ctypes.c_float * 3)(0, 0, 0)
    cnt, tot = ctypes.c_int64(), ctypes.c_int64()
    # NULL context
    st = L.pnm_extract_pores_d(None, ctypes.c_void_p(1), dims, f3, f3, ctypes.c_void_p(1), 10,
                               ctypes.byref(cnt), ctypes.byref(tot))
    assert st == capi.PNM_EINVAL
    assert b"context" in L.pnm_last_error()
    # negative dims
    bad = (ctypes.c_int32 * 3)(4, -1, 4)
    st = L.pnm_extract_pores_d(None, ctypes.c_void_p(1), bad, f3, f3, ctypes.c_void_p(1), 10,
                               ctypes.byref(cnt), ctypes.byref(tot))
    assert st == capi.PNM_EINVAL and b"non-negative" in L.pnm_last_error()
    # n >= 2^31
    big = (ctypes.c_int32 * 3)(2048, 1024, 1024)
    st = L.pnm_segment_volume_d(None, ctypes.c_void_p(1), big, f3, ctypes.c_void_p(1), None)
    assert st == capi.PNM_EINVAL and b"2^31" in L.pnm_last_error()
    # NULL sdf for a non-empty volume
    st = L.pnm_segment_volume_d(None, None, dims, f3, ctypes.c_void_p(1), None)
    assert st == capi.PNM_EINVAL
    # openness arity
    r = capi.FlowResult()
    gp = (ctypes.c_double * 3)(0, 0, 0)
    one = ctypes.c_void_p(1)
    st = L.pnm_extract_network_flow_d(None, one, dims, f3, f3, one, one, one, one, one, None,
                                      None, gp, ctypes.byref(r))
    assert st == capi.PNM_EINVAL and b"openness" in L.pnm_last_error()


def test_context_without_gpu_fails_loudly():
    if capi.lib().pnm_device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    st = capi.lib().pnm_ctx_create(ctypes.byref(h), 0)
    assert st == capi.PNM_EHIP
    assert b"no HIP device" in capi.lib().pnm_last_error()


def test_binding_errors_match_reference():
    import peclet.pnm as pnm

    assert pnm.execution_space == "HIP"
    with pytest.raises(RuntimeError, match=r"SDF array must be 3D \(Nz,Ny,Nx\)"):
        pnm.extract_pores(np.zeros((4, 4), np.float32), [0, 0, 0], [1, 1, 1])
    with pytest.raises(RuntimeError, match=r"SDF array must be 3D"):
        pnm.segment_volume(np.zeros((2, 2, 2, 2), np.float32), [1, 1, 1])
    sdf = np.zeros((3, 4, 5), np.float32)
    f = np.zeros((3, 4, 5))
    with pytest.raises(RuntimeError, match=r"pass all three openness arrays \(ox,oy,oz\) or none"):
        pnm.extract_network_flow(sdf, [0, 0, 0], [1, 1, 1], f, f, f, f, ox=f)
    with pytest.raises(RuntimeError, match=r"u must be \(Nz,Ny,Nx\) matching the SDF"):
        pnm.extract_network_flow(sdf, [0, 0, 0], [1, 1, 1], np.zeros((5, 4, 3)), f, f, f)
    with pytest.raises(RuntimeError, match=r"oz must be \(Nz,Ny,Nx\) matching the SDF"):
        pnm.extract_network_flow(sdf, [0, 0, 0], [1, 1, 1], f, f, f, f, ox=f, oy=f,
                                 oz=np.zeros(3))
    with pytest.raises(RuntimeError, match="does not match prod"):
        pnm.extract_topology_gpu(np.zeros(10, np.int32), [2, 2, 2])
    with pytest.raises(RuntimeError, match="3 entries"):
        pnm.extract_pores(sdf, [0, 0], [1, 1, 1])


def test_empty_volume_returns_empty_without_device():
    import peclet.pnm as pnm

    e = np.zeros((0, 4, 4), np.float32)
    assert pnm.extract_pores(e, [0, 0, 0], [1, 1, 1]) == []
    assert len(pnm.segment_volume(e, [1, 1, 1])) == 0
    pores, seg, conns = pnm.extract_pore_network(e, [0, 0, 0], [1, 1, 1])
    assert pores == [] and len(seg) == 0 and conns == []
    assert pnm.extract_topology_gpu([], [0, 3, 3]) == []
    z = np.zeros((0, 4, 4))
    d = pnm.extract_network_flow(e, [0, 0, 0], [1, 1, 1], z, z, z, z)
    assert d["pores"] == [] and d["throats"] == []
    assert set(d) == {"pores", "pore_pressure", "pore_residual", "throats", "throat_flow",
                      "throat_area", "throat_dp"}


def test_pore_type():
    import peclet.pnm as pnm

    p = pnm.Pore()
    p.x, p.y, p.z, p.radius = 1.0, 2.0, 3.0, 4.5
    assert (p.x, p.y, p.z, p.radius) == (1.0, 2.0, 3.0, 4.5)
    assert ctypes.sizeof(capi.Pore) == 16


def test_single_hip_runtime_in_process():
    """libpnm_hip.so and torch must share one HIP runtime: two runtimes in one process cannot
    both bring up the GPU."""
    capi.lib()
    import torch  # noqa: F401

    libs = {line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line}
    assert len(libs) == 1, libs
