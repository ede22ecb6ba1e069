"""Throat geometry without fields (pnm_throat_geometry_d) against extract_network_flow under the
project solver's binary openness (a face is open iff both of its cells have sdf > 0).

Pores, throats, throat_area and throat_center must equal the flow path's bit for bit, whether the
segmentation is computed inside the call, given from segment_volume_d or given from one volume
of a batched segmentation; the throat list and areas are also checked against the oracle. The
surrogate inputs the call writes on the device must equal the host statement
(surrogate.throat_inputs) to within one FP32 ulp (FP64 log on the device vs numpy's).
"""
from __future__ import annotations

import numpy as np
import pytest

import flowcases as fc
import pnm_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def xyz(v):
    return (v[2], v[1], v[0])


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def binary_openness(sdf):
    """ox, oy, oz as the solver writes them: o_d[c] = 1 iff c and its -d neighbour are fluid."""
    fl = sdf > 0
    return [(fl & np.roll(fl, 1, axis=a)).astype(np.float64) for a in (2, 1, 0)]


def cases():
    from peclet.pnm import synth

    return {
        "chain": (fc.chain_x(3, 24, 7.0, 3.0), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
        "two_patches": (fc.two_chambers_two_tubes(), (0.5, 1.0, -2.0), (1.5, 1.0, 0.5)),
        "side_by_side": (fc.two_chambers_side_by_side_tubes(), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
        "lattice48": (synth.chamber_sdf((48, 48, 48), 2, 7.0, 2.0, 4.0, seed=5),
                      (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
        "packing": (synth.packing_sdf((40, 44, 48), "uniform", 3.0, 6.0, 0.5, seed=101)[0],
                    (2.0, -1.0, 0.5), (1.5, 1.0, 0.5)),
    }


CASES = ["chain", "two_patches", "side_by_side", "lattice48", "packing"]


def flow_binary(ctx, sdf, org, spc, seed=0):
    rng = np.random.default_rng(seed)
    u, v, w, p = (rng.standard_normal(sdf.shape) for _ in range(4))
    ox, oy, oz = binary_openness(sdf)
    g = ctx.extract_network_flow_d(dev(sdf), xyz(sdf.shape), xyz(org), xyz(spc), dev(u), dev(v),
                                   dev(w), dev(p), dev(ox), dev(oy), dev(oz), (0.1, 0.2, 0.3))
    o = orc.extract_network_flow(sdf, org, spc, u, v, w, p, ox, oy, oz, (0.3, 0.2, 0.1))
    return g, o


def same_geometry(a, b):
    np.testing.assert_array_equal(np.ascontiguousarray(a["pores"]).view(np.uint32),
                                  np.ascontiguousarray(b["pores"]).view(np.uint32))
    np.testing.assert_array_equal(a["throats"], b["throats"])
    for k in ("throat_area", "throat_center"):
        np.testing.assert_array_equal(np.asarray(a[k]).view(np.uint64),
                                      np.asarray(b[k]).view(np.uint64), err_msg=k)


@pytest.mark.parametrize("case", CASES)
def test_geometry_equals_flow_under_binary_openness(capi_ctx, case):
    sdf, org, spc = cases()[case]
    g, o = flow_binary(capi_ctx, sdf, org, spc)
    assert len(g["throats"]) > 0
    geo = capi_ctx.throat_geometry_d(dev(sdf), xyz(sdf.shape), xyz(org), xyz(spc))
    same_geometry(geo, g)
    np.testing.assert_array_equal(geo["throats"], o["throats"])
    np.testing.assert_array_equal(geo["throat_area"].view(np.uint64),
                                  o["throat_area"].view(np.uint64))
    assert set(geo) == {"pores", "throats", "throat_area", "throat_center"}


@pytest.mark.parametrize("case", CASES)
def test_geometry_from_given_segmentation(capi_ctx, case):
    sdf, org, spc = cases()[case]
    s = dev(sdf)
    seg = torch.empty(sdf.size, dtype=torch.int32, device="cuda")
    capi_ctx.segment_volume_d(s, xyz(sdf.shape), xyz(spc), seg)
    a = capi_ctx.throat_geometry_d(s, xyz(sdf.shape), xyz(org), xyz(spc))
    b = capi_ctx.throat_geometry_d(s, xyz(sdf.shape), xyz(org), xyz(spc), seg=seg)
    same_geometry(a, b)


def test_geometry_from_batched_segmentation(capi_ctx):
    """One batched segmentation of several volumes, then each volume's geometry from its slice
    of the batch: equal to the geometry computed from scratch."""
    from peclet.pnm import synth

    sdfs = [synth.packing_sdf((40 + 8 * k, 44, 48 - 4 * k), "uniform", 3.0, 6.0, 0.5,
                              seed=200 + k)[0] for k in range(4)]
    sdfs.append(synth.chamber_sdf((48, 48, 48), 2, 7.0, 2.0, 4.0, seed=7))
    vols = [dev(a).view(-1) for a in sdfs]
    segs = [torch.empty(a.size, dtype=torch.int32, device="cuda") for a in sdfs]
    capi_ctx.segment_batch_d(vols, segs, [xyz(a.shape) for a in sdfs], (1.0, 1.0, 1.0),
                             graph=False)
    for a, v, sg in zip(sdfs, vols, segs):
        r0 = capi_ctx.throat_geometry_d(v, xyz(a.shape), (0, 0, 0), (1, 1, 1))
        r1 = capi_ctx.throat_geometry_d(v, xyz(a.shape), (0, 0, 0), (1, 1, 1), seg=sg)
        assert len(r0["throats"]) > 0
        same_geometry(r0, r1)


@pytest.mark.parametrize("case", ["lattice48", "packing"])
def test_device_features_match_host(capi_ctx, case):
    from peclet.pnm import surrogate as S

    sdf, _, _ = cases()[case]
    s = dev(sdf)
    geo, cen, raw = S.throat_geometry(capi_ctx, s, sdf.shape)
    hc, hr = S.throat_inputs(geo, sdf.shape)
    assert cen.shape == hc.shape and raw.shape == hr.shape and len(hc) > 0
    for d, h in ((cen.cpu().numpy(), hc), (raw.cpu().numpy(), hr)):
        ulp = np.abs(d.view(np.int32).astype(np.int64) - h.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1, ulp.max()


def test_features_not_written_beyond_capacity(capi_ctx):
    """A feature buffer smaller than the throat count is left untouched (the host falls back to
    the returned arrays)."""
    sdf, org, spc = cases()["lattice48"]
    cen = torch.full((3,), -5.0, device="cuda")
    raw = torch.full((4,), -5.0, device="cuda")
    geo = capi_ctx.throat_geometry_d(dev(sdf), xyz(sdf.shape), xyz(org), xyz(spc), centers=cen,
                                     raw=raw)
    assert len(geo["throats"]) > 1
    assert bool((cen == -5.0).all()) and bool((raw == -5.0).all())


def test_empty_inputs(capi_ctx):
    """All-solid volume, and a given segmentation without pore labels: no pores, no throats."""
    shape = (16, 16, 16)
    solid = np.full(shape, -1.0, np.float32)
    r = capi_ctx.throat_geometry_d(dev(solid), xyz(shape), (0, 0, 0), (1, 1, 1))
    assert len(r["pores"]) == 0 and len(r["throats"]) == 0
    sdf, _, _ = cases()["lattice48"]
    seg = torch.full((sdf.size,), -1, dtype=torch.int32, device="cuda")
    r = capi_ctx.throat_geometry_d(dev(sdf), xyz(sdf.shape), (0, 0, 0), (1, 1, 1), seg=seg)
    assert len(r["pores"]) == 0 and len(r["throats"]) == 0
