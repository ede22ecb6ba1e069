"""Floods that need many Jacobi sweeps (pore_extraction.hpp:237-269 loops until a sweep changes
nothing): the HIP path must run them to the end and match the oracle bit for bit.

* thin plates (tests/deepflood.py): 219 and 59 sweeps, oracle computed live;
* the reference's trig field with its spacing at 256^3 and 512^3 (37 and about 74 sweeps),
  against the digests of tests/golden/deep_flood.json (written by
  tests/golden/make_deep_flood.py from the oracle).
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import deepflood as dfl
import pnm_oracle as orc
from test_gpu_plain import check_against_oracle, xyz

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("kind", ["plate", "two_grains"])
def test_plate_flood_to_convergence(capi_ctx, kind):
    sdf = dfl.plate_sdf() if kind == "plate" else dfl.two_grain_plate_sdf()
    spc = (1.0, 1.0, 1.0)
    _, sweeps = orc.flood(sdf, orc.markers_ccl(sdf, spc))
    assert sweeps > 54  # more sweeps than the old fixed counter block allowed
    print(f"[{kind}] shape {sdf.shape}, oracle flood sweeps {sweeps}", flush=True)
    check_against_oracle(capi_ctx, sdf, (0.0, 0.0, 0.0), spc)
    # staged path: segment_volume alone and the flood labels
    n = sdf.size
    s = torch.from_numpy(sdf).cuda()
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    capi_ctx.segment_volume_d(s, xyz(sdf.shape), xyz(spc), seg)
    np.testing.assert_array_equal(seg.cpu().numpy(), orc.segment_volume(sdf, spc))
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    capi_ctx.stage_labels_d(s, xyz(sdf.shape), xyz(spc), lab)
    np.testing.assert_array_equal(lab.cpu().numpy(),
                                  orc.flood(sdf, orc.markers_ccl(sdf, spc))[0])


@pytest.mark.parametrize("name", ["trig256", "trig512"])
def test_trig_field_deep_flood(capi_ctx, name):
    with open(os.path.join(HERE, "golden", "deep_flood.json")) as f:
        exp = json.load(f)[name]
    from peclet.pnm import synth

    shape = tuple(exp["shape"])
    org, spc = tuple(exp["origin_zyx"]), tuple(exp["spacing_zyx"])
    print(f"[{name}] oracle flood sweeps {exp['flood_sweeps']}", flush=True)
    sdf = np.ascontiguousarray(synth.trig_sdf(shape), dtype=np.float32)
    assert digest(sdf) == exp["sdf"], "SDF generator output differs from the fixture's input"
    n = sdf.size
    s = torch.from_numpy(sdf).cuda()
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    pores = torch.empty(4 * 1_000_000, dtype=torch.float32, device="cuda")
    pairs = torch.empty(2 * 3 * n, dtype=torch.int32, device="cuda")
    npores, _, npairs = capi_ctx.extract_pore_network_d(s, xyz(shape), xyz(org), xyz(spc), pores,
                                                        seg, pairs)
    assert npores == exp["npores"]
    assert digest(pores[: 4 * npores].view(-1, 4).cpu().numpy()) == exp["pores"]
    segh = seg.cpu().numpy()
    assert int(segh.max()) == exp["n_pore_labels"]
    assert int(-segh.min()) == exp["n_solid_labels"]
    assert digest(segh) == exp["seg"]
    assert npairs == exp["nconn"]
    assert digest(pairs[: 2 * npairs].view(-1, 2).cpu().numpy()) == exp["conns"]
    # the staged segment_volume runs the same flood
    seg2 = torch.empty(n, dtype=torch.int32, device="cuda")
    capi_ctx.segment_volume_d(s, xyz(shape), xyz(spc), seg2)
    assert digest(seg2.cpu().numpy()) == exp["seg"]
