"""The batched-segmentation configuration at full size: 4096 periodic cubic crops with edges
U[48, 160] of the 512^3 polydisperse grain pack (5.97 Gvox), cut on the GPU and segmented in one
persistent launch (replayed from its hipGraph, and with the autotuner), against the serial
oracle's digests of a fixed sample that includes the ten largest crops
(tests/golden/batch_config.json, written by tests/golden/make_batch_config.py); per-crop label
counts of the whole batch against per-volume segment_volume calls on a further sample."""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def batch():
    import make_batch_config as mbc
    from peclet.pnm import synth

    with open(os.path.join(HERE, "golden", "batch_config.json")) as f:
        exp = json.load(f)
    src = synth.polydisperse_sdf((512, 512, 512), 9.0, 0.3, 0.35, seed=2001)[0]
    assert digest(src) == exp["source_sdf"], "SDF generator output differs from the fixture's"
    edges, offs = mbc.crop_specs(src.shape)
    dev = torch.device("cuda")
    s = torch.from_numpy(src).to(dev)
    vols, segs, dims = [], [], []
    for e, o in zip(edges, offs):
        idx = [(torch.arange(int(e), device=dev) + int(o[a])) % src.shape[a] for a in range(3)]
        v = s.index_select(0, idx[0]).index_select(1, idx[1]).index_select(2, idx[2])
        vols.append(v.contiguous().view(-1))
        segs.append(torch.empty(int(e) ** 3, dtype=torch.int32, device=dev))
        dims.append((int(e),) * 3)
    del s
    assert sum(int(e) ** 3 for e in edges) == exp["total_voxels"]
    return exp, edges, vols, segs, dims


@pytest.mark.parametrize("mode", ["graph", "autotune"])
def test_batch_config_against_oracle(capi_ctx, batch, mode):
    exp, edges, vols, segs, dims = batch
    npo = torch.empty(len(vols), dtype=torch.int32, device="cuda")
    nso = torch.empty(len(vols), dtype=torch.int32, device="cuda")
    for s in segs:
        s.fill_(-7)
    print(f"[batch_config] {len(vols)} crops, mode {mode}", flush=True)
    capi_ctx.segment_batch_d(vols, segs, dims, (1.0, 1.0, 1.0), npo, nso, graph=True,
                             autotune=(mode == "autotune"))
    torch.cuda.synchronize()
    npo, nso = npo.cpu().numpy(), nso.cpu().numpy()
    for k, want in exp["sample"].items():
        k = int(k)
        assert int(edges[k]) == want["edge"]
        assert digest(vols[k].cpu().numpy()) == want["sdf"]
        assert digest(segs[k].cpu().numpy()) == want["seg"], f"crop {k} (edge {edges[k]})"
        assert npo[k] == want["n_pores"] and nso[k] == want["n_solids"]
    assert max(int(edges[int(k)]) for k in exp["sample"]) >= 159


def test_batch_config_counts_match_single_calls(capi_ctx, batch):
    """Every crop's label counts from the batch equal its own segment_volume call's, on 64 crops
    spread over the batch (the largest included)."""
    exp, edges, vols, segs, dims = batch
    npo = torch.empty(len(vols), dtype=torch.int32, device="cuda")
    nso = torch.empty(len(vols), dtype=torch.int32, device="cuda")
    capi_ctx.segment_batch_d(vols, segs, dims, (1.0, 1.0, 1.0), npo, nso, graph=True)
    torch.cuda.synchronize()
    npo, nso = npo.cpu().numpy(), nso.cpu().numpy()
    pick = list(np.linspace(0, len(vols) - 1, 63).astype(int)) + [int(np.argmax(edges))]
    for k in pick:
        one = torch.empty_like(segs[k])
        capi_ctx.segment_volume_d(vols[k], dims[k], (1.0, 1.0, 1.0), one)
        assert torch.equal(one, segs[k]), f"crop {k}"
        o = one.cpu().numpy()
        assert npo[k] == max(0, int(o.max())) and nso[k] == max(0, -int(o.min()))
