"""Batched persistent segmentation vs per-volume segment_volume (GPU): bit-exact per volume,
identical across launch shapes (autotuner candidates) and with / without the hipGraph."""
from __future__ import annotations

import numpy as np
import pytest

import pnm_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def quantized(shape, rng, step=0.5):
    return (np.round(rng.uniform(-3, 3, size=shape) / step) * step).astype(np.float32)


def crops(src, rng, count, lo, hi):
    """Cubic crops with random edges and offsets (periodic wrap) of a source volume."""
    out = []
    nz, ny, nx = src.shape
    for _ in range(count):
        e = int(rng.integers(lo, hi + 1))
        z0, y0, x0 = (int(rng.integers(0, k)) for k in (nz, ny, nx))
        idx = [np.arange(o, o + e) % k for o, k in ((z0, nz), (y0, ny), (x0, nx))]
        out.append(np.ascontiguousarray(src[np.ix_(*idx)]))
    return out


def run_batch(ctx, vols, spc_zyx, **kw):
    sd = [torch.from_numpy(v).cuda() for v in vols]
    sg = [torch.empty(v.size, dtype=torch.int32, device="cuda") for v in vols]
    npo = torch.empty(len(vols), dtype=torch.int32, device="cuda")
    nso = torch.empty(len(vols), dtype=torch.int32, device="cuda")
    dims = [(v.shape[2], v.shape[1], v.shape[0]) for v in vols]
    ctx.segment_batch_d(sd, sg, dims, spc_zyx[::-1], npo, nso, **kw)
    return [g.cpu().numpy() for g in sg], npo.cpu().numpy(), nso.cpu().numpy()


def test_batch_mixed_vs_oracle(capi_ctx):
    rng = np.random.default_rng(42)
    from peclet.pnm import synth

    src, _ = synth.packing_sdf((64, 64, 64), "uniform", 3.0, 6.0, 0.55, seed=9)
    vols = crops(src, rng, 40, 5, 40)
    vols += [quantized(s, rng) for s in [(1, 1, 1), (1, 2, 3), (2, 2, 2), (3, 5, 70), (7, 3, 65),
                                         (9, 9, 9), (1, 1, 130)]]
    spc = (1.0, 0.75, 0.5)
    segs, npo, nso = run_batch(capi_ctx, vols, spc)
    for k, v in enumerate(vols):
        want = orc.segment_volume(v, spc)
        np.testing.assert_array_equal(segs[k], want, err_msg=f"volume {k} {v.shape}")
        assert npo[k] == max(0, int(want.max()))
        assert nso[k] == max(0, -int(want.min()))


def test_batch_capped_walk_volume(capi_ctx):
    """A 1100-voxel tent whose walks hit the 512-step cap, inside a batch."""
    x = np.arange(1100)
    tent = (1.0 + np.minimum(x, 1099 - x)).astype(np.float32).reshape(1, 1, 1100)
    rng = np.random.default_rng(1)
    vols = [tent, quantized((6, 6, 6), rng), tent.copy()]
    segs, _, _ = run_batch(capi_ctx, vols, (1, 1, 1))
    for k, v in enumerate(vols):
        np.testing.assert_array_equal(segs[k], orc.segment_volume(v, (1, 1, 1)))


def test_batch_launch_shapes_agree(capi_ctx):
    """Same outputs for every autotuner candidate, graph or not; equals single-volume calls."""
    from peclet.pnm import synth

    rng = np.random.default_rng(7)
    src = synth.boolean_sdf((128, 128, 128), 5.0, 0.3, 0.5, seed=3)
    vols = crops(src, rng, 96, 16, 72)
    ref, _, _ = run_batch(capi_ctx, vols, (1, 1, 1), graph=False)
    for v, r in zip(vols, ref):
        s = torch.from_numpy(v).cuda()
        g = torch.empty(v.size, dtype=torch.int32, device="cuda")
        capi_ctx.segment_volume_d(s, (v.shape[2], v.shape[1], v.shape[0]), (1, 1, 1), g)
        np.testing.assert_array_equal(g.cpu().numpy(), r)
    tuned, _, _ = run_batch(capi_ctx, vols, (1, 1, 1), graph=True, autotune=True)
    cfg = capi_ctx.batch_config()
    assert cfg["block"] in (256, 512, 1024) and cfg["groups"] > 0
    for a, b in zip(ref, tuned):
        np.testing.assert_array_equal(a, b)
    again, _, _ = run_batch(capi_ctx, vols, (1, 1, 1), graph=True)
    for a, b in zip(ref, again):
        np.testing.assert_array_equal(a, b)


def test_batch_rejects_bad_spacing(capi_ctx):
    v = np.zeros((4, 4, 4), np.float32)
    with pytest.raises(RuntimeError, match="spacing"):
        run_batch(capi_ctx, [v], (1.0, 0.0, 1.0))
