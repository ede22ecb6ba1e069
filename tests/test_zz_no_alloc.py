"""The no-allocation entry point (pnm_segment_volume_noalloc_d) against pnm_segment_volume_d,
through the C-ABI: bit-exact segmentation and roots on consecutive calls (the second smaller
than the first, so stale scratch from the first must not leak into it), the largest accepted
volume, the walk-cap path, the size limit, and no growth of the context's arena."""
from __future__ import annotations

import numpy as np
import pytest

import deepflood as dfl
from peclet.pnm import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _quantized(shape, seed):
    rng = np.random.default_rng(seed)
    return (np.round(rng.uniform(-3, 3, size=shape) / 0.5) * 0.5).astype(np.float32)


def _tent(nx):
    x = np.arange(nx)
    return (1.0 + np.minimum(x, nx - 1 - x)).astype(np.float32).reshape(1, 1, nx)


def _xyz(shape):
    return (shape[2], shape[1], shape[0])


@pytest.fixture(scope="module")
def ctxs():
    torch.zeros(1, device="cuda")
    from peclet.pnm import capi

    fixed, ref = capi.Context(0), capi.Context(0)
    yield fixed, ref
    fixed.close()
    ref.close()


def _both(ctxs, sdf, spc, with_roots):
    fixed, ref = ctxs
    n = sdf.size
    s = torch.from_numpy(np.ascontiguousarray(sdf)).cuda()
    seg_a = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    seg_b = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    rt_a = torch.full((n,), 3, dtype=torch.int32, device="cuda") if with_roots else None
    rt_b = torch.full((n,), -3, dtype=torch.int32, device="cuda") if with_roots else None
    print(f"[no_alloc] shape {sdf.shape} roots={with_roots}", flush=True)
    fixed.segment_volume_noalloc_d(s, _xyz(sdf.shape), _xyz(spc), seg_a, rt_a)
    ref.segment_volume_d(s, _xyz(sdf.shape), _xyz(spc), seg_b, rt_b)
    np.testing.assert_array_equal(seg_a.cpu().numpy(), seg_b.cpu().numpy())
    if with_roots:
        np.testing.assert_array_equal(rt_a.cpu().numpy(), rt_b.cpu().numpy())
    assert fixed.arena_bytes() == 0  # the fixed path never touched the context's arena


def test_noalloc_two_calls_second_smaller(ctxs):
    big = synth.packing_sdf((128, 128, 128), "uniform", 6.0, 10.0, 0.6, seed=1001)[0]
    _both(ctxs, big, (1.0, 1.0, 1.0), with_roots=True)    # the largest accepted cube
    small = _quantized((20, 17, 65), 5)
    _both(ctxs, small, (1.0, 0.5, 0.75), with_roots=False)


@pytest.mark.parametrize("case", ["plate", "tent1100", "tiny", "quantized"])
def test_noalloc_matches_arena_path(ctxs, case):
    spc = (1.0, 1.0, 1.0)
    if case == "plate":          # 219 flood sweeps
        sdf = dfl.plate_sdf()
    elif case == "tent1100":     # walks stop at the 512-step cap: full first-appearance path
        sdf = _tent(1100)
    elif case == "tiny":
        sdf = _quantized((1, 1, 2), 3)
    else:
        sdf = _quantized((33, 2, 130), 9)
        spc = (1.0, 0.5, 0.75)
    _both(ctxs, sdf, spc, with_roots=True)


def test_noalloc_rejects_too_large(ctxs):
    from peclet.pnm import capi

    fixed, _ = ctxs
    dims = (129, 128, 128)  # X padded to 192: 192*128*128 > 2^21
    assert capi.noalloc_padded_voxels(dims) > capi.PNM_NOALLOC_MAX_PADDED
    s = torch.zeros(1, dtype=torch.float32, device="cuda")  # never read: rejected first
    seg = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="no-allocation entry point"):
        fixed.segment_volume_noalloc_d(s, dims, (1, 1, 1), seg)
