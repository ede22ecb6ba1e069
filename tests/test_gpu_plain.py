"""HIP path vs the oracle: pores, segmentation stages, topology, fused pipeline (GPU).

Bar: bit-exact. Pores compare as raw float32 bits in ascending peak-voxel order (the oracle's
emission order; the reference's is an atomic-slot race and its own tests compare sets).
"""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

import bruteforce as bf
import pnm_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def xyz(v):
    return (v[2], v[1], v[0])


def bits(pores):
    """Structured pore array -> uint32 view (bitwise compare, NaN-safe)."""
    return np.ascontiguousarray(pores).view(np.uint32)


def run_fused_d(ctx, sdf, org, spc):
    n = sdf.size
    s = dev(sdf)
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    pores = torch.empty(4 * max(1, min(n, 1000000)), dtype=torch.float32, device="cuda")
    pairs = torch.empty(2 * 3 * n + 2, dtype=torch.int32, device="cuda")
    npores, tot, npairs = ctx.extract_pore_network_d(s, xyz(sdf.shape), xyz(org), xyz(spc), pores,
                                                     seg, pairs)
    p = pores[: 4 * npores].cpu().numpy().view(orc.PORE_DTYPE)
    return p, tot, seg.cpu().numpy(), pairs[: 2 * npairs].cpu().numpy().reshape(-1, 2)


def check_against_oracle(ctx, sdf, org, spc, golden=None):
    o_pores, o_tot = orc.extract_pores(sdf, org, spc)
    o_seg = orc.segment_volume(sdf, spc)
    o_top = orc.extract_topology(o_seg, sdf.shape)
    p, tot, seg, top = run_fused_d(ctx, sdf, org, spc)
    assert tot == o_tot
    np.testing.assert_array_equal(bits(p), bits(o_pores))
    np.testing.assert_array_equal(seg, o_seg)
    np.testing.assert_array_equal(top, o_top)
    if golden is not None:
        np.testing.assert_array_equal(seg, golden["seg"])
        np.testing.assert_array_equal(top, golden["topology"])
    return p, seg, top


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_fused_device(capi_ctx, path):
    g = load(path)
    check_against_oracle(capi_ctx, g["sdf"], g["origin_zyx"], g["spacing_zyx"], g)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_python_api(path):
    """The drop-in module (host buffers), staged and fused, against the oracle."""
    import peclet.pnm as pnm

    g = load(path)
    sdf, org, spc = g["sdf"], list(g["origin_zyx"]), list(g["spacing_zyx"])
    o_pores, _ = orc.extract_pores(sdf, org, spc)
    o_seg = orc.segment_volume(sdf, spc)
    o_top = [tuple(r) for r in orc.extract_topology(o_seg, sdf.shape).tolist()]
    pores = pnm.extract_pores(sdf, org, spc)
    got = np.array([(q.x, q.y, q.z, q.radius) for q in pores], dtype=np.float32)
    want = np.stack([o_pores[k] for k in ("x", "y", "z", "radius")], axis=1)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    seg = pnm.segment_volume(sdf, spc)
    np.testing.assert_array_equal(seg, o_seg)
    assert pnm.extract_topology_gpu(seg, list(sdf.shape)) == o_top
    assert pnm.extract_topology_gpu(seg.tolist(), list(sdf.shape)) == o_top
    fp, fseg, fconn = pnm.extract_pore_network(sdf, org, spc)
    assert len(fp) == len(pores)
    np.testing.assert_array_equal(fseg, o_seg)
    assert fconn == o_top


def quantized(shape, seed, step=0.5, nan_frac=0.0, inf_frac=0.0):
    rng = np.random.default_rng(seed)
    v = np.round(rng.uniform(-3, 3, size=shape) / step) * step
    if nan_frac:
        v[rng.random(shape) < nan_frac] = np.nan
    if inf_frac:
        v[rng.random(shape) < inf_frac] = np.inf
    return v.astype(np.float32)


# (20, 17, 65) and (34, 25, 129): several tiles in y and z with a last row chunk one column wide
# (its first lane is also its last one in the seam passes).
SHAPES = [(1, 1, 1), (1, 1, 2), (2, 2, 2), (1, 3, 70), (3, 5, 65), (5, 7, 63), (4, 9, 128),
          (17, 13, 11), (33, 2, 130), (8, 64, 8), (20, 17, 65), (34, 25, 129)]


@pytest.mark.parametrize("shape", SHAPES)
def test_random_shapes(capi_ctx, shape):
    sdf = quantized(shape, sum(shape) * 7 + 1)
    check_against_oracle(capi_ctx, sdf, (0.5, -1.0, 2.0), (1.0, 0.5, 0.75))


# Coarse levels make most neighbourhoods hold ties, including ties with wrapped neighbours at
# every face of the volume (the tie goes to the larger flat index, whose order wraps there).
TIE_SHAPES = [(3, 3, 3), (2, 1, 65), (2, 130, 3), (16, 8, 64), (17, 9, 66), (32, 16, 128)]


@pytest.mark.parametrize("shape", TIE_SHAPES)
def test_heavy_ties(capi_ctx, shape):
    sdf = quantized(shape, sum(shape) * 11 + 5, step=1.5)
    check_against_oracle(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))


def test_pore_heavy_rows(capi_ctx):
    """Mostly sdf>0: tile rows with more sdf>0 voxels than the per-wave walk lists hold take
    the per-lane walk loops."""
    rng = np.random.default_rng(77)
    sdf = (np.round(rng.uniform(-1.0, 5.0, size=(24, 20, 130)) / 0.5) * 0.5).astype(np.float32)
    assert (sdf > 0).mean() > 0.75
    check_against_oracle(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))


def test_band_heavy_rows(capi_ctx):
    """Mostly unlabelled solid (-1.5 < sdf < 0) with sparse markers: tile rows with more band
    candidates than the flood tile's per-wave list holds sweep their row slice by slice, and the
    flood needs many global sweeps."""
    rng = np.random.default_rng(78)
    sdf = np.round(rng.uniform(-1.25, -0.25, size=(24, 20, 130)) / 0.25) * 0.25
    sdf[rng.random(sdf.shape) < 0.03] = -3.0
    sdf[rng.random(sdf.shape) < 0.03] = 2.0
    check_against_oracle(capi_ctx, sdf.astype(np.float32), (0, 0, 0), (1, 1, 1))


def test_wide_plane_general_multiply(capi_ctx):
    """ny * nx >= 2^24: the tile pass computes flat indices with 32-bit multiplies instead of
    24-bit ones."""
    shape = (1, 4100, 4100)
    assert shape[1] * shape[2] >= 1 << 24
    check_against_oracle(capi_ctx, quantized(shape, 79), (0, 0, 0), (1, 1, 1))


@pytest.mark.parametrize("seed", [1, 2])
def test_nan_inf(capi_ctx, seed):
    sdf = quantized((10, 12, 70), seed, nan_frac=0.05, inf_frac=0.01)
    check_against_oracle(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))


def test_uniform_volumes(capi_ctx):
    for val in (-5.0, -0.5, 0.0, 0.5, 3.0):
        sdf = np.full((6, 7, 9), val, np.float32)
        check_against_oracle(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_stages(capi_ctx, path):
    g = load(path)
    sdf, spc = g["sdf"], g["spacing_zyx"]
    n = sdf.size
    s = dev(sdf)
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    ccl = torch.empty(n, dtype=torch.int32, device="cuda")
    capi_ctx.stage_labels_d(s, xyz(sdf.shape), xyz(spc), lab, ccl)
    np.testing.assert_array_equal(ccl.cpu().numpy(), orc.markers_ccl(sdf, spc))
    np.testing.assert_array_equal(lab.cpu().numpy(), g["labels"])
    r = torch.empty(n, dtype=torch.int32, device="cuda")
    capi_ctx.stage_roots_d(s, xyz(sdf.shape), False, r)
    np.testing.assert_array_equal(r.cpu().numpy(), orc.gradient_roots(sdf))
    capi_ctx.stage_roots_d(s, xyz(sdf.shape), True, r)
    np.testing.assert_array_equal(r.cpu().numpy(), orc.gradient_roots(sdf, all_voxels=True))


@pytest.fixture(scope="module")
def packing128():
    from peclet.pnm import synth

    sdf, _ = synth.packing_sdf((128, 128, 128), "uniform", 6.0, 10.0, 0.60, seed=1001)
    return sdf


def test_packing128_bit_exact(capi_ctx, packing128):
    """The 128^3 sphere-packing workload (radii U[6,10], porosity ~0.40): bit-exact."""
    p, seg, top = check_against_oracle(capi_ctx, packing128, (0, 0, 0), (1, 1, 1))
    assert len(p) > 100 and seg.max() > 100 and seg.min() < -10 and len(top) > 1000


def test_boolean128_bit_exact(capi_ctx):
    from peclet.pnm import synth

    sdf = synth.boolean_sdf((128, 128, 128), 9.0, 0.3, 0.5, seed=2002)
    check_against_oracle(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))


def test_deterministic_and_staged_equals_fused(capi_ctx, packing128):
    sdf = packing128
    a = run_fused_d(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))
    b = run_fused_d(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            np.testing.assert_array_equal(np.ascontiguousarray(x).view(np.uint8),
                                          np.ascontiguousarray(y).view(np.uint8))
    n = sdf.size
    s = dev(sdf)
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    capi_ctx.segment_volume_d(s, xyz(sdf.shape), (1, 1, 1), seg)
    np.testing.assert_array_equal(seg.cpu().numpy(), a[2])
    pairs = torch.empty(6 * n, dtype=torch.int32, device="cuda")
    k = capi_ctx.extract_topology_d(seg, xyz(sdf.shape), pairs)
    np.testing.assert_array_equal(pairs[: 2 * k].cpu().numpy().reshape(-1, 2), a[3])
    pores = torch.empty(4 * 1000000, dtype=torch.float32, device="cuda")
    cnt, tot = capi_ctx.extract_pores_d(s, xyz(sdf.shape), (0, 0, 0), (1, 1, 1), pores)
    np.testing.assert_array_equal(pores[: 4 * cnt].cpu().numpy().view(np.uint32),
                                  bits(a[0]))


def test_on_caller_stream(capi_ctx, packing128):
    """The context on a caller's (non-default, default-priority) torch stream: the pipeline's
    second stream is forked from and joined back to that stream, so two calls in a row, read
    back on the same stream without any other synchronisation, equal the own-stream result."""
    sdf = packing128
    a = run_fused_d(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))
    st = torch.cuda.Stream()
    try:
        with torch.cuda.stream(st):
            capi_ctx.set_stream(st.cuda_stream)
            for _ in range(2):
                b = run_fused_d(capi_ctx, sdf, (0, 0, 0), (1, 1, 1))
                assert b[1] == a[1]
                np.testing.assert_array_equal(bits(b[0]), bits(a[0]))
                np.testing.assert_array_equal(b[2], a[2])
                np.testing.assert_array_equal(b[3], a[3])
    finally:
        capi_ctx.set_stream(None)
        torch.cuda.synchronize()


def test_capacity_errors(capi_ctx):
    sdf = quantized((8, 8, 8), 5)
    s = dev(sdf)
    seg = torch.empty(sdf.size, dtype=torch.int32, device="cuda")
    small = torch.empty(4, dtype=torch.float32, device="cuda")  # one pore record
    pairs = torch.empty(2, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="too small"):
        capi_ctx.extract_pore_network_d(s, (8, 8, 8), (0, 0, 0), (1, 1, 1), small, seg, pairs)


def test_pairs_capacity_writes_nothing(capi_ctx):
    """A pairs buffer below the unique pair count: PNM_ECAPACITY, and the buffer is not written
    (the decode reads the count on the device and skips when it exceeds the capacity)."""
    sdf = quantized((8, 8, 8), 5)
    s = dev(sdf)
    seg = torch.empty(sdf.size, dtype=torch.int32, device="cuda")
    pores = torch.empty(4 * sdf.size, dtype=torch.float32, device="cuda")
    pairs = torch.full((2,), -7, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="too small"):
        capi_ctx.extract_pore_network_d(s, (8, 8, 8), (0, 0, 0), (1, 1, 1), pores, seg, pairs)
    assert pairs.cpu().tolist() == [-7, -7]


def _torch_topology(seg3):
    """Independent full-size topology on the GPU with torch ops."""
    keys = []
    for dim in (2, 1, 0):
        nb = torch.roll(seg3, shifts=-1, dims=dim)
        d = seg3 != nb
        lo = torch.minimum(seg3, nb)[d].to(torch.int64)
        hi = torch.maximum(seg3, nb)[d].to(torch.int64)
        # signed (lo, hi) lexicographic order: lo * 2^32 keeps lo's sign, hi is biased unsigned
        keys.append((lo << 32) + (hi + 2 ** 31))
    k = torch.unique(torch.cat(keys))
    return torch.stack([k >> 32, (k & 0xFFFFFFFF) - 2 ** 31], dim=1).to(torch.int32)


@pytest.mark.parametrize("kind", ["polydisperse", "boolean"])
def test_512_properties(capi_ctx, kind):
    """512^3 workloads: properties that pin the result without the oracle.
    - topology equals an independent torch computation of the full-size face pairs;
    - labels are 1..np and -1..-ns, ranked by first appearance (torch scatter_reduce amin);
    - every pore voxel shares its label with its ascent successor (torch 26-roll lexmax);
    - 26-adjacent markers share a label; fused == staged; repeated runs identical."""
    from peclet.pnm import synth

    shape = (512, 512, 512)
    if kind == "polydisperse":
        sdf, _, _ = synth.polydisperse_sdf(shape, 9.0, 0.3, 0.35, seed=2001)
    else:
        sdf = synth.boolean_sdf(shape, 9.0, 0.3, 0.5, seed=2002)
    n = sdf.size
    s = dev(sdf)
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    pores = torch.empty(4 * 1000000, dtype=torch.float32, device="cuda")
    pairs = torch.empty(2 * 20_000_000, dtype=torch.int32, device="cuda")
    npores, tot, npairs = capi_ctx.extract_pore_network_d(s, xyz(shape), (0, 0, 0), (1, 1, 1),
                                                          pores, seg, pairs)
    seg3 = seg.view(shape)
    top = pairs[: 2 * npairs].view(-1, 2)
    ref = _torch_topology(seg3)
    if not torch.equal(top, ref):
        a = set(map(tuple, top.cpu().tolist()))
        b = set(map(tuple, ref.cpu().tolist()))
        sorted_ok = bool((top[1:, 0] > top[:-1, 0]).logical_or(
            (top[1:, 0] == top[:-1, 0]) & (top[1:, 1] > top[:-1, 1])).all())
        raise AssertionError(f"topology differs: {top.shape[0]} pairs vs {ref.shape[0]}; "
                             f"sorted unique {sorted_ok}; only in result {sorted(a - b)[:8]} "
                             f"({len(a - b)}); only in reference {sorted(b - a)[:8]} "
                             f"({len(b - a)}); seg range [{int(seg.min())}, {int(seg.max())}]")
    npl, nsl = int(seg.max()), int(-seg.min())
    assert npl == tot  # no NaN and no capped walks: one pore per peak
    idx = torch.arange(n, device="cuda", dtype=torch.int64)
    first = torch.full((npl + nsl + 1,), n, dtype=torch.int64, device="cuda")
    first.scatter_reduce_(0, (seg.to(torch.int64) + nsl), idx, reduce="amin")
    fp = first[nsl + 1:]
    fs = first[:nsl].flip(0)  # labels -1, -2, ...
    assert int((fp == n).sum()) == 0 and int((fs == n).sum()) == 0
    assert bool((fp[1:] > fp[:-1]).all()) and bool((fs[1:] > fs[:-1]).all())
    # ascent successor shares the label
    sv = s.view(shape)
    iv = idx.view(shape)
    bv, bi = sv.clone(), iv.clone()
    for dz, dy, dx in bf.OFFSETS26:
        nv = torch.roll(sv, shifts=(-dz, -dy, -dx), dims=(0, 1, 2))
        ni = torch.roll(iv, shifts=(-dz, -dy, -dx), dims=(0, 1, 2))
        better = (nv > bv) | ((nv == bv) & (ni > bi))
        bv = torch.where(better, nv, bv)
        bi = torch.where(better, ni, bi)
        del nv, ni, better
    pos = sv > 0
    segf = seg.to(torch.int64)
    assert torch.equal(segf[iv[pos]], segf[bi[pos]])
    del bv, bi
    # adjacent markers share a label
    mk = sv < -1.5
    for dz, dy, dx in bf.OFFSETS26[:13]:
        nm = torch.roll(mk, shifts=(-dz, -dy, -dx), dims=(0, 1, 2))
        ns = torch.roll(seg3, shifts=(-dz, -dy, -dx), dims=(0, 1, 2))
        both = mk & nm
        assert torch.equal(seg3[both], ns[both])
        del nm, ns, both
    # determinism and staged == fused
    seg2 = torch.empty_like(seg)
    capi_ctx.segment_volume_d(s, xyz(shape), (1, 1, 1), seg2)
    assert torch.equal(seg, seg2)
    pairs2 = torch.empty_like(pairs)
    k = capi_ctx.extract_topology_d(seg2, xyz(shape), pairs2)
    assert k == npairs and torch.equal(pairs2[: 2 * k], pairs[: 2 * npairs])


@pytest.mark.parametrize("case", ["sphere_lattice", "trig", "tent1100", "quantized", "random"])
def test_roots_out_is_reference_roots(capi_ctx, case):
    """segment_volume_d's roots_d is rootsOut (pore_extraction.hpp:383-384): ascent walkers of
    sdf>0 voxels and descent walkers of the others, each capped at 512 steps."""
    if case == "random":
        sdf = quantized((20, 17, 65), 77, nan_frac=0.02)
    else:
        sdf = load(os.path.join(os.path.dirname(__file__), "golden", case + ".npz"))["sdf"]
    n = sdf.size
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    rts = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    capi_ctx.segment_volume_d(dev(sdf), xyz(sdf.shape), (1.0, 1.0, 1.0), seg, rts)
    o_seg, o_rts, _ = orc.segment_volume(sdf, (1.0, 1.0, 1.0), return_stages=True)
    np.testing.assert_array_equal(rts.cpu().numpy(), o_rts)
    np.testing.assert_array_equal(seg.cpu().numpy(), o_seg)
