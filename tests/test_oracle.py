"""The oracle, pinned: golden fixtures, brute force, closed forms (CPU only)."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

import bruteforce as bf
import pnm_oracle as orc

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module", autouse=True)
def _built():
    orc.lib()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_golden(path):
    g = load(path)
    sdf, org, spc = g["sdf"], g["origin_zyx"], g["spacing_zyx"]
    pores, total = orc.extract_pores(sdf, org, spc)
    # peaks: same voxels, ascending order; radius is the SDF value bitwise
    assert total == g["peak_idx"].size
    nz, ny, nx = sdf.shape
    flat = sdf.ravel()
    np.testing.assert_array_equal(pores["radius"], flat[g["peak_idx"]])
    # positions: FP32 restatement within 1e-5 relative of the float64 centroid
    xyz = np.stack([pores["x"], pores["y"], pores["z"]], axis=1).astype(np.float64)
    np.testing.assert_allclose(xyz, g["peak_xyz"], rtol=1e-5, atol=1e-5)
    # stages
    lab_ccl = orc.markers_ccl(sdf, spc)
    np.testing.assert_array_equal(lab_ccl, g["labels_ccl"])
    lab, _ = orc.flood(sdf, lab_ccl)
    np.testing.assert_array_equal(lab, g["labels"])
    rts = orc.gradient_roots(sdf)
    pos = flat > 0
    np.testing.assert_array_equal(rts[pos], g["roots"][pos])
    np.testing.assert_array_equal(orc.gradient_roots(sdf, all_voxels=True), g["flow_roots"])
    seg = orc.segment_volume(sdf, spc)
    np.testing.assert_array_equal(seg, g["seg"])
    np.testing.assert_array_equal(orc.extract_topology(seg, sdf.shape), g["topology"])


def quantized(shape, seed, step=0.5, nan_frac=0.0):
    rng = np.random.default_rng(seed)
    v = np.round(rng.uniform(-3, 3, size=shape) / step) * step
    if nan_frac:
        v[rng.random(shape) < nan_frac] = np.nan
    return v.astype(np.float32)


SHAPES = [(1, 1, 1), (1, 2, 3), (2, 1, 2), (3, 3, 3), (4, 5, 6), (7, 3, 9), (9, 8, 10)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_vs_bruteforce_random(shape, seed):
    sdf = quantized(shape, 100 * seed + sum(shape))
    spc = (1.0, 0.5, 0.75)
    seg, rts, lab = orc.segment_volume(sdf, spc, return_stages=True)
    ref_lab = bf.flood_layers(sdf, bf.ccl(bf.markers(sdf, spc)))
    np.testing.assert_array_equal(lab, ref_lab)
    np.testing.assert_array_equal(rts, bf.roots(sdf))  # ascent and descent walkers
    np.testing.assert_array_equal(seg, bf.segment(sdf, spc))
    np.testing.assert_array_equal(orc.extract_topology(seg, shape), bf.topology(seg, shape))
    pores, tot = orc.extract_pores(sdf, (0, 0, 0), spc)
    np.testing.assert_array_equal(np.flatnonzero(bf.peaks(sdf).ravel()).size, tot)


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_nan_semantics(seed):
    """NaN voxels: extract_pores peaks (no comparison dominates them), floodable solid for the
    segmentation, and a NaN first voxel leaves its solid label at id 0."""
    sdf = quantized((6, 7, 8), seed, nan_frac=0.08)
    spc = (1.0, 1.0, 1.0)
    seg = orc.segment_volume(sdf, spc)
    np.testing.assert_array_equal(seg, bf.segment(sdf, spc))
    _, tot = orc.extract_pores(sdf, (0, 0, 0), spc)
    assert tot == int(bf.peaks(sdf).sum())
    assert tot >= int(np.isnan(sdf).sum())


def test_single_sphere_closed_form():
    """One pore sphere centred on a voxel: exactly one peak, zero centroid offset, so the
    position is o + i*s bitwise and the radius is sdf[peak]."""
    n = 21
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float32),) * 3, indexing="ij")
    c = np.float32(10)
    sdf = (np.float32(6.0) - np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)).astype(np.float32)
    org = (0.25, -1.0, 3.0)
    spc = (0.5, 2.0, 1.0)
    pores, tot = orc.extract_pores(sdf, org, spc)
    assert tot == 1
    p = pores[0]
    assert p["radius"] == sdf[10, 10, 10]
    assert p["x"] == np.float32(3.0) + np.float32(10) * np.float32(1.0)
    assert p["y"] == np.float32(-1.0) + np.float32(10) * np.float32(2.0)
    assert p["z"] == np.float32(0.25) + np.float32(10) * np.float32(0.5)
    seg = orc.segment_volume(sdf, spc)
    assert seg.max() == 1  # one pore basin
    # solid shell outside the sphere: markers where sdf < -1.5*min(spc) = -0.75


def test_empty_and_uniform():
    assert orc.extract_pores(np.zeros((0, 3, 3), np.float32), (0, 0, 0), (1, 1, 1))[1] == 0
    solid = np.full((4, 4, 4), -5.0, np.float32)
    seg = orc.segment_volume(solid, (1, 1, 1))
    assert (seg == -1).all()  # one periodic marker component
    assert orc.extract_topology(seg, solid.shape).shape == (0, 2)
    pore = np.full((4, 4, 4), 2.0, np.float32)  # plateau: ties to the largest index
    seg = orc.segment_volume(pore, (1, 1, 1))
    assert (seg == 1).all()
    pores, tot = orc.extract_pores(pore, (0, 0, 0), (1, 1, 1))
    assert tot == 1 and pores[0]["radius"] == 2.0


def test_walk_cap_tent():
    """Walks deeper than 512 steps stop at the cap (closed form on a 1-D tent)."""
    nx = 1100
    x = np.arange(nx)
    sdf = (1.0 + np.minimum(x, nx - 1 - x)).astype(np.float32).reshape(1, 1, nx)
    r = orc.gradient_roots(sdf)
    peak = 550  # values tie at 549/550; the larger index wins
    for start in (0, 10, 37, 38, 100, 600, 1099):
        depth = abs(peak - start)
        want = start + min(depth, 512) * (1 if start < peak else -1)
        assert r[start] == want, start


@pytest.mark.parametrize("kind", ["plate", "two_grains"])
def test_flood_deep_plate_closed_form(kind):
    """A thin plate fed by one grain floods one voxel per sweep (plate length + 1 sweeps, the
    last one changing nothing); between two grains the equidistant middle voxel takes the
    smaller label. Also equal to the BFS-layer flood."""
    import deepflood as dfl

    sdf = dfl.plate_sdf() if kind == "plate" else dfl.two_grain_plate_sdf()
    lab0 = orc.markers_ccl(sdf, (1, 1, 1))
    lab, sweeps = orc.flood(sdf, lab0)
    np.testing.assert_array_equal(lab, bf.flood_layers(sdf, lab0))
    lab = lab.reshape(sdf.shape)
    plate = lab[3:5, 7:9, :]
    if kind == "plate":
        assert sweeps == (230 - 12) + 1
        assert (plate[:, :, 12:230] == 0).all()  # grain A's label: its first voxel, index 0
    else:
        assert sweeps == 58 + 1
        assert (plate[:, :, 12:70] == 0).all()      # x <= 69: grain A (tie at 69 -> smaller)
        assert (plate[:, :, 70:127] == 127).all()   # grain B's first voxel is (0, 0, 127)
    assert (lab[sdf >= 0] == -1).all()


def test_descent_walk_cap_valley():
    """rootsOut of sdf <= 0 voxels is the descent walker (pore_extraction.hpp:283-301): on a 1-D
    valley it runs to the lowest voxel (ties to the larger index) or stops at the 512-step cap."""
    nx = 1100
    x = np.arange(nx)
    sdf = (-1.0 - np.minimum(x, nx - 1 - x)).astype(np.float32).reshape(1, 1, nx)
    r = orc.gradient_roots(sdf)
    low = 550  # values tie at 549/550; the larger index wins
    for start in (0, 10, 37, 38, 100, 600, 1099):
        depth = abs(low - start)
        want = start + min(depth, 512) * (1 if start < low else -1)
        assert r[start] == want, start
    np.testing.assert_array_equal(r, bf.roots(sdf))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_openmp_baseline_equals_oracle(path):
    """The OpenMP restatement timed as the CPU baseline computes what the serial oracle does:
    seg and connections bit for bit, pores as a set (atomic-slot order, as the reference)."""
    g = load(path)
    sdf, org, spc = g["sdf"], g["origin_zyx"], g["spacing_zyx"]
    p, seg, conns, _ = orc.ref_omp_extract_pore_network(sdf, org, spc)
    o_pores, o_seg, o_conns = orc.extract_pore_network(sdf, org, spc)
    np.testing.assert_array_equal(seg, o_seg)
    np.testing.assert_array_equal(conns, o_conns)
    key = lambda a: np.sort(np.ascontiguousarray(a).view(np.uint32).reshape(-1, 4).view(
        "u4,u4,u4,u4").ravel())
    np.testing.assert_array_equal(key(p), key(o_pores))


def test_oracle_under_sanitizers():
    """The oracle built with AddressSanitizer + UndefinedBehaviourSanitizer (no recovery, leak
    check on) runs every function on 40 small and degenerate cases without a report."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    here = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle")
    r = subprocess.run(["make", "-s", "-C", here, "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "40 cases clean" in r.stdout
