"""Network-flow oracle pinned by closed forms and invariants (CPU only).

  - divergence-free MAC field  -> every pore residual ~ 0 (the residual is the basin's summed
    discrete divergence); pins the face -> velocity mapping (velocity at the +d neighbour)
  - uniform x-flow on an x-aligned chamber chain -> Q/A = +-U on every throat, #throats = #pores
  - two chambers joined directly and through the periodic wrap -> the pair appears twice
  - p = const -> pore_pressure = const; p linear in the cell index -> exact trilinear value
  - p = 0 with grad_p along x -> dp + gx (xj - xi) is gx * L * (integer image count)
"""
from __future__ import annotations

import numpy as np
import pytest

import flowcases as fc
import pnm_oracle as orc


def lattice48():
    from peclet.pnm import synth

    return synth.chamber_sdf((48, 48, 48), m=2, chamber_r=7.0, tube_rmin=2.0, tube_rmax=4.0,
                             seed=5)


def test_divergence_free_field_has_zero_residual():
    sdf = lattice48()
    spc = (1.5, 1.0, 0.5)
    f = fc.divergence_free(sdf.shape, spc, seed=1)
    div = fc.mac_divergence(f, spc)
    assert np.abs(div).max() < 1e-12
    r = orc.extract_network_flow(sdf, (0, 0, 0), spc, f["u"], f["v"], f["w"], f["p"])
    scale = np.abs(f["u"]).max() * spc[0] * spc[1]
    assert r["pore_residual"].size >= 8
    assert np.abs(r["pore_residual"]).max() < 1e-11 * scale * sdf.size ** 0.5
    # a mis-mapped velocity (u at the cell instead of its +x neighbour) breaks it
    g = dict(f)
    g["u"] = np.roll(f["u"], 1, axis=2)
    r2 = orc.extract_network_flow(sdf, (0, 0, 0), spc, g["u"], g["v"], g["w"], g["p"])
    assert np.abs(r2["pore_residual"]).max() > 1e-6 * scale


def test_uniform_flow_on_chain():
    sdf = fc.chain_x(3, 24, 7.0, 3.0)
    U = 2.5
    f = fc.uniform_x(sdf.shape, U)
    r = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), f["u"], f["v"], f["w"], f["p"])
    assert len(r["pores"]) == 3
    assert len(r["throats"]) == 3  # periodic ring: #throats = #pores
    np.testing.assert_array_equal(np.abs(r["throat_flow"] / r["throat_area"]), np.full(3, U))
    assert np.all(r["pore_residual"] == 0.0)  # uniform flux: exact cancellation
    # flow runs +x: through each throat from the lower-x chamber to the next
    for (li, lj), q in zip(r["throats"], r["throat_flow"]):
        xi, xj = r["pores"][li - 1]["x"], r["pores"][lj - 1]["x"]
        dx = (xj - xi + 36) % 72 - 36
        assert np.sign(q) == np.sign(dx)


def test_parallel_patches_repeat_the_pair():
    sdf = fc.two_chambers_two_tubes()
    f = fc.uniform_x(sdf.shape, 1.0)
    r = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), f["u"], f["v"], f["w"], f["p"])
    assert len(r["pores"]) == 2
    assert [tuple(t) for t in r["throats"]] == [(1, 2), (1, 2)]
    q = r["throat_flow"]
    assert q[0] == -q[1] and abs(q[0]) > 0  # one patch carries 1->2, the other 2->1


def test_pressure_constant_and_linear():
    sdf = fc.chain_x(3, 24, 7.0, 3.0)
    shape = sdf.shape
    f = fc.uniform_x(shape, 0.0)
    c = 3.25
    r = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), f["u"], f["v"], f["w"],
                                 np.full(shape, c))
    np.testing.assert_allclose(r["pore_pressure"], c, rtol=0, atol=4 * np.spacing(c))
    z, y, x = fc.grid(shape)
    al, be, ga = 0.125, -0.5, 0.25
    p = al * x + be * y + ga * z
    org, spc = (2.0, -1.0, 0.5), (1.5, 1.0, 0.75)  # zyx
    r = orc.extract_network_flow(sdf, org, spc, f["u"], f["v"], f["w"], p)
    for k, po in enumerate(r["pores"]):
        gx = np.float32(np.float32(po["x"]) - np.float32(org[2])) / np.float32(spc[2])
        gy = np.float32(np.float32(po["y"]) - np.float32(org[1])) / np.float32(spc[1])
        gz = np.float32(np.float32(po["z"]) - np.float32(org[0])) / np.float32(spc[0])
        if not (0 <= gx < shape[2] - 1 and 0 <= gy < shape[1] - 1 and 0 <= gz < shape[0] - 1):
            continue  # the trilinear stencil wraps: linear p is not linear across the seam
        want = al * float(gx) + be * float(gy) + ga * float(gz)
        assert abs(r["pore_pressure"][k] - want) < 1e-12 * (1 + abs(want))


def test_dp_image_count_is_integer():
    sdf = fc.chain_x(3, 24, 7.0, 3.0)
    f = fc.uniform_x(sdf.shape, 1.0)
    gx = 0.01
    spc = (1.0, 1.0, 0.5)  # zyx
    r = orc.extract_network_flow(sdf, (0, 0, 0), spc, f["u"], f["v"], f["w"], f["p"],
                                 grad_p_zyx=(0.0, 0.0, gx))
    Lx = sdf.shape[2] * spc[2]
    for (li, lj), dp in zip(r["throats"], r["throat_dp"]):
        xi = float(r["pores"][li - 1]["x"])
        xj = float(r["pores"][lj - 1]["x"])
        k = (dp / gx + (xj - xi)) / Lx  # dp = -gx ((xj - xi) - L k)
        assert abs(k - round(k)) < 1e-9
    # the wrap-around throat of the ring has a non-zero image count
    ks = [round((dp / gx + float(r["pores"][lj - 1]["x"]) - float(r["pores"][li - 1]["x"])) / Lx)
          for (li, lj), dp in zip(r["throats"], r["throat_dp"])]
    assert sorted(abs(k) for k in ks) == [0, 0, 1]


def test_empty_when_no_pores():
    sdf = np.full((6, 6, 6), -2.0, np.float32)
    z = np.zeros(sdf.shape)
    r = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), z, z, z, z)
    assert len(r["pores"]) == 0 and len(r["throats"]) == 0


@pytest.mark.parametrize("case", ["side_by_side", "pack2", "pack3"])
def test_face_tiers_against_graph_components(case):
    """Every face's patch label (core CCL, film attachment to the minimum reachable core patch,
    leftover films) equals an independent graph-component computation, on geometries where a
    film component touches two or three core patches of its pore pair, so that a film taking
    any patch other than the minimum one fails."""
    import bruteforce as bf
    from peclet.pnm import synth

    if case == "side_by_side":
        sdf = fc.two_chambers_side_by_side_tubes()
    else:
        sdf, _ = synth.packing_sdf((32, 32, 32), "uniform", 3.0, 5.0, 0.55, seed=int(case[-1]))
    f = fc.uniform_x(sdf.shape, 1.0)
    r = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), f["u"], f["v"], f["w"], f["p"],
                                 debug=True)
    flow_lab = bf.segment(sdf, (1, 1, 1))[bf.roots(sdf, all_voxels=True)]
    np.testing.assert_array_equal(r["flow_lab"], flow_lab)
    want, touch = fc.face_tiers(sdf, flow_lab)
    assert touch >= 2
    np.testing.assert_array_equal(r["fpar"], want)


@pytest.mark.parametrize("alpha", [0.5, 0.25, 0.3])
def test_fractional_openness_scales_flux_and_area(alpha):
    """Openness alpha in (0, 1) on every face: the same interface patches, throat flow and area
    alpha times their fully open values (bit for bit for powers of two), residual still zero
    under uniform flow; openness 0 on x faces removes exactly the x-faces from the interfaces."""
    sdf = fc.chain_x(3, 24, 7.0, 3.0)
    U = 2.5
    f = fc.uniform_x(sdf.shape, U)
    one = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), f["u"], f["v"], f["w"], f["p"])
    o = np.full(sdf.shape, alpha)
    r = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), f["u"], f["v"], f["w"], f["p"],
                                 ox=o, oy=o, oz=o)
    np.testing.assert_array_equal(r["throats"], one["throats"])
    if alpha in (0.5, 0.25):
        np.testing.assert_array_equal(r["throat_flow"], alpha * one["throat_flow"])
        np.testing.assert_array_equal(r["throat_area"], alpha * one["throat_area"])
    else:
        np.testing.assert_allclose(r["throat_flow"], alpha * one["throat_flow"], rtol=1e-14)
        np.testing.assert_allclose(r["throat_area"], alpha * one["throat_area"], rtol=1e-14)
    assert np.all(r["pore_residual"] == 0.0)
    # closing every x face: no x-face interface is left, so the x-flux through throats is zero
    z = np.zeros(sdf.shape)
    r0 = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), f["u"], f["v"], f["w"], f["p"],
                                  ox=z, oy=o, oz=o)
    assert np.all(r0["throat_flow"] == 0.0)
    assert np.all(r0["pore_residual"] == 0.0)


def _blocked_sum(t):
    """The fixed blocked summation order, restated: blocks of 2048; lane l adds terms l, l + 64,
    ... serially from +0.0; pairwise tree over the 64 lanes; block totals added serially."""
    acc = 0.0
    for b in range(0, len(t), 2048):
        blk = t[b:b + 2048]
        p = []
        for lane in range(64):
            s = 0.0
            for v in blk[lane::64]:
                s += v
            p.append(s)
        sh = 1
        while sh < 64:
            for lane in range(0, 64, 2 * sh):
                p[lane] = p[lane] + p[lane + sh]
            sh *= 2
        acc += p[0]
    return acc


def test_residual_sum_order_and_accuracy():
    """The pore residuals equal an independent numpy statement of the per-face terms (q of the
    +d face, read at the +d neighbour, + for the pore on the cell side, - on the neighbour side)
    summed per pore in face order in the blocked order, bit for bit, and the exactly rounded sum
    (math.fsum) within 1e-12 of the largest term."""
    import math

    sdf = lattice48()
    spc = (1.5, 1.0, 0.5)
    f = fc.divergence_free(sdf.shape, spc, seed=7)
    r = orc.extract_network_flow(sdf, (0, 0, 0), spc, f["u"], f["v"], f["w"], f["p"],
                                 debug=True)
    lab = r["flow_lab"].reshape(sdf.shape)
    sz, sy, sx = spc
    area = (sy * sz, sx * sz, sx * sy)
    vel = (f["u"], f["v"], f["w"])
    axis = (2, 1, 0)
    n = sdf.size
    keys, terms, order = [], [], []
    for d in range(3):
        nbl = np.roll(lab, -1, axis=axis[d]).ravel()
        q = (1.0 * np.roll(vel[d], -1, axis=axis[d]) * area[d]).ravel()
        a = lab.ravel()
        face = 3 * np.arange(n) + d
        m = (a != nbl) & (a > 0)
        keys.append(a[m] - 1), terms.append(q[m]), order.append(2 * face[m])
        m = (a != nbl) & (nbl > 0)
        keys.append(nbl[m] - 1), terms.append(-q[m]), order.append(2 * face[m] + 1)
    keys, terms, order = (np.concatenate(x) for x in (keys, terms, order))
    srt = np.lexsort((order, keys))
    keys, terms = keys[srt], terms[srt]
    bounds = np.searchsorted(keys, np.arange(len(r["pore_residual"]) + 1))
    big = np.abs(terms).max()
    for k, got in enumerate(r["pore_residual"]):
        t = terms[bounds[k]:bounds[k + 1]].tolist()
        assert got == _blocked_sum(t), k
        assert abs(got - math.fsum(t)) <= 1e-12 * big
