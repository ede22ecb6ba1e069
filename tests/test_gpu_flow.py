"""HIP extract_network_flow vs the oracle (GPU).

Discrete outputs (pores, throat list) must be bit-exact. The FP64 sums are taken per throat /
pore over the faces in ascending face order, in the fixed blocked order the oracle uses too, so
Q, A, the residual, pore pressures and dp are expected bit-identical; the stated bar is 1e-12
relative (to the largest |Q| for the flux sums), checked first, then exactness.
"""
from __future__ import annotations

import numpy as np
import pytest

import flowcases as fc
import pnm_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KEYS = ("pore_pressure", "pore_residual", "throat_flow", "throat_area", "throat_dp")


def xyz(v):
    return (v[2], v[1], v[0])


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run_gpu(ctx, sdf, org, spc, f, gp=(0.0, 0.0, 0.0)):
    return ctx.extract_network_flow_d(dev(sdf), xyz(sdf.shape), xyz(org), xyz(spc), dev(f["u"]),
                                      dev(f["v"]), dev(f["w"]), dev(f["p"]), dev(f.get("ox")),
                                      dev(f.get("oy")), dev(f.get("oz")), xyz(gp))


def compare(g, o, exact=True):
    np.testing.assert_array_equal(np.ascontiguousarray(g["pores"]).view(np.uint32),
                                  np.ascontiguousarray(o["pores"]).view(np.uint32))
    np.testing.assert_array_equal(g["throats"], o["throats"])
    qs = max(1e-30, float(np.abs(o["throat_flow"]).max()) if len(o["throat_flow"]) else 1e-30)
    for k in ("throat_flow", "throat_area", "pore_residual"):
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=1e-12 * qs, err_msg=k)
    for k in ("pore_pressure", "throat_dp"):
        np.testing.assert_allclose(g[k], o[k], rtol=1e-12, atol=1e-15, err_msg=k)
    if exact:
        for k in KEYS:
            np.testing.assert_array_equal(g[k].view(np.uint64), o[k].view(np.uint64), err_msg=k)


def oracle(sdf, org, spc, f, gp=(0.0, 0.0, 0.0)):
    return orc.extract_network_flow(sdf, org, spc, f["u"], f["v"], f["w"], f["p"], f.get("ox"),
                                    f.get("oy"), f.get("oz"), gp)


@pytest.mark.parametrize("with_open", [True, False], ids=["openness", "fully_open"])
def test_reference_flow_configuration(capi_ctx, with_open):
    """The reference's distributed-flow test case (test_pnm_flow_mpi.cpp:58-61): 36x30x24
    sphere lattice, origin (0.5,-1,2), spacing (0.5,1,1.5), grad_p (1e-3,-2e-3,0.5e-3) xyz."""
    from peclet.pnm import synth

    shape = (24, 30, 36)
    sdf = synth.sphere_lattice_sdf(shape)
    f = synth.lattice_flow_fields(shape, with_openness=with_open)
    org, spc, gp = (2.0, -1.0, 0.5), (1.5, 1.0, 0.5), (0.5e-3, -2e-3, 1e-3)
    o = oracle(sdf, org, spc, f, gp)
    assert len(o["throats"]) > 0
    compare(run_gpu(capi_ctx, sdf, org, spc, f, gp), o)


@pytest.mark.parametrize("case", ["chain", "two_patches", "lattice48"])
def test_closed_form_geometries(capi_ctx, case):
    from peclet.pnm import synth

    if case == "chain":
        sdf = fc.chain_x(3, 24, 7.0, 3.0)
    elif case == "two_patches":
        sdf = fc.two_chambers_two_tubes()
    else:
        sdf = synth.chamber_sdf((48, 48, 48), 2, 7.0, 2.0, 4.0, seed=5)
    spc = (1.5, 1.0, 0.5)
    for f in (fc.uniform_x(sdf.shape, 2.5), fc.divergence_free(sdf.shape, spc, seed=3)):
        f["p"] = np.cos(np.arange(sdf.size).reshape(sdf.shape) * 0.01)
        compare(run_gpu(capi_ctx, sdf, (0.5, 1.0, -2.0), spc, f, (1e-3, 0.0, -2e-3)),
                oracle(sdf, (0.5, 1.0, -2.0), spc, f, (1e-3, 0.0, -2e-3)))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_packings_with_openness(capi_ctx, seed):
    from peclet.pnm import synth

    shape = (40, 44, 48)
    sdf, _ = synth.packing_sdf(shape, "uniform", 3.0, 6.0, 0.5, seed=100 + seed)
    f = synth.lattice_flow_fields(shape, with_openness=True)
    rng = np.random.default_rng(seed)
    f["u"] = f["u"] + rng.standard_normal(shape) * 0.1
    compare(run_gpu(capi_ctx, sdf, (0, 0, 0), (1, 1, 1), f, (0.1, 0.2, 0.3)),
            oracle(sdf, (0, 0, 0), (1, 1, 1), f, (0.1, 0.2, 0.3)))


def test_python_api_flow(capi_ctx):
    import peclet.pnm as pnm
    from peclet.pnm import synth

    shape = (24, 30, 36)
    sdf = synth.sphere_lattice_sdf(shape)
    f = synth.lattice_flow_fields(shape)
    d = pnm.extract_network_flow(sdf, [2.0, -1.0, 0.5], [1.5, 1.0, 0.5], f["u"], f["v"], f["w"],
                                 f["p"], f["ox"], f["oy"], f["oz"], [0.5e-3, -2e-3, 1e-3])
    o = oracle(sdf, (2.0, -1.0, 0.5), (1.5, 1.0, 0.5), f, (0.5e-3, -2e-3, 1e-3))
    assert [tuple(t) for t in o["throats"].tolist()] == d["throats"]
    np.testing.assert_array_equal(np.array(d["throat_flow"]), o["throat_flow"])
    np.testing.assert_array_equal(np.array(d["throat_dp"]), o["throat_dp"])
    assert len(d["pores"]) == len(o["pores"])


def test_no_pores_gives_empty(capi_ctx):
    sdf = np.full((8, 8, 8), -3.0, np.float32)
    f = fc.uniform_x(sdf.shape)
    g = run_gpu(capi_ctx, sdf, (0, 0, 0), (1, 1, 1), f)
    assert len(g["pores"]) == 0 and len(g["throats"]) == 0


def test_flow_256_chamber_lattice_properties(capi_ctx):
    """256^3 chamber-tube lattice (chambers r=12 on a 5^3 lattice, pitch 51.2, tubes
    r~U[3,8]) with a divergence-free field: residuals at rounding level; 125 chambers found;
    determinism across two runs."""
    from peclet.pnm import synth

    shape = (256, 256, 256)
    sdf = synth.chamber_sdf(shape, 5, 12.0, 3.0, 8.0, seed=3001)
    spc = (1.0, 1.0, 1.0)
    f = fc.divergence_free(shape, spc, seed=7)
    a = run_gpu(capi_ctx, sdf, (0, 0, 0), spc, f)
    b = run_gpu(capi_ctx, sdf, (0, 0, 0), spc, f)
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k])
    assert len(a["pores"]) >= 125
    scale = float(np.abs(f["u"]).max())
    assert np.abs(a["pore_residual"]).max() < 1e-9 * scale


@pytest.mark.parametrize("case", ["side_by_side", "pack2", "pack3"])
@pytest.mark.parametrize("alpha", [1.0, 0.3])
def test_films_bridging_core_patches(capi_ctx, case, alpha):
    """Geometries where one film component touches two or three core patches of its pore pair
    (the film must attach to the smallest), with full or fractional openness: bit-exact."""
    from peclet.pnm import synth

    if case == "side_by_side":
        sdf = fc.two_chambers_side_by_side_tubes()
    else:
        sdf, _ = synth.packing_sdf((32, 32, 32), "uniform", 3.0, 5.0, 0.55, seed=int(case[-1]))
    f = fc.divergence_free(sdf.shape, (1.0, 1.0, 1.0), seed=3)
    if alpha != 1.0:
        rng = np.random.default_rng(4)
        for k in ("ox", "oy", "oz"):
            f[k] = np.where(rng.random(sdf.shape) < 0.1, 0.0, alpha)
    o = oracle(sdf, (0, 0, 0), (1, 1, 1), f)
    assert len(o["throats"]) > 0
    compare(run_gpu(capi_ctx, sdf, (0, 0, 0), (1, 1, 1), f), o)
