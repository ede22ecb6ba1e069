"""extract_network_flow at its configuration: the 256^3 chamber lattice (5^3 chambers r = 12 at
pitch 51.2, tubes r ~ U[3, 8], seed 3001) with FP64 MAC fields from the project Stokes solver,
body force along x and binary openness. At full size the checks are size-independent:

- the Stokes fields converge (relative residual below 1e-9);
- per-pore mass balance of the network: each pore's residual (net outflow over its boundary
  faces) equals the signed sum of its throats' flows, to rounding (faces to non-pore basins are
  closed, so every open boundary face belongs to a throat);
- the residual itself sits at the solver tolerance relative to the largest throat flow;
- the network's dissipation sum(Q * dp) is positive and every throat has g = Q/dp > 0;
- two calls are bit-identical;
- the throat geometry without fields (pnm_throat_geometry_d) equals the flow path's.

A 96^3 chamber lattice with Stokes fields and openness is compared with the oracle bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import pnm_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def stokes(ctx, sdf, force_x=1.0, rtol=1e-10):
    n = sdf.size
    dims = sdf.shape[::-1]
    s = torch.from_numpy(np.ascontiguousarray(sdf)).cuda()
    f = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(7)]
    it, rr = ctx.stokes_solve_d(s, dims, (1, 1, 1), (force_x, 0.0, 0.0), *f, mu=1.0, rtol=rtol,
                                max_iter=40000)
    return s, f, it, rr


def flow(ctx, s, dims, f, force_x=1.0):
    return ctx.extract_network_flow_d(s, dims, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), *f[:4],
                                      ox=f[4], oy=f[5], oz=f[6], grad_p_xyz=(-force_x, 0.0, 0.0))


def test_chamber_lattice_256_stokes_openness(capi_ctx):
    from peclet.pnm import synth

    shape = (256, 256, 256)
    sdf = synth.chamber_sdf(shape, m=5, chamber_r=12.0, tube_rmin=3.0, tube_rmax=8.0, seed=3001)
    print("[flow_config] 256^3 chamber lattice: Stokes solve", flush=True)
    s, f, it, rr = stokes(capi_ctx, sdf)
    assert rr < 1e-9, (it, rr)
    # binary openness as the solver defines it: open iff both cells are fluid
    ox = f[4].view(shape)
    fl = s.view(shape) > 0
    assert torch.equal(ox > 0, fl & torch.roll(fl, 1, dims=2))
    r = flow(capi_ctx, s, shape[::-1], f)
    r2 = flow(capi_ctx, s, shape[::-1], f)
    for k in ("pore_residual", "throat_flow", "throat_area", "throat_dp", "pore_pressure"):
        np.testing.assert_array_equal(r[k].view(np.uint64), r2[k].view(np.uint64), err_msg=k)
    np.testing.assert_array_equal(r["throats"], r2["throats"])
    q, dp, th = r["throat_flow"], r["throat_dp"], r["throats"]
    npores = len(r["pores"])
    assert npores >= 125 and len(q) >= 300
    qmax = np.abs(q).max()
    balance = np.zeros(npores)
    np.add.at(balance, th[:, 0] - 1, q)
    np.add.at(balance, th[:, 1] - 1, -q)
    assert np.abs(balance - r["pore_residual"]).max() <= 1e-10 * qmax
    assert np.abs(r["pore_residual"]).max() <= 1e-8 * qmax
    g = q / dp
    assert (q * dp).sum() > 0
    print(f"[flow_config] {len(q)} throats, g > 0 on {(g > 0).mean():.4f}, "
          f"max residual / max Q {np.abs(r['pore_residual']).max() / qmax:.2e}", flush=True)
    assert (g > 0).all(), (g > 0).mean()
    # the throat geometry without fields is the flow path's under this openness, bit for bit
    geo = capi_ctx.throat_geometry_d(s, shape[::-1], (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    np.testing.assert_array_equal(geo["throats"], r["throats"])
    np.testing.assert_array_equal(np.ascontiguousarray(geo["pores"]).view(np.uint32),
                                  np.ascontiguousarray(r["pores"]).view(np.uint32))
    for k in ("throat_area", "throat_center"):
        np.testing.assert_array_equal(geo[k].view(np.uint64), r[k].view(np.uint64), err_msg=k)


def test_chamber_lattice_96_stokes_vs_oracle(capi_ctx):
    from peclet.pnm import synth

    shape = (96, 96, 96)
    sdf = synth.chamber_sdf(shape, m=2, chamber_r=12.0, tube_rmin=3.0, tube_rmax=8.0, seed=3002)
    s, f, it, rr = stokes(capi_ctx, sdf)
    assert rr < 1e-9
    g = flow(capi_ctx, s, shape[::-1], f)
    u, v, w, p, ox, oy, oz = (a.view(shape).cpu().numpy() for a in f)
    o = orc.extract_network_flow(sdf, (0, 0, 0), (1, 1, 1), u, v, w, p, ox, oy, oz,
                                 grad_p_zyx=(0.0, 0.0, -1.0))
    assert len(o["throats"]) > 0
    np.testing.assert_array_equal(np.ascontiguousarray(g["pores"]).view(np.uint32),
                                  np.ascontiguousarray(o["pores"]).view(np.uint32))
    np.testing.assert_array_equal(g["throats"], o["throats"])
    for k in ("pore_pressure", "pore_residual", "throat_flow", "throat_area", "throat_dp"):
        np.testing.assert_array_equal(g[k].view(np.uint64), o[k].view(np.uint64), err_msg=k)
