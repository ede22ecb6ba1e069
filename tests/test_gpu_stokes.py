"""Project MAC Stokes solver (GPU) and the physics validation of extract_network_flow on its fields.

- closed form: plane Poiseuille flow in a periodic slit (the discrete solution of the staircase
  MAC system is the parabola through the two solid rows, exactly);
- an independent numpy statement of the discrete operator gives the residual of the GPU solution;
- the chamber-tube chain, the asymmetric tube lattice and the parallel capsules of
  scripts/verify_network_flow.py:23-96, :100-141, :184-249 (geometry as there; our solver uses
  binary openness, so every flux balance closes to the solver tolerance).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def solve(ctx, sdf_zyx, spacing_zyx=(1.0, 1.0, 1.0), force_zyx=(0.0, 0.0, 1.0), mu=1.0,
          rtol=1e-11, max_iter=50000):
    nz, ny, nx = sdf_zyx.shape
    s = torch.from_numpy(np.ascontiguousarray(sdf_zyx, dtype=np.float32)).cuda()
    f = [torch.empty(sdf_zyx.size, dtype=torch.float64, device="cuda") for _ in range(7)]
    it, rr = ctx.stokes_solve_d(s, (nx, ny, nz), spacing_zyx[::-1], force_zyx[::-1], *f, mu=mu,
                                rtol=rtol, max_iter=max_iter)
    u, v, w, p, ox, oy, oz = (a.view(sdf_zyx.shape).cpu().numpy() for a in f)
    return dict(u=u, v=v, w=w, p=p, ox=ox, oy=oy, oz=oz, iterations=it, rel_residual=rr)


def numpy_operator(sdf, x, spacing_zyx, mu, force_zyx):
    """b - M x of the discrete Stokes system, restated with numpy rolls (axes z, y, x)."""
    fl = sdf > 0
    hz, hy, hx = spacing_zyx
    h = {0: hz, 1: hy, 2: hx}
    comps = {2: "u", 1: "v", 0: "w"}  # x-face velocity u lives on axis 2 (x)
    opn = {a: fl & np.roll(fl, 1, axis=a) for a in (0, 1, 2)}  # -a face of each cell
    out = {}
    fz, fy, fx = force_zyx
    fvec = {2: fx, 1: fy, 0: fz}
    p = x["p"]
    for a, name in comps.items():
        q = x[name]
        lap = sum((2 * q - np.roll(q, 1, b) - np.roll(q, -1, b)) / h[b] ** 2 for b in (0, 1, 2))
        grad = (p - np.roll(p, 1, a)) / h[a]
        r = np.where(opn[a], fvec[a] - (mu * lap + grad), 0.0)
        out[name] = r
    div = sum((x[comps[a]] - np.roll(x[comps[a]], -1, a)) / h[a] for a in (0, 1, 2))
    out["p"] = np.where(fl, -div, 0.0)
    bnorm = np.sqrt(sum(np.where(opn[a], fvec[a], 0.0) ** 2 for a in (0, 1, 2)).sum())
    rnorm = np.sqrt(sum((out[k] ** 2).sum() for k in out))
    return rnorm / bnorm


@pytest.mark.parametrize("spacing", [(1.0, 1.0, 1.0), (2.0, 0.5, 1.0)])
def test_slit_poiseuille_closed_form(capi_ctx, spacing):
    H = 9
    nz, ny, nx = 3, H + 1, 5
    sdf = np.full((nz, ny, nx), 1.0, np.float32)
    sdf[:, H, :] = -1.0  # one solid row; periodic in y, so the walls sit at y = -1 and y = H
    mu, fx = 0.7, 0.3
    r = solve(capi_ctx, sdf, spacing, (0.0, 0.0, fx), mu=mu)
    hy = spacing[1]
    j = np.arange(H)
    expect = fx * hy * hy / (2 * mu) * (j + 1) * (H - j)
    u = r["u"]
    assert r["rel_residual"] < 1e-9
    np.testing.assert_allclose(u[:, :H, :], np.broadcast_to(expect[None, :, None], (nz, H, nx)),
                               rtol=1e-8, atol=0)
    assert np.all(u[:, H, :] == 0.0)
    assert np.abs(r["v"]).max() < 1e-9 * expect.max()
    assert np.abs(r["w"]).max() < 1e-9 * expect.max()
    assert np.ptp(r["p"][:, :H, :]) < 1e-8 * fx * H * spacing[2]


def test_residual_by_independent_operator(capi_ctx):
    from peclet.pnm import synth

    sdf, _ = synth.packing_sdf((24, 28, 32), "uniform", 3.0, 5.0, 0.55, seed=11)
    spacing = (1.0, 1.5, 0.75)
    force = (0.2, -0.1, 1.0)
    r = solve(capi_ctx, sdf, spacing, force, mu=0.5, rtol=1e-11)
    rel = numpy_operator(sdf, r, spacing, 0.5, force)
    assert rel < 1e-8, rel
    assert abs(rel - r["rel_residual"]) < 1e-6 + 0.5 * rel
    # openness outputs are the binary masks
    fl = sdf > 0
    np.testing.assert_array_equal(r["ox"], (fl & np.roll(fl, 1, 2)).astype(np.float64))
    np.testing.assert_array_equal(r["oz"], (fl & np.roll(fl, 1, 0)).astype(np.float64))


def chamber_tube_chain():
    NX, NY, NZ = 48, 32, 32
    z, y, x = np.meshgrid(*(np.arange(d, dtype=np.float64) for d in (NZ, NY, NX)), indexing="ij")
    sdf = np.full(x.shape, -1e30)
    for cx in (12.0, 36.0):
        dx = np.abs(x - cx)
        dx = np.minimum(dx, NX - dx)
        sdf = np.maximum(sdf, 9.0 - np.sqrt(dx ** 2 + (y - 16.0) ** 2 + (z - 16.0) ** 2))
    sdf = np.maximum(sdf, 4.0 - np.sqrt((y - 16.0) ** 2 + (z - 16.0) ** 2))
    return sdf.astype(np.float32)


def tube_lattice():
    N = 48
    z, y, x = np.meshgrid(*(np.arange(N, dtype=np.float64) for _ in range(3)), indexing="ij")
    sdf = np.full(x.shape, -1e30)
    for cy, cz, rt in ((12.0, 12.0, 3.0), (12.0, 36.0, 4.0), (36.0, 12.0, 5.0), (36.0, 36.0, 6.0)):
        for cx in (12.0, 36.0):
            dx = np.abs(x - cx)
            dx = np.minimum(dx, N - dx)
            sdf = np.maximum(sdf, 9.0 - np.sqrt(dx ** 2 + (y - cy) ** 2 + (z - cz) ** 2))
        sdf = np.maximum(sdf, rt - np.sqrt((y - cy) ** 2 + (z - cz) ** 2))
    return sdf.astype(np.float32)


def parallel_capsules():
    NX, NY, NZ = 48, 32, 32
    z, y, x = np.meshgrid(*(np.arange(d, dtype=np.float64) for d in (NZ, NY, NX)), indexing="ij")
    sdf = np.full(x.shape, -1e30)
    for cx in (12.0, 36.0):
        dx = np.abs(x - cx)
        dx = np.minimum(dx, NX - dx)
        sdf = np.maximum(sdf, 8.0 - np.sqrt(dx ** 2 + (y - 16.0) ** 2 + (z - 16.0) ** 2))
    for cy, rt in ((12.0, 3.0), (20.0, 4.2)):
        xs = np.clip(x, 12.0, 36.0)
        sdf = np.maximum(sdf, rt - np.sqrt((x - xs) ** 2 + (y - cy) ** 2 + (z - 16.0) ** 2))
    xw = np.where(x < 12.0, x + NX, x)
    xs = np.clip(xw, 36.0, 12.0 + NX)
    sdf = np.maximum(sdf, 4.0 - np.sqrt((xw - xs) ** 2 + (y - 16.0) ** 2 + (z - 16.0) ** 2))
    return sdf.astype(np.float32)


MU, FX = 0.05, 1.0e-3


def flow_network(capi_ctx, sdf):
    import peclet.pnm as pnm

    r = solve(capi_ctx, sdf, (1.0, 1.0, 1.0), (0.0, 0.0, FX), mu=MU, rtol=1e-12)
    assert r["rel_residual"] < 1e-9
    net = pnm.extract_network_flow(sdf, [0, 0, 0], [1, 1, 1], r["u"], r["v"], r["w"], r["p"],
                                   ox=r["ox"], oy=r["oy"], oz=r["oz"], grad_p_zyx=[0, 0, -FX])
    planes = np.array([(r["ox"][:, :, i] * r["u"][:, :, i]).sum() for i in range(sdf.shape[2])])
    return r, net, planes


def test_chamber_tube_chain(capi_ctx):
    sdf = chamber_tube_chain()
    r, net, planes = flow_network(capi_ctx, sdf)
    F = planes.mean()
    assert F > 0
    assert np.abs(planes - F).max() < 1e-8 * F  # divergence-free: the plane flux is constant
    Q, dp = np.array(net["throat_flow"]), np.array(net["throat_dp"])
    res = np.array(net["pore_residual"])
    assert len(net["pores"]) == 3
    assert len(net["throats"]) == len(net["pores"])  # a single loop
    assert np.abs(res).max() < 1e-7 * F
    assert np.abs(np.abs(Q) - F).max() < 1e-7 * F
    assert np.all(Q * dp > 0)


def test_tube_lattice(capi_ctx):
    sdf = tube_lattice()
    r, net, planes = flow_network(capi_ctx, sdf)
    F = planes.mean()
    Q, dp = np.array(net["throat_flow"]), np.array(net["throat_dp"])
    res = np.array(net["pore_residual"])
    assert np.abs(res).max() < 1e-7 * F
    assert np.all(Q * dp > 0)
    # four independent tube columns of three throats each, each throat carrying its column flux
    assert abs(np.abs(Q).sum() / 3.0 - F) < 1e-6 * F


def test_parallel_capsules(capi_ctx):
    sdf = parallel_capsules()
    r, net, planes = flow_network(capi_ctx, sdf)
    F = planes.mean()
    Q, dp = np.array(net["throat_flow"]), np.array(net["throat_dp"])
    res = np.array(net["pore_residual"])
    th = [tuple(t) for t in net["throats"]]
    assert len(net["pores"]) == 3
    par = [k for k, ab in enumerate(th) if ab == (1, 2)]
    assert len(par) == 2  # two disjoint interfaces between the chambers: two throats
    assert np.abs(res).max() < 1e-7 * F
    qa, qb = Q[par[0]], Q[par[1]]
    assert abs(qa + qb - F) < 1e-7 * F
    assert qa > 0 and qb > 0 and max(qa, qb) > 1.5 * min(qa, qb)
    wrap = [k for k, ab in enumerate(th) if ab != (1, 2)]
    assert np.all(np.abs(np.abs(Q[wrap]) - F) < 1e-7 * F)
    assert np.all(Q * dp > 0)
