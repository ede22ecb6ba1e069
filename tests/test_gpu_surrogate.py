"""Throat surrogate kernels on the GPU against the FP64 oracle (oracle/surrogate_oracle.py):
BF16-MFMA convolutions forward, backward (weight and input gradients), dense head, loss, Adam and
the patch gather."""
from __future__ import annotations

import numpy as np
import pytest

import surrogate_oracle as so

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def group_errors(g_gpu, g_ref):
    off, _ = so.layout()
    out = {}
    for name, (o, shape) in off.items():
        n = int(np.prod(shape))
        a, b = g_gpu[o:o + n], g_ref[o:o + n]
        out[name] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    return out


@pytest.mark.parametrize("B", [3, 16])
def test_forward_backward_vs_oracle(capi_ctx, B):
    from peclet.pnm import capi

    rng = np.random.default_rng(100 + B)
    p = capi.sur_init(11 + B)
    P = so.unpack(p)
    P["b0"] += rng.normal(0, 0.05, 16)
    P["b1"] += rng.normal(0, 0.05, 32)
    P["b4"] += rng.normal(0, 0.05, 64)
    p = so.pack(P).astype(np.float32)
    x = rng.normal(0, 1, size=(B, 16, 16, 16)).astype(np.float32)
    s = rng.normal(0, 1, size=(B, 4)).astype(np.float32)
    dy = rng.normal(0, 1, size=B).astype(np.float32)
    pd, xd, sd, dyd = dev(p), dev(x), dev(s), dev(dy)
    y = torch.empty(B, device="cuda")
    g = torch.zeros(p.size, device="cuda")
    capi_ctx.sur_forward_d(B, xd, sd, pd, y)
    capi_ctx.sur_backward_d(B, dyd, pd, g)
    y_ref, cache = so.forward(p.astype(np.float64), x, s, rounding=True)
    g_ref = so.backward(cache, dy)
    yg = y.cpu().numpy()
    assert np.abs(yg - y_ref).max() <= 2e-3 * max(1.0, np.abs(y_ref).max()), (yg, y_ref)
    errs = group_errors(g.cpu().numpy().astype(np.float64), g_ref)
    assert max(errs.values()) < 2e-2, errs
    # padded K columns of the conv weights get no gradient
    gg = g.cpu().numpy()
    off, _ = so.layout()
    for l, (_, ci, co, kp) in enumerate(so.CONV):
        o, _ = off[f"w{l}"]
        assert np.all(gg[o:o + co * kp].reshape(co, kp)[:, 27 * ci:] == 0.0)


def test_loss_and_adam(capi_ctx):
    from peclet.pnm import capi

    rng = np.random.default_rng(3)
    B = 37
    yv = rng.normal(size=B).astype(np.float32)
    tv = rng.normal(size=B).astype(np.float32)
    dy = torch.empty(B, device="cuda")
    loss = torch.empty(1, device="cuda")
    capi_ctx.sur_loss_d(B, dev(yv), dev(tv), dy, loss)
    l_ref, dy_ref = so.mse(yv, tv)
    assert abs(loss.item() - l_ref) < 1e-5 * max(1.0, l_ref)
    np.testing.assert_allclose(dy.cpu().numpy(), dy_ref, rtol=1e-5, atol=1e-7)
    capi_ctx.sur_loss_d(B, dev(yv), dev(tv), dy, loss, delta=0.5)
    l_ref, dy_ref = so.huber(yv, tv, 0.5)
    assert abs(loss.item() - l_ref) < 1e-5 * max(1.0, l_ref)
    np.testing.assert_allclose(dy.cpu().numpy(), dy_ref, rtol=1e-5, atol=1e-7)
    n = capi.sur_param_count()
    p0 = rng.normal(size=n).astype(np.float32)
    gv = rng.normal(size=n).astype(np.float32)
    p, m, v = dev(p0), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    for step in (1, 2):
        capi_ctx.sur_adam_d(p, dev(gv), m, v, step, lr=1e-2)
    pe, me, ve = p0.astype(np.float64), np.zeros(n), np.zeros(n)
    for step in (1, 2):
        me = 0.9 * me + 0.1 * gv
        ve = 0.999 * ve + 0.001 * gv.astype(np.float64) ** 2
        pe = pe - 1e-2 * (me / (1 - 0.9 ** step)) / (np.sqrt(ve / (1 - 0.999 ** step)) + 1e-8)
    np.testing.assert_allclose(p.cpu().numpy(), pe, rtol=1e-5, atol=1e-6)


def test_patches_gather(capi_ctx):
    rng = np.random.default_rng(9)
    shape = (20, 24, 28)  # z, y, x
    sdf = rng.normal(size=shape).astype(np.float32)
    centers = np.array([[0.2, 0.0, 0.0], [27.9, 23.5, 19.99], [13.0, 5.5, 2.0]], np.float32)
    out = torch.empty(3 * 4096, device="cuda")
    capi_ctx.sur_patches_d(dev(sdf), (28, 24, 20), dev(centers), 3, 4.0, out)
    got = out.cpu().numpy().reshape(3, 16, 16, 16)
    for b, (cx, cy, cz) in enumerate(centers):
        ix = (np.floor(cx).astype(int) - 7 + np.arange(16)) % 28
        iy = (np.floor(cy).astype(int) - 7 + np.arange(16)) % 24
        iz = (np.floor(cz).astype(int) - 7 + np.arange(16)) % 20
        want = sdf[np.ix_(iz, iy, ix)] / 4.0
        np.testing.assert_array_equal(got[b], want)


def test_stream_switch_orders_pending_work():
    """Forward on the context's own stream (asynchronous), then a switch to another stream and
    the backward pass there: the switch must make the new stream wait for the queued forward, so
    outputs and gradients equal the single-stream run bit for bit."""
    from peclet.pnm import capi

    rng = np.random.default_rng(21)
    B = 64
    p = capi.sur_init(5)
    x = rng.normal(0, 1, size=(B, 16, 16, 16)).astype(np.float32)
    s = rng.normal(0, 1, size=(B, 4)).astype(np.float32)
    dy = rng.normal(0, 1, size=B).astype(np.float32)
    pd, xd, sd, dyd = dev(p), dev(x), dev(s), dev(dy)
    ctx = capi.Context(0)
    try:
        y1, g1 = torch.empty(B, device="cuda"), torch.zeros(p.size, device="cuda")
        y2, g2 = torch.empty(B, device="cuda"), torch.zeros(p.size, device="cuda")
        torch.cuda.synchronize()
        ctx.sur_forward_d(B, xd, sd, pd, y1)
        ctx.sur_backward_d(B, dyd, pd, g1)
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        for _ in range(3):
            ctx.sur_forward_d(B, xd, sd, pd, y2)
            ctx.set_stream(side.cuda_stream)
            ctx.sur_backward_d(B, dyd, pd, g2)
            ctx.set_stream(None)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(y2.cpu().numpy(), y1.cpu().numpy())
            np.testing.assert_array_equal(g2.cpu().numpy(), g1.cpu().numpy())
    finally:
        ctx.close()


@pytest.mark.parametrize("B,scale", [(300, 8.0), (1000, 8.0), (300, 6.0)])
def test_fused_inference_matches_forward(capi_ctx, B, scale):
    """pnm_sur_infer_d (patch gather + network in one kernel) against pnm_sur_patches_d +
    pnm_sur_forward_d on the same throats, centres spread over the whole periodic volume
    (patches wrap), more throats than workgroups (the persistent loop), and against the FP64
    oracle on a subset. Both round the same values to BF16 at the same points; the FP32 sums
    differ in order (K-steps of 32 in the fused kernel, 16 in the forward GEMM; the global
    average per z-slice), so the bar is 5e-5 of the output scale. scale 8 takes the exact
    multiply-by-inverse path, scale 6 the division."""
    from peclet.pnm import capi

    rng = np.random.default_rng(7 + B)
    shape = (40, 48, 56)  # z, y, x
    sdf = (rng.normal(0, 4, size=shape)).astype(np.float32)
    p = capi.sur_init(3)
    P = so.unpack(p)
    for k in ("b0", "b1", "b2", "b4"):
        P[k] += rng.normal(0, 0.05, P[k].shape)
    p = so.pack(P).astype(np.float32)
    centers = (rng.random((B, 3)) * np.array([56, 48, 40])).astype(np.float32)
    scal = rng.normal(0, 1, size=(B, 4)).astype(np.float32)
    sd, cd, scd, pd = dev(sdf), dev(centers), dev(scal), dev(p)
    fused = torch.empty(B, device="cuda")
    capi_ctx.sur_infer_d(sd, (56, 48, 40), cd, scd, B, scale, pd, fused)
    patches = torch.empty(B * 4096, device="cuda")
    capi_ctx.sur_patches_d(sd, (56, 48, 40), cd, B, scale, patches)
    ref = torch.empty(B, device="cuda")
    capi_ctx.sur_forward_d(B, patches, scd, pd, ref)
    f, r = fused.cpu().numpy(), ref.cpu().numpy()
    ys = max(1.0, float(np.abs(r).max()))
    print(f"[fused B={B}] max |fused - forward| = {np.abs(f - r).max():.3e}, "
          f"bitwise equal {np.array_equal(f, r)}", flush=True)
    assert np.abs(f - r).max() <= 5e-5 * ys
    k = 24
    y_ref, _ = so.forward(p.astype(np.float64), patches[: k * 4096].cpu().numpy().reshape(
        k, 16, 16, 16), scal[:k], rounding=True)
    assert np.abs(f[:k] - y_ref).max() <= 2e-3 * max(1.0, np.abs(y_ref).max())
