"""FP64 CPU oracle of the throat-conductance surrogate (test infrastructure only: the tests and
the benchmark's checks call it; the product never does).

It restates the network the GPU kernels compute (peclet-pnm_amd/csrc/k_surrogate.hip) in numpy:
16^3 patch -> conv3^3 1->16, ReLU, avg-pool 2 -> conv3^3 16->32, ReLU, avg-pool 2 -> conv3^3
32->32, ReLU, global average -> concat 4 scalars -> dense 36->64, ReLU -> dense 64->1, and its
backward pass by the chain rule. With `bf16=True` the operands of every convolution GEMM are
rounded to BF16 where the GPU rounds them (conv inputs, conv weights, and the pre-activation
gradients that feed the weight and input gradients); everything else is FP64.

The oracle's own backward pass is pinned by central finite differences in FP64
(tests/test_surrogate_oracle.py), which fixes it by the mathematics alone.

Parameter layout (flat float32, same as the GPU): for each conv l: W_l [CO][Kp] (k = tap*CI+ci,
tap = (dz+1)*9 + (dy+1)*3 + (dx+1), columns k >= 27*CI are zero padding), b_l [CO]; then W4 [64][36],
b4 [64], W5 [64], b5 [1].
"""
from __future__ import annotations

import numpy as np

CONV = [(16, 1, 16, 32), (8, 16, 32, 432), (4, 32, 32, 864)]  # (D, CI, CO, Kp)


def layout():
    off, o = {}, 0
    for l, (_, ci, co, kp) in enumerate(CONV):
        off[f"w{l}"] = (o, (co, kp))
        o += co * kp
        off[f"b{l}"] = (o, (co,))
        o += co
    for name, shape in (("W4", (64, 36)), ("b4", (64,)), ("W5", (64,)), ("b5", (1,))):
        off[name] = (o, shape)
        o += int(np.prod(shape))
    return off, o


def unpack(flat):
    off, n = layout()
    assert flat.size == n, (flat.size, n)
    return {k: flat[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for k, (o, s) in off.items()}


def pack(d):
    off, n = layout()
    out = np.zeros(n, dtype=np.float64)
    for k, (o, s) in off.items():
        out[o:o + int(np.prod(s))] = np.asarray(d[k], dtype=np.float64).reshape(-1)
    return out


def bf16(x):
    """Round to the nearest BF16 (ties to even) through float32, returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = u.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(nan, np.nan, r)


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def im2col(x, D, CI, Kp):
    """x [B, D, D, D, CI] -> col [B, D, D, D, Kp] with zero padding."""
    B = x.shape[0]
    xp = np.zeros((B, D + 2, D + 2, D + 2, CI))
    xp[:, 1:-1, 1:-1, 1:-1, :] = x
    col = np.zeros((B, D, D, D, Kp))
    for t, (dz, dy, dx) in enumerate(TAPS):
        col[..., t * CI:(t + 1) * CI] = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + D,
                                           1 + dx:1 + dx + D, :]
    return col


def col2im(dcol, D, CI):
    """Adjoint of im2col: dcol [B, D, D, D, Kp] -> dx [B, D, D, D, CI]."""
    B = dcol.shape[0]
    dxp = np.zeros((B, D + 2, D + 2, D + 2, CI))
    for t, (dz, dy, dx) in enumerate(TAPS):
        dxp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + D, 1 + dx:1 + dx + D, :] += \
            dcol[..., t * CI:(t + 1) * CI]
    return dxp[:, 1:-1, 1:-1, 1:-1, :]


def pool(a):
    B, D, _, _, C = a.shape
    h = D // 2
    return a.reshape(B, h, 2, h, 2, h, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(dp):
    return np.repeat(np.repeat(np.repeat(dp, 2, 1), 2, 2), 2, 3) / 8.0


def forward(params, patches, scalars, rounding=True):
    """patches [B, 16, 16, 16] (z, y, x), scalars [B, 4] -> (y [B], cache)."""
    P = unpack(params) if not isinstance(params, dict) else params
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    B = patches.shape[0]
    x = np.asarray(patches, dtype=np.float64).reshape(B, 16, 16, 16, 1)
    cache = {"B": B, "rounding": rounding}
    for l, (D, CI, CO, Kp) in enumerate(CONV):
        col = im2col(r(x), D, CI, Kp)
        W = r(P[f"w{l}"])
        z = col @ W.T + P[f"b{l}"]
        a = np.maximum(z, 0.0)
        cache[f"col{l}"], cache[f"a{l}"] = col, a
        x = pool(a) if l < 2 else a
    g = cache["a2"].reshape(B, 64, 32).mean(axis=1)
    h = np.concatenate([g, np.asarray(scalars, dtype=np.float64)], axis=1)
    z4 = h @ P["W4"].T + P["b4"]
    r4 = np.maximum(z4, 0.0)
    y = r4 @ P["W5"] + P["b5"][0]
    cache.update(h=h, r4=r4, P=P)
    return y, cache


def backward(cache, dy):
    """Gradients of sum_b dy[b] * y[b] as a flat float64 vector (GPU layout)."""
    P, B = cache["P"], cache["B"]
    r = bf16 if cache["rounding"] else (lambda v: np.asarray(v, dtype=np.float64))
    dy = np.asarray(dy, dtype=np.float64)
    G = {}
    r4, h = cache["r4"], cache["h"]
    G["W5"] = dy @ r4
    G["b5"] = np.array([dy.sum()])
    dz4 = np.outer(dy, P["W5"]) * (r4 > 0)
    G["W4"] = dz4.T @ h
    G["b4"] = dz4.sum(axis=0)
    dh = dz4 @ P["W4"]
    dg = dh[:, :32]
    da = np.broadcast_to((dg / 64.0)[:, None, None, None, :], (B, 4, 4, 4, 32))
    for l in (2, 1, 0):
        D, CI, CO, Kp = CONV[l]
        dz = da * (cache[f"a{l}"] > 0)
        G[f"b{l}"] = dz.reshape(-1, CO).sum(axis=0)
        dzr = r(dz).reshape(-1, CO)
        G[f"w{l}"] = dzr.T @ cache[f"col{l}"].reshape(-1, Kp)
        if l > 0:
            dcol = (dzr @ r(P[f"w{l}"])).reshape(B, D, D, D, Kp)
            dx = col2im(dcol, D, CI)
            da = pool_bwd(dx)
    return pack(G)


def mse(y, t):
    e = np.asarray(y, dtype=np.float64) - np.asarray(t, dtype=np.float64)
    return float((e * e).mean()), 2.0 * e / e.size


def huber(y, t, delta):
    """Mean Huber loss and its gradient (threshold delta)."""
    e = np.asarray(y, dtype=np.float64) - np.asarray(t, dtype=np.float64)
    a = np.abs(e)
    l = np.where(a <= delta, 0.5 * e * e, delta * (a - 0.5 * delta))
    g = np.where(a <= delta, e, delta * np.sign(e))
    return float(l.mean()), g / e.size
