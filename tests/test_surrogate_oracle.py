"""The surrogate's FP64 oracle, pinned without the GPU:
- its backward pass against central finite differences (FP64, no rounding);
- im2col/col2im are adjoint; pooling forward/backward are adjoint;
- its BF16 rounding against torch's float32 -> bfloat16 conversion (round to nearest even);
- its parameter layout against the library's (pnm_sur_param_count / pnm_sur_init)."""
from __future__ import annotations

import numpy as np
import pytest

import surrogate_oracle as so


def random_params(rng, scale=0.3):
    off, n = so.layout()
    p = rng.normal(0, scale, size=n)
    for l, (_, ci, co, kp) in enumerate(so.CONV):
        o, _ = off[f"w{l}"]
        w = p[o:o + co * kp].reshape(co, kp)
        w[:, 27 * ci:] = 0.0
        w *= np.sqrt(2.0 / (27 * ci)) / scale
    return p


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = random_params(rng)
    B = 2
    x = rng.normal(0, 1, size=(B, 16, 16, 16))
    s = rng.normal(0, 1, size=(B, 4))
    dy = rng.normal(0, 1, size=B)
    _, cache = so.forward(p, x, s, rounding=False)
    g = so.backward(cache, dy)

    def L(q):
        y, _ = so.forward(q, x, s, rounding=False)
        return float(dy @ y)

    off, _ = so.layout()
    idx = []
    for name, (o, shape) in off.items():
        n = int(np.prod(shape))
        cand = np.arange(o, o + n)
        if name.startswith("w"):
            l = int(name[1])
            ci, kp = so.CONV[l][1], so.CONV[l][3]
            cand = cand[(cand - o) % kp < 27 * ci]
        idx += list(rng.choice(cand, size=min(6, cand.size), replace=False))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = L(q)
        q[i] -= 2 * h
        lm = L(q)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_im2col_col2im_adjoint():
    rng = np.random.default_rng(1)
    for D, CI, _, Kp in so.CONV:
        x = rng.normal(size=(2, D, D, D, CI))
        c = rng.normal(size=(2, D, D, D, Kp))
        c[..., 27 * CI:] = 0.0
        lhs = float((so.im2col(x, D, CI, Kp) * c).sum())
        rhs = float((x * so.col2im(c, D, CI)).sum())
        assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs))


def test_pool_adjoint():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(3, 8, 8, 8, 5))
    d = rng.normal(size=(3, 4, 4, 4, 5))
    assert abs(float((so.pool(a) * d).sum()) - float((a * so.pool_bwd(d)).sum())) < 1e-10


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 10, size=20000).astype(np.float32),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, -0.0,
                                  3.3895314e38], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(so.bf16(x), ref)


def test_layout_matches_library():
    from peclet.pnm import capi

    _, n = so.layout()
    assert capi.sur_param_count() == n
    p = capi.sur_init(7)
    assert p.dtype == np.float32 and p.size == n
    P = so.unpack(p)
    for l, (_, ci, co, kp) in enumerate(so.CONV):
        w = P[f"w{l}"]
        assert np.all(w[:, 27 * ci:] == 0.0)  # K padding
        sd = w[:, :27 * ci].std()
        assert 0.7 * np.sqrt(2.0 / (27 * ci)) < sd < 1.3 * np.sqrt(2.0 / (27 * ci))
        assert np.all(P[f"b{l}"] == 0.0)
    np.testing.assert_array_equal(capi.sur_init(7), p)  # deterministic
