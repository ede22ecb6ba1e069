"""Throat-conductance surrogate: learns log g, g = Q/dp, from extract_network_flow's own output.

The network and its training run on the GPU through the C-ABI (pnm_sur_*: BF16-MFMA convolutions,
FP32 dense head, Adam). The plain bookkeeping (extract_network_flow on solved fields) produces
the targets and remains the fallback and the yardstick of the surrogate's error.

Per throat the inputs are
  - a 16^3 SDF patch centred on the interface centroid (throat_center), scaled by 1/8;
  - 4 scalars: log throat area, log radius of each pore (SDF at its peak), log min-image pore
    separation, standardised with the training set's mean and spread.
The throat geometry (throats, areas, centroids, pore positions and radii) comes from
pnm_throat_geometry_d: the segmentation (computed there, or one volume of a batched segmentation)
and the interface faces under the solver's binary openness, with no fields, so inference needs
only the SDF. Its throats are exactly those extract_network_flow returns on the solved fields.
"""
from __future__ import annotations

import numpy as np

from . import capi

PATCH_SCALE = 8.0


def _torch():
    import torch

    return torch


def throat_inputs(net, shape_zyx, spacing_zyx=(1.0, 1.0, 1.0), origin_zyx=(0.0, 0.0, 0.0)):
    """Host arrays (centers_vox [nt,3] xyz, raw scalars [nt,4]) from a flow-result dict."""
    th = np.asarray(net["throats"], dtype=np.int64).reshape(-1, 2)
    nt = th.shape[0]
    pores = net["pores"]
    pos = np.stack([pores["x"], pores["y"], pores["z"]], axis=1).astype(np.float64)
    rad = np.asarray(pores["radius"], dtype=np.float64)
    org = np.asarray(origin_zyx[::-1], dtype=np.float64)
    spc = np.asarray(spacing_zyx[::-1], dtype=np.float64)
    L = np.asarray(shape_zyx[::-1], dtype=np.float64) * spc
    centers = (np.asarray(net["throat_center"], dtype=np.float64).reshape(-1, 3) - org) / spc
    if nt == 0:
        return centers.astype(np.float32), np.zeros((0, 4), np.float32)
    d = pos[th[:, 1] - 1] - pos[th[:, 0] - 1]
    d -= L * np.round(d / L)
    sep = np.sqrt((d * d).sum(axis=1))
    A = np.asarray(net["throat_area"], dtype=np.float64)
    raw = np.stack([np.log(np.maximum(A, 1e-12)),
                    np.log(np.maximum(rad[th[:, 0] - 1], 1e-6)),
                    np.log(np.maximum(rad[th[:, 1] - 1], 1e-6)),
                    np.log(np.maximum(sep, 1e-6))], axis=1)
    return centers.astype(np.float32), raw.astype(np.float32)


def throat_geometry(ctx, sdf_dev, shape_zyx, seg=None):
    """Throat geometry of a device SDF (zyx, unit spacing, origin 0) without fields, with the
    surrogate inputs computed on the device: (geometry dict, centers [nt,3] and raw scalars
    [nt,4] as device float32 tensors). seg: optional device segmentation of the same SDF (e.g.
    one volume of a batched segmentation)."""
    torch = _torch()
    n = int(np.prod(shape_zyx))
    cap = max(4096, n // 256)
    centers = torch.empty(3 * cap, dtype=torch.float32, device="cuda")
    raw = torch.empty(4 * cap, dtype=torch.float32, device="cuda")
    dims = tuple(shape_zyx)[::-1]
    geo = ctx.throat_geometry_d(sdf_dev, dims, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), seg=seg,
                                centers=centers, raw=raw)
    nt = len(geo["throats"])
    if nt > cap:  # more throats than the buffers hold: the same inputs from the host arrays
        c, r = throat_inputs(geo, shape_zyx)
        return geo, torch.from_numpy(c).cuda(), torch.from_numpy(r).cuda()
    return geo, centers[: 3 * nt].view(nt, 3), raw[: 4 * nt].view(nt, 4)


def conductance_targets(net):
    """(log g, valid mask) with g = Q/dp; throats with Q*dp <= 0 are flagged invalid."""
    q = np.asarray(net["throat_flow"], dtype=np.float64)
    dp = np.asarray(net["throat_dp"], dtype=np.float64)
    ok = (q * dp > 0) & np.isfinite(q) & np.isfinite(dp)
    g = np.where(ok, q / np.where(dp == 0, 1.0, dp), 1.0)
    return np.log(np.abs(g)).astype(np.float32), ok


def geometric_baseline_fit(raw, logg):
    """The method without the network: g = c * A^2 / L (log g = 2 log A - log L + log c)."""
    base = 2.0 * raw[:, 0] - raw[:, 3]
    return float(np.median(logg - base))


def geometric_baseline_predict(raw, c):
    return 2.0 * raw[:, 0] - raw[:, 3] + c


def gather_patches(ctx, sdf_dev, shape_zyx, centers):
    """Device float32 [nt, 4096] patches of the device SDF around host centres [nt, 3] (xyz)."""
    torch = _torch()
    nt = centers.shape[0]
    out = torch.empty(max(nt, 1) * 4096, device="cuda")
    if nt:
        c = torch.from_numpy(np.ascontiguousarray(centers, dtype=np.float32)).cuda()
        ctx.sur_patches_d(sdf_dev, tuple(shape_zyx)[::-1], c, nt, PATCH_SCALE, out)
    return out[: nt * 4096].view(nt, 4096)


class ThroatSurrogate:
    """Parameters, Adam state and feature standardisation of one surrogate network."""

    def __init__(self, ctx, seed=0):
        torch = _torch()
        self.ctx = ctx
        self.params = torch.from_numpy(capi.sur_init(seed)).cuda()
        n = self.params.numel()
        self.m = torch.zeros(n, device="cuda")
        self.v = torch.zeros(n, device="cuda")
        self.grads = torch.zeros(n, device="cuda")
        self.step = 0
        self.mu = np.zeros(4, np.float32)
        self.sd = np.ones(4, np.float32)
        self.t_mean = 0.0

    # ---- features ----
    def standardize(self, raw):
        return (raw - self.mu) / self.sd

    def fit_standardization(self, raw, logg):
        self.mu = raw.mean(axis=0).astype(np.float32)
        self.sd = np.maximum(raw.std(axis=0), 1e-6).astype(np.float32)
        self.t_mean = float(np.median(logg))

    def patches(self, sdf_dev, shape_zyx, centers):
        return gather_patches(self.ctx, sdf_dev, shape_zyx, centers)

    # ---- network ----
    def predict(self, patches, scalars_std, batch=4096):
        """log g for device patches [N,4096] and host standardised scalars [N,4]."""
        torch = _torch()
        N = patches.shape[0]
        out = torch.empty(max(N, 1), device="cuda")
        s = torch.from_numpy(np.ascontiguousarray(scalars_std, dtype=np.float32)).cuda()
        for a in range(0, N, batch):
            b = min(N, a + batch)
            self.ctx.sur_forward_d(b - a, patches[a:b].contiguous(), s[a:b].contiguous(),
                                   self.params, out[a:b])
        return out[:N] + self.t_mean

    def predict_sdf(self, sdf_dev, shape_zyx, centers, scalars_std):
        """log g straight from the device SDF (fused kernel: patch gather, convolutions and head
        in one launch; no patches or activations in HBM). centers: host [N, 3] xyz voxels."""
        torch = _torch()
        N = centers.shape[0]
        out = torch.empty(max(N, 1), device="cuda")
        if N:
            c = torch.from_numpy(np.ascontiguousarray(centers, dtype=np.float32)).cuda()
            s = torch.from_numpy(np.ascontiguousarray(scalars_std, dtype=np.float32)).cuda()
            self.ctx.sur_infer_d(sdf_dev, tuple(shape_zyx)[::-1], c, s, N, PATCH_SCALE,
                                 self.params, out)
        return out[:N] + self.t_mean

    def predict_device(self, sdf_dev, shape_zyx, centers_d, raw_d):
        """log g from device inputs (throat_geometry's centers and raw scalars): standardised on
        the device, then the fused kernel. Returns a device tensor."""
        torch = _torch()
        N = centers_d.shape[0]
        out = torch.empty(max(N, 1), device="cuda")
        if N:
            mu = torch.from_numpy(self.mu).cuda()
            sd = torch.from_numpy(self.sd).cuda()
            s = ((raw_d - mu) / sd).contiguous()
            self.ctx.sur_infer_d(sdf_dev, tuple(shape_zyx)[::-1], centers_d.contiguous(), s, N,
                                 PATCH_SCALE, self.params, out)
        return out[:N] + self.t_mean

    def train_step(self, patches, scalars_std, targets, lr, delta=0.25):
        """One Adam step on a device minibatch; returns the loss tensor (device)."""
        torch = _torch()
        B = patches.shape[0]
        y = torch.empty(B, device="cuda")
        dy = torch.empty(B, device="cuda")
        loss = torch.empty(1, device="cuda")
        self.ctx.sur_forward_d(B, patches, scalars_std, self.params, y)
        self.ctx.sur_loss_d(B, y, targets - self.t_mean, dy, loss, delta=delta)
        self.ctx.sur_backward_d(B, dy, self.params, self.grads)
        self.step += 1
        self.ctx.sur_adam_d(self.params, self.grads, self.m, self.v, self.step, lr=lr)
        return loss

    def fit(self, patches, scalars_raw, logg, steps=2000, batch=256, lr=2e-3, seed=0,
            log_every=0, delta=0.25, log=None):
        """Train on device patches [N,4096], host raw scalars [N,4] and host log g [N] with a
        Huber loss in log g (threshold delta; a few throats with near-zero Q or dp have very
        large |log g| errors, and the error measure is a median)."""
        torch = _torch()
        self.fit_standardization(scalars_raw, logg)
        s = torch.from_numpy(self.standardize(scalars_raw).astype(np.float32)).cuda()
        t = torch.from_numpy(np.asarray(logg, dtype=np.float32)).cuda()
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        N = patches.shape[0]
        hist = []
        for k in range(steps):
            idx = torch.randint(0, N, (batch,), device="cuda", generator=gen)
            lr_k = lr * 0.5 * (1.0 + np.cos(np.pi * k / steps))  # cosine decay
            loss = self.train_step(patches.index_select(0, idx).contiguous(),
                                   s.index_select(0, idx).contiguous(),
                                   t.index_select(0, idx).contiguous(), lr_k, delta)
            if log_every and (k % log_every == 0 or k == steps - 1):
                hist.append((k, float(loss.item())))
                if log:
                    log({"step": k, "loss": hist[-1][1]})
        return hist

    def state(self):
        return {"params": self.params.cpu().numpy(), "mu": self.mu, "sd": self.sd,
                "t_mean": np.float64(self.t_mean)}

    def load(self, st):
        torch = _torch()
        self.params = torch.from_numpy(np.asarray(st["params"], np.float32)).cuda()
        self.mu = np.asarray(st["mu"], np.float32)
        self.sd = np.asarray(st["sd"], np.float32)
        self.t_mean = float(st["t_mean"])


def median_rel_error(logg_pred, logg_true):
    """Median over throats of |g_pred - g| / |g|."""
    return float(np.median(np.abs(np.expm1(np.asarray(logg_pred, np.float64)
                                           - np.asarray(logg_true, np.float64)))))


def flow_sample(ctx, sdf, mu=1.0, fx=1.0, rtol=1e-9, max_iter=40000, geometry_only=False,
                return_fields=False):
    """Stokes solve (unless geometry_only) + extract_network_flow on a host SDF (zyx).
    Returns (device sdf, result dict, seconds of the Stokes solve), and the seven device fields
    (u, v, w, p, ox, oy, oz) after them when return_fields."""
    import time

    torch = _torch()
    n = sdf.size
    dims = sdf.shape[::-1]
    s = torch.from_numpy(np.ascontiguousarray(sdf, dtype=np.float32)).cuda()
    f = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(7)]
    # geometry only: zero fields with the solver's binary openness (max_iter = 0), so the
    # throats are exactly those of the solved case
    torch.cuda.synchronize()
    t0 = time.time()
    ctx.stokes_solve_d(s, dims, (1, 1, 1), (fx, 0.0, 0.0), *f, mu=mu, rtol=rtol,
                       max_iter=0 if geometry_only else max_iter)
    torch.cuda.synchronize()
    t_st = time.time() - t0
    ox, oy, oz = f[4], f[5], f[6]
    net = ctx.extract_network_flow_d(s, dims, (0, 0, 0), (1, 1, 1), *f[:4], ox=ox, oy=oy, oz=oz,
                                     grad_p_xyz=(-fx, 0.0, 0.0))
    if return_fields:
        return s, net, t_st, f
    return s, net, t_st


def direction_spread(ctx, seeds, size=128, rtol=1e-9):
    """How far g = Q/dp of one throat depends on the flow around it: each packing is solved with
    the body force along x, y and z (same throats, since the geometry is the same), and the
    median relative differences of the throats' g between directions are returned. A model that
    sees only the throat's geometry gives one g per throat, so its median error for the x-driven
    g cannot go far below that of the mean over the y and z directions ("x_vs_mean_yz")."""
    torch = _torch()
    from . import synth

    G = []
    for seed in seeds:
        prm = packing_params(seed)
        sdf, _, _ = synth.polydisperse_sdf((size, size, size), prm["median"], prm["sigma"],
                                           prm["porosity"], seed=int(seed))
        s = torch.from_numpy(np.ascontiguousarray(sdf, dtype=np.float32)).cuda()
        n, dims = sdf.size, sdf.shape[::-1]
        cols, th0 = [], None
        for ax in range(3):
            f = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(7)]
            force = [0.0, 0.0, 0.0]
            force[ax] = 1.0
            ctx.stokes_solve_d(s, dims, (1, 1, 1), tuple(force), *f, mu=1.0, rtol=rtol,
                               max_iter=40000)
            net = ctx.extract_network_flow_d(s, dims, (0, 0, 0), (1, 1, 1), *f[:4], ox=f[4],
                                             oy=f[5], oz=f[6],
                                             grad_p_xyz=tuple(-v for v in force))
            th = np.asarray(net["throats"], np.int64).reshape(-1, 2)
            if th0 is not None and not np.array_equal(th, th0):
                raise RuntimeError("throat sets differ between flow directions")
            th0 = th
            logg, ok = conductance_targets(net)
            cols.append(np.where(ok, logg.astype(np.float64), np.nan))
        G.append(np.stack(cols, axis=1))
    G = np.concatenate(G) if G else np.zeros((0, 3))

    def mre(a, b):
        m = np.isfinite(a) & np.isfinite(b)
        return float(np.median(np.abs(np.expm1(a[m] - b[m])))) if m.any() else None

    return {"throats": int(G.shape[0]), "x_vs_y": mre(G[:, 0], G[:, 1]),
            "x_vs_z": mre(G[:, 0], G[:, 2]), "y_vs_z": mre(G[:, 1], G[:, 2]),
            "x_vs_mean_yz": mre(G[:, 0], 0.5 * (G[:, 1] + G[:, 2]))}


def packing_params(seed):
    """Grain-pack parameters of training/test packing `seed` (varied so the model generalises)."""
    rng = np.random.default_rng(seed)
    return {"median": float(rng.uniform(4.0, 8.0)), "sigma": float(rng.uniform(0.2, 0.35)),
            "porosity": float(rng.uniform(0.33, 0.45))}


def build_dataset(ctx, seeds, size, rtol=1e-9, log=None, max_throats=0):
    """Solve and extract every packing; returns device patches [N,4096] and host arrays
    (raw scalars [N,4], log g [N], packing index [N]) of the throats with Q*dp > 0, plus a
    per-packing summary. max_throats > 0 keeps a random sample of that many of each packing's
    valid throats (seeded by the packing's seed)."""
    import time

    torch = _torch()
    from . import synth

    P, R, T, K, info = [], [], [], [], []
    for k, seed in enumerate(seeds):
        prm = packing_params(seed)
        t0 = time.time()
        sdf, phi, _ = synth.polydisperse_sdf((size, size, size), prm["median"], prm["sigma"],
                                             prm["porosity"], seed=int(seed))
        s, net, t_st = flow_sample(ctx, sdf, rtol=rtol)
        centers, raw = throat_inputs(net, sdf.shape)
        logg, ok = conductance_targets(net)
        if max_throats > 0 and ok.sum() > max_throats:
            keep = np.random.default_rng(int(seed)).choice(np.flatnonzero(ok), max_throats,
                                                           replace=False)
            ok = np.zeros_like(ok)
            ok[np.sort(keep)] = True
        P.append(gather_patches(ctx, s, sdf.shape, centers[ok]))
        R.append(raw[ok])
        T.append(logg[ok])
        K.append(np.full(int(ok.sum()), k, np.int32))
        info.append({"seed": int(seed), **prm, "porosity_achieved": round(phi, 4),
                     "throats": int(len(ok)), "kept": int(ok.sum()),
                     "stokes_s": round(t_st, 3), "total_s": round(time.time() - t0, 3)})
        if log:
            log(info[-1])
        del s
    return (torch.cat(P) if P else torch.empty(0, 4096, device="cuda"),
            np.concatenate(R) if R else np.zeros((0, 4), np.float32),
            np.concatenate(T) if T else np.zeros(0, np.float32),
            np.concatenate(K) if K else np.zeros(0, np.int32), info)


def solve_samples(ctx, seeds, size, out_dir, max_throats=0, rtol=1e-9, log=None):
    """Solve and extract each packing and write its valid throats' inputs and targets to
    out_dir/pack_<seed>_<size>.npz (centers, raw scalars, log g; a random sample of max_throats
    when > 0), so that solving can be split over several runs and the patches regathered later
    from the regenerated SDF (dataset_from_samples)."""
    import json
    import os
    import time

    from . import synth

    os.makedirs(out_dir, exist_ok=True)
    for seed in seeds:
        prm = packing_params(seed)
        t0 = time.time()
        sdf, phi, _ = synth.polydisperse_sdf((size, size, size), prm["median"], prm["sigma"],
                                             prm["porosity"], seed=int(seed))
        _, net, t_st = flow_sample(ctx, sdf, rtol=rtol)
        centers, raw = throat_inputs(net, sdf.shape)
        logg, ok = conductance_targets(net)
        n_ok = int(ok.sum())
        if max_throats > 0 and n_ok > max_throats:
            keep = np.random.default_rng(int(seed)).choice(np.flatnonzero(ok), max_throats,
                                                           replace=False)
            ok = np.zeros_like(ok)
            ok[np.sort(keep)] = True
        info = {"seed": int(seed), **prm, "porosity_achieved": round(phi, 4),
                "throats": int(len(ok)), "valid": n_ok, "kept": int(ok.sum()),
                "stokes_s": round(t_st, 3), "total_s": round(time.time() - t0, 3)}
        np.savez(os.path.join(out_dir, f"pack_{int(seed)}_{size}.npz"), centers=centers[ok],
                 raw=raw[ok], logg=logg[ok], info=json.dumps(info))
        if log:
            log(info)


def dataset_from_samples(ctx, seeds, size, in_dir, log=None):
    """build_dataset's result from files written by solve_samples: the SDF of each packing is
    regenerated from its seed (the generator is deterministic) and the patches gathered at the
    saved centres."""
    import json
    import os

    torch = _torch()
    from . import synth

    P, R, T, K, info = [], [], [], [], []
    for k, seed in enumerate(seeds):
        d = np.load(os.path.join(in_dir, f"pack_{int(seed)}_{size}.npz"))
        prm = packing_params(seed)
        sdf, _, _ = synth.polydisperse_sdf((size, size, size), prm["median"], prm["sigma"],
                                           prm["porosity"], seed=int(seed))
        s = torch.from_numpy(np.ascontiguousarray(sdf, dtype=np.float32)).cuda()
        P.append(gather_patches(ctx, s, sdf.shape, d["centers"]))
        R.append(d["raw"])
        T.append(d["logg"])
        K.append(np.full(len(d["logg"]), k, np.int32))
        info.append(json.loads(str(d["info"])))
        if log:
            log({"loaded": int(seed), "throats": int(len(d["logg"]))})
        del s
    return (torch.cat(P) if P else torch.empty(0, 4096, device="cuda"),
            np.concatenate(R) if R else np.zeros((0, 4), np.float32),
            np.concatenate(T) if T else np.zeros(0, np.float32),
            np.concatenate(K) if K else np.zeros(0, np.int32), info)
