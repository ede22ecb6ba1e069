"""Train the throat-conductance surrogate on GPU-generated data and evaluate it.

    python -m peclet.pnm.surrogate_train --train 56 --test 8 --size 128 --steps 4000 \\
        --out <dir> [--weights <file.npz>] [--direction-spread N] [--max-throats K]

Large runs can solve the packings in parts first (--samples DIR --solve-train A:B /
--solve-test A:B write each packing's throat inputs and targets), then train from DIR
(--samples DIR): the patches are regathered from the regenerated SDFs.

Every packing is a polydisperse grain pack (packing_params(seed)); its fields come from the
project Stokes solver and its throats and g = Q/dp from extract_network_flow. Train and test
throats come from different packings. Writes <dir>/surrogate_report.json and the weights
(<dir>/surrogate_weights.npz, or --weights).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=56)
    ap.add_argument("--test", type=int, default=8)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--delta", type=float, default=0.25, help="Huber threshold (<= 0: MSE)")
    ap.add_argument("--direction-spread", type=int, default=0, metavar="N",
                    help="also solve N test packings with the flow along x, y and z and report "
                         "how much each throat's g changes with the direction")
    ap.add_argument("--max-throats", type=int, default=0, metavar="K",
                    help="keep a random sample of K valid throats per training packing (0: all)")
    ap.add_argument("--samples", default=None, metavar="DIR",
                    help="per-packing throat samples (solve_samples files): with --solve-train /"
                         " --solve-test, solve those packings into DIR and stop; otherwise train"
                         " from the samples in DIR instead of solving")
    ap.add_argument("--solve-train", default=None, metavar="A:B",
                    help="training packings A..B-1 to solve into --samples")
    ap.add_argument("--solve-test", default=None, metavar="A:B",
                    help="test packings A..B-1 to solve into --samples (all their throats)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--weights", default=None)
    args = ap.parse_args(argv)

    import torch

    from . import capi
    from . import surrogate as S

    torch.zeros(1, device="cuda")
    os.makedirs(args.out, exist_ok=True)
    ctx = capi.Context(0)

    def log(d):
        print(json.dumps(d), flush=True)

    def span(text):
        a, b = (int(v) for v in text.split(":"))
        return range(a, b)

    if args.solve_train or args.solve_test:
        if not args.samples:
            raise SystemExit("--solve-train / --solve-test need --samples DIR")
        if args.solve_train:
            S.solve_samples(ctx, [9000 + k for k in span(args.solve_train)], args.size,
                            args.samples, max_throats=args.max_throats, log=log)
        if args.solve_test:
            S.solve_samples(ctx, [9500 + k for k in span(args.solve_test)], args.size,
                            args.samples, log=log)
        ctx.close()
        return 0

    t0 = time.time()
    if args.samples:
        tr = S.dataset_from_samples(ctx, [9000 + k for k in range(args.train)], args.size,
                                    args.samples, log=log)
        te = S.dataset_from_samples(ctx, [9500 + k for k in range(args.test)], args.size,
                                    args.samples, log=log)
    else:
        tr = S.build_dataset(ctx, [9000 + k for k in range(args.train)], args.size, log=log,
                             max_throats=args.max_throats)
        te = S.build_dataset(ctx, [9500 + k for k in range(args.test)], args.size, log=log)
    t_data = time.time() - t0
    Ptr, Rtr, Ttr, _, info_tr = tr
    Pte, Rte, Tte, _, info_te = te
    log({"train_throats": int(len(Ttr)), "test_throats": int(len(Tte)),
         "data_s": round(t_data, 1)})

    untrained = S.ThroatSurrogate(ctx, seed=args.seed)
    untrained.fit_standardization(Rtr, Ttr)
    err_untrained = S.median_rel_error(
        untrained.predict(Pte, untrained.standardize(Rte)).cpu().numpy(), Tte)
    c_geo = S.geometric_baseline_fit(Rtr, Ttr)
    err_geo = S.median_rel_error(S.geometric_baseline_predict(Rte, c_geo), Tte)

    model = S.ThroatSurrogate(ctx, seed=args.seed)
    torch.cuda.synchronize()
    t1 = time.time()
    hist = model.fit(Ptr, Rtr, Ttr, steps=args.steps, batch=args.batch, lr=args.lr,
                     seed=args.seed, log_every=max(1, args.steps // 20), delta=args.delta,
                     log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    pred_te = model.predict(Pte, model.standardize(Rte)).cpu().numpy()
    pred_tr = model.predict(Ptr, model.standardize(Rtr)).cpu().numpy()
    err = S.median_rel_error(pred_te, Tte)
    err_train = S.median_rel_error(pred_tr, Ttr)
    # training throughput: FLOPs of one forward pass per throat (convolutions + dense head)
    conv_flops = sum(2 * co * 27 * ci * d ** 3 for d, ci, co in ((16, 1, 16), (8, 16, 32),
                                                                   (4, 32, 32)))
    fwd_flops = conv_flops + 2 * (36 * 64 + 64)
    report = {
        "train_packings": args.train, "test_packings": args.test, "size": args.size,
        "max_throats_per_train_packing": args.max_throats,
        "train_throats": int(len(Ttr)), "test_throats": int(len(Tte)),
        "steps": args.steps, "batch": args.batch, "lr": args.lr, "huber_delta": args.delta,
        "data_seconds": round(t_data, 1), "train_seconds": round(t_train, 2),
        "train_tflops": round(3 * fwd_flops * args.batch * args.steps / t_train / 1e12, 3),
        "median_rel_error_test": err, "median_rel_error_train": err_train,
        "median_rel_error_geometric_A2_over_L": err_geo,
        "median_rel_error_untrained": err_untrained,
        "loss_history": hist, "geometric_c": c_geo,
        "packings_train": info_tr, "packings_test": info_te,
    }
    if args.direction_spread > 0:
        report["direction_spread"] = S.direction_spread(
            ctx, [9500 + k for k in range(args.direction_spread)], args.size)
    wpath = args.weights or os.path.join(args.out, "surrogate_weights.npz")
    np.savez(wpath, **model.state(), geometric_c=np.float64(c_geo))
    with open(os.path.join(args.out, "surrogate_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    log({k: v for k, v in report.items() if not isinstance(v, list)})
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
