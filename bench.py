#!/usr/bin/env python3
"""Benchmark: fused extract_pore_network on a 512^3 fp32 SDF on one MI355X.

One step = one pass of the fused pipeline (pores, watershed segmentation, topology) over the
device-resident SDF, exactly what pnm_extract_pore_network_d does: the SDF is in HBM before the
timed region, the segmentation stays in HBM, the pore list and the unique connection list are
written to device buffers. Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--workload polydisperse|boolean] [--size 512]
                    [--no-batched] [--no-flow] [--no-cpu-baseline]

`value` is voxels per second of the whole step (Gvox/s). `roofline` is the dominant kernel's
algorithmic bytes per launch over its average duration, timed live with HIP events on the
launch stream. `cpu_baseline` times the OpenMP restatement of the reference's
extract_pore_network_k (oracle/pnm_ref_omp.cpp, OMP_NUM_THREADS threads) on a bounded z-slab of
the same volume (--cpu-full: the whole volume), median of 5 after a warm-up, with the serial C
oracle (1 core) beside it. Two more objects on the same line:
  batched: 4096 cubic crops, edges U[48,160] at seeded offsets of the same 512^3 volume (periodic
           wrap), segmented in one persistent launch replayed from a hipGraph with the autotuned
           shape, against per-volume segment_volume calls on a sample of the same crops;
  flow:    extract_network_flow on the 256^3 chamber lattice (5^3 chambers r=12 at pitch 51.2,
           tubes r ~ U[3,8], seed 3001) with FP64 fields from the project Stokes solver;
  surrogate: throat conductance g = Q/dp on a held-out 256^3 grain pack: the plain path (Stokes
           solve + extract_network_flow) against the surrogate path (throat geometry from the
           SDF alone, then the fused CNN inference with the packaged trained weights).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "peclet-pnm_amd"), os.path.join(ROOT, "oracle")]

METRIC = "extract_pore_network Gvox/s @512³ (% HBM BW); batched & surrogate speedup"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory, one streaming pass) bytes per voxel of each per-voxel kernel.
ALGO_BYTES_PER_VOXEL = {
    "k_tile": 12,         # read sdf 4; write marker label 4, gradient-path root 4
    "k_flood_tile": 12,   # read labels 4, sdf 4; write labels 4
    "k_seg_topo": 12,     # read sdf 4, then root (pore) or label (solid) 4; write seg 4
    "k_pre_peaks": 4,     # read sdf 4
    "k_ccl_flatten": 8,   # read + write labels
    "k_seg": 16,          # read sdf 4, roots 4, labels 4; write seg 4
    "k_topo_cand": 4,     # read seg 4
}
PIPELINE_BYTES_PER_VOXEL = 64  # S1..S6 compulsory traffic of extract_pore_network

# HBM traffic per launch measured with rocprofv3 counters on this workload (written by
# --pmc-summary); rocprof kernel symbols differ from the library's timing names for two kernels.
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_extract_pore_network_512_polydisperse.md")
PROFILER_NAMES = {"k_seg_topo": "k_topo_cand<true>",
                  "k_tile": "k_tile<true>"}  # the 24-bit index variant (ny*nx, nz*ny < 2^24)


def make_sdf(workload, size, seed):
    from peclet.pnm import synth

    shape = (size, size, size)
    t = time.time()
    if workload == "polydisperse":
        sdf, phi, info = synth.polydisperse_sdf(shape, 9.0, 0.3, 0.35, seed=seed)
        desc = {"workload": "polydisperse grain pack, lognormal radii (median 9 vox, sigma 0.3)"
                            ", RSA + uniform overgrowth",
                "porosity": round(phi, 4), "spheres": info["spheres"]}
    else:
        sdf = synth.boolean_sdf(shape, 9.0, 0.3, 0.5, seed=seed)
        desc = {"workload": "Boolean model of overlapping spheres, lognormal radii "
                            "(median 9 vox, sigma 0.3)",
                "porosity": round(float((sdf > 0).mean()), 4)}
    desc["gen_s"] = round(time.time() - t, 2)
    return sdf, desc


def _host_cpu():
    """CPU model and the logical CPUs this process may use (the OpenMP threads are set apart)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(sdf, slab, serial_slab, reps=5, full=False):
    """The OpenMP restatement of the reference's extract_pore_network_k (oracle/pnm_ref_omp.cpp,
    threads from OMP_NUM_THREADS) on the first `slab` z-slices of the same SDF as their own
    periodic volume (the whole volume with full=True): one warm-up run, then the median of `reps`.
    The serial C oracle (1 core) on `serial_slab` slices is reported beside it."""
    import numpy as np
    import pnm_oracle as orc

    crop = np.ascontiguousarray(sdf if full else sdf[:slab])
    orc.ref_omp_extract_pore_network(crop, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))  # warm-up
    times, stages = [], []
    for i in range(reps):
        t = time.time()
        _, _, _, ms = orc.ref_omp_extract_pore_network(crop, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
        times.append(time.time() - t)
        stages.append(ms)
        print(f"[cpu_baseline] run {i + 1}/{reps}: {times[-1]:.2f} s", file=sys.stderr, flush=True)
    k = int(np.argsort(times)[len(times) // 2])
    dt = times[k]
    threads = orc.ref_omp().pnmr_threads()
    model, cpus = _host_cpu()
    small = np.ascontiguousarray(sdf[:serial_slab])
    t = time.time()
    orc.extract_pore_network(small, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ts = time.time() - t
    what = "whole volume" if full else f"z-slab {crop.shape} of the same SDF as its own periodic volume"
    return {"value": round(crop.size / dt / 1e9, 6), "unit": "Gvox/s", "cores": threads,
            "kind": "port",
            "sample": (f"OpenMP restatement of the reference's extract_pore_network_k "
                       f"(oracle/pnm_ref_omp.cpp), {what} ({crop.size} voxels), {threads} "
                       f"threads, median of {reps} runs after one warm-up: {dt:.2f} s"),
            "stage_ms": {name: round(float(v), 1) for name, v in
                         zip(("upload", "pores", "segment", "topology", "download"), stages[k])},
            "runs_s": [round(x, 3) for x in times],
            "host_cpu": model, "host_logical_cpus": cpus,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
            "serial_oracle": {"value": round(small.size / ts / 1e9, 6), "unit": "Gvox/s",
                              "cores": 1,
                              "sample": f"serial C oracle on z-slab {small.shape}, {ts:.2f} s"}}


def bench_batched(ctx, src, reps, seed=4001, count=4096, lo=48, hi=160, sample=256):
    """Batched segmentation of `count` periodic cubic crops of the device volume `src`."""
    import numpy as np
    import torch

    dev = src.device
    N = src.shape
    rng = np.random.default_rng(seed)
    edges = rng.integers(lo, hi + 1, size=count)
    offs = np.stack([rng.integers(0, k, size=count) for k in N], axis=1)
    vols, segs, dims = [], [], []
    for e, o in zip(edges, offs):
        idx = [(torch.arange(int(e), device=dev) + int(o[a])) % N[a] for a in range(3)]
        v = src.index_select(0, idx[0]).index_select(1, idx[1]).index_select(2, idx[2])
        vols.append(v.contiguous().view(-1))
        segs.append(torch.empty(int(e) ** 3, dtype=torch.int32, device=dev))
        dims.append((int(e), int(e), int(e)))
    total = int(sum(int(e) ** 3 for e in edges))
    spc = (1.0, 1.0, 1.0)
    stream = torch.cuda.current_stream(dev)
    ctx.segment_batch_d(vols, segs, dims, spc, graph=True, autotune=True)  # tune + capture
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        ctx.segment_batch_d(vols, segs, dims, spc, graph=True)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # per-volume calls on a sample of the same crops; their labels must equal the batch's
    pick = rng.choice(count, size=min(sample, count), replace=False)
    one = [torch.empty_like(segs[k]) for k in pick]
    torch.cuda.synchronize()
    t = time.time()
    for k, out in zip(pick, one):
        ctx.segment_volume_d(vols[k], dims[k], spc, out)
    torch.cuda.synchronize()
    dt = time.time() - t
    sample_vox = int(sum(int(edges[k]) ** 3 for k in pick))
    same = all(torch.equal(segs[k], out) for k, out in zip(pick, one))
    per_vol = sample_vox / dt / 1e9
    value = total / (ms / 1e3) / 1e9
    return {"workload": f"{count} periodic cubic crops, edge U[{lo},{hi}] (seed {seed}), of the "
                        "512^3 volume above", "voxels": total, "ms": round(ms, 3),
            "value": round(value, 3), "unit": "Gvox/s",
            "hbm_frac": round(56 * total / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "per_volume_calls": {"value": round(per_vol, 3), "unit": "Gvox/s",
                                 "sample_volumes": int(len(pick)), "sample_voxels": sample_vox},
            "speedup_vs_per_volume": round(value / per_vol, 2),
            "sample_labels_identical": bool(same), "launch": ctx.batch_config()}


def bench_flow(ctx, reps, size=256, stokes_iters=40000):
    """extract_network_flow on the chamber lattice with fields from the project Stokes solver
    (at most `stokes_iters` MINRES iterations: a profiling run may cut the solve short, which
    changes the field values, not the work of the extraction)."""
    import numpy as np
    import torch

    from peclet.pnm import synth

    dev = torch.device("cuda:0")
    shape = (size, size, size)
    sdf = synth.chamber_sdf(shape, m=5, chamber_r=12.0, tube_rmin=3.0, tube_rmax=8.0, seed=3001)
    n = sdf.size
    dims = (size, size, size)
    s = torch.from_numpy(sdf).to(dev)
    f = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(7)]
    mu, fx = 1.0, 1.0
    torch.cuda.synchronize()
    t = time.time()
    it, rr = ctx.stokes_solve_d(s, dims, (1, 1, 1), (fx, 0.0, 0.0), *f, mu=mu, rtol=1e-10,
                                max_iter=stokes_iters)
    torch.cuda.synchronize()
    t_stokes = time.time() - t
    args = (s, dims, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), *f[:4])
    kw = dict(ox=f[4], oy=f[5], oz=f[6], grad_p_xyz=(-fx, 0.0, 0.0))
    res = ctx.extract_network_flow_d(*args, **kw)  # warm-up
    torch.cuda.synchronize()
    t = time.time()
    for _ in range(reps):
        res = ctx.extract_network_flow_d(*args, **kw)
    torch.cuda.synchronize()
    ms = (time.time() - t) / reps * 1e3
    q = np.abs(res["throat_flow"])
    return {"workload": "256^3 chamber lattice: 5^3 chambers r=12 at pitch 51.2, capsule tubes "
                        "r ~ U[3,8] (seed 3001); binary openness",
            "ms": round(ms, 3), "value": round(n / (ms / 1e3) / 1e9, 3), "unit": "Gvox/s",
            "hbm_frac": round(172 * n / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "pores": int(len(res["pores"])), "throats": int(len(q)),
            "max_residual_over_max_q": float(np.abs(res["pore_residual"]).max() / q.max())
            if len(q) else None,
            "stokes": {"solver": "MAC, FP64 preconditioned MINRES", "iterations": it,
                       "rel_residual": rr, "seconds": round(t_stokes, 2), "mu": mu,
                       "body_force_x": fx}}


def bench_surrogate(ctx, size=256, seed=9999):
    """Plain (Stokes + bookkeeping) vs surrogate (geometry + CNN) throat conductance."""
    import numpy as np
    import torch

    from peclet.pnm import surrogate as S
    from peclet.pnm import synth

    wpath = os.path.join(ROOT, "peclet-pnm_amd", "peclet", "pnm", "data",
                         "surrogate_weights.npz")
    if not os.path.exists(wpath):
        return None
    prm = S.packing_params(seed)
    sdf, phi, _ = synth.polydisperse_sdf((size,) * 3, prm["median"], prm["sigma"],
                                         prm["porosity"], seed=seed)
    model = S.ThroatSurrogate(ctx)
    model.load(np.load(wpath))
    # plain path: Stokes solve + bookkeeping
    torch.cuda.synchronize()
    t0 = time.time()
    s_dev, net, t_stokes, fields = S.flow_sample(ctx, sdf, return_fields=True)
    torch.cuda.synchronize()
    logg, ok = S.conductance_targets(net)

    # bookkeeping alone on the solved fields (the plain path minus the solve), from the same
    # host SDF: upload, extract_network_flow_d, results on the host (one warm-up, median of 5)
    def bookkeeping():
        s2 = torch.from_numpy(np.ascontiguousarray(sdf, dtype=np.float32)).cuda()
        return ctx.extract_network_flow_d(s2, sdf.shape[::-1], (0, 0, 0), (1, 1, 1), *fields[:4],
                                          ox=fields[4], oy=fields[5], oz=fields[6],
                                          grad_p_xyz=(-1.0, 0.0, 0.0))
    bookkeeping()
    tb = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.time()
        bookkeeping()
        tb.append(time.time() - t0)
    t_book = float(np.median(tb))
    t_plain = t_stokes + t_book
    del fields
    # surrogate path from the same host SDF: upload, throat geometry without fields (segmentation
    # and interface faces, surrogate inputs computed on the device), fused inference (one
    # warm-up, then the median of 5)
    def surrogate_path():
        s2 = torch.from_numpy(np.ascontiguousarray(sdf, dtype=np.float32)).cuda()
        geo, cen, raw = S.throat_geometry(ctx, s2, sdf.shape)
        pred = model.predict_device(s2, sdf.shape, cen, raw).cpu()
        return geo, pred
    surrogate_path()
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        t0 = time.time()
        geo, pred = surrogate_path()
        ts.append(time.time() - t0)
    t_sur = float(np.median(ts))
    same = np.array_equal(np.asarray(geo["throats"]), np.asarray(net["throats"]))
    same_geo = same and all(np.array_equal(np.asarray(geo[k]).view(np.uint64),
                                           np.asarray(net[k]).view(np.uint64))
                            for k in ("throat_area", "throat_center"))
    pred = pred.cpu().numpy()
    err = S.median_rel_error(pred[ok], logg[ok]) if same and ok.any() else None
    c = float(np.load(wpath)["geometric_c"])
    _, raw = S.throat_inputs(net, sdf.shape)
    err_geo = S.median_rel_error(S.geometric_baseline_predict(raw[ok], c), logg[ok])
    # inference alone (the fused kernel: patches gathered from the SDF inside it), timed with
    # HIP events around its launches on the context's stream
    nt = len(pred)
    _, cen, raw = S.throat_geometry(ctx, s_dev, sdf.shape)
    model.predict_device(s_dev, sdf.shape, cen, raw)
    ctx.profile(True)
    ctx.profile_only(None)
    ctx.profile_read()
    for _ in range(5):
        S.throat_geometry(ctx, s_dev, sdf.shape)
        model.predict_device(s_dev, sdf.shape, cen, raw)
    torch.cuda.synchronize()
    prof = ctx.profile_read()
    ctx.profile(False)
    t_inf = prof["k_sur_infer"][0] / prof["k_sur_infer"][1] / 1e3
    t_geo = sum(v[0] for k, v in prof.items() if k != "k_sur_infer") / 5 / 1e3
    conv_flops = sum(2 * co * 27 * ci * d ** 3 for d, ci, co in ((16, 1, 16), (8, 16, 32),
                                                                   (4, 32, 32)))
    tflops = conv_flops * nt / t_inf / 1e12
    return {"workload": f"held-out {size}^3 polydisperse grain pack (seed {seed}, median radius "
                        f"{prm['median']:.2f}, porosity {phi:.3f})",
            "throats": int(nt), "throats_with_q_dp_positive": int(ok.sum()),
            "plain_s": round(t_plain, 3), "stokes_s": round(t_stokes, 3),
            "surrogate_s": round(t_sur, 4),
            "surrogate_device_s": {"geometry": round(t_geo, 5), "inference": round(t_inf, 5)},
            "bookkeeping_s": round(t_book, 4),
            "speedup_vs_plain": round(t_plain / t_sur, 1),
            "speedup_vs_bookkeeping_only": round(t_book / t_sur, 2),
            "median_rel_error_g": err, "median_rel_error_g_geometric_A2_over_L": err_geo,
            "same_throats": bool(same), "same_areas_and_centroids": bool(same_geo),
            "inference": {"seconds": round(t_inf, 5), "throats_per_s": round(nt / t_inf),
                          "roofline": {"bound": "mfma", "achieved": round(tflops, 3),
                                       "peak": 2500.0, "unit": "TFLOP/s",
                                       "frac": round(tflops / 2500.0, 6)}}}


def _kernel_symbol(name):
    """'void pnm::k_topo_cand<true>(int const*, ...)' -> 'k_topo_cand<true>'."""
    head = name[5:] if name.startswith("void ") else name
    head = head.split("(", 1)[0]
    return head.split("::")[-1] if "<" not in head else head.split("<", 1)[0].split("::")[-1] + \
        "<" + head.split("<", 1)[1]


def _counter_rows(path):
    import csv

    with open(path) as f:
        return [(_kernel_symbol(r["Kernel_Name"]), float(r["Counter_Value"]))
                for r in csv.DictReader(f)]


def pmc_summary(root, copy_bytes, voxels, workload="extract_pore_network"):
    """Markdown summary of four rocprofv3 --pmc runs under `root` (pmc_copy_fetch,
    pmc_copy_write: the calibration kernel; pmc_net_fetch, pmc_net_write: the pipeline), each with
    run_counter_collection.csv. FETCH_SIZE and WRITE_SIZE are KiB per dispatch; each counter is
    scaled by (true bytes of the calibration kernel) / (what the counter reads for it)."""
    from collections import defaultdict

    def calib(name, counter):
        vals = [v for k, v in _counter_rows(os.path.join(root, name, "run_counter_collection.csv"))
                if "neg_kernel" in k]
        if not vals:
            raise SystemExit(f"no calibration dispatches in {name}")
        reading = sorted(vals)[len(vals) // 2] * 1024.0
        return reading, copy_bytes / reading

    rf, kf = calib("pmc_copy_fetch", "FETCH_SIZE")
    rw, kw = calib("pmc_copy_write", "WRITE_SIZE")
    per = {}
    for name, scale, key in (("pmc_net_fetch", kf, "fetch"), ("pmc_net_write", kw, "write")):
        acc = defaultdict(list)
        for k, v in _counter_rows(os.path.join(root, name, "run_counter_collection.csv")):
            acc[k].append(v * 1024.0 * scale)
        for k, vals in acc.items():
            per.setdefault(k, {})[key] = vals
    steps = len(per.get(PROFILER_NAMES["k_tile"], {}).get("fetch", [])) or 1
    rows = []
    for k, d in per.items():
        if k.startswith(("vectorized_elementwise", "reduce_kernel", "distribution", "elementwise")):
            continue  # torch's own set-up kernels outside the pipeline
        if workload != "extract_pore_network" and k.startswith(("k_ms_", "k_st_",
                                                                 "__amd_rocclr_fillBuffer")):
            continue  # the Stokes solve (its kernels and memsets) that generates the flow fields
        f, w = d.get("fetch", []), d.get("write", [])
        rows.append((sum(f) / steps + sum(w) / steps, k, len(f) / steps,
                     sum(f) / max(1, len(f)), sum(w) / max(1, len(w)),
                     sum(f) / steps, sum(w) / steps))
    rows.sort(reverse=True)
    lines = ["| kernel | launches/step | fetch B/launch | write B/launch | fetch B/voxel/step "
             "| write B/voxel/step | algorithmic B/voxel |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    inv = {v: k for k, v in PROFILER_NAMES.items()}
    machine = {}
    for _, k, nl, fa, wa, fs, ws in rows:
        alg = ALGO_BYTES_PER_VOXEL.get(inv.get(k, k))
        lines.append(f"| `{k}` | {nl:g} | {fa:.4g} | {wa:.4g} | {fs / voxels:.3f} | "
                     f"{ws / voxels:.3f} | {alg if alg is not None else ''} |")
        machine[k] = {"fetch_bytes": round(fa), "write_bytes": round(wa),
                      "launches_per_step": nl}
    tf = sum(r[5] for r in rows) / voxels
    tw = sum(r[6] for r in rows) / voxels
    out = ["## Calibration", "",
           f"Elementwise negation of {copy_bytes} bytes (reads and writes that many bytes):", "",
           "| counter | reads (KiB) | true (KiB) | scale factor |", "|---|---:|---:|---:|",
           f"| FETCH_SIZE | {rf / 1024:.1f} | {copy_bytes / 1024:.1f} | {kf:.4f} |",
           f"| WRITE_SIZE | {rw / 1024:.1f} | {copy_bytes / 1024:.1f} | {kw:.4f} |", "",
           f"## Pipeline ({steps} passes of {workload} over {voxels} voxels; "
           "bytes are scaled)", ""] + lines + [
           "", f"Total per step: fetch {tf:.2f} B/voxel, write {tw:.2f} B/voxel "
               f"({tf + tw:.2f} B/voxel; {(tf + tw) * voxels / 1e9:.3f} GB).", "",
           "## Machine-readable (bench.py reads `traffic` from this block)", "", "```json",
           json.dumps({"voxels": voxels, "fetch_scale": kf, "write_scale": kw,
                       "kernels": machine}), "```"]
    return "\n".join(out)


def sq_summary(root, voxels, workload="extract_pore_network"):
    """Markdown summary of two rocprofv3 --pmc runs of SQ counters under `root` (pmc_sq_a:
    SQ_WAVES, SQ_INSTS_VALU, SQ_INSTS_SALU, SQ_INSTS_LDS, SQ_INSTS_VMEM_RD, SQ_INSTS_VMEM_WR,
    SQ_WAVE_CYCLES, SQ_BUSY_CYCLES; pmc_sq_b: SQ_WAIT_ANY, SQ_WAIT_INST_ANY, SQ_ACTIVE_INST_VALU,
    SQ_ACTIVE_INST_LDS, SQ_LDS_BANK_CONFLICT, SQ_LDS_IDX_ACTIVE, SQ_WAIT_INST_LDS,
    SQ_ACTIVE_INST_ANY): per kernel, instructions per wave and per launch, and the VALU issue
    time per launch at one wave64 VALU instruction per CU per clock."""
    import csv
    from collections import defaultdict

    agg = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    for name in ("pmc_sq_a", "pmc_sq_b"):
        with open(os.path.join(root, name, "run_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                k = _kernel_symbol(r["Kernel_Name"])
                agg[k][r["Counter_Name"]] += float(r["Counter_Value"])
                if name == "pmc_sq_a":
                    disp[k].add(r["Dispatch_Id"])
    cus, clock_ghz = 256, 2.4
    rows = []
    for k, c in agg.items():
        if k.startswith(("vectorized_elementwise", "reduce_kernel", "distribution", "elementwise")):
            continue  # torch's own set-up kernels outside the pipeline
        if workload != "extract_pore_network" and k.startswith(("k_ms_", "k_st_",
                                                                 "__amd_rocclr_fillBuffer")):
            continue  # the Stokes solve (its kernels and memsets) that generates the flow fields
        w, nd = c.get("SQ_WAVES", 0.0), max(1, len(disp[k]))
        if w <= 0:
            continue
        valu = c.get("SQ_INSTS_VALU", 0.0) / nd
        rows.append((valu, k, nd, w / nd, c.get("SQ_INSTS_VALU", 0.0) / w,
                     c.get("SQ_INSTS_SALU", 0.0) / w, c.get("SQ_INSTS_LDS", 0.0) / w,
                     c.get("SQ_INSTS_VMEM_RD", 0.0) / w, c.get("SQ_INSTS_VMEM_WR", 0.0) / w,
                     c.get("SQ_WAIT_ANY", 0.0) / max(c.get("SQ_WAVE_CYCLES", 0.0), 1.0),
                     c.get("SQ_LDS_BANK_CONFLICT", 0.0) / max(c.get("SQ_LDS_IDX_ACTIVE", 0.0), 1.0),
                     c.get("SQ_INSTS_SALU", 0.0) / nd))
    rows.sort(reverse=True)
    lines = ["| kernel | waves/launch | VALU/wave | SALU/wave | LDS/wave | VMEM rd/wave "
             "| VMEM wr/wave | wait share | LDS conflicts | VALU/launch (M) | VALU issue ms "
             "| SALU/launch (M) |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for valu, k, nd, wpl, vw, sw, lw, rw, ww, wait, bank, salu in rows[:16]:
        lines.append(f"| `{k}` | {wpl:.0f} | {vw:.0f} | {sw:.0f} | {lw:.0f} | {rw:.1f} | "
                     f"{ww:.1f} | {wait:.2f} | {bank:.2f} | {valu / 1e6:.1f} | "
                     f"{valu / cus / clock_ghz / 1e6:.3f} | {salu / 1e6:.1f} |")
    return "\n".join(
        [f"SQ counters per kernel ({voxels} voxels; wait share = SQ_WAIT_ANY / SQ_WAVE_CYCLES; LDS "
         "conflicts = SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE; VALU issue ms = VALU "
         f"wave-instructions per launch / ({cus} CUs x {clock_ghz} GHz), one wave64 VALU "
         "instruction per CU per clock).", ""] + lines)


def load_traffic(kernel):
    """HBM bytes per launch (fetch + write) of `kernel` from the committed counter summary, or
    None when it is absent."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        d = json.loads(block)["kernels"][PROFILER_NAMES.get(kernel, kernel)]
        return d["fetch_bytes"] + d["write_bytes"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--workload", default="polydisperse", choices=["polydisperse", "boolean"])
    ap.add_argument("--seed", type=int, default=2001)
    ap.add_argument("--cpu-slab", type=int, default=64,
                    help="z-slices of the SDF timed by the OpenMP CPU baseline")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the CPU baseline on the whole volume instead of a slab")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--serial-slab", type=int, default=32,
                    help="z-slices timed by the serial oracle reported beside the baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-flow", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched-segmentation section (profiling runs)")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="only the surrogate section (profiling runs)")
    ap.add_argument("--flow-only", action="store_true",
                    help="run only the extract_network_flow section (profiling)")
    ap.add_argument("--flow-stokes-iters", type=int, default=40000,
                    help="MINRES iteration limit of the flow section's Stokes solve")
    ap.add_argument("--pmc-workload", default="extract_pore_network",
                    help="workload named in the --pmc-summary table title")
    ap.add_argument("--pmc-summary", default=None, metavar="DIR",
                    help="print the markdown HBM traffic summary of the rocprofv3 --pmc runs under "
                         "DIR (pmc_copy_fetch, pmc_copy_write, pmc_net_fetch, pmc_net_write; "
                         "--copy-bytes gives the calibration size) and exit")
    ap.add_argument("--sq-summary", default=None, metavar="DIR",
                    help="print the markdown SQ instruction-counter summary of the rocprofv3 --pmc "
                         "runs under DIR (pmc_sq_a, pmc_sq_b) and exit")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run an elementwise negation that reads and writes this many bytes, "
                         "5 times (a known-size copy kernel to calibrate rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU only (the 512^3 volume is one GPU's block); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.pmc_summary:
        print(pmc_summary(args.pmc_summary, args.copy_bytes, args.size ** 3, args.pmc_workload))
        return
    if args.sq_summary:
        print(sq_summary(args.sq_summary, args.size ** 3, args.pmc_workload))
        return
    import torch

    from peclet.pnm import capi

    torch.zeros(1, device="cuda:0")  # bring up the GPU context before the OpenMP generators
    if args.copy_bytes:
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.neg(src, out=dst)  # a shader kernel; copy_ may go to a DMA engine
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4, "reps": 5}), flush=True)
        return
    if args.flow_only:  # profiling runs of the flow section alone
        ctx = capi.Context(0)
        print(json.dumps({"flow": bench_flow(ctx, reps=3, stokes_iters=args.flow_stokes_iters)}),
              flush=True)
        ctx.close()
        return
    if args.surrogate_only:
        ctx = capi.Context(0)
        print(json.dumps({"surrogate": bench_surrogate(ctx)}), flush=True)
        ctx.close()
        return
    sdf, desc = make_sdf(args.workload, args.size, args.seed)
    if args.batched_only:
        ctx = capi.Context(0)
        src = torch.from_numpy(sdf).to("cuda:0")
        print(json.dumps({"batched": bench_batched(ctx, src, reps=3)}), flush=True)
        ctx.close()
        return
    shape = sdf.shape
    n = sdf.size
    dims = (shape[2], shape[1], shape[0])
    dev = torch.device("cuda:0")
    s = torch.from_numpy(sdf).to(dev)
    seg = torch.empty(n, dtype=torch.int32, device=dev)
    pores = torch.empty(4 * 1_000_000, dtype=torch.float32, device=dev)
    pairs = torch.empty(2 * 40_000_000, dtype=torch.int32, device=dev)
    ctx = capi.Context(0)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)

    def step():
        return ctx.extract_pore_network_d(s, dims, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), pores, seg,
                                          pairs)

    for _ in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    # Per-kernel breakdown from an untimed pass with every launch bracketed by HIP events (the
    # events add a few microseconds between launches, so the timed steps below carry events
    # around the dominant kernel only).
    ctx.profile(True)
    ctx.profile_only(None)
    ctx.profile_read()
    nbk = max(1, min(args.steps, 5))
    for _ in range(nbk):
        step()
    prof = ctx.profile_read()
    kernels = {k: {"ms_per_step": round(v[0] / nbk, 4),
                   "avg_ms": round(v[0] / v[1], 5), "launches_per_step": v[1] / nbk}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    dom = next((k for k in kernels if k in ALGO_BYTES_PER_VOXEL), None)
    ctx.profile_only([dom] if dom is not None else ["-"])
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        res = step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    prof_dom = ctx.profile_read()
    ctx.profile_only(None)
    ctx.profile(False)
    npores, total_pores, npairs = res

    roof = None
    if dom is not None:
        bpl = ALGO_BYTES_PER_VOXEL[dom] * n
        dom_ms, dom_launches = prof_dom[dom]
        avg_s = dom_ms / dom_launches / 1e3
        ach = bpl / avg_s / 1e9
        measured = (args.workload == "polydisperse" and args.size == 512 and args.seed == 2001)
        traffic = load_traffic(dom) if measured else None
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": (os.path.relpath(TRAFFIC_SUMMARY, ROOT)
                                   if traffic is not None else None),
                "algorithmic_bytes_per_launch": bpl,
                "avg_ms": round(avg_s * 1e3, 5),
                "share_of_step": round(dom_ms / args.steps / ms, 3)}
    gvox = n / (ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gvox, 3),
        "unit": "Gvox/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(desc, shape=list(shape), spacing=[1.0, 1.0, 1.0],
                       pores=npores, pore_labels=int(seg.max().item()),
                       connections=npairs),
        "hbm_frac_pipeline": round(PIPELINE_BYTES_PER_VOXEL * n / (ms / 1e3) / 1e9
                                   / HBM_PEAK_GBS, 4),
        "roofline": roof,
        "kernels": kernels,
        "kernels_source": f"{nbk} untimed steps with every launch bracketed by HIP events",
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sdf, min(args.cpu_slab, shape[0]),
                                           min(args.serial_slab, shape[0]),
                                           reps=args.cpu_reps, full=args.cpu_full)
    else:
        out["cpu_baseline"] = None
    if not args.no_batched:
        seg = pairs = pores = None
        torch.cuda.empty_cache()
        out["batched"] = bench_batched(ctx, s.view(shape), reps=3)
    if not args.no_flow:
        torch.cuda.empty_cache()
        out["flow"] = bench_flow(ctx, reps=3, stokes_iters=args.flow_stokes_iters)
    if not args.no_surrogate:
        torch.cuda.empty_cache()
        out["surrogate"] = bench_surrogate(ctx)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
