"""Write the golden fixtures under tests/golden/ (inputs + expected outputs).

The expected outputs come from tests/bruteforce.py — formulations independent of the oracle's
restatement — not from the oracle itself, so the fixtures pin the oracle (and through it the
HIP path). The reference ships no golden vectors of its own (it has no committed test data);
its test geometries are restated as inputs here:
  ci_lattice     .github/workflows/ci.yml:61-68
  sphere_lattice tests/kokkos_mpi/test_pnm_mpi.cpp:29-45, 167-168 (36x30x24, unit spacing)
  trig           tests/kokkos_mpi/test_pnm_mpi.cpp:47-51, 169-170 (32^3, origin (-1,2,0.5),
                 spacing (0.5,1,2) in xyz)

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "peclet-pnm_amd")]

import bruteforce as bf  # noqa: E402
from peclet.pnm import synth  # noqa: E402


def quantized_field(shape, seed, step=0.5, lo=-3.0, hi=3.0):
    """Random field rounded to a coarse step: many exact ties and plateaus."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(lo, hi, size=shape)
    return (np.round(v / step) * step).astype(np.float32)


def tent_ramp(nx=1100):
    """(1,1,nx) tent: values rise by 1 per voxel to the middle. Walks from the ends are deeper
    than 512 steps, so their roots are fixed by the step cap."""
    x = np.arange(nx)
    v = 1.0 + np.minimum(x, nx - 1 - x)
    return v.astype(np.float32).reshape(1, 1, nx)


def cases():
    yield "ci_lattice", synth.ci_lattice_sdf(32, 10.0), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    yield "sphere_lattice", synth.sphere_lattice_sdf((24, 30, 36)), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    yield "trig", synth.trig_sdf((32, 32, 32)), (0.5, 2.0, -1.0), (2.0, 1.0, 0.5)
    sdf, _ = synth.packing_sdf((40, 40, 40), "uniform", 3.0, 5.0, 0.55, seed=11)
    yield "packing40", sdf, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    yield "boolean40", synth.boolean_sdf((40, 40, 40), 4.0, 0.3, 0.5, seed=12), (1.5, -2.0, 0.25), (1.0, 1.0, 1.0)
    yield "quantized", quantized_field((16, 18, 20), 13), (0.0, 0.0, 0.0), (1.0, 0.75, 1.25)
    yield "tiny_1x3x5", quantized_field((1, 3, 5), 14), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    yield "tiny_2x2x2", quantized_field((2, 2, 2), 15), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    yield "tiny_1x1x7", quantized_field((1, 1, 7), 16), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    yield "tent1100", tent_ramp(1100), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)


def expected(sdf, origin_zyx, spacing_zyx):
    pk = bf.peaks(sdf)
    peak_idx = np.flatnonzero(pk.ravel())
    nz, ny, nx = sdf.shape
    cen = np.zeros((peak_idx.size, 3))
    for k, i in enumerate(peak_idx):
        z, y, x = i // (ny * nx), (i // nx) % ny, i % nx
        f = bf.centroid64(sdf, z, y, x)
        o = np.array([origin_zyx[2], origin_zyx[1], origin_zyx[0]], dtype=np.float64)
        s = np.array([spacing_zyx[2], spacing_zyx[1], spacing_zyx[0]], dtype=np.float64)
        cen[k] = o + (np.array([x, y, z]) + f) * s
    lab_ccl = bf.ccl(bf.markers(sdf, spacing_zyx))
    lab = bf.flood_layers(sdf, lab_ccl)
    rts = bf.roots(sdf)
    seg = bf.relabel(sdf, rts, lab)
    topo = bf.topology(seg, sdf.shape)
    flow_roots = bf.roots(sdf, all_voxels=True)
    return dict(peak_idx=peak_idx.astype(np.int64), peak_xyz=cen,
                peak_radius=sdf.ravel()[peak_idx], labels_ccl=lab_ccl, labels=lab, roots=rts,
                flow_roots=flow_roots, seg=seg, topology=topo)


def main():
    for name, sdf, org, spc in cases():
        exp = expected(sdf, org, spc)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, sdf=sdf, origin_zyx=np.array(org), spacing_zyx=np.array(spc),
                            **exp)
        print(f"{name}: shape={sdf.shape} peaks={exp['peak_idx'].size} "
              f"pores={int(exp['seg'].max())} solids={int(-exp['seg'].min())} "
              f"pairs={len(exp['topology'])}")


if __name__ == "__main__":
    main()
