"""Expected results of the reference's trig field (tests/kokkos_mpi/test_pnm_mpi.cpp:47-51, with
its origin and spacing, :169-170) at 256^3 and 512^3, from the serial C oracle.

The marker threshold at spacing (0.5, 1, 2) is -0.75, so the floodable band is a thick shell
and the Jacobi flood (pore_extraction.hpp:237-269) runs about 37 sweeps at 256^3 and 74 at
512^3. The fixture keeps the sweep count, counts and SHA-256 digests of the input SDF, the
pore records, the label volume and the connection list of extract_pore_network;
tests/test_gpu_deep_flood.py recomputes them from the GPU results.

    python tests/golden/make_deep_flood.py            # both sizes (minutes at 512^3)
    python tests/golden/make_deep_flood.py 256        # one size
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "peclet-pnm_amd"),
                os.path.dirname(HERE)]

import pnm_oracle as orc  # noqa: E402
from deepflood import TRIG_ORIGIN_ZYX, TRIG_SPACING_ZYX  # noqa: E402
from peclet.pnm import synth  # noqa: E402

OUT = os.path.join(HERE, "deep_flood.json")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pores_digest(pores):
    arr = np.stack([pores[k] for k in ("x", "y", "z", "radius")], axis=1).astype(np.float32)
    return digest(arr)


def main(sizes):
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    for n in sizes:
        t = time.time()
        sdf = np.ascontiguousarray(synth.trig_sdf((n, n, n)), dtype=np.float32)
        _, sweeps = orc.flood(sdf, orc.markers_ccl(sdf, TRIG_SPACING_ZYX))
        pores, seg, conns = orc.extract_pore_network(sdf, TRIG_ORIGIN_ZYX, TRIG_SPACING_ZYX)
        out[f"trig{n}"] = {
            "shape": [n, n, n],
            "origin_zyx": list(TRIG_ORIGIN_ZYX),
            "spacing_zyx": list(TRIG_SPACING_ZYX),
            "sdf": digest(sdf),
            "flood_sweeps": int(sweeps),
            "npores": int(len(pores)),
            "pores": pores_digest(pores),
            "n_pore_labels": int(seg.max()),
            "n_solid_labels": int(-seg.min()),
            "seg": digest(seg.astype(np.int32)),
            "nconn": int(len(conns)),
            "conns": digest(np.asarray(conns, dtype=np.int32).reshape(-1, 2)),
        }
        print(f"trig{n}", json.dumps(out[f"trig{n}"]), f"{time.time() - t:.1f} s", flush=True)
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [256, 512])
