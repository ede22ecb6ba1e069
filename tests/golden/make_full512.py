"""Expected results of the two 512^3 workloads, from the serial C oracle (oracle/pnm_oracle.c).

At full size the fixture keeps counts and SHA-256 digests instead of the arrays (the label
volume alone is 512 MiB): the input SDF (so that a generator difference shows up as such), the
pore records, the label volume and the connection list of extract_pore_network. The GPU test
(tests/test_gpu_full512.py) recomputes the digests from the device results.

    python tests/golden/make_full512.py      # about a minute per workload on one core
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
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "peclet-pnm_amd")]

import pnm_oracle as orc  # noqa: E402
from peclet.pnm import synth  # noqa: E402

SHAPE = (512, 512, 512)
WORKLOADS = {
    "polydisperse": lambda: synth.polydisperse_sdf(SHAPE, 9.0, 0.3, 0.35, seed=2001)[0],
    "boolean": lambda: synth.boolean_sdf(SHAPE, 9.0, 0.3, 0.5, seed=2002),
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pores_digest(pores):
    """Digest of the pore records' exact bytes (x, y, z, radius as float32 per pore)."""
    arr = np.stack([pores[k] for k in ("x", "y", "z", "radius")], axis=1).astype(np.float32)
    return digest(arr)


def main():
    out = {}
    for name, make in WORKLOADS.items():
        t = time.time()
        sdf = np.ascontiguousarray(make(), dtype=np.float32)
        pores, seg, conns = orc.extract_pore_network(sdf, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
        out[name] = {
            "sdf": digest(sdf),
            "npores": int(len(pores)),
            "pores": pores_digest(pores),
            "n_pore_labels": int(seg.max()),
            "n_solid_labels": int(-seg.min()),
            "seg": digest(seg.astype(np.int32)),
            "nconn": int(len(conns)),
            "conns": digest(np.asarray(conns, dtype=np.int32).reshape(-1, 2)),
        }
        print(name, json.dumps(out[name]), f"{time.time() - t:.1f} s", flush=True)
    with open(os.path.join(HERE, "full512.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
