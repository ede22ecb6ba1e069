"""The two 512^3 workloads against the serial C oracle's results, bit-exact at full size:
pore records, label volume and connection list of extract_pore_network compared by SHA-256
digest with tests/golden/full512.json (written by tests/golden/make_full512.py from
oracle/pnm_oracle.c). The input SDF's digest is checked first, so a generator difference is
reported as such and not as a pipeline difference."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPE = (512, 512, 512)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_sdf(kind):
    from peclet.pnm import synth

    if kind == "polydisperse":
        return synth.polydisperse_sdf(SHAPE, 9.0, 0.3, 0.35, seed=2001)[0]
    return synth.boolean_sdf(SHAPE, 9.0, 0.3, 0.5, seed=2002)


@pytest.mark.parametrize("kind", ["polydisperse", "boolean"])
def test_full512_bit_exact(capi_ctx, kind):
    with open(os.path.join(HERE, "golden", "full512.json")) as f:
        exp = json.load(f)[kind]
    sdf = np.ascontiguousarray(make_sdf(kind), dtype=np.float32)
    assert digest(sdf) == exp["sdf"], "SDF generator output differs from the fixture's input"
    n = sdf.size
    s = torch.from_numpy(sdf).cuda()
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    pores = torch.empty(4 * 1_000_000, dtype=torch.float32, device="cuda")
    pairs = torch.empty(2 * 20_000_000, dtype=torch.int32, device="cuda")
    npores, tot, npairs = capi_ctx.extract_pore_network_d(s, SHAPE, (0, 0, 0), (1, 1, 1), pores,
                                                          seg, pairs)
    assert npores == exp["npores"]
    assert digest(pores[: 4 * npores].view(-1, 4).cpu().numpy()) == exp["pores"]
    segh = seg.cpu().numpy()
    assert int(segh.max()) == exp["n_pore_labels"]
    assert int(-segh.min()) == exp["n_solid_labels"]
    assert digest(segh) == exp["seg"]
    assert npairs == exp["nconn"]
    assert digest(pairs[: 2 * npairs].view(-1, 2).cpu().numpy()) == exp["conns"]


def test_full512_repeated_alternating(capi_ctx):
    """Both workloads alternately on one context, six times each: every label volume and
    connection list must still equal the oracle's (a timing-dependent race in a flatten once
    left an intermediate union-find node as a label in about one run in ten to thirty)."""
    with open(os.path.join(HERE, "golden", "full512.json")) as f:
        exp = json.load(f)
    vols = {k: torch.from_numpy(np.ascontiguousarray(make_sdf(k), dtype=np.float32)).cuda()
            for k in ("polydisperse", "boolean")}
    n = 512 ** 3
    seg = torch.empty(n, dtype=torch.int32, device="cuda")
    pores = torch.empty(4 * 1_000_000, dtype=torch.float32, device="cuda")
    pairs = torch.empty(2 * 20_000_000, dtype=torch.int32, device="cuda")
    bad = []
    for it in range(6):
        for kind, s in vols.items():
            _, _, npairs = capi_ctx.extract_pore_network_d(s, SHAPE, (0, 0, 0), (1, 1, 1), pores,
                                                           seg, pairs)
            ok = (digest(seg.cpu().numpy()) == exp[kind]["seg"] and npairs == exp[kind]["nconn"]
                  and digest(pairs[: 2 * npairs].view(-1, 2).cpu().numpy()) == exp[kind]["conns"])
            if not ok:
                bad.append((it, kind, npairs))
    assert not bad, bad
