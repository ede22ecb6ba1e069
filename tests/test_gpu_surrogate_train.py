"""The surrogate trained on the GPU by its own kernels, against the method without it.

Targets g = Q/dp come from extract_network_flow on project-Stokes fields; train and test throats
come from different packings. The trained network must beat both the geometric baseline
g = c A^2/L (c fitted on the training throats) and the untrained network on the held-out
packings."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_trained_network_beats_geometric_baseline(capi_ctx):
    from peclet.pnm import surrogate as S

    seeds_train = list(range(7001, 7009))
    seeds_test = [7101, 7102, 7103]
    Ptr, Rtr, Ttr, _, info = S.build_dataset(capi_ctx, seeds_train, 72)
    Pte, Rte, Tte, _, _ = S.build_dataset(capi_ctx, seeds_test, 72)
    assert len(Ttr) > 200 and len(Tte) > 60, (len(Ttr), len(Tte), info)

    untrained = S.ThroatSurrogate(capi_ctx, seed=1)
    untrained.fit_standardization(Rtr, Ttr)
    err_untrained = S.median_rel_error(
        untrained.predict(Pte, untrained.standardize(Rte)).cpu().numpy(), Tte)

    c = S.geometric_baseline_fit(Rtr, Ttr)
    err_geo = S.median_rel_error(S.geometric_baseline_predict(Rte, c), Tte)

    model = S.ThroatSurrogate(capi_ctx, seed=1)
    hist = model.fit(Ptr, Rtr, Ttr, steps=600, batch=128, lr=3e-3, log_every=100)
    assert hist[-1][1] < hist[0][1]  # the training loss falls
    pred = model.predict(Pte, model.standardize(Rte)).cpu().numpy()
    err = S.median_rel_error(pred, Tte)
    print(f"median |g_pred-g|/g on held-out packings: trained {err:.3f}, "
          f"geometric A^2/L {err_geo:.3f}, untrained {err_untrained:.3f}")
    assert err < err_geo and err < err_untrained
