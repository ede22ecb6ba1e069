"""Shared test setup: import paths, the `gpu` marker, and small helpers."""
from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "peclet-pnm_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def capi_ctx():
    """A libpnm_hip context on device 0 (GPU tests only). torch's GPU context is brought up
    first, before any OpenMP region of the geometry generators runs in this process."""
    import torch

    torch.zeros(1, device="cuda")
    from peclet.pnm import capi

    ctx = capi.Context(0)
    yield ctx
    ctx.close()
