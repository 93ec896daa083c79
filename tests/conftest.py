import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np

    d = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(d, "golden.json")) as f:
        man = json.load(f)
    npz = np.load(os.path.join(d, "golden.npz"), allow_pickle=False)
    cases, pad_cases = [], []
    for c in man["cases"]:
        n = c["name"]
        d = dict(c, x=npz[n + "__x"], archive=npz[n + "__archive"].tobytes(), out=npz[n + "__out"])
        (pad_cases if "pad_mode" in c else cases).append(d)
    return {"cases": cases, "pad_cases": pad_cases, "digests": man["digests"]}


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context on cuda:0 for all GPU tests."""
    import torch

    from vecsz_amd import Context

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return Context(0)
