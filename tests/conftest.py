import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def load_golden(family: str) -> dict:
    with np.load(os.path.join(GOLDEN, f"{family}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_cases(g: dict, prefix: str):
    """Group fixture keys "<fn>/<case>/<field>" into {(fn, case): {field: arr}}."""
    out: dict = {}
    for k, v in g.items():
        parts = k.split("/")
        if not parts[0].startswith(prefix):
            continue
        key = (parts[0], "/".join(parts[1:-1]))
        out.setdefault(key, {})[parts[-1]] = v
    return out


@pytest.fixture(scope="session")
def golden():
    return load_golden
