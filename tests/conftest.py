"""Shared test setup: the `gpu` marker, the repository on sys.path, and fixtures for the checkers and a context.

CPU tests (`-m "not gpu"`) pin the oracle against the reference's golden vectors and check host logic and the
C ABI's exported symbols. GPU tests (`-m gpu`) run the HIP path through the C ABI and compare it with the oracle.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libiscg.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.Oracle()


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name))
    return load


@pytest.fixture(scope="session")
def ctx():
    from insiliscope_amd.host import Context
    c = Context(0)
    yield c
    c.close()


def record_accuracy(case: str, **values):
    """Appends one measured accuracy record (max abs and relative L2 errors against the FP64 oracle, ...) as a JSON
    line to the file named by ISCG_ACCURACY_OUT, when set; the GPU tests call it where they measure an error."""
    path = os.environ.get("ISCG_ACCURACY_OUT")
    if path:
        import json
        with open(path, "a") as f:
            f.write(json.dumps({"case": case, **{k: float(v) for k, v in values.items()}}) + "\n")
