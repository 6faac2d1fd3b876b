import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GRAPHS = os.path.join(GOLDEN, "graphs")

# fixture name -> (one_based, auto_n); matches tests/golden/make_golden.py
FIXTURES = {
    "tree": (False, False),
    "test4": (False, False),
    "test": (True, False),
    "test2": (True, False),
    "test3": (True, False),
    "test_isolated_node": (True, False),
    "mig_tree": (True, False),
    "wiki-Vote": (True, True),
}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-size workload (seconds to minutes)")


@pytest.fixture(scope="session")
def golden_levels():
    return dict(np.load(os.path.join(GOLDEN, "fixture_levels.npz")))


def fixture_path(name):
    return os.path.join(GRAPHS, f"{name}.txt")


def load_fixture(name):
    import gpu_graphs_amd as gg
    one_based, auto_n = FIXTURES[name]
    return gg.csr(fixture_path(name), one_based=one_based, auto_n=auto_n)


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when no GPU is visible: GPU tests must run on the GPU."""
    import gpu_graphs_amd as gg
    n = gg.device_count()
    assert n > 0, "no GPU visible to libgpugraphs.so"
    return 0
