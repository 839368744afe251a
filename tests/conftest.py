import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")
    config.addinivalue_line("markers", "slow: long-running parity case")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def cf():
    import risdsdgcfem_thesis_amd as m
    return m


@pytest.fixture(scope="session")
def gpu(cf):
    if cf.device_count() < 1:
        pytest.skip("no GPU visible")
    return cf


def rel_node_err(a, b, floor=1e-3):
    """max_i |a_i - b_i| / max(|b_i|, floor) over 3-vectors."""
    a = np.asarray(a).reshape(-1, 3)
    b = np.asarray(b).reshape(-1, 3)
    return float((np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), floor)).max())


def beam_config1(mod):
    """Config 1: 10x10x40 cube beam (11x11x41 vertices), half sizes (0.05,0.05,0.2) m,
    z-min face fixed."""
    pos, tets = mod.mesh_box((0.05, 0.05, 0.2), (11, 11, 41))
    knc = np.nonzero(np.isclose(pos[2::3], -0.2))[0].astype(np.uint32)
    return pos, tets, knc
