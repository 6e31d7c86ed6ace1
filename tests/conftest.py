import glob
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP engine")


def load_fixture(name):
    with np.load(os.path.join(HERE, "golden", name), allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    meta = json.loads(str(d.pop("meta")))
    meta.update(d)
    return meta


def fixture_names():
    return sorted(os.path.basename(p) for p in glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        pytest.fail("oracle/liboracle.so is not built (run __graft_entry__.build())")
    o = Oracle()
    o.set_threads(min(8, os.cpu_count() or 1))
    return o


@pytest.fixture(scope="session")
def engine():
    """The HIP engine; GPU tests fail loudly (not skip) when it is unusable."""
    import seis_amd
    seis_amd.lib()
    n = seis_amd.device_count()
    if n < 1:
        pytest.fail("no GPU visible to the HIP engine")
    return seis_amd


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
