import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import Reference, reference_available
    if not reference_available():
        pytest.skip("oracle/_ref (the compiled reference) is not built here")
    return Reference()


@pytest.fixture(scope="session")
def tk():
    import tensorkyber_amd
    tensorkyber_amd.load_library()
    return tensorkyber_amd


@pytest.fixture(scope="session")
def gpu(tk):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    rc = tk.lib().tk_init()
    assert rc == 0, f"tk_init failed: {rc}"
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def digests():
    """Full-size digests of the compiled reference (tests/golden/digests.json)."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "digests.json")) as f:
        return json.load(f)


def negacyclic(a, b):
    """a * b in Z_q[x]/(x^256 + 1), schoolbook (int64), canonical."""
    a = np.asarray(a, dtype=np.int64) % 3329
    b = np.asarray(b, dtype=np.int64) % 3329
    full = np.convolve(a, b)
    r = full[:256].copy()
    r[: len(full) - 256] -= full[256:]
    return r % 3329


def canon(a):
    return np.mod(np.asarray(a, dtype=np.int32), 3329).astype(np.int16)


def centred(a):
    c = np.mod(np.asarray(a, dtype=np.int32), 3329)
    return np.where(c > 1664, c - 3329, c).astype(np.int16)
