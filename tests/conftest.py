import ctypes
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


class Hip:
    """The HIP runtime already loaded into this process (torch's), through
    ctypes, so that pointers, streams and graphs are shared with libtkyber:
    exact-size allocations and graph calls that torch does not expose."""

    ALLOC = 2 << 20  # bytes per at_end allocation (a multiple of any page size)

    def __init__(self):
        import torch
        torch.cuda.init()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        assert path, "HIP runtime not loaded"
        self.lib = ctypes.CDLL(path)
        vp, sz, pvp = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)
        for name, args in (("hipMalloc", [pvp, sz]), ("hipFree", [vp]),
                           ("hipHostMalloc", [pvp, sz, ctypes.c_uint]),
                           ("hipHostFree", [vp]), ("hipMemcpy", [vp, vp, sz, ctypes.c_int]),
                           ("hipDeviceSynchronize", []),
                           ("hipGraphCreate", [pvp, ctypes.c_uint]),
                           ("hipGraphDestroy", [vp]),
                           ("hipGraphAddChildGraphNode", [pvp, vp, vp, sz, vp]),
                           ("hipGraphInstantiate", [pvp, vp, vp, vp, sz]),
                           ("hipGraphLaunch", [vp, vp]),
                           ("hipGraphExecDestroy", [vp])):
            getattr(self.lib, name).argtypes = args
            getattr(self.lib, name).restype = ctypes.c_int
        self.live = []

    def alloc(self, nbytes, pinned=False):
        p = ctypes.c_void_p()
        rc = (self.lib.hipHostMalloc(ctypes.byref(p), nbytes, 0) if pinned
              else self.lib.hipMalloc(ctypes.byref(p), nbytes))
        assert rc == 0 and p.value, f"allocation of {nbytes} bytes failed: {rc}"
        self.live.append((p.value, pinned))
        return p.value

    def at_end(self, nbytes, pinned=False, size=ALLOC):
        """Address of nbytes that end at the last byte of a fresh allocation."""
        assert nbytes <= size
        return self.alloc(size, pinned) + size - nbytes

    def put(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.lib.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def get(self, src, shape, dtype=np.int16):
        a = np.empty(shape, dtype)
        assert self.lib.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0
        return a

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def instantiate(self, graph):
        ex = ctypes.c_void_p()
        assert self.lib.hipGraphInstantiate(ctypes.byref(ex), graph, None, None, 0) == 0
        return ex.value

    def compose(self, children):
        """A new graph running the child graphs one after another."""
        g = ctypes.c_void_p()
        assert self.lib.hipGraphCreate(ctypes.byref(g), 0) == 0
        prev = None
        for child in children:
            node = ctypes.c_void_p()
            deps = (ctypes.c_void_p * 1)(prev) if prev else None
            assert self.lib.hipGraphAddChildGraphNode(ctypes.byref(node), g, deps,
                                                      1 if prev else 0, child) == 0
            prev = node.value
        return g.value

    def free_all(self):
        self.sync()
        for p, pinned in self.live:
            (self.lib.hipHostFree if pinned else self.lib.hipFree)(p)
        self.live = []


@pytest.fixture
def hip(gpu):
    h = Hip()
    yield h
    h.free_all()
