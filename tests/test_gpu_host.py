"""GPU parity of host-resident batches streamed through the device
(tk_host_pipeline_*, ntt_batch_host): pinned and pageable host buffers,
chunk tails, in place and out of place, against the oracle."""
import numpy as np
import pytest

from conftest import canon, centred

pytestmark = pytest.mark.gpu

N = 256


@pytest.mark.parametrize("npoly,chunk,depth", [(1, 1024, 2), (5000, 1024, 3), (4096, 1024, 1),
                                               (70000, 16384, 4)])
def test_pipeline_pageable(tk, gpu, oracle, npoly, chunk, depth):
    x = tk.wl_uniform(npoly * N, 500 + npoly, -32768, 32768).reshape(npoly, N)
    p = tk.HostPipeline(chunk, depth)
    out = np.empty_like(x)
    p.run(x, out, direction=0)
    # any int16 input: the reference's int16 butterflies need canonical inputs
    assert np.array_equal(out, canon(oracle.ntt(canon(x)))), f"forward npoly {npoly}"
    back = out.copy()
    p.run(back, direction=1)  # in place
    assert np.array_equal(back, centred(oracle.invntt(out))), f"inverse npoly {npoly}"
    p.close()


def test_pipeline_pinned(tk, gpu, oracle):
    import torch
    npoly = 50000
    x = tk.wl_uniform(npoly * N, 510, 0, 3329).reshape(npoly, N)
    hin = torch.from_numpy(x).pin_memory()
    hout = torch.empty((npoly, N), dtype=torch.int16, pin_memory=True)
    p = tk.HostPipeline(8192, 3)
    p.run(hin, hout, direction=0, method="mfma32")
    assert np.array_equal(hout.numpy(), canon(oracle.ntt(x)))
    p.run(hout, hin, direction=1, form="canonical")
    assert np.array_equal(hin.numpy(), canon(oracle.invntt(hout.numpy())))
    p.close()


def test_ntt_batch_host_default_pipeline(tk, gpu, oracle):
    """The reference's host loop shape: in place over a host array larger
    than one default chunk."""
    import ctypes
    npoly = 70000
    x = tk.wl_edge_mix(npoly, 520)
    y = x.copy()
    assert tk.lib().ntt_batch_host(y.ctypes.data, npoly) == 0
    assert np.array_equal(y, canon(oracle.ntt(x)))
    z = y.copy()
    assert tk.lib().invntt_batch_host(z.ctypes.data, npoly) == 0
    assert np.array_equal(z, centred(oracle.invntt(y)))
    assert tk.lib().ntt_batch_host(ctypes.c_void_p(0), 0) == 0


def test_pipeline_rejects_bad_runs(tk, gpu):
    p = tk.HostPipeline(1024, 2)
    x = np.zeros((8, N), dtype=np.int16)
    L = tk.lib()
    # partial overlap of in and out; bad direction; bad form for the method
    assert L.tk_host_pipeline_run(p._p, x.ctypes.data + 512, x.ctypes.data, 4, 0, 0, 0) == tk.EINVAL
    assert L.tk_host_pipeline_run(p._p, x.ctypes.data, x.ctypes.data, 4, 2, 0, 0) == tk.EINVAL
    assert L.tk_host_pipeline_run(p._p, x.ctypes.data, x.ctypes.data, 4, 0, tk.METHOD_MFMA,
                                  tk.FORM_REF) == tk.ENOTSUP
    # device memory on either side (the staged path's host copies cannot touch it)
    import torch
    d = torch.zeros(8 * N, dtype=torch.int16, device="cuda:0")
    before = d.clone()
    assert L.tk_host_pipeline_run(p._p, d.data_ptr(), x.ctypes.data, 4, 0, 0, 0) == tk.EINVAL
    assert L.tk_host_pipeline_run(p._p, x.ctypes.data, d.data_ptr(), 4, 0, 0, 0) == tk.EINVAL
    assert L.tk_host_pipeline_run(p._p, d.data_ptr(), d.data_ptr(), 4, 0, 0, 0) == tk.EINVAL
    torch.cuda.synchronize()
    assert torch.equal(d, before)
    # the pipeline still runs after the rejected calls
    y = np.arange(8 * N, dtype=np.int16).reshape(8, N) % 3329
    assert L.tk_host_pipeline_run(p._p, y.ctypes.data, y.ctypes.data, 8, 0, 0, 0) == tk.OK
    p.close()


@pytest.mark.parametrize("method", ["mfma16", "mfma32", "butterfly"])
@pytest.mark.parametrize("npoly", [1, 17, 33, 20000])
def test_pinned_buffers_transformed_over_pcie(tk, gpu, oracle, method, npoly):
    """Pinned buffers are read and written by the transform kernel itself
    (no staging): out of place, in place, and a view inside a larger pinned
    allocation; then a pinned input with a pageable output takes the copy
    path."""
    import torch
    x = tk.wl_uniform(npoly * N, 530 + npoly, 0, 3329).reshape(npoly, N)
    want = canon(oracle.ntt(x))
    p = tk.HostPipeline(4096, 2)
    hin = torch.from_numpy(x).pin_memory()
    hout = torch.empty((npoly, N), dtype=torch.int16, pin_memory=True)
    p.run(hin, hout, direction=0, method=method)
    assert np.array_equal(hout.numpy(), want), "out of place"
    p.run(hout, direction=1, method=method)  # in place
    assert np.array_equal(hout.numpy(), centred(oracle.invntt(want))), "in place"
    big = torch.full((npoly + 10, N), 77, dtype=torch.int16).pin_memory()
    big[5:5 + npoly] = torch.from_numpy(x)
    p.run(big[5:5 + npoly], direction=0, method=method)
    host = big.numpy()
    assert np.array_equal(host[5:5 + npoly], want), "view inside a pinned allocation"
    assert (host[:5] == 77).all() and (host[5 + npoly:] == 77).all(), "written outside the view"
    pageable = np.empty_like(x)
    p.run(hin, pageable, direction=0, method=method)
    assert np.array_equal(pageable, want), "pinned in, pageable out"
    p.close()
