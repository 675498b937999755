"""Partial last tiles that end exactly at the end of their allocation.

The MFMA and fused kernels stage a tile through a buffer descriptor whose
size is the tile's valid bytes, so that a partial last tile needs no per-lane
checks; the staging loads pass the chunk offset as the scalar offset, which
the gfx950 range check covers for these loads.  Here every buffer of a call
is placed so that the batch ends at the last byte of its own hipMalloc (or
hipHostMalloc, for the kernels that read pinned host memory over PCIe), with
batch sizes that leave the last tile partial, and the results are checked
against the oracle on every coefficient.  The other tests leave guard polys
or allocator slack after the batch.
"""
import ctypes

import numpy as np
import pytest

from conftest import canon, centred

pytestmark = pytest.mark.gpu

N, Q = 256, 3329
PB = N * 2  # bytes per poly


def _uniform(tk, npoly, seed):
    return tk.wl_uniform(npoly * N, seed).reshape(npoly, N)


@pytest.mark.parametrize("method", ["mfma16", "mfma32", "butterfly"])
def test_transforms_ending_at_allocation_end(tk, hip, oracle, method):
    L = tk.lib()
    m = tk.METHODS[method]
    for npoly in (1, 17, 31, 33, 4095):
        print(f"{method} npoly={npoly}", flush=True)
        x = _uniform(tk, npoly, 900 + npoly)
        src, dst = hip.at_end(npoly * PB), hip.at_end(npoly * PB)
        hip.put(src, x)
        assert L.ntt_batch_ex(dst, src, npoly, m, tk.FORM_CANONICAL, None) == 0
        hip.sync()
        fwd = canon(oracle.ntt(x))
        assert np.array_equal(hip.get(dst, x.shape), fwd), f"forward {method} {npoly}"
        assert L.invntt_batch_ex(dst, dst, npoly, m, tk.FORM_CENTERED, None) == 0
        hip.sync()
        assert np.array_equal(hip.get(dst, x.shape), centred(oracle.invntt(fwd))), \
            f"inverse {method} {npoly}"
        assert np.array_equal(hip.get(src, x.shape), x), "input changed"


@pytest.mark.parametrize("k,ninst", [(2, 1), (3, 17), (4, 17), (4, 33)])
def test_fused_buffers_ending_at_allocation_end(tk, hip, oracle, k, ninst):
    """matvec (int16 and 3-bit packed s) and decryption with every buffer at
    the end of its own allocation; a k = 4 matrix tile is 128 KB."""
    L = tk.lib()
    print(f"fused k={k} ninst={ninst}", flush=True)
    A = tk.wl_uniform(ninst * k * k * N, 910 + k, 0, Q).reshape(ninst, k, k, N)
    s = tk.wl_cbd2(ninst * k * N, 911 + k).reshape(ninst, k, N)
    dA, ds, dt = hip.at_end(A.nbytes), hip.at_end(s.nbytes), hip.at_end(s.nbytes)
    hip.put(dA, A)
    hip.put(ds, s)
    assert L.matvec_ntt_fused(dt, dA, ds, k, ninst, None) == 0
    hip.sync()
    want = centred(oracle.matvec(A, s, k))
    assert np.array_equal(hip.get(dt, s.shape), want), f"matvec k {k} ninst {ninst}"

    ppb = tk.packed_poly_bytes(3)
    sp = oracle.pack_small(s.reshape(-1, N), 3)
    dsp, dt2 = hip.at_end(sp.nbytes), hip.at_end(s.nbytes)
    hip.put(dsp, sp)
    assert sp.nbytes == ninst * k * ppb
    assert L.matvec_ntt_fused_packed(dt2, dA, dsp, k, ninst, 3, None) == 0
    hip.sync()
    assert np.array_equal(hip.get(dt2, s.shape), want), f"packed matvec k {k} ninst {ninst}"

    w = tk.kpke_workload(ninst, k, seed=920 + k)
    _, s_hat = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    u = w["r"]  # any small polys serve as u
    v = w["e2"]
    want_w = oracle.kpke_decrypt(s_hat, u, v, k)
    dsh, du, dv, dw = (hip.at_end(s_hat.nbytes), hip.at_end(u.nbytes), hip.at_end(v.nbytes),
                       hip.at_end(v.nbytes))
    hip.put(dsh, s_hat)
    hip.put(du, u)
    hip.put(dv, v)
    assert L.kpke_decrypt_core(dw, dsh, du, dv, k, ninst, None) == 0
    hip.sync()
    assert np.array_equal(hip.get(dw, v.shape), want_w), f"decrypt k {k} ninst {ninst}"


@pytest.mark.parametrize("bits", [2, 3])
def test_packed_forward_ending_at_allocation_end(tk, hip, oracle, bits):
    L = tk.lib()
    for method in ("mfma16", "mfma32"):
        for npoly in (1, 17, 33):
            print(f"packed {bits} bits {method} npoly={npoly}", flush=True)
            x = tk.packed_workload(bits, npoly)
            p = oracle.pack_small(x, bits)
            dp, dout = hip.at_end(p.nbytes), hip.at_end(npoly * PB)
            hip.put(dp, p)
            assert L.ntt_batch_packed(dout, dp, npoly, bits, tk.METHODS[method],
                                      tk.FORM_CANONICAL, None) == 0
            hip.sync()
            assert np.array_equal(hip.get(dout, x.shape), canon(oracle.ntt(x))), \
                f"packed {bits} {method} {npoly}"


@pytest.mark.parametrize("method", ["mfma16", "mfma32"])
def test_pinned_host_batch_ending_at_allocation_end(tk, hip, oracle, method):
    """The host pipeline runs the transform kernel straight over pinned
    buffers (PCIe): here each batch ends at the last byte of its own
    hipHostMalloc."""
    L = tk.lib()
    size = 64 << 10  # 128 polys per pinned allocation
    p = ctypes.c_void_p()
    assert L.tk_host_pipeline_create(ctypes.byref(p), 1024, 2) == 0
    try:
        for npoly in (1, 17, 127):
            print(f"pinned {method} npoly={npoly}", flush=True)
            x = _uniform(tk, npoly, 930 + npoly)
            src = hip.at_end(npoly * PB, pinned=True, size=size)
            dst = hip.at_end(npoly * PB, pinned=True, size=size)
            hin = np.ctypeslib.as_array((ctypes.c_int16 * (npoly * N)).from_address(src))
            hout = np.ctypeslib.as_array((ctypes.c_int16 * (npoly * N)).from_address(dst))
            hin[:] = x.reshape(-1)
            assert L.tk_host_pipeline_run(p, dst, src, npoly, 0, tk.METHODS[method],
                                          tk.FORM_CANONICAL) == 0
            fwd = canon(oracle.ntt(x))
            assert np.array_equal(hout.reshape(npoly, N), fwd), f"pinned forward {npoly}"
            assert L.tk_host_pipeline_run(p, dst, dst, npoly, 1, tk.METHODS[method],
                                          tk.FORM_CENTERED) == 0
            assert np.array_equal(hout.reshape(npoly, N), centred(oracle.invntt(fwd)))
    finally:
        L.tk_host_pipeline_destroy(p)


@pytest.mark.parametrize("method", ["mfma16", "mfma32"])
def test_ragged_requests_ending_at_allocation_end(tk, hip, oracle, method):
    """A ragged plan whose requests each end at the last byte of their own
    allocation, with sizes that leave every request's last tile partial."""
    L = tk.lib()
    sizes = (1, 17, 33, 47)
    reqs = (tk.TkRequest * len(sizes))()
    xs, bases = [], []
    for r, n in enumerate(sizes):
        x = _uniform(tk, n, 940 + n)
        b = hip.at_end(n * PB)
        hip.put(b, x)
        reqs[r].base, reqs[r].npoly, reqs[r].k, reqs[r].kind = b, n, 2, 0
        xs.append(x)
        bases.append(b)
    for direction in (0, 1):
        print(f"ragged {method} direction {direction}", flush=True)
        plan = ctypes.c_void_p()
        assert L.tk_ragged_plan_create(ctypes.byref(plan), ctypes.addressof(reqs), len(sizes),
                                       direction, tk.METHODS[method], None) == 0
        assert L.tk_ragged_plan_execute(plan, None) == 0
        hip.sync()
        for x, b in zip(xs, bases):
            got = hip.get(b, x.shape)
            want = canon(oracle.ntt(x)) if direction == 0 else centred(oracle.invntt(x))
            assert np.array_equal(got, want), f"{method} dir {direction} request of {len(x)}"
            hip.put(b, got)
        xs = [hip.get(b, x.shape) for x, b in zip(xs, bases)]
        assert L.tk_ragged_plan_destroy(plan) == 0
