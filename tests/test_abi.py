"""C ABI without a GPU: the library loads, exports every declared symbol, and
rejects bad arguments before anything touches the device; synthetic workload
generators are deterministic and have the documented shapes."""
import ctypes
import os
import re

import numpy as np
import pytest

N, Q = 256, 3329
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in ("tkyber.h", "tkyber_workload.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;]*\)\s*;", text, flags=re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol(tk):
    names = declared_functions()
    assert "ntt_batch" in names and "invntt_batch" in names and "basemul_batch" in names
    lib = ctypes.CDLL(tk.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_version_and_strerror(tk):
    L = tk.lib()
    assert b"gfx950" in L.tk_version()
    for code in (0, -22, -12, -5, -95, -19):
        assert L.tk_strerror(code) and L.tk_strerror(code) != b"unknown status"


def test_validation_before_device(tk):
    L = tk.lib()
    buf = (ctypes.c_int16 * (N * 8 + 16))()
    base = ctypes.addressof(buf)
    al = (base + 15) & ~15
    assert L.ntt_batch(None, 5, None) == tk.EINVAL
    assert L.invntt_batch(None, 5, None) == tk.EINVAL
    assert L.ntt_batch(al + 2, 1, None) == tk.EINVAL            # misaligned
    assert L.ntt_batch(None, 0, None) == tk.OK                  # empty batch: no-op
    assert L.ntt_batch_ex(al, al, 1, 99, 0, None) == tk.EINVAL  # bad method
    assert L.ntt_batch_ex(al, al, 1, 0, 99, None) == tk.EINVAL  # bad form
    assert L.ntt_batch_ex(al, al, 1, tk.METHOD_MFMA, tk.FORM_REF, None) == tk.ENOTSUP
    assert L.invntt_batch_ex(al + 32, al, 2, 0, 0, None) == tk.EINVAL  # partial overlap
    assert L.basemul_batch(None, al, al, 1, None) == tk.EINVAL
    assert L.basemul_batch_ex(al, al, al, 1, 99, None) == tk.EINVAL
    assert L.basemul_batch(al + 32, al, al, 2, None) == tk.EINVAL


def test_ragged_and_fused_validation(tk):
    L = tk.lib()
    if not hasattr(L, "matvec_ntt_fused"):
        pytest.skip("fused entry point not built")
    assert L.matvec_ntt_fused(None, None, None, 3, 4, None) == tk.EINVAL
    assert L.matvec_ntt_fused(None, None, None, 3, 0, None) == tk.OK
    buf = (ctypes.c_int16 * (N * 64 + 16))()
    al = (ctypes.addressof(buf) + 15) & ~15
    assert L.matvec_ntt_fused(al, al, al, 5, 1, None) == tk.EINVAL    # k out of range
    plan = ctypes.c_void_p()
    assert L.tk_ragged_plan_create(ctypes.byref(plan), None, 3, 0, 0, None) == tk.EINVAL
    assert L.tk_ragged_plan_destroy(None) == tk.EINVAL


def test_packed_validation(tk):
    L = tk.lib()
    assert [L.tk_packed_poly_bytes(b) for b in (1, 2, 3, 4, 5, 8)] == [0, 64, 128, 128, 0, 0]
    buf = (ctypes.c_int16 * (N * 64 + 16))()
    al = (ctypes.addressof(buf) + 15) & ~15
    assert L.pack_small_batch(al, al + 4096, 1, 5, None) == tk.EINVAL       # bad width
    assert L.pack_small_batch(None, al, 1, 2, None) == tk.EINVAL
    assert L.pack_small_batch(al + 8, al + 4096, 1, 2, None) == tk.EINVAL   # misaligned
    assert L.pack_small_batch(al, al + 32, 2, 2, None) == tk.EINVAL         # overlap
    assert L.pack_small_batch(None, None, 0, 2, None) == tk.OK              # empty: no-op
    assert L.ntt_batch_packed(al, al + 4096, 1, 7, 0, 0, None) == tk.EINVAL
    assert L.ntt_batch_packed(al, al + 4096, 1, 2, 99, 0, None) == tk.EINVAL  # bad method
    assert L.ntt_batch_packed(al, al + 4096, 1, 2, 0, 99, None) == tk.EINVAL  # bad form
    assert L.ntt_batch_packed(al, al + 4096, 1, 2, tk.METHOD_BUTTERFLY, 0, None) == tk.ENOTSUP
    assert L.ntt_batch_packed(al, al + 4096, 1, 2, 0, tk.FORM_REF, None) == tk.ENOTSUP
    assert L.ntt_batch_packed(al, al + 256, 1, 2, 0, 0, None) == tk.EINVAL  # overlap
    assert L.matvec_ntt_fused_packed(al, al, al, 3, 1, 5, None) == tk.EINVAL
    assert L.matvec_ntt_fused_packed(al, al, al, 5, 1, 3, None) == tk.EINVAL
    assert L.matvec_ntt_fused_packed(al, al + 20000, al, 2, 1, 3, None) == tk.EINVAL  # t over s
    assert L.matvec_ntt_fused_packed(None, None, None, 3, 0, 3, None) == tk.OK


def test_kpke_validation(tk):
    L = tk.lib()
    buf = (ctypes.c_int16 * (N * 256 + 16))()
    al = (ctypes.addressof(buf) + 15) & ~15
    P = N * 2  # bytes per poly
    # k out of range; empty batch; null inputs
    assert L.kpke_keygen_core(al, al + 10 * P, al + 20 * P, al + 60 * P, al + 70 * P, 5, 1,
                              None) == tk.EINVAL
    assert L.kpke_keygen_core(None, None, None, None, None, 3, 0, None) == tk.OK
    assert L.kpke_keygen_core(al, None, al + 20 * P, al + 60 * P, al + 70 * P, 3, 1,
                              None) == tk.EINVAL
    # t_hat overlapping A_hat (not an allowed alias), misaligned output
    assert L.kpke_keygen_core(al + 20 * P, al + 10 * P, al + 20 * P, al + 60 * P, al + 70 * P, 3,
                              1, None) == tk.EINVAL
    assert L.kpke_keygen_core(al + 2, al + 10 * P, al + 20 * P, al + 60 * P, al + 70 * P, 3, 1,
                              None) == tk.EINVAL
    # u partially overlapping e1; v aliasing m
    assert L.kpke_encrypt_core(al + P, al + 10 * P, al + 20 * P, al + 40 * P, al + 50 * P, al,
                               al + 60 * P, al + 70 * P, 3, 1, None) == tk.EINVAL
    assert L.kpke_encrypt_core(al, al + 70 * P, al + 20 * P, al + 40 * P, al + 50 * P, al + 80 * P,
                               al + 60 * P, al + 70 * P, 3, 1, None) == tk.EINVAL
    assert L.kpke_encrypt_core(None, None, None, None, None, None, None, None, 2, 0, None) == tk.OK
    # w partially overlapping u
    assert L.kpke_decrypt_core(al + 21 * P, al, al + 20 * P, al + 60 * P, 3, 1, None) == tk.EINVAL
    assert L.kpke_decrypt_core(al, al, al, al, 1, 1, None) == tk.EINVAL


def test_host_pipeline_validation(tk):
    L = tk.lib()
    p = ctypes.c_void_p()
    assert L.tk_host_pipeline_create(ctypes.byref(p), 0, 3) == tk.EINVAL
    assert L.tk_host_pipeline_create(ctypes.byref(p), 1024, 0) == tk.EINVAL
    assert L.tk_host_pipeline_create(ctypes.byref(p), 1024, 5) == tk.EINVAL
    assert L.tk_host_pipeline_create(None, 1024, 2) == tk.EINVAL
    assert L.tk_host_pipeline_run(None, None, None, 4, 0, 0, 0) == tk.EINVAL
    assert L.tk_host_pipeline_destroy(None) == tk.EINVAL
    assert L.ntt_batch_host(None, 5) == tk.EINVAL
    assert L.invntt_batch_host(None, 5) == tk.EINVAL
    assert L.ntt_batch_host(None, 0) == tk.OK


def test_fused_plan_validation(tk):
    L = tk.lib()
    plan = ctypes.c_void_p()
    assert L.tk_fused_plan_create(ctypes.byref(plan), None, 1) == tk.EINVAL
    buf = (ctypes.c_int16 * (N * 64 + 16))()
    al = (ctypes.addressof(buf) + 15) & ~15
    reqs = (tk.TkFusedRequest * 2)()
    reqs[0].mode, reqs[0].k, reqs[0].ninst = tk.FUSED_MATVEC, 3, 1
    reqs[0].out0, reqs[0].mat, reqs[0].x = al, al + 3 * 512, al + 12 * 512
    reqs[1].mode, reqs[1].k, reqs[1].ninst = tk.FUSED_DECRYPT, 2, 0  # empty: skipped
    table = ctypes.addressof(reqs)
    reqs[0].mode = 7
    assert L.tk_fused_plan_create(ctypes.byref(plan), table, 2) == tk.EINVAL   # bad mode
    reqs[0].mode, reqs[0].k = tk.FUSED_MATVEC, 5
    assert L.tk_fused_plan_create(ctypes.byref(plan), table, 2) == tk.EINVAL   # bad rank
    reqs[0].k, reqs[0].x = 3, None
    assert L.tk_fused_plan_create(ctypes.byref(plan), table, 2) == tk.EINVAL   # missing x
    reqs[0].x, reqs[0].out0 = al + 12 * 512, al + 4 * 512
    assert L.tk_fused_plan_create(ctypes.byref(plan), table, 2) == tk.EINVAL   # out over mat
    reqs[0].ninst = 0
    assert L.tk_fused_plan_create(ctypes.byref(plan), table, 2) == tk.EINVAL   # nothing to do
    assert L.tk_fused_plan_destroy(None) == tk.EINVAL
    assert L.tk_fused_plan_execute(None, None) == tk.EINVAL


def test_workload_determinism(tk):
    a = tk.wl_uniform(1 << 17, 42)
    b = tk.wl_uniform(1 << 17, 42)
    assert np.array_equal(a, b) and a.min() >= 0 and a.max() < Q
    # counter-based: a prefix equals the start of a longer draw
    assert np.array_equal(tk.wl_uniform(1000, 42), a[:1000])
    assert not np.array_equal(tk.wl_uniform(1000, 43), a[:1000])
    c = tk.wl_uniform(100000, 7, -3328, 3329)
    assert c.min() == -3328 and c.max() == 3328
    # the documented draw: ((v >> 32) * n) >> 32
    v = tk.lib().tk_wl_draw(42, tk.WL_COEF, 5)
    assert a[5] == ((v >> 32) * Q) >> 32


def test_edge_mix_categories(tk):
    x = tk.wl_edge_mix(20000, 3)
    zero = (x == 0).all(axis=1)
    hi = (x == 3328).all(axis=1)
    lo = (x == -3328).all(axis=1)
    imp = ((x == 1).sum(axis=1) == 1) & ((x == 0).sum(axis=1) == 255)
    uni = ~(zero | hi | lo | imp)
    frac = np.array([uni.mean(), zero.mean(), hi.mean(), lo.mean(), imp.mean()])
    assert np.allclose(frac, [0.70, 0.10, 0.10, 0.05, 0.05], atol=0.015)
    assert x[uni].min() >= -3328 and x[uni].max() <= 3328


def test_cbd2(tk):
    s = tk.wl_cbd2(1 << 16, 5)
    counts = np.bincount(s + 2, minlength=5) / s.size
    assert s.min() >= -2 and s.max() <= 2
    assert np.allclose(counts, [1 / 16, 4 / 16, 6 / 16, 4 / 16, 1 / 16], atol=0.01)


def test_ragged_layout(tk):
    reqs, pool = tk.wl_ragged_layout(1 << 20, 4)
    off, n, k, kind = reqs.T
    assert n.sum() == 1 << 20 and n.min() >= 1 and n.max() <= 4096
    assert set(np.unique(k)) <= {2, 3, 4} and set(np.unique(kind)) <= {0, 1}
    assert np.all(off[1:] >= off[:-1] + n[:-1]), "requests overlap"
    assert off[-1] + n[-1] <= pool
    obj = np.where(kind == 1, k * k, k)
    assert np.all(n[:-1] % obj[:-1] == 0), "requests hold whole vectors/matrices"
    assert 0.18 < (kind == 1).mean() < 0.32


def test_canon_digest_is_representative_free(tk):
    x = tk.wl_uniform(4096, 9, -32768, 32768)
    y = (x.astype(np.int32) + Q * np.where(x > 0, -3, 3)).astype(np.int16)
    assert tk.canon_digest(x) == tk.canon_digest(y)


def _created_or_nodev(tk, L, rc, plan, destroy):
    """A table that passes validation reaches the device: ENODEV here, a plan
    on a GPU (destroyed at once)."""
    assert rc in (tk.ENODEV, tk.OK), rc
    if rc == tk.OK:
        assert destroy(plan) == tk.OK


def test_ragged_plan_rejects_overlapping_requests(tk):
    L = tk.lib()
    buf = (ctypes.c_int16 * (N * 64 + 16))()
    al = (ctypes.addressof(buf) + 15) & ~15
    plan = ctypes.c_void_p()
    reqs = (tk.TkRequest * 3)()
    for r, (off, n) in enumerate(((0, 6), (6, 4), (20, 8))):
        reqs[r].base, reqs[r].npoly, reqs[r].k, reqs[r].kind = al + off * 512, n, 2, 0

    def create():
        return L.tk_ragged_plan_create(ctypes.byref(plan), ctypes.addressof(reqs), 3, 0,
                                       tk.METHOD_MFMA16, None)
    _created_or_nodev(tk, L, create(), plan, L.tk_ragged_plan_destroy)  # touching, disjoint
    reqs[1].base = al + 5 * 512                                        # one poly shared
    assert create() == tk.EINVAL
    reqs[1].base = al + 20 * 512                                       # same base as request 2
    assert create() == tk.EINVAL
    reqs[1].npoly = 0                                                  # empty: no range
    _created_or_nodev(tk, L, create(), plan, L.tk_ragged_plan_destroy)


def test_fused_plan_rejects_cross_request_overlap(tk):
    L = tk.lib()
    buf = (ctypes.c_int16 * (N * 128 + 16))()
    al = (ctypes.addressof(buf) + 15) & ~15
    P = 512
    plan = ctypes.c_void_p()
    reqs = (tk.TkFusedRequest * 2)()
    # two rank-2 matvecs: A_hat 4 polys, s 2 polys, t 2 polys each
    for r, base in enumerate((0, 16)):
        reqs[r].mode, reqs[r].k, reqs[r].ninst = tk.FUSED_MATVEC, 2, 1
        reqs[r].mat, reqs[r].x, reqs[r].out0 = (al + (base + 0) * P, al + (base + 4) * P,
                                                al + (base + 6) * P)

    def create():
        return L.tk_fused_plan_create(ctypes.byref(plan), ctypes.addressof(reqs), 2)
    _created_or_nodev(tk, L, create(), plan, L.tk_fused_plan_destroy)
    reqs[1].mat = reqs[0].mat                              # a shared input is fine
    _created_or_nodev(tk, L, create(), plan, L.tk_fused_plan_destroy)
    reqs[1].x = reqs[0].out0                               # reads another request's output
    assert create() == tk.EINVAL
    reqs[1].x = al + 20 * P
    reqs[1].out0 = al + 7 * P                              # outputs overlap by one poly
    assert create() == tk.EINVAL
    reqs[1].out0 = al + 1 * P                              # writes another request's input
    assert create() == tk.EINVAL
    reqs[1].out0 = reqs[1].x                               # its own input, exactly: allowed
    _created_or_nodev(tk, L, create(), plan, L.tk_fused_plan_destroy)


def test_noalloc_validation(tk):
    """The no-allocation ragged transform rejects tables beyond its fixed
    buffer, and bad requests, before touching the device."""
    L = tk.lib()
    buf = (ctypes.c_int16 * (N * 64 + 16))()
    al = (ctypes.addressof(buf) + 15) & ~15
    one = (tk.TkRequest * 1)()
    one[0].base, one[0].npoly, one[0].k, one[0].kind = al, 64, 3, 0
    f = L.tk_ragged_transform_noalloc
    assert f(None, 1, 0, None) == tk.EINVAL
    assert f(ctypes.addressof(one), 0, 0, None) == tk.EINVAL
    assert f(ctypes.addressof(one), 1, 2, None) == tk.EINVAL             # direction
    one[0].k = 5
    assert f(ctypes.addressof(one), 1, 0, None) == tk.EINVAL             # rank
    one[0].k, one[0].base = 3, al + 2
    assert f(ctypes.addressof(one), 1, 0, None) == tk.EINVAL             # misaligned
    one[0].base, one[0].npoly = al, 16 * 131072 + 1
    assert f(ctypes.addressof(one), 1, 0, None) == tk.EINVAL             # 2^17 + 1 tiles
    one[0].npoly = 0
    assert f(ctypes.addressof(one), 1, 0, None) == tk.EINVAL             # nothing to do
    many = (tk.TkRequest * 8193)()
    for r in range(8193):
        many[r].base, many[r].npoly, many[r].k, many[r].kind = al, 1, 2, 0
    assert f(ctypes.addressof(many), 8193, 0, None) == tk.EINVAL         # 8193 requests
    one[0].npoly = 64
    assert f(ctypes.addressof(one), 1, 0, None) in (tk.ENODEV, tk.EINVAL, tk.OK)
