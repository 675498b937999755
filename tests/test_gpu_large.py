"""Largest sizes and concurrent callers.

* 2^23 polys (4 GiB per buffer): every method's forward and inverse, with
  rows sampled across the whole range (both ends, around 2^22, random)
  checked against the oracle, and the round trip invntt(ntt(x)) == 2^16 x
  (mod q) checked on every coefficient, chunk by chunk, on the device.
* basemul over 2^23 pairs: every coefficient through the Montgomery identity,
  sampled rows against the oracle.
* Matvec and K-PKE batches (ranks 2-4) larger than the fused kernels' grids:
  sampled instances against the oracle.
* Several host threads calling the library at once, each on its own stream
  and buffers, through the transforms, the fused product and a ragged plan.
"""
import ctypes
import threading

import numpy as np
import pytest

from conftest import canon, centred

pytestmark = pytest.mark.gpu

N, Q = 256, 3329


@pytest.mark.parametrize("method", ["mfma16", "mfma32", "butterfly"])
def test_four_gib_batch(tk, gpu, oracle, method):
    import torch
    npoly = 1 << 23
    g = torch.Generator(device=gpu)
    g.manual_seed(23)
    x = torch.randint(0, Q, (npoly, N), dtype=torch.int16, device=gpu, generator=g)
    print(f"4 GiB batch, {method}", flush=True)
    y = torch.empty_like(x)
    tk.ntt(x, out=y, method=method)
    rng = np.random.default_rng(5)
    rows = np.unique(np.concatenate([
        [0, 1, 15, 16, (1 << 22) - 1, 1 << 22, (1 << 22) + 1, npoly - 17, npoly - 1],
        rng.integers(0, npoly, 2000)]))
    ridx = torch.from_numpy(rows).to(gpu)
    xs = x.index_select(0, ridx).cpu().numpy()
    ys = y.index_select(0, ridx).cpu().numpy()
    assert np.array_equal(ys, canon(oracle.ntt(xs))), "forward rows"
    tk.invntt(y, method=method)
    zs = y.index_select(0, ridx).cpu().numpy()
    assert np.array_equal(zs, centred(oracle.invntt(ys))), "inverse rows"
    # every coefficient: invntt(ntt(x)) == 2^16 x (mod q)
    step = 1 << 20
    for c in range(0, npoly, step):
        want = (x[c:c + step].to(torch.int32) * 65536) % Q
        got = y[c:c + step].to(torch.int32) % Q
        assert torch.equal(got, want), f"round trip, polys {c}.."
    del x, y
    torch.cuda.empty_cache()


def test_four_gib_basemul(tk, gpu, oracle):
    """basemul_batch over 2^23 poly pairs (a grid of 2^20 workgroups): every
    coefficient through the Montgomery identity element (b0 = 2^16 mod q,
    b1 = 0 on every pair gives basemul(a, b) == a mod q), and sampled rows
    with random b against the oracle."""
    import torch
    npoly = 1 << 23
    g = torch.Generator(device=gpu)
    g.manual_seed(29)
    a = torch.randint(-32768, 32768, (npoly, N), dtype=torch.int16, device=gpu, generator=g)
    one = torch.zeros((npoly, N), dtype=torch.int16, device=gpu)
    one[:, 0::2] = 65536 % Q
    print("4 GiB basemul, identity", flush=True)
    r = tk.basemul(a, one, form="canonical")
    step = 1 << 20
    for c in range(0, npoly, step):
        assert torch.equal(r[c:c + step].to(torch.int32), a[c:c + step].to(torch.int32) % Q), \
            f"identity, polys {c}.."
    del one
    print("4 GiB basemul, random rows", flush=True)
    b = torch.randint(0, Q, (npoly, N), dtype=torch.int16, device=gpu, generator=g)
    tk.basemul(a, b, out=r)
    rng = np.random.default_rng(7)
    rows = np.unique(np.concatenate([[0, 1, (1 << 22) - 1, 1 << 22, npoly - 2, npoly - 1],
                                     rng.integers(0, npoly, 2000)]))
    ridx = torch.from_numpy(rows).to(gpu)
    want = oracle.basemul(a.index_select(0, ridx).cpu().numpy(), b.index_select(0, ridx).cpu().numpy())
    assert np.array_equal(r.index_select(0, ridx).cpu().numpy(), want)
    del a, b, r
    torch.cuda.empty_cache()


def test_concurrent_host_threads(tk, gpu, oracle):
    import torch
    nthreads, reps = 4, 6
    errors = []
    barrier = threading.Barrier(nthreads)

    def worker(tid):
        try:
            stream = torch.cuda.Stream(device=gpu)
            k, ninst, npoly = 2 + tid % 3, 97 + tid, 4096 + 17 * tid
            x = tk.wl_uniform(npoly * N, 700 + tid, 0, Q).reshape(npoly, N)
            A = tk.wl_uniform(ninst * k * k * N, 710 + tid, 0, Q).reshape(ninst, k, k, N)
            s = tk.wl_cbd2(ninst * k * N, 720 + tid).reshape(ninst, k, N)
            want_f = canon(oracle.ntt(x))
            want_t = centred(oracle.matvec(A, s, k)).reshape(ninst * k, N)
            reqs, pool = tk.wl_ragged_layout(3000, 730 + tid)
            rp = tk.wl_uniform(pool * N, 740 + tid, 0, Q).reshape(pool, N)
            want_r = rp.copy()
            for off, n, _k, _kind in reqs:
                want_r[off:off + n] = canon(oracle.ntt(rp[off:off + n]))
            with torch.cuda.stream(stream):
                dx = torch.from_numpy(x).to(gpu, non_blocking=False)
                dA = torch.from_numpy(A).to(gpu)
                ds = torch.from_numpy(s).to(gpu)
                drp = torch.from_numpy(rp).to(gpu)
                src = drp.clone()
                table = (tk.TkRequest * len(reqs))()
                for r, (off, n, kk, kind) in enumerate(reqs):
                    table[r].base = drp.data_ptr() + int(off) * N * 2
                    table[r].npoly, table[r].k, table[r].kind = int(n), int(kk), int(kind)
                plan = ctypes.c_void_p()
                sp = ctypes.c_void_p(stream.cuda_stream)
                rc = tk.lib().tk_ragged_plan_create(ctypes.byref(plan), ctypes.addressof(table),
                                                    len(reqs), 0, tk.METHOD_MFMA16, sp)
                assert rc == 0, \
                    f"thread {tid}: plan create {rc} (HIP error {tk.lib().tk_last_hip_error()})"
                barrier.wait()
                for rep in range(reps):
                    out = torch.empty_like(dx)
                    tk.ntt(dx, out=out, stream=stream)
                    t = tk.matvec(dA, ds, k, stream=stream)
                    drp.copy_(src)
                    assert tk.lib().tk_ragged_plan_execute(plan, sp) == 0
                    stream.synchronize()
                    assert np.array_equal(out.cpu().numpy(), want_f), f"thread {tid} rep {rep} ntt"
                    assert np.array_equal(t.cpu().numpy().reshape(ninst * k, N), want_t), \
                        f"thread {tid} rep {rep} matvec"
                    assert np.array_equal(drp.cpu().numpy(), want_r), f"thread {tid} rep {rep} plan"
                tk.lib().tk_ragged_plan_destroy(plan)
        except Exception as e:  # noqa: BLE001  (reported below with the thread id)
            errors.append(f"thread {tid}: {e!r}")
            try:
                barrier.abort()
            except threading.BrokenBarrierError:
                pass

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(nthreads)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors


@pytest.mark.parametrize("mode,k,n", [("matvec", 3, 120000), ("matvec_packed3", 3, 120000),
                                      ("matvec_packed2", 2, 120000), ("matvec_packed4", 4, 120000),
                                      ("keygen", 2, 120000), ("keygen", 3, 120000),
                                      ("keygen", 4, 120000), ("encrypt", 2, 120000),
                                      ("encrypt", 3, 100000), ("encrypt", 4, 120000),
                                      ("decrypt", 2, 300000), ("decrypt", 3, 300000),
                                      ("decrypt", 4, 300000)])
def test_fused_batches_larger_than_the_grid(tk, gpu, oracle, mode, k, n):
    """Batches of more tiles than the fused kernels' grids (4x their resident
    workgroups), so that each workgroup (each wave, for decryption) loops
    over several tiles: sampled instances across the batch against the
    oracle (raw int16 for the K-PKE cores, which end in poly_reduce; centred
    for the matvec)."""
    import torch
    g = torch.Generator(device=gpu)
    g.manual_seed(91)

    def rnd(*shape, lo=0, hi=Q):
        return torch.randint(lo, hi, shape, dtype=torch.int16, device=gpu, generator=g)

    rng = np.random.default_rng(13)
    rows = np.unique(np.concatenate([[0, 1, 15, 16, n // 2, n - 17, n - 1],
                                     rng.integers(0, n, 41)]))
    ridx = torch.from_numpy(rows).to(gpu)

    def pick(t, *shape):
        return t.reshape(n, *shape).index_select(0, ridx).cpu().numpy()

    print(f"fused batch larger than the grid: {mode}, k = {k}, {n} instances", flush=True)
    if mode.startswith("matvec"):
        # packed s must fit its width: [-2, 1] at 2 bits, CBD [-2, 2] otherwise
        A, s = rnd(n, k, k, N), rnd(n, k, N, lo=-2, hi=2 if mode.endswith("2") else 3)
        if mode == "matvec":
            got = pick(tk.matvec(A, s, k), k, N)
        else:
            bits = int(mode[-1])
            got = pick(tk.matvec_packed(A, tk.pack_small(s, bits), k, bits), k, N)
        want = oracle.matvec(pick(A, k, k, N), pick(s, k, N), k).reshape(got.shape)
        assert np.array_equal(got, centred(want))  # the oracle returns invntt's (-q, q) values
    elif mode == "keygen":
        A, s, e = rnd(n, k, k, N), rnd(n, k, N, lo=-2, hi=3), rnd(n, k, N, lo=-2, hi=3)
        t_hat, s_hat = tk.kpke_keygen(A, s, e, k)
        want_t, want_s = oracle.kpke_keygen(pick(A, k, k, N), pick(s, k, N), pick(e, k, N), k)
        assert np.array_equal(pick(t_hat, k, N), want_t.reshape(-1, k, N))
        assert np.array_equal(pick(s_hat, k, N), want_s.reshape(-1, k, N))
    elif mode == "encrypt":
        At, th, r = rnd(n, k, k, N), rnd(n, k, N), rnd(n, k, N, lo=-2, hi=3)
        e1, e2, m = rnd(n, k, N, lo=-2, hi=3), rnd(n, N, lo=-2, hi=3), rnd(n, N, lo=0, hi=2) * 1665
        u, v = tk.kpke_encrypt(At, th, r, e1, e2, m, k)
        want_u, want_v = oracle.kpke_encrypt(pick(At, k, k, N), pick(th, k, N), pick(r, k, N),
                                             pick(e1, k, N), pick(e2, N), pick(m, N), k)
        assert np.array_equal(pick(u, k, N), want_u.reshape(-1, k, N))
        assert np.array_equal(pick(v, N), want_v.reshape(-1, N))
    else:
        sh, u, v = rnd(n, k, N, lo=-32768, hi=32768), rnd(n, k, N), rnd(n, N)
        w = tk.kpke_decrypt(sh, u, v, k)
        want = oracle.kpke_decrypt(pick(sh, k, N), pick(u, k, N), pick(v, N), k)
        assert np.array_equal(pick(w, N), want.reshape(-1, N))
    torch.cuda.empty_cache()
