"""The no-allocation ragged transform (tk_ragged_transform_noalloc) against
the entry point it stands in for (tk_ragged_plan_create +
tk_ragged_plan_execute), through the C ABI: bit-exact on every coefficient,
gaps between requests untouched, on two calls in a row of which the second
is smaller (the device buffer keeps the first call's larger table), in both
directions; and the same outputs as the oracle on the requests."""
import ctypes

import numpy as np
import pytest

from conftest import canon, centred

pytestmark = pytest.mark.gpu

N = 256


def _table(tk, d, reqs):
    t = (tk.TkRequest * len(reqs))()
    for r, (off, n, k, kind) in enumerate(reqs):
        t[r].base = d.data_ptr() + int(off) * N * 2
        t[r].npoly, t[r].k, t[r].kind = int(n), int(k), int(kind)
    return t


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_noalloc_matches_plan_on_two_calls(tk, gpu, oracle):
    import torch
    L = tk.lib()
    for npoly, seed in ((60000, 61), (3000, 62)):
        reqs, pool = tk.wl_ragged_layout(npoly, seed)
        x = tk.wl_uniform(pool * N, seed, -32768, 32768).reshape(pool, N)
        d_plan = torch.from_numpy(x).to(gpu)
        d_free = torch.from_numpy(x).to(gpu)
        t_plan, t_free = _table(tk, d_plan, reqs), _table(tk, d_free, reqs)
        for direction in (0, 1):
            print(f"no-alloc {npoly} polys ({len(reqs)} requests), direction {direction}",
                  flush=True)
            plan = ctypes.c_void_p()
            assert L.tk_ragged_plan_create(ctypes.byref(plan), ctypes.addressof(t_plan),
                                           len(reqs), direction, tk.METHOD_AUTO, None) == 0
            before = d_free.cpu().numpy()
            assert L.tk_ragged_plan_execute(plan, _stream()) == 0
            assert L.tk_ragged_transform_noalloc(ctypes.addressof(t_free), len(reqs), direction,
                                                 _stream()) == 0
            torch.cuda.synchronize()
            got, want = d_free.cpu().numpy(), d_plan.cpu().numpy()
            assert np.array_equal(got, want), f"differs from the plan: {npoly} dir {direction}"
            # against the oracle on every request, and the gaps unchanged
            mask = np.zeros(pool, bool)
            for off, n, _k, _kind in reqs:
                seg = before[off:off + n]
                ref = (canon(oracle.ntt(canon(seg))) if direction == 0
                       else centred(oracle.invntt(canon(seg))))
                assert np.array_equal(got[off:off + n], ref), f"request at {off}"
                mask[off:off + n] = True
            assert np.array_equal(got[~mask], x[~mask]), "a gap was written"
            assert L.tk_ragged_plan_destroy(plan) == 0


def test_noalloc_rejects_tables_beyond_its_buffer(tk, gpu):
    L = tk.lib()
    import torch
    d = torch.zeros((32, N), dtype=torch.int16, device=gpu)
    one = (tk.TkRequest * 1)()
    one[0].base, one[0].npoly, one[0].k, one[0].kind = d.data_ptr(), 32, 2, 0
    assert L.tk_ragged_transform_noalloc(ctypes.addressof(one), 1, 0, None) == 0
    # too many requests
    many = (tk.TkRequest * 8193)()
    for r in range(8193):
        many[r].base, many[r].npoly, many[r].k, many[r].kind = d.data_ptr(), 0, 2, 0
    many[0].npoly = 1
    assert L.tk_ragged_transform_noalloc(ctypes.addressof(many), 8193, 0, None) == tk.EINVAL
    # too many tiles: 2^17 + 1 tiles of 16 polys (the table is checked before any access)
    big = (tk.TkRequest * 1)()
    big[0].base, big[0].npoly, big[0].k, big[0].kind = d.data_ptr(), 16 * 131072 + 1, 2, 0
    assert L.tk_ragged_transform_noalloc(ctypes.addressof(big), 1, 0, None) == tk.EINVAL
    torch.cuda.synchronize()
