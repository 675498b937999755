"""GPU parity of rank-aware fused plans (tk_fused_plan_*): a mixed batch of
matvec / keygen / encryption / decryption requests at ranks 2, 3 and 4, each
request's outputs against the oracle, through graph replays (with eager work
between them) and direct launches."""
import ctypes

import numpy as np
import pytest

from conftest import centred

pytestmark = pytest.mark.gpu

N = 256


def _dev(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int16)).to(device)


def _transpose(A):
    return np.ascontiguousarray(A.transpose(0, 2, 1, 3))


def _mixed_requests(tk, oracle, gpu, nreq, seed):
    """Requests with their device buffers and expected outputs."""
    import torch
    modes = tk.wl_uniform(nreq, seed, 0, 4)
    ranks = tk.wl_uniform(nreq, seed + 1, 2, 5)
    sizes = tk.wl_uniform(nreq, seed + 2, 1, 41)
    out = []
    for r in range(nreq):
        mode, k, n = int(modes[r]), int(ranks[r]), int(sizes[r])
        w = tk.kpke_workload(n, k, seed=seed + 10 * r)
        q = {"mode": mode, "k": k, "n": n, "keep": []}
        if mode == tk.FUSED_MATVEC:
            ins = {"mat": w["A_hat"], "x": w["s"]}
            want = {"out0": centred(oracle.matvec(w["A_hat"], w["s"], k))}
        elif mode == tk.FUSED_KEYGEN:
            ins = {"mat": w["A_hat"], "x": w["s"], "add0": w["e"]}
            t_hat, s_hat = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
            want = {"out0": t_hat, "out1": s_hat}
        else:
            t_hat, s_hat = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
            At = _transpose(w["A_hat"])
            m = w["m"] if r % 3 else None
            u, v = oracle.kpke_encrypt(At, t_hat, w["r"], w["e1"], w["e2"], m, k)
            if mode == tk.FUSED_ENCRYPT:
                ins = {"mat": At, "mat2": t_hat, "x": w["r"], "add0": w["e1"], "add1": w["e2"]}
                if m is not None:
                    ins["add2"] = m
                want = {"out0": u, "out1": v}
            else:
                ins = {"mat": s_hat, "x": u, "add0": v}
                want = {"out0": oracle.kpke_decrypt(s_hat, u, v, k)}
        q["dev_in"] = {name: _dev(a, gpu) for name, a in ins.items()}
        q["dev_out"] = {name: torch.full(a.shape, 12345, dtype=torch.int16, device=gpu)
                        for name, a in want.items()}
        q["want"] = want
        out.append(q)
    table = (tk.TkFusedRequest * nreq)()
    for i, q in enumerate(out):
        table[i].mode, table[i].k, table[i].ninst = q["mode"], q["k"], q["n"]
        for name, t in list(q["dev_in"].items()) + list(q["dev_out"].items()):
            setattr(table[i], name, t.data_ptr())
    return out, table


def _check(reqs, tag):
    for i, q in enumerate(reqs):
        for name, want in q["want"].items():
            got = q["dev_out"][name].cpu().numpy().reshape(want.shape)
            assert np.array_equal(got, want), f"{tag}: request {i} mode {q['mode']} " \
                                              f"k {q['k']} n {q['n']} {name}"


def _reset(reqs):
    for q in reqs:
        for t in q["dev_out"].values():
            t.fill_(12345)


def test_mixed_plan(tk, gpu, oracle):
    import torch
    reqs, table = _mixed_requests(tk, oracle, gpu, 36, 600)
    plan = ctypes.c_void_p()
    assert tk.lib().tk_fused_plan_create(ctypes.byref(plan), ctypes.addressof(table), 36) == 0
    ng, ni = ctypes.c_int(), ctypes.c_size_t()
    assert tk.lib().tk_fused_plan_info(plan, ctypes.byref(ng), ctypes.byref(ni)) == 0
    assert ni.value == sum(q["n"] for q in reqs)
    assert ng.value == len({(q["mode"], q["k"]) for q in reqs})
    print(f"fused plan: {ng.value} groups, {ni.value} instances", flush=True)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        _reset(reqs)
        tk.ntt(torch.zeros((64, N), dtype=torch.int16, device=gpu))  # eager work between
        torch.cuda.synchronize()
        assert tk.lib().tk_fused_plan_execute(plan, s) == 0
        torch.cuda.synchronize()
        _check(reqs, f"graph replay {rep}")
    _reset(reqs)
    assert tk.lib().tk_fused_plan_launch(plan, s) == 0
    torch.cuda.synchronize()
    _check(reqs, "direct launch")
    assert tk.lib().tk_fused_plan_destroy(plan) == 0


def test_single_group_plan_matches_single_call(tk, gpu, oracle):
    """One big encryption request: the table-driven kernel against the
    contiguous one."""
    import torch
    k, n = 3, 2000
    w = tk.kpke_workload(n, k, seed=650)
    t_hat, _ = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    args = [_dev(a, gpu) for a in (_transpose(w["A_hat"]), t_hat, w["r"], w["e1"], w["e2"],
                                   w["m"])]
    u1, v1 = tk.kpke_encrypt(*args, k)
    u2, v2 = torch.empty_like(u1), torch.empty_like(v1)
    table = (tk.TkFusedRequest * 1)()
    t = table[0]
    t.mode, t.k, t.ninst = tk.FUSED_ENCRYPT, k, n
    t.out0, t.out1 = u2.data_ptr(), v2.data_ptr()
    t.mat, t.mat2, t.x, t.add0, t.add1, t.add2 = (a.data_ptr() for a in args)
    plan = ctypes.c_void_p()
    assert tk.lib().tk_fused_plan_create(ctypes.byref(plan), ctypes.addressof(table), 1) == 0
    assert tk.lib().tk_fused_plan_execute(plan, None) == 0
    torch.cuda.synchronize()
    assert torch.equal(u1, u2) and torch.equal(v1, v2)
    tk.lib().tk_fused_plan_destroy(plan)


@pytest.mark.parametrize("mode,k,n", [("encrypt", 4, 100000), ("decrypt", 2, 240000)])
def test_plan_groups_larger_than_the_grid(tk, gpu, mode, k, n):
    """Groups of more tiles than the table-driven kernels' grid (4x their
    resident workgroups), so that each workgroup loops over several tiles
    with the pointer rows double-buffered by tile parity: the plan's outputs
    against the contiguous kernels' (themselves pinned to the oracle), bit
    for bit."""
    import torch
    g = torch.Generator(device=gpu)
    g.manual_seed(77 + k)

    def rnd(rows, lo=0, hi=3329):
        return torch.randint(lo, hi, (rows, 256), dtype=torch.int16, device=gpu, generator=g)

    print(f"plan group larger than the grid: {mode}, k = {k}, {n} instances", flush=True)
    table = (tk.TkFusedRequest * 1)()
    t = table[0]
    t.k, t.ninst = k, n
    if mode == "encrypt":
        At, th, r = rnd(n * k * k), rnd(n * k), rnd(n * k, -3, 4)
        e1, e2, m = rnd(n * k, -3, 4), rnd(n, -3, 4), rnd(n, 0, 2) * 1665
        u1, v1 = tk.kpke_encrypt(At, th, r, e1, e2, m, k)
        u2, v2 = torch.empty_like(u1), torch.empty_like(v1)
        t.mode = tk.FUSED_ENCRYPT
        t.out0, t.out1 = u2.data_ptr(), v2.data_ptr()
        t.mat, t.mat2, t.x = At.data_ptr(), th.data_ptr(), r.data_ptr()
        t.add0, t.add1, t.add2 = e1.data_ptr(), e2.data_ptr(), m.data_ptr()
    else:
        sh, u, v = rnd(n * k, -32768, 32768), rnd(n * k), rnd(n)
        w1 = tk.kpke_decrypt(sh, u, v, k)
        w2 = torch.empty_like(w1)
        t.mode = tk.FUSED_DECRYPT
        t.out0, t.mat, t.x, t.add0 = w2.data_ptr(), sh.data_ptr(), u.data_ptr(), v.data_ptr()
    plan = ctypes.c_void_p()
    assert tk.lib().tk_fused_plan_create(ctypes.byref(plan), ctypes.addressof(table), 1) == 0
    assert tk.lib().tk_fused_plan_execute(plan, None) == 0
    torch.cuda.synchronize()
    if mode == "encrypt":
        assert torch.equal(u1, u2) and torch.equal(v1, v2)
    else:
        assert torch.equal(w1, w2)
    tk.lib().tk_fused_plan_destroy(plan)
