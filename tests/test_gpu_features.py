"""GPU parity of the ragged persistent path (hipGraph plan + autotuner) and of
the fused NTT -> basemul -> iNTT matrix-vector product, against the oracle
and the full-size digests of the compiled reference."""
import ctypes

import numpy as np
import pytest

from conftest import canon, centred

pytestmark = pytest.mark.gpu

N, Q = 256, 3329



def _dev(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int16)).to(device)


def _make_plan(tk, pool_dev, reqs, direction, method):
    table = (tk.TkRequest * len(reqs))()
    base = pool_dev.data_ptr()
    for r, (off, n, k, kind) in enumerate(reqs):
        table[r].base = base + int(off) * N * 2
        table[r].npoly, table[r].k, table[r].kind = int(n), int(k), int(kind)
    plan = ctypes.c_void_p()
    rc = tk.lib().tk_ragged_plan_create(ctypes.byref(plan), ctypes.addressof(table), len(reqs),
                                        direction, method, None)
    assert rc == 0, f"plan create failed: {rc}"
    return plan


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("total", [20000, 150000])
@pytest.mark.parametrize("method", ["auto", "mfma16", "mfma32"])
def test_ragged_plan_small(tk, gpu, oracle, method, total):
    """Forward and inverse plans against the oracle, gaps untouched; at
    150,000 polys the plans store their output with the default cache policy
    (kept in the Infinity Cache), at 20,000 nontemporally."""
    import torch
    reqs, pool = tk.wl_ragged_layout(total, 21)
    x = tk.wl_uniform(pool * N, 21).reshape(pool, N)
    d = _dev(x, gpu)
    plan = _make_plan(tk, d, reqs, 0, tk.METHODS[method])
    info = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()]
    assert tk.lib().tk_ragged_plan_info(plan, *[ctypes.byref(v) for v in info]) == 0
    assert info[3].value == total and info[1].value in (16, 32)
    print(f"ragged plan {method}: shape {info[1].value} oversub {info[2].value}", flush=True)
    assert tk.lib().tk_ragged_plan_execute(plan, _stream()) == 0
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    want = x.copy()
    for off, n, _k, _kind in reqs:
        want[off:off + n] = canon(oracle.ntt(x[off:off + n]))
    assert np.array_equal(got, want), "ragged forward mismatch (or gap modified)"
    # inverse plan on the forward output, launched without the graph
    iplan = _make_plan(tk, d, reqs, 1, tk.METHODS[method])
    assert tk.lib().tk_ragged_plan_launch(iplan, _stream()) == 0
    torch.cuda.synchronize()
    got2 = d.cpu().numpy()
    want2 = want.copy()
    for off, n, _k, _kind in reqs:
        want2[off:off + n] = centred(oracle.invntt(want[off:off + n]))
    assert np.array_equal(got2, want2)
    assert tk.lib().tk_ragged_plan_destroy(plan) == 0
    assert tk.lib().tk_ragged_plan_destroy(iplan) == 0


def test_ragged_graph_replays(tk, gpu, oracle, hip):
    """The plan's recorded graph, instantiated and replayed: each replay
    matches the eager result, with eager work and synchronisation between the
    replays; then the forward and inverse plans as child nodes of one graph
    of the caller's."""
    import torch
    reqs, pool = tk.wl_ragged_layout(5000, 22)
    x = tk.wl_uniform(pool * N, 22).reshape(pool, N)
    d = _dev(x, gpu)
    plan = _make_plan(tk, d, reqs, 0, tk.METHOD_MFMA16)
    iplan = _make_plan(tk, d, reqs, 1, tk.METHOD_MFMA32)
    want = x.copy()
    for off, n, _k, _kind in reqs:
        want[off:off + n] = canon(oracle.ntt(x[off:off + n]))
    back = want.copy()
    for off, n, _k, _kind in reqs:
        back[off:off + n] = centred(oracle.invntt(want[off:off + n]))
    g, gi = ctypes.c_void_p(), ctypes.c_void_p()
    assert tk.lib().tk_ragged_plan_graph(plan, ctypes.byref(g)) == 0
    assert tk.lib().tk_ragged_plan_graph(iplan, ctypes.byref(gi)) == 0
    ex = hip.instantiate(g.value)
    src = _dev(x, gpu)
    for rep in range(4):
        d.copy_(src)
        tk.ntt(torch.empty((64, N), dtype=torch.int16, device=gpu).zero_())  # eager work
        torch.cuda.synchronize()
        assert hip.lib.hipGraphLaunch(ex, _stream()) == 0
        torch.cuda.synchronize()
        assert np.array_equal(d.cpu().numpy(), want), f"replay {rep}"
    hip.lib.hipGraphExecDestroy(ex)
    both = hip.compose([g.value, gi.value])
    exb = hip.instantiate(both)
    for rep in range(3):
        d.copy_(src)
        torch.cuda.synchronize()
        assert hip.lib.hipGraphLaunch(exb, _stream()) == 0
        torch.cuda.synchronize()
        assert np.array_equal(d.cpu().numpy(), back), f"composed replay {rep}"
        assert tk.lib().tk_ragged_plan_execute(plan, _stream()) == 0  # eager, in between
        torch.cuda.synchronize()
    hip.lib.hipGraphExecDestroy(exb)
    hip.lib.hipGraphDestroy(both)
    tk.lib().tk_ragged_plan_destroy(plan)
    tk.lib().tk_ragged_plan_destroy(iplan)


@pytest.mark.parametrize("method", ["auto", "mfma16", "mfma32"])
def test_ragged_full_size_digest(tk, gpu, digests, method):
    """Configuration 4: 2^20 polys in ragged requests, forward then inverse,
    with the autotuner's choice and with each tile shape."""
    import torch
    from golden.make_golden import ragged_pool
    reqs, pool = ragged_pool()
    d = _dev(pool, gpu)
    dg = digests["config4"]
    print(f"config 4 digest, method {method}", flush=True)
    fplan = _make_plan(tk, d, reqs, 0, tk.METHODS[method])
    iplan = _make_plan(tk, d, reqs, 1, tk.METHODS[method])
    assert tk.lib().tk_ragged_plan_execute(fplan, _stream()) == 0
    torch.cuda.synchronize()
    assert tk.canon_digest(d.cpu().numpy()) == dg["ntt"]
    assert tk.lib().tk_ragged_plan_execute(iplan, _stream()) == 0
    torch.cuda.synchronize()
    assert tk.canon_digest(d.cpu().numpy()) == dg["invntt_of_canonical_ntt"]
    tk.lib().tk_ragged_plan_destroy(fplan)
    tk.lib().tk_ragged_plan_destroy(iplan)


@pytest.mark.parametrize("direction", [0, 1])
def test_autotune(tk, gpu, direction):
    m, w, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert tk.lib().tk_autotune(1 << 18, direction, ctypes.byref(m), ctypes.byref(w),
                                ctypes.byref(g), _stream()) == 0
    assert m.value in (tk.METHOD_MFMA16, tk.METHOD_MFMA32)
    assert w.value in (2, 4) and g.value in (4, 8, 12)
    print(f"autotune dir {direction} 2^18: method {m.value} waves/WG {w.value} "
          f"oversubscription {g.value}", flush=True)


@pytest.mark.parametrize("k,ninst", [(2, 333), (3, 1000), (4, 257), (3, 1), (3, 17)])
def test_matvec_vs_oracle(tk, gpu, oracle, k, ninst):
    A = tk.wl_uniform(ninst * k * k * N, 30 + k, 0, Q).reshape(ninst, k, k, N)
    s = tk.wl_cbd2(ninst * k * N, 30 + k).reshape(ninst, k, N)
    t = tk.matvec(_dev(A, gpu), _dev(s, gpu), k).cpu().numpy().reshape(ninst, k, N)
    assert np.array_equal(t, centred(oracle.matvec(A, s, k)))


def test_matvec_full_range_inputs(tk, gpu, oracle):
    """A_hat and s over all of int16: exact mod q (the oracle runs on canonical
    inputs, where its int16 arithmetic cannot wrap)."""
    k, ninst = 3, 200
    A = tk.wl_uniform(ninst * k * k * N, 40, -32768, 32768).reshape(ninst, k, k, N)
    s = tk.wl_uniform(ninst * k * N, 41, -32768, 32768).reshape(ninst, k, N)
    t = tk.matvec(_dev(A, gpu), _dev(s, gpu), k).cpu().numpy().reshape(ninst, k, N)
    assert np.array_equal(t, centred(oracle.matvec(canon(A), canon(s), k)))


def test_matvec_in_place_over_s(tk, gpu, oracle):
    k, ninst = 3, 100
    A = tk.wl_uniform(ninst * k * k * N, 42, 0, Q).reshape(ninst, k, k, N)
    s = tk.wl_cbd2(ninst * k * N, 43).reshape(ninst, k, N)
    ds = _dev(s, gpu)
    tk.matvec(_dev(A, gpu), ds, k, out=ds)
    assert np.array_equal(ds.cpu().numpy().reshape(ninst, k, N), centred(oracle.matvec(A, s, k)))


@pytest.mark.parametrize("k", [2, 3, 4])
def test_matvec_full_size_digest(tk, gpu, digests, k):
    """Configuration 5's product, 65,536 instances, at rank k (the workload of
    configuration 5 itself at k = 3)."""
    from golden.make_golden import matvec_inputs
    d = digests["config5"] if k == 3 else digests["config5_by_rank"][str(k)]
    A, s = matvec_inputs(k=k)
    assert (tk.canon_digest(A), tk.canon_digest(s)) == (d["A"], d["s"])
    t = tk.matvec(_dev(A, gpu), _dev(s, gpu), k).cpu().numpy()
    assert tk.canon_digest(t) == d["t"]


@pytest.mark.parametrize("cfg", [1, 2, 3])
@pytest.mark.parametrize("method", ["mfma16", "mfma32", "butterfly"])
def test_full_size_digests(tk, gpu, digests, cfg, method):
    """Configurations 1-3 at full size: forward, then inverse of the forward."""
    import torch
    x = tk.workload(cfg)
    d = _dev(x, gpu)
    tk.ntt(d, method=method)
    torch.cuda.synchronize()
    f = d.cpu().numpy()
    assert tk.canon_digest(f) == digests[f"config{cfg}"]["ntt"]
    assert f.min() >= 0 and f.max() < Q
    tk.invntt(d, method=method)
    assert tk.canon_digest(d.cpu().numpy()) == digests[f"config{cfg}"]["invntt_of_canonical_ntt"]
