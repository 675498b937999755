"""GPU parity of the fused K-PKE arithmetic cores (keygen, encrypt, decrypt)
against the oracle's pqcrystals composition: raw int16 equality at ragged
instance counts and every rank, the allowed in-place aliases, the message
round trip, and the full-size digests of the compiled reference."""
import numpy as np
import pytest

from conftest import centred

pytestmark = pytest.mark.gpu

N, Q = 256, 3329


def _dev(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int16)).to(device)


def _transpose(A):
    return np.ascontiguousarray(A.transpose(0, 2, 1, 3))


def _np(t, shape):
    return t.cpu().numpy().reshape(shape)


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("ninst", [1, 17, 300])
def test_kpke_cores_vs_oracle(tk, gpu, oracle, k, ninst):
    w = tk.kpke_workload(ninst, k, seed=400 + 7 * k + ninst)
    want_t, want_s = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    t_hat, s_hat = tk.kpke_keygen(_dev(w["A_hat"], gpu), _dev(w["s"], gpu), _dev(w["e"], gpu), k)
    assert np.array_equal(_np(s_hat, want_s.shape), want_s), f"s_hat k {k} ninst {ninst}"
    assert np.array_equal(_np(t_hat, want_t.shape), want_t), f"t_hat k {k} ninst {ninst}"

    At = _transpose(w["A_hat"])
    want_u, want_v = oracle.kpke_encrypt(At, want_t, w["r"], w["e1"], w["e2"], w["m"], k)
    u, v = tk.kpke_encrypt(_dev(At, gpu), t_hat, _dev(w["r"], gpu), _dev(w["e1"], gpu),
                           _dev(w["e2"], gpu), _dev(w["m"], gpu), k)
    assert np.array_equal(_np(u, want_u.shape), want_u), f"u k {k} ninst {ninst}"
    assert np.array_equal(_np(v, want_v.shape), want_v), f"v k {k} ninst {ninst}"
    # without a message poly
    want_u0, want_v0 = oracle.kpke_encrypt(At, want_t, w["r"], w["e1"], w["e2"], None, k)
    _, v0 = tk.kpke_encrypt(_dev(At, gpu), t_hat, _dev(w["r"], gpu), _dev(w["e1"], gpu),
                            _dev(w["e2"], gpu), None, k)
    assert np.array_equal(_np(v0, want_v0.shape), want_v0)

    want_w = oracle.kpke_decrypt(want_s, want_u, want_v, k)
    dec = tk.kpke_decrypt(s_hat, u, v, k)
    assert np.array_equal(_np(dec, want_w.shape), want_w), f"w k {k} ninst {ninst}"


def test_kpke_large_addends(tk, gpu, oracle):
    """e1, e2, m and v far outside ML-KEM's ranges but short of overflowing
    the reference's int16 adds: the sums stay exact."""
    k, ninst = 3, 40
    w = tk.kpke_workload(ninst, k, seed=450)
    e1 = tk.wl_uniform(ninst * k * N, 451, -29000, 29000).reshape(ninst, k, N)
    e2 = tk.wl_uniform(ninst * N, 452, -14000, 14000).reshape(ninst, N)
    m = tk.wl_uniform(ninst * N, 453, -14000, 14000).reshape(ninst, N)
    v_in = tk.wl_uniform(ninst * N, 454, -29000, 29000).reshape(ninst, N)
    t_hat, s_hat = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    At = _transpose(w["A_hat"])
    want_u, want_v = oracle.kpke_encrypt(At, t_hat, w["r"], e1, e2, m, k)
    u, v = tk.kpke_encrypt(_dev(At, gpu), _dev(t_hat, gpu), _dev(w["r"], gpu), _dev(e1, gpu),
                           _dev(e2, gpu), _dev(m, gpu), k)
    assert np.array_equal(_np(u, want_u.shape), want_u)
    assert np.array_equal(_np(v, want_v.shape), want_v)
    want_w = oracle.kpke_decrypt(s_hat, want_u, v_in, k)
    dec = tk.kpke_decrypt(_dev(s_hat, gpu), u, _dev(v_in, gpu), k)
    assert np.array_equal(_np(dec, want_w.shape), want_w)


def test_kpke_in_place_aliases(tk, gpu, oracle):
    """s_hat over s, t_hat over e, u over e1, v over e2, w over v."""
    k, ninst = 3, 100
    w = tk.kpke_workload(ninst, k, seed=460)
    want_t, want_s = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    ds, de = _dev(w["s"], gpu), _dev(w["e"], gpu)
    tk.kpke_keygen(_dev(w["A_hat"], gpu), ds, de, k, t_hat=de, s_hat=ds)
    assert np.array_equal(_np(ds, want_s.shape), want_s)
    assert np.array_equal(_np(de, want_t.shape), want_t)
    At = _transpose(w["A_hat"])
    want_u, want_v = oracle.kpke_encrypt(At, want_t, w["r"], w["e1"], w["e2"], w["m"], k)
    de1, de2 = _dev(w["e1"], gpu), _dev(w["e2"], gpu)
    tk.kpke_encrypt(_dev(At, gpu), de, _dev(w["r"], gpu), de1, de2, _dev(w["m"], gpu), k,
                    u=de1, v=de2)
    assert np.array_equal(_np(de1, want_u.shape), want_u)
    assert np.array_equal(_np(de2, want_v.shape), want_v)
    want_w = oracle.kpke_decrypt(want_s, want_u, want_v, k)
    tk.kpke_decrypt(ds, de1, de2, k, w=de2)
    assert np.array_equal(_np(de2, want_w.shape), want_w)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_kpke_message_round_trip(tk, gpu, k):
    w = tk.kpke_workload(1000, k, seed=470 + k)
    t_hat, s_hat = tk.kpke_keygen(_dev(w["A_hat"], gpu), _dev(w["s"], gpu), _dev(w["e"], gpu), k)
    u, v = tk.kpke_encrypt(_dev(_transpose(w["A_hat"]), gpu), t_hat, _dev(w["r"], gpu),
                           _dev(w["e1"], gpu), _dev(w["e2"], gpu), _dev(w["m"], gpu), k)
    dec = np.mod(tk.kpke_decrypt(s_hat, u, v, k).cpu().numpy().astype(np.int64), Q)
    assert np.array_equal(((dec * 2 + Q // 2) // Q) & 1, w["bits"])


@pytest.mark.parametrize("k", [2, 3, 4])
def test_kpke_full_size_digests(tk, gpu, digests, k):
    """65,536 instances of rank k: every output centred and equal, mod q, to
    the compiled reference's (so raw-equal)."""
    d = digests["kpke"] if k == 3 else digests["kpke_by_rank"][str(k)]
    assert d["k"] == k
    w = tk.kpke_workload(d["ninst"], k)
    t_hat, s_hat = tk.kpke_keygen(_dev(w["A_hat"], gpu), _dev(w["s"], gpu), _dev(w["e"], gpu), k)
    u, v = tk.kpke_encrypt(_dev(_transpose(w["A_hat"]), gpu), t_hat, _dev(w["r"], gpu),
                           _dev(w["e1"], gpu), _dev(w["e2"], gpu), _dev(w["m"], gpu), k)
    dec = tk.kpke_decrypt(s_hat, u, v, k)
    for name, t in (("t_hat", t_hat), ("s_hat", s_hat), ("u", u), ("v", v), ("w", dec)):
        x = t.cpu().numpy()
        assert np.array_equal(x, centred(x)), f"{name} not centred"
        assert tk.canon_digest(x) == d[name], name


@pytest.mark.parametrize("k,ninst", [(2, 1), (3, 17), (4, 33)])
def test_fused_outputs_stop_at_the_batch(tk, gpu, oracle, k, ninst):
    """Partial last tiles: every output right and nothing written past any of
    them (the kernels rely on the buffer range check for the lanes of
    instances past the batch)."""
    import torch
    G = 8  # guard polys after each output

    def guarded(npolys):
        buf = torch.full((npolys + G, N), 1234, dtype=torch.int16, device=gpu)
        return buf, buf[:npolys]

    def untouched(buf, npolys, name):
        assert (buf[npolys:] == 1234).all().item(), f"{name}: written past the batch"

    w = tk.kpke_workload(ninst, k, seed=470 + k + ninst)
    want_t, want_s = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    tb, t_hat = guarded(ninst * k)
    sb, s_hat = guarded(ninst * k)
    tk.kpke_keygen(_dev(w["A_hat"], gpu), _dev(w["s"], gpu), _dev(w["e"], gpu), k,
                   t_hat=t_hat, s_hat=s_hat)
    assert np.array_equal(_np(t_hat, want_t.shape), want_t)
    assert np.array_equal(_np(s_hat, want_s.shape), want_s)
    untouched(tb, ninst * k, "t_hat")
    untouched(sb, ninst * k, "s_hat")

    At = _transpose(w["A_hat"])
    want_u, want_v = oracle.kpke_encrypt(At, want_t, w["r"], w["e1"], w["e2"], w["m"], k)
    ub, u = guarded(ninst * k)
    vb, v = guarded(ninst)
    tk.kpke_encrypt(_dev(At, gpu), t_hat, _dev(w["r"], gpu), _dev(w["e1"], gpu),
                    _dev(w["e2"], gpu), _dev(w["m"], gpu), k, u=u, v=v)
    assert np.array_equal(_np(u, want_u.shape), want_u)
    assert np.array_equal(_np(v, want_v.shape), want_v)
    untouched(ub, ninst * k, "u")
    untouched(vb, ninst, "v")

    want_w = oracle.kpke_decrypt(want_s, want_u, want_v, k)
    wb, dec = guarded(ninst)
    tk.kpke_decrypt(s_hat, u, v, k, w=dec)
    assert np.array_equal(_np(dec, want_w.shape), want_w)
    untouched(wb, ninst, "w")

    # the matrix-vector product, int16 and packed secrets
    A = w["A_hat"]
    s = w["s"]
    want_mv = centred(oracle.matvec(A, s, k)).reshape(ninst * k, N)
    mb, mv = guarded(ninst * k)
    tk.matvec(_dev(A, gpu), _dev(s, gpu), k, out=mv)
    assert np.array_equal(mv.cpu().numpy(), want_mv)
    untouched(mb, ninst * k, "matvec t")
    pb, pmv = guarded(ninst * k)
    packed = torch.from_numpy(oracle.pack_small(s.reshape(ninst * k, N), 3)).to(gpu)
    tk.matvec_packed(_dev(A, gpu), packed, k, 3, out=pmv)
    assert np.array_equal(pmv.cpu().numpy(), want_mv)
    untouched(pb, ninst * k, "packed matvec t")
