"""Pins the oracle's K-PKE arithmetic cores (keygen, encrypt, decrypt as the
pqcrystals-kyber reference composes them in ref/indcpa.c) before they check
the fused device kernels:

* raw int16 equality with the same composition built from the compiled
  reference's ntt, invntt, basemul and reductions (oracle/_ref);
* the module identities they stand for, by schoolbook negacyclic products:
  t_hat = NTT(A s + e), u = A^T r + e1, v = t^T r + e2 + m, w = v - s^T u;
* an end-to-end message round trip: keygen, encrypt a 256-bit message,
  decrypt, decode.
"""
import numpy as np
import pytest

from conftest import canon, centred, negacyclic

N, Q = 256, 3329
R2INV = 169  # 2^-16 mod q


def _inputs(tk, k, ninst, seed):
    A = tk.wl_uniform(ninst * k * k * N, seed, 0, Q).reshape(ninst, k, k, N)
    s = tk.wl_cbd2(ninst * k * N, seed + 1).reshape(ninst, k, N)
    e = tk.wl_cbd2(ninst * k * N, seed + 2).reshape(ninst, k, N)
    r = tk.wl_cbd2(ninst * k * N, seed + 3).reshape(ninst, k, N)
    e1 = tk.wl_cbd2(ninst * k * N, seed + 4).reshape(ninst, k, N)
    e2 = tk.wl_cbd2(ninst * N, seed + 5).reshape(ninst, N)
    bits = tk.wl_uniform(ninst * N, seed + 6, 0, 2).reshape(ninst, N)
    m = (bits * ((Q + 1) // 2)).astype(np.int16)  # poly_frommsg: bit -> (q+1)/2
    return A, s, e, r, e1, e2, m, bits


def _transpose(A):
    return np.ascontiguousarray(A.transpose(0, 2, 1, 3))


@pytest.mark.parametrize("k", [2, 3, 4])
def test_kpke_matches_compiled_reference(oracle, reference, tk, k):
    A, s, e, r, e1, e2, m, _ = _inputs(tk, k, 200, 300 + k)
    t_hat, s_hat = oracle.kpke_keygen(A, s, e, k)
    want = reference.kpke_keygen(A, s, e, k)
    assert np.array_equal(t_hat, want[0]) and np.array_equal(s_hat, want[1])
    At = _transpose(A)
    u, v = oracle.kpke_encrypt(At, t_hat, r, e1, e2, m, k)
    wu, wv = reference.kpke_encrypt(At, t_hat, r, e1, e2, m, k)
    assert np.array_equal(u, wu) and np.array_equal(v, wv)
    u0, v0 = oracle.kpke_encrypt(At, t_hat, r, e1, e2, None, k)
    assert np.array_equal(canon(v0.astype(np.int32) + m), canon(v))
    w = oracle.kpke_decrypt(s_hat, u, v, k)
    assert np.array_equal(w, reference.kpke_decrypt(s_hat, u, v, k))
    # every output is a poly_reduce output: the centred representative
    for x in (t_hat, s_hat, u, v, w):
        assert np.array_equal(x, centred(x))


def test_kpke_module_identities(oracle, tk):
    k, ninst = 3, 2
    a = tk.wl_uniform(ninst * k * k * N, 320, -3328, 3329).reshape(ninst, k, k, N)
    A_hat = oracle.ntt(a.reshape(-1, N)).reshape(ninst, k, k, N)
    _, s, e, r, e1, e2, m, _ = _inputs(tk, k, ninst, 321)
    t_hat, s_hat = oracle.kpke_keygen(A_hat, s, e, k)
    assert np.array_equal(canon(s_hat), canon(oracle.ntt(s.reshape(-1, N))).reshape(ninst, k, N))
    # t = NTT^-1(t_hat) = A s + e  (invntt carries 2^16)
    t = canon(oracle.invntt(t_hat.reshape(-1, N)).astype(np.int64) * R2INV).reshape(ninst, k, N)
    for inst in range(ninst):
        for i in range(k):
            want = (sum(negacyclic(a[inst, i, j], s[inst, j]) for j in range(k)) + e[inst, i]) % Q
            assert np.array_equal(t[inst, i], want), f"keygen inst {inst} row {i}"
    At_hat = _transpose(A_hat)
    u, v = oracle.kpke_encrypt(At_hat, t_hat, r, e1, e2, m, k)
    for inst in range(ninst):
        for i in range(k):
            want = (sum(negacyclic(a[inst, j, i], r[inst, j]) for j in range(k))
                    + e1[inst, i]) % Q
            assert np.array_equal(canon(u[inst, i]), want), f"encrypt u inst {inst} row {i}"
        want = (sum(negacyclic(t[inst, j], r[inst, j]) for j in range(k))
                + e2[inst] + m[inst]) % Q
        assert np.array_equal(canon(v[inst]), want), f"encrypt v inst {inst}"
    sv = s.astype(np.int64)
    w = oracle.kpke_decrypt(s_hat, u, v, k)
    for inst in range(ninst):
        want = (v[inst] - sum(negacyclic(sv[inst, j], u[inst, j]) for j in range(k))) % Q
        assert np.array_equal(canon(w[inst]), want), f"decrypt inst {inst}"


@pytest.mark.parametrize("k", [2, 3, 4])
def test_kpke_message_round_trip(oracle, tk, k):
    """Without ciphertext compression the noise e^T r - s^T e1 + e2 stays far
    below q/4, so every message bit decodes (poly_tomsg rounding)."""
    A, s, e, r, e1, e2, m, bits = _inputs(tk, k, 64, 340 + k)
    t_hat, s_hat = oracle.kpke_keygen(A, s, e, k)
    u, v = oracle.kpke_encrypt(_transpose(A), t_hat, r, e1, e2, m, k)
    w = canon(oracle.kpke_decrypt(s_hat, u, v, k)).astype(np.int64)
    decoded = ((w * 2 + Q // 2) // Q) & 1
    assert np.array_equal(decoded, bits)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_kpke_full_size_digests(oracle, tk, digests, k):
    """65,536 instances of rank k against the compiled reference's digests."""
    d = digests["kpke"] if k == 3 else digests["kpke_by_rank"][str(k)]
    assert d["k"] == k
    w = tk.kpke_workload(d["ninst"], k)
    assert tk.canon_digest(w["A_hat"]) == d["A_hat"]
    t_hat, s_hat = oracle.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    assert (tk.canon_digest(t_hat), tk.canon_digest(s_hat)) == (d["t_hat"], d["s_hat"])
    u, v = oracle.kpke_encrypt(_transpose(w["A_hat"]), t_hat, w["r"], w["e1"], w["e2"], w["m"], k)
    assert (tk.canon_digest(u), tk.canon_digest(v)) == (d["u"], d["v"])
    assert tk.canon_digest(oracle.kpke_decrypt(s_hat, u, v, k)) == d["w"]
    assert d["message_bits_recovered"]
