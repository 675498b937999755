#!/usr/bin/env python3
"""Writes the golden vectors of the NTT path from the compiled reference.

Runs in the development container only, where the reference checkout exists
and ``make -C oracle`` has compiled its ``ntt_bench/ref/ntt.c`` and
``ntt_bench/ref/reduce.c`` into ``oracle/_ref/libkyber_ref.so``.  The
reference ships no vectors of its own (its only check compares 64 polys drawn
with glibc rand(), ntt_bench.cu:941-961), so these are produced here:

* ``fixtures.npz``: small inputs and the reference's raw int16 outputs of
  ntt(), invntt(), invntt(ntt()), whole-poly basemul and the k = 2, 3, 4
  module matrix-vector product (composed from the reference's functions).
* ``digests.json``: FNV-1a digests (tk_wl_canon_digest) of the canonical
  outputs of the full-size benchmark configurations, inputs from the
  project's SplitMix64 streams (include/tkyber_workload.h).

Usage: python tests/golden/make_golden.py [--digests-only | --ranks-only]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import tensorkyber_amd as tk  # noqa: E402  (host-only workload generators)
from oracle.oracle import Reference, canon  # noqa: E402

N, Q = 256, 3329
NFIX = 32


def fixtures(ref: Reference) -> dict:
    cats = {
        "uniform": tk.wl_uniform(NFIX * N, 101).reshape(NFIX, N),
        "full": tk.wl_uniform(NFIX * N, 102, -32768, 32768).reshape(NFIX, N),
        "edge": tk.wl_edge_mix(NFIX, 103),
        "small": tk.wl_uniform(NFIX * N, 104, -16383, 16384).reshape(NFIX, N),
    }
    cats["full"][0, :] = 32767
    cats["full"][1, :] = -32768
    out = {}
    for name, x in cats.items():
        out[f"in_{name}"] = x
        out[f"ntt_{name}"] = ref.ntt(x)
        out[f"invntt_{name}"] = ref.invntt(x)
        out[f"rt_{name}"] = ref.roundtrip(x)
    a = tk.wl_uniform(NFIX * N, 105, -32768, 32768).reshape(NFIX, N)
    b = tk.wl_uniform(NFIX * N, 106, -32768, 32768).reshape(NFIX, N)
    out["bm_a"], out["bm_b"], out["bm_r"] = a, b, ref.basemul(a, b)
    for k in (2, 3, 4):
        A = tk.wl_uniform(4 * k * k * N, 107 + k, 0, Q).reshape(4, k, k, N)
        s = tk.wl_cbd2(4 * k * N, 107 + k).reshape(4, k, N)
        out[f"mv{k}_A"], out[f"mv{k}_s"], out[f"mv{k}_t"] = A, s, ref.matvec(A, s, k)
    return out


def config_inputs(cfg: int) -> np.ndarray:
    return tk.workload(cfg)


def ragged_pool():
    reqs, pool = tk.wl_ragged_layout(1 << 20, 4)
    data = tk.wl_uniform(pool * N, 4).reshape(pool, N)
    return reqs, data


def matvec_inputs(ninst: int = 65536, k: int = 3):
    return tk.matvec_workload(ninst, k)


def packed_inputs(bits: int, npoly: int = 1 << 20) -> np.ndarray:
    return tk.packed_workload(bits, npoly)


def packed_digests(ref: Reference) -> dict:
    from oracle.oracle import Oracle
    o = Oracle()
    d = {}
    for bits in (2, 3, 4):
        x = packed_inputs(bits)
        b = o.pack_small(x, bits)
        assert np.array_equal(o.unpack_small(b, bits), x)
        d[f"bits{bits}"] = {"npoly": int(x.shape[0]), "input": tk.canon_digest(x),
                            "packed_sha256": hashlib.sha256(b.tobytes()).hexdigest(),
                            "ntt": tk.canon_digest(ref.ntt(x))}
        print(f"packed {bits}: {d[f'bits{bits}']}", flush=True)
    return d


def kpke_digests(ref: Reference, ninst: int = 65536, k: int = 3) -> dict:
    """K-PKE cores on kpke_workload (seed 9): keygen, encryption with the
    transposed matrix, decryption of that ciphertext."""
    w = tk.kpke_workload(ninst, k)
    t_hat, s_hat = ref.kpke_keygen(w["A_hat"], w["s"], w["e"], k)
    At = np.ascontiguousarray(w["A_hat"].transpose(0, 2, 1, 3))
    u, v = ref.kpke_encrypt(At, t_hat, w["r"], w["e1"], w["e2"], w["m"], k)
    dec = ref.kpke_decrypt(s_hat, u, v, k)
    bits = ((np.mod(dec.astype(np.int64), Q) * 2 + Q // 2) // Q) & 1
    d = {"ninst": ninst, "k": k, "A_hat": tk.canon_digest(w["A_hat"]),
         "t_hat": tk.canon_digest(t_hat), "s_hat": tk.canon_digest(s_hat),
         "u": tk.canon_digest(u), "v": tk.canon_digest(v), "w": tk.canon_digest(dec),
         "message_bits_recovered": bool(np.array_equal(bits, w["bits"]))}
    print(f"kpke: {d}", flush=True)
    return d


def digests(ref: Reference) -> dict:
    d = {}
    for cfg in (1, 2, 3):
        x = config_inputs(cfg)
        f = ref.ntt(x)
        inv = ref.invntt(canon(f))
        rt = ref.invntt(f)
        d[f"config{cfg}"] = {
            "npoly": int(x.shape[0]),
            "input": tk.canon_digest(x),
            "ntt": tk.canon_digest(f),
            "invntt_of_canonical_ntt": tk.canon_digest(inv),
            "invntt_of_raw_ntt": tk.canon_digest(rt),
            "raw_and_canonical_chains_agree": bool(np.array_equal(canon(inv), canon(rt))),
        }
        print(f"config {cfg}: {d[f'config{cfg}']}", flush=True)
    reqs, pool = ragged_pool()
    fwd = pool.copy()
    for off, n, _k, _kind in reqs:
        fwd[off:off + n] = ref.ntt(pool[off:off + n])
    inv = fwd.copy()
    for off, n, _k, _kind in reqs:
        inv[off:off + n] = ref.invntt(canon(fwd[off:off + n]))
    d["config4"] = {"nreq": int(len(reqs)), "pool_polys": int(pool.shape[0]),
                    "npoly": int(reqs[:, 1].sum()), "input": tk.canon_digest(pool),
                    "ntt": tk.canon_digest(fwd), "invntt_of_canonical_ntt": tk.canon_digest(inv)}
    print(f"config 4: {d['config4']}", flush=True)
    A, s = matvec_inputs()
    t = ref.matvec(A, s, 3)
    d["config5"] = {"ninst": int(s.shape[0]), "k": 3, "A": tk.canon_digest(A),
                    "s": tk.canon_digest(s), "t": tk.canon_digest(t)}
    print(f"config 5: {d['config5']}", flush=True)
    d["packed"] = packed_digests(ref)
    d["kpke"] = kpke_digests(ref)
    d.update(rank_digests(ref))
    return d


def rank_digests(ref: Reference) -> dict:
    """Configuration 5's product and the K-PKE cores at ranks 2 and 4, on the
    same 65,536-instance workloads as rank 3."""
    d = {"config5_by_rank": {}, "kpke_by_rank": {}}
    for k in (2, 4):
        A, s = matvec_inputs(k=k)
        t = ref.matvec(A, s, k)
        d["config5_by_rank"][str(k)] = {"ninst": int(s.shape[0]), "k": k,
                                        "A": tk.canon_digest(A), "s": tk.canon_digest(s),
                                        "t": tk.canon_digest(t)}
        print(f"config 5, k = {k}: {d['config5_by_rank'][str(k)]}", flush=True)
        d["kpke_by_rank"][str(k)] = kpke_digests(ref, k=k)
    return d


def main():
    ref = Reference()
    path = os.path.join(HERE, "digests.json")
    if "--ranks-only" in sys.argv:  # add the rank 2 and 4 digests to the existing file
        with open(path) as f:
            d = json.load(f)
        d.update(rank_digests(ref))
        with open(path, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        print("wrote digests.json")
        return
    if "--digests-only" not in sys.argv:
        fx = fixtures(ref)
        np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **fx)
        print("wrote fixtures.npz")
    d = digests(ref)
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    print("wrote digests.json")


if __name__ == "__main__":
    main()
