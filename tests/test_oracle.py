"""Pins the oracle (oracle/kyber_oracle.c) before it is trusted as the checker.

* raw int16 equality with the committed golden vectors of the compiled
  reference (tests/golden/fixtures.npz) and, where it is built here, with the
  compiled reference itself (oracle/_ref);
* whole-domain contracts of the reductions (ref/reduce.c:5-15, :25-34);
* the negacyclic NTT closed form by brute force, invntt(ntt(x)) == 2^16 x,
  and the schoolbook product through basemul, all mod q;
* full-size digests of every benchmark configuration.
"""
import os

import numpy as np
import pytest

from conftest import canon, negacyclic

N, Q = 256, 3329
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def fx():
    with np.load(os.path.join(GOLD, "fixtures.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def brv7(i):
    return int(f"{i:07b}"[::-1], 2)


def ntt_closed_form(f):
    """Pair p of ntt(f) is f mod (X^2 - g_p), g_p = 17^(2 brv7(p) + 1) (mod q)."""
    f = np.asarray(f, dtype=np.int64) % Q
    out = np.zeros(N, dtype=np.int64)
    for p in range(128):
        g = pow(17, 2 * brv7(p) + 1, Q)
        pw = np.array([pow(g, e, Q) for e in range(128)], dtype=np.int64)
        out[2 * p] = int(np.dot(f[0::2], pw) % Q)
        out[2 * p + 1] = int(np.dot(f[1::2], pw) % Q)
    return out


@pytest.mark.parametrize("cat", ["uniform", "full", "edge", "small"])
def test_golden_raw(oracle, fx, cat):
    x = fx[f"in_{cat}"]
    assert np.array_equal(oracle.ntt(x), fx[f"ntt_{cat}"])
    assert np.array_equal(oracle.invntt(x), fx[f"invntt_{cat}"])
    assert np.array_equal(oracle.roundtrip(x), fx[f"rt_{cat}"])


def test_golden_basemul_and_matvec(oracle, fx):
    assert np.array_equal(oracle.basemul(fx["bm_a"], fx["bm_b"]), fx["bm_r"])
    for k in (2, 3, 4):
        t = oracle.matvec(fx[f"mv{k}_A"], fx[f"mv{k}_s"], k)
        assert np.array_equal(t, fx[f"mv{k}_t"])


def test_zetas_closed_form(oracle, tk):
    z = oracle.zetas()
    want = []
    for i in range(128):
        v = pow(17, brv7(i), Q) * 65536 % Q
        want.append(v - Q if v > Q // 2 else v)
    assert z.tolist() == want
    assert np.array_equal(tk.host_tables()["zetas"], z)


def test_zetas_match_compiled_reference(oracle, reference):
    assert np.array_equal(oracle.zetas(), reference.zetas())


def test_against_compiled_reference(oracle, reference, tk):
    for seed, lo, hi in [(201, 0, Q), (202, -32768, 32768), (203, -3328, 3329)]:
        x = tk.wl_uniform(4096 * N, seed, lo, hi).reshape(4096, N)
        assert np.array_equal(oracle.ntt(x), reference.ntt(x))
        assert np.array_equal(oracle.invntt(x), reference.invntt(x))
    a = tk.wl_uniform(512 * N, 204, -32768, 32768).reshape(512, N)
    b = tk.wl_uniform(512 * N, 205, -32768, 32768).reshape(512, N)
    assert np.array_equal(oracle.basemul(a, b), reference.basemul(a, b))
    for v in (-32768, -1, 0, 1, 32767, 12345, -3329):
        assert oracle.barrett_reduce(v) == reference.barrett_reduce(v)
    for v in (-(Q << 15), (Q << 15) - 1, 0, 123456789 % (Q << 15), -77777777):
        assert oracle.montgomery_reduce(v) == reference.montgomery_reduce(v)


def test_oracle_under_sanitizers():
    """The restatement built with AddressSanitizer and UndefinedBehaviorSanitizer
    (gcc, host only, single-threaded) runs every exported function over the
    int16 range and checks its invariants (oracle/sanitize_check.c)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc or make not available")
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle")
    r = subprocess.run(["make", "-s", "-C", here, "sanitize-check"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitize-check ok" in r.stdout


def test_reductions_whole_domain(oracle):
    """montgomery: all of [-q 2^15, q 2^15); barrett: every int16."""
    assert oracle.check_reductions() == 0


def test_ntt_brute_force(oracle, tk):
    x = tk.wl_uniform(6 * N, 206, -3328, 3329).reshape(6, N)
    x[5] = 0
    x[5, 7] = 1
    got = canon(oracle.ntt(x))
    for p in range(6):
        assert np.array_equal(got[p], ntt_closed_form(x[p])), f"poly {p}"


def test_invntt_is_scaled_inverse(oracle, tk):
    x = tk.wl_uniform(64 * N, 207, -3328, 3329).reshape(64, N)
    rt = canon(oracle.invntt(oracle.ntt(x)))
    assert np.array_equal(rt, canon(x.astype(np.int64) * 65536))  # 2^16 x


def test_basemul_is_negacyclic_product(oracle, tk):
    a = tk.wl_uniform(4 * N, 208, -3328, 3329).reshape(4, N)
    b = tk.wl_uniform(4 * N, 209, -3328, 3329).reshape(4, N)
    prod = oracle.invntt(oracle.basemul(oracle.ntt(a), oracle.ntt(b)))
    for p in range(4):
        assert np.array_equal(canon(prod[p]), negacyclic(a[p], b[p]))


def test_matvec_is_module_product(oracle, tk):
    k = 3
    a = tk.wl_uniform(2 * k * k * N, 210, -3328, 3329).reshape(2, k, k, N)
    A_hat = oracle.ntt(a.reshape(-1, N)).reshape(2, k, k, N)
    s = tk.wl_cbd2(2 * k * N, 211).reshape(2, k, N)
    t = oracle.matvec(A_hat, s, k)
    for inst in range(2):
        for i in range(k):
            want = sum(negacyclic(a[inst, i, j], s[inst, j]) for j in range(k)) % Q
            assert np.array_equal(canon(t[inst, i]), want)


@pytest.mark.parametrize("cfg", [1, 2, 3])
def test_full_size_digests(oracle, tk, digests, cfg):
    d = digests[f"config{cfg}"]
    x = tk.workload(cfg)
    assert tk.canon_digest(x) == d["input"]
    f = oracle.ntt(x)
    assert tk.canon_digest(f) == d["ntt"]
    assert tk.canon_digest(oracle.invntt(canon(f))) == d["invntt_of_canonical_ntt"]


def test_ragged_and_matvec_digests(oracle, tk, digests):
    from golden.make_golden import matvec_inputs, ragged_pool
    reqs, pool = ragged_pool()
    d = digests["config4"]
    assert (len(reqs), pool.shape[0], int(reqs[:, 1].sum())) == (d["nreq"], d["pool_polys"],
                                                                  d["npoly"])
    fwd = pool.copy()
    for off, n, _k, _kind in reqs:
        fwd[off:off + n] = oracle.ntt(pool[off:off + n])
    assert tk.canon_digest(fwd) == d["ntt"]
    A, s = matvec_inputs()
    dm = digests["config5"]
    assert (tk.canon_digest(A), tk.canon_digest(s)) == (dm["A"], dm["s"])
    assert tk.canon_digest(oracle.matvec(A, s, 3)) == dm["t"]
    for k in (2, 4):  # the same product at the other ranks
        A, s = matvec_inputs(k=k)
        dm = digests["config5_by_rank"][str(k)]
        assert (tk.canon_digest(A), tk.canon_digest(s)) == (dm["A"], dm["s"])
        assert tk.canon_digest(oracle.matvec(A, s, k)) == dm["t"]
