"""Packed small-coefficient input (the reference's TC-PACK format): the
oracle's pack/unpack pinned by hand-computed bytes, round trips and the
wrap-around of out-of-range values."""
import numpy as np
import pytest

N = 256


def _poly(vals):
    p = np.zeros((1, N), dtype=np.int16)
    p[0, :len(vals)] = vals
    return p


def test_packed_sizes(oracle):
    assert [oracle.packed_poly_bytes(b) for b in (1, 2, 3, 4, 5)] == [0, 64, 128, 128, 0]


@pytest.mark.parametrize("bits,vals,want", [
    # 2 bits, 4 per byte: fields 1 | -1 -> 0b11 | 0 | -2 -> 0b10
    (2, [1, -1, 0, -2], [0b10_00_11_01]),
    (2, [0, 0, 0, 0, 1, 1, 1, 1], [0, 0b01_01_01_01]),
    # 3 bits, 2 per byte at bits 0 and 3: -2 -> 6, 1 -> 1 ; 3, -3 -> 5
    (3, [-2, 1, 3, -3], [6 | 1 << 3, 3 | 5 << 3]),
    (3, [-4, 3], [4 | 3 << 3]),
    # 4 bits, nibbles: -8 -> 8, 7
    (4, [-8, 7, -1, 0], [8 | 7 << 4, 15]),
])
def test_pack_hand_vectors(oracle, bits, vals, want):
    b = oracle.pack_small(_poly(vals), bits)
    assert list(b[0, :len(want)]) == want
    assert not b[0, len(want):].any()
    assert list(oracle.unpack_small(b, bits)[0, :len(vals)]) == vals


def test_last_coefficient_position(oracle):
    # coefficient 255 sits in the last byte at the last field of that byte
    for bits, per in ((2, 4), (3, 2), (4, 2)):
        p = np.zeros((1, N), dtype=np.int16)
        p[0, 255] = 1
        b = oracle.pack_small(p, bits)
        assert b[0, -1] == 1 << ((per - 1) * bits) and not b[0, :-1].any()


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_round_trips(oracle, bits):
    rng = np.random.default_rng(bits)
    lo, hi = -(1 << (bits - 1)), 1 << (bits - 1)
    x = rng.integers(lo, hi, size=(300, N)).astype(np.int16)
    assert np.array_equal(oracle.unpack_small(oracle.pack_small(x, bits), bits), x)
    per = 8 // bits
    used = sum(((1 << bits) - 1) << (bits * f) for f in range(per))
    b = (rng.integers(0, 256, size=(300, oracle.packed_poly_bytes(bits))) & used).astype(np.uint8)
    assert np.array_equal(oracle.pack_small(oracle.unpack_small(b, bits), bits), b)


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_out_of_range_keeps_low_bits(oracle, bits):
    x = np.arange(-40, 40, dtype=np.int16)
    x = np.resize(x, (1, N)).astype(np.int16)
    got = oracle.unpack_small(oracle.pack_small(x, bits), bits).astype(np.int32)
    m = 1 << bits
    want = ((x.astype(np.int32) % m) + m) % m
    want = np.where(want >= m // 2, want - m, want)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_packed_full_size_digests(oracle, tk, digests, bits):
    """The 2^20-poly packed workloads: input, packed bytes and the forward NTT
    of the compiled reference, recomputed with the oracle."""
    import hashlib
    from golden.make_golden import packed_inputs
    d = digests["packed"][f"bits{bits}"]
    x = packed_inputs(bits)
    assert tk.canon_digest(x) == d["input"]
    assert hashlib.sha256(oracle.pack_small(x, bits).tobytes()).hexdigest() == d["packed_sha256"]
    assert tk.canon_digest(oracle.ntt(x)) == d["ntt"]
