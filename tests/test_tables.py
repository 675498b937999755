"""Host constant builder (tensorkyber_amd/csrc/tk_tables.cpp), CPU only.

The stage matrices are checked against their closed forms, the packed MFMA
operands are checked by emulating the documented lane layouts of
v_mfma_i32_16x16x32_i8 and v_mfma_i32_32x32x32_i8 in integer numpy (the GPU
tests confirm the layouts on the hardware) against the oracle, and the
24-bit Barrett epilogue is checked exhaustively over its input range.
"""
import numpy as np
import pytest

from conftest import canon

N, Q = 256, 3329
BIAS_Q, CENTRE, BARRETT_M = 2500, 1664, 1290167


def brv(x, bits):
    return int(f"{x:0{bits}b}"[::-1], 2)


@pytest.fixture(scope="module")
def T(tk):
    return {k: v.astype(np.int64) for k, v in tk.host_tables().items()}


def test_fwd1_is_vandermonde(T):
    # layers 1-4 leave block i holding f mod (X^16 - b_i), b_i = 17^(8 (2 brv4(i) + 1))
    for i in range(16):
        b = pow(17, 8 * (2 * brv(i, 4) + 1), Q)
        assert T["fwd1"][i].tolist() == [pow(b, m, Q) for m in range(16)]


def test_fwd2_parity_blocks(T):
    # block m, output pair s: reduce mod (X^2 - d), d = 17^(2 brv7(8m + s) + 1)
    for m in range(16):
        for c in range(16):
            d = pow(17, 2 * brv(8 * m + c // 2, 7) + 1, Q)
            for j in range(16):
                want = pow(d, (j - c % 2) // 2, Q) if (j - c) % 2 == 0 else 0
                assert T["fwd2"][m, c, j] == want


def test_inverse_scales(T):
    """inv2 = 8 fwd2^-1 per block and inv1 = 2^13 fwd1^-1 (2^13 * 8 = 2^16)."""
    eye = np.eye(16, dtype=np.int64)
    assert np.array_equal(T["inv1"] @ T["fwd1"] % Q, eye * pow(2, 13, Q) % Q)
    for m in range(16):
        assert np.array_equal(T["inv2"][m] @ T["fwd2"][m] % Q, eye * 8)


def test_stage_product_is_reference_ntt(T, oracle, tk):
    x = tk.wl_uniform(16 * N, 301, -32768, 32768).reshape(16, N)
    X = canon(x).astype(np.int64).reshape(16, 16, 16)
    Y = np.einsum("im,pmc->pic", T["fwd1"], X) % Q
    Z = np.einsum("bcj,pbj->pbc", T["fwd2"], Y) % Q
    assert np.array_equal(Z.reshape(16, N), canon(oracle.ntt(canon(x))))
    Ya = np.einsum("bcj,pbj->pbc", T["inv2"], X) % Q
    Za = np.einsum("im,pmc->pic", T["inv1"], Ya) % Q
    assert np.array_equal(Za.reshape(16, N), canon(oracle.invntt(canon(x))))


# --- emulation of the packed operands ---------------------------------------

def digits(x):
    """x = 128 + x0' + 256 x1': x0' = (x & 255) - 128, x1' = x >> 8."""
    x = np.asarray(x, dtype=np.int64)
    return (x & 255) - 128, x >> 8


def sbytes(words):
    return np.ascontiguousarray(words, dtype=np.uint32).view(np.uint8).astype(np.int8).astype(
        np.int64)


def emu16(ops, bias, data, d01=None):
    """One 16x16x32 MFMA pair: ops[2][64][2], data[16 cols][16 u] -> S[16][16]
    (d01: the two data digits, default the byte digits of `data`)."""
    A = np.zeros((2, 16, 32), np.int64)
    for lane in range(64):
        g, i = lane >> 4, lane & 15
        for limb in range(2):
            A[limb, i, 8 * g:8 * g + 8] = sbytes(ops[limb, lane])
    B = np.zeros((32, 16), np.int64)
    d0, d1 = digits(data) if d01 is None else d01
    for lane in range(64):
        g, j = lane >> 4, lane & 15
        B[8 * g:8 * g + 4, j] = d0[j, 4 * g:4 * g + 4]
        B[8 * g + 4:8 * g + 8, j] = d1[j, 4 * g:4 * g + 4]
    return A[0] @ B + bias[:, None] + 256 * (A[1] @ B)


def emu32(ops, bias, data):
    """One 32x32x32 MFMA: ops[64][4], data[32 cols][16 u] -> S[16][32]."""
    A = np.zeros((32, 32), np.int64)
    for lane in range(64):
        h, rho = lane >> 5, lane & 31
        A[rho, 16 * h:16 * h + 16] = sbytes(ops[lane])
    uset = lambda h: [4 * h, 4 * h + 1, 4 * h + 2, 4 * h + 3, 8 + 4 * h, 9 + 4 * h, 10 + 4 * h,
                      11 + 4 * h]
    B = np.zeros((32, 32), np.int64)
    d0, d1 = digits(data)
    for lane in range(64):
        h, j = lane >> 5, lane & 31
        B[16 * h:16 * h + 8, j] = d0[j, uset(h)]
        B[16 * h + 8:16 * h + 16, j] = d1[j, uset(h)]
    D = A @ B
    S = np.zeros((16, 32), np.int64)
    for h in range(2):
        for q, row in enumerate(uset(h)):
            # rows of the accumulator as each lane half sees them: lo row U_h[q],
            # hi row U_h[q] + 16
            S[row] = D[row] + 256 * D[row + 16] + bias[row]
    return S


def mod_ok(S, W, data, extra=0):
    want = (W @ np.asarray(data, dtype=np.int64).T + extra) % Q
    return np.array_equal(S % Q, want)


@pytest.mark.parametrize("stage,name", [(0, "fwd1"), (1, "fwd2"), (2, "inv2"), (3, "inv1")])
def test_mfma16_operands(tk, T, stage, name):
    ops, bias = tk.host_mfma_operands(16, stage)
    extra = CENTRE if stage == 3 else 0
    W = T[name] if stage in (0, 3) else None
    for mat in range(ops.shape[0]):
        Wm = W if W is not None else T[name][mat]
        for seed, lo, hi in [(310 + mat, -32768, 32768), (330 + mat, 0, 2 * Q)]:
            data = tk.wl_uniform(16 * 16, seed, lo, hi).reshape(16, 16)
            S = emu16(ops[mat], bias[mat], data)
            assert S.min() >= 0 and S.max() < (1 << 24)
            assert mod_ok(S, Wm, data, extra)


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_mfma16_packed_first_pass_operands(tk, T, bits):
    """Stage 2+bits: forward layers 1-4 on packed digits (field f, sign bit s),
    x = f - 2^bits s, over every value of the field and random tiles."""
    ops, bias = tk.host_mfma_operands(16, 2 + bits)
    lo, hi = -(1 << (bits - 1)), 1 << (bits - 1)
    tiles = [tk.wl_uniform(256, 360 + bits + t, lo, hi).reshape(16, 16) for t in range(4)]
    tiles.append(np.resize(np.arange(lo, hi), 256).reshape(16, 16))
    for x in tiles:
        field = np.asarray(x, dtype=np.int64) & ((1 << bits) - 1)
        f, sgn = field, field >> (bits - 1)
        assert np.array_equal(f - (sgn << bits), x)
        S = emu16(ops[0], bias[0], x, (f, sgn))
        assert S.min() >= 0 and S.max() < (1 << 24)
        assert mod_ok(S, T["fwd1"], x)


@pytest.mark.parametrize("stage,name", [(0, "fwd1"), (1, "fwd2"), (2, "inv2"), (3, "inv1")])
def test_mfma32_operands(tk, T, stage, name):
    ops, bias = tk.host_mfma_operands(32, stage)
    extra = CENTRE if stage == 3 else 0
    for mat in range(ops.shape[0]):
        Wm = T[name] if stage in (0, 3) else T[name][mat]
        data = tk.wl_uniform(32 * 16, 350 + mat, -32768, 32768).reshape(32, 16)
        S = emu32(ops[mat], bias[mat], data)
        assert S.min() >= 0 and S.max() < (1 << 24)
        assert mod_ok(S, Wm, data, extra)


def test_worst_case_accumulator_bounds(tk):
    """|C_lo + 256 C_hi| over all int16 data stays inside the bias margin."""
    for shape in (16, 32):
        for stage in range(4):
            ops, bias = tk.host_mfma_operands(shape, stage)
            by = sbytes(ops)
            assert by.min() >= -128 and by.max() <= 127
            if shape == 16:
                lo = np.abs(sbytes(ops[:, 0])).reshape(ops.shape[0], 64, 8)
                hi = np.abs(sbytes(ops[:, 1])).reshape(ops.shape[0], 64, 8)
                assert hi.max() <= 7
            assert bias.min() >= BIAS_Q * Q and bias.max() < BIAS_Q * Q + Q + CENTRE
    assert 32 * 128 * 128 + 256 * 32 * 7 * 128 < BIAS_Q * Q
    assert 32 * 128 * 128 + 256 * 32 * 7 * 128 + BIAS_Q * Q + Q + CENTRE < (1 << 24)


def test_barrett24_exhaustive():
    """For every 0 <= S < 2^24: t = (S*m) >> 32 with m = floor(2^32/q) gives
    S - t q in [0, 2q) and congruent to S; one conditional subtract -> [0, q)."""
    S = np.arange(1 << 24, dtype=np.uint64)
    t = (S * np.uint64(BARRETT_M)) >> np.uint64(32)
    r = S - t * np.uint64(Q)
    assert r.max() < 2 * Q
    assert np.array_equal(r % Q, S % Q)
    r1 = np.minimum(r, (r - np.uint64(Q)) & np.uint64(0xFFFFFFFF))
    assert r1.max() < Q
