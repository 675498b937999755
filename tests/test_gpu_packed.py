"""GPU parity of the packed small-coefficient paths (the reference's TC-PACK
input format): the device packer against the oracle's pack_small, the packed
forward NTT against the oracle's ntt of the unpacked polys, and the fused
matrix-vector product with packed secrets against the int16 one and the
oracle, including the full-size digests of the compiled reference."""
import numpy as np
import pytest

from conftest import canon, centred

pytestmark = pytest.mark.gpu

N = 256


def _dev(a, device, dtype=np.int16):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(device)


def _small(tk, npoly, bits, seed):
    lo, hi = -(1 << (bits - 1)), 1 << (bits - 1)
    return tk.wl_uniform(npoly * N, seed, lo, hi).reshape(npoly, N)


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_pack_kernel_vs_oracle(tk, gpu, oracle, bits):
    x = _small(tk, 777, bits, 60 + bits)
    got = tk.pack_small(_dev(x, gpu), bits).cpu().numpy().reshape(777, -1)
    assert np.array_equal(got, oracle.pack_small(x, bits))


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_pack_kernel_keeps_low_bits(tk, gpu, oracle, bits):
    """Values outside the field's range are wrapped as pack_small does."""
    x = tk.wl_uniform(64 * N, 70 + bits, -300, 300).reshape(64, N)
    got = tk.pack_small(_dev(x, gpu), bits).cpu().numpy().reshape(64, -1)
    assert np.array_equal(got, oracle.pack_small(x, bits))


@pytest.mark.parametrize("bits", [2, 3, 4])
@pytest.mark.parametrize("method", ["mfma16", "mfma32"])
@pytest.mark.parametrize("npoly", [1, 15, 17, 33, 1000])
def test_ntt_packed_vs_oracle(tk, gpu, oracle, bits, method, npoly):
    x = _small(tk, npoly, bits, 80 + bits + npoly)
    packed = _dev(oracle.pack_small(x, bits), gpu, np.uint8)
    want = canon(oracle.ntt(x))
    got = tk.ntt_packed(packed, bits, method=method).cpu().numpy()
    assert np.array_equal(got, want), f"bits {bits} {method} npoly {npoly}"
    got_c = tk.ntt_packed(packed, bits, method=method, form="centered").cpu().numpy()
    assert np.array_equal(got_c, centred(want))
    # partial last tile: nothing written past the batch
    import torch
    buf = torch.full((npoly + 8, N), 1234, dtype=torch.int16, device=gpu)
    tk.ntt_packed(packed, bits, out=buf[:npoly], method=method)
    host = buf.cpu().numpy()
    assert np.array_equal(host[:npoly], want)
    assert (host[npoly:] == 1234).all(), "kernel wrote past the batch"


def test_ntt_packed_extremes(tk, gpu, oracle):
    """Every poly at a field extreme: all min, all max, alternating."""
    for bits in (2, 3, 4):
        lo, hi = -(1 << (bits - 1)), (1 << (bits - 1)) - 1
        x = np.zeros((4, N), dtype=np.int16)
        x[0], x[1], x[2, ::2], x[2, 1::2], x[3, 0] = lo, hi, lo, hi, hi
        got = tk.ntt_packed(_dev(oracle.pack_small(x, bits), gpu, np.uint8), bits).cpu().numpy()
        assert np.array_equal(got, canon(oracle.ntt(x))), f"bits {bits}"


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_ntt_packed_full_size_digest(tk, gpu, digests, bits):
    """2^20 polys: the device packer's bytes and the packed forward NTT against
    the digests of the compiled reference."""
    import hashlib
    import torch
    from golden.make_golden import packed_inputs
    d = digests["packed"][f"bits{bits}"]
    x = _dev(packed_inputs(bits), gpu)
    packed = tk.pack_small(x, bits)
    torch.cuda.synchronize()
    assert hashlib.sha256(packed.cpu().numpy().tobytes()).hexdigest() == d["packed_sha256"]
    out = tk.ntt_packed(packed, bits)
    assert tk.canon_digest(out.cpu().numpy()) == d["ntt"]


@pytest.mark.parametrize("k,ninst", [(2, 333), (3, 1000), (4, 257), (3, 1), (3, 17)])
def test_matvec_packed_vs_oracle(tk, gpu, oracle, k, ninst):
    A = tk.wl_uniform(ninst * k * k * N, 90 + k, 0, 3329).reshape(ninst, k, k, N)
    s = tk.wl_cbd2(ninst * k * N, 95 + k).reshape(ninst, k, N)
    want = centred(oracle.matvec(A, s, k)).reshape(ninst * k, N)
    dA = _dev(A, gpu)
    for bits in (3, 4):
        sp = _dev(oracle.pack_small(s.reshape(-1, N), bits), gpu, np.uint8)
        got = tk.matvec_packed(dA, sp, k, bits).cpu().numpy()
        assert np.array_equal(got, want), f"k {k} ninst {ninst} bits {bits}"
    # ternary secrets in 2 bits against the int16 path
    s2 = tk.wl_uniform(ninst * k * N, 99, -1, 2).reshape(ninst, k, N)
    sp2 = _dev(oracle.pack_small(s2.reshape(-1, N), 2), gpu, np.uint8)
    got2 = tk.matvec_packed(dA, sp2, k, 2).cpu().numpy()
    ref2 = tk.matvec(dA, _dev(s2, gpu), k).cpu().numpy().reshape(ninst * k, N)
    assert np.array_equal(got2, ref2)


def test_matvec_packed_full_size_digest(tk, gpu, digests):
    """Configuration 5 with the CBD secrets packed in 3 bits."""
    from golden.make_golden import matvec_inputs
    A, s = matvec_inputs()
    sp = tk.pack_small(_dev(s.reshape(-1, N), gpu), 3)
    t = tk.matvec_packed(_dev(A, gpu), sp, 3, 3).cpu().numpy()
    assert tk.canon_digest(t) == digests["config5"]["t"]
