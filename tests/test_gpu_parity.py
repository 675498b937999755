"""GPU parity: every kernel through the C ABI against the oracle.

Bit-exact bar: canonical (mod q) equality on every coefficient of every poly
for the MFMA method and the canonical/centred forms (the reference's own
criterion, ntt_bench.cu:397-421, applied to all polys), and raw int16 equality
with ref ntt()/invntt()/basemul() for TK_FORM_REF.
"""
import numpy as np
import pytest

from conftest import canon, centred

pytestmark = pytest.mark.gpu

N, Q = 256, 3329
METHODS = ["butterfly", "mfma16", "mfma32"]


def _dev(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int16)).to(device)


def _host(t):
    return t.cpu().numpy()


def _inputs(tk, kind, npoly, seed):
    if kind == "uniform":
        return tk.wl_uniform(npoly * N, seed).reshape(npoly, N)
    if kind == "edge":
        return tk.wl_edge_mix(npoly, seed)
    if kind == "small":   # |x| <= 16383: no int16 wrap inside ref ntt()/invntt()
        return tk.wl_uniform(npoly * N, seed, -16383, 16384).reshape(npoly, N)
    raise ValueError(kind)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("kind,npoly", [("uniform", 4096), ("edge", 4099), ("small", 1000)])
def test_forward_canonical(tk, gpu, oracle, method, kind, npoly):
    x = _inputs(tk, kind, npoly, seed=11)
    print(f"forward {method} {kind} npoly={npoly}", flush=True)
    out = _host(tk.ntt(_dev(x, gpu), method=method, form="canonical"))
    want = canon(oracle.ntt(x))
    assert out.min() >= 0 and out.max() < Q
    bad = np.count_nonzero(out != want)
    assert bad == 0, f"{bad} coefficients differ"


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("kind,npoly", [("uniform", 4096), ("edge", 4099), ("small", 1000)])
def test_inverse_centred(tk, gpu, oracle, method, kind, npoly):
    x = _inputs(tk, kind, npoly, seed=12)
    print(f"inverse {method} {kind} npoly={npoly}", flush=True)
    out = _host(tk.invntt(_dev(x, gpu), method=method, form="centered"))
    want = centred(oracle.invntt(x))
    assert out.min() >= -1664 and out.max() <= 1664
    bad = np.count_nonzero(out != want)
    assert bad == 0, f"{bad} coefficients differ"


@pytest.mark.parametrize("npoly", [1, 2, 15, 16, 17, 31, 33, 63, 65, 257])
@pytest.mark.parametrize("method", METHODS)
def test_ragged_tail_sizes(tk, gpu, oracle, method, npoly):
    """Partial tiles: batches that are not multiples of 4/16/32 polys."""
    x = _inputs(tk, "uniform", npoly, seed=13)
    guard = np.full((8, N), 1234, np.int16)
    buf = _dev(np.concatenate([x, guard]), gpu)
    view = buf[: npoly * N] if buf.dim() == 1 else buf[:npoly]
    tk.ntt(view, method=method)
    got = _host(buf)
    fwd = canon(oracle.ntt(x))
    assert np.array_equal(got[:npoly], fwd)
    assert np.array_equal(got[npoly:], guard), "kernel wrote past the batch"
    tk.invntt(view, method=method)
    got = _host(buf)
    assert np.array_equal(got[:npoly], centred(oracle.invntt(fwd)))
    assert np.array_equal(got[npoly:], guard), "inverse wrote past the batch"


@pytest.mark.parametrize("kind", ["uniform", "edge", "full"])
def test_butterfly_ref_form_raw(tk, gpu, oracle, kind):
    """TK_FORM_REF is raw-int16 identical to ref ntt()/invntt(), wraps included."""
    npoly = 2048
    if kind == "full":
        x = tk.wl_uniform(npoly * N, 14, -32768, 32768).reshape(npoly, N)
    else:
        x = _inputs(tk, kind, npoly, seed=14)
    f = _host(tk.ntt(_dev(x, gpu), method="butterfly", form="ref"))
    assert np.array_equal(f, oracle.ntt(x))
    i = _host(tk.invntt(_dev(x, gpu), method="butterfly", form="ref"))
    assert np.array_equal(i, oracle.invntt(x))


@pytest.mark.parametrize("method", ["mfma16", "mfma32"])
def test_mfma_full_int16_domain(tk, gpu, oracle, method):
    """Any int16 input: the MFMA result is the exact transform mod q.

    ntt/invntt are linear mod q, so the transform of x equals the transform of
    canon(x); the oracle on canon(x) never wraps (|x| < q).
    """
    npoly = 1024
    x = tk.wl_uniform(npoly * N, 15, -32768, 32768).reshape(npoly, N)
    x[0, :] = 32767
    x[1, :] = -32768
    xc = canon(x)
    f = _host(tk.ntt(_dev(x, gpu), method=method))
    assert np.array_equal(f, canon(oracle.ntt(xc)))
    i = _host(tk.invntt(_dev(x, gpu), method=method))
    assert np.array_equal(i, centred(oracle.invntt(xc)))


@pytest.mark.parametrize("method", METHODS)
def test_out_of_place_and_roundtrip(tk, gpu, oracle, method):
    import torch
    npoly = 3000
    x = _inputs(tk, "uniform", npoly, seed=16)
    src = _dev(x, gpu)
    dst = torch.empty_like(src)
    tk.ntt(src, out=dst, method=method)
    assert np.array_equal(_host(src), x), "input modified by out-of-place call"
    tk.invntt(dst, method=method)
    # invntt(ntt(x)) == 2^16 x == -1044 x (mod q)
    assert np.array_equal(canon(_host(dst)), canon(x.astype(np.int64) * -1044))


@pytest.mark.parametrize("form", ["ref", "canonical", "centered"])
def test_basemul(tk, gpu, oracle, form):
    npoly = 1500
    a = _inputs(tk, "uniform", npoly, seed=17)
    b = tk.wl_uniform(npoly * N, 18, -3328, 3329).reshape(npoly, N)
    r = _host(tk.basemul(_dev(a, gpu), _dev(b, gpu), form=form))
    want = oracle.basemul(a, b)
    if form == "ref":
        assert np.array_equal(r, want)
    elif form == "canonical":
        assert np.array_equal(r, canon(want))
    else:
        assert np.array_equal(r, centred(want))


def test_default_entry_points(tk, gpu, oracle):
    """ntt_batch / invntt_batch / basemul_batch with the reference signatures."""
    import torch
    npoly = 512
    x = _inputs(tk, "uniform", npoly, seed=19)
    d = _dev(x, gpu)
    s = torch.cuda.current_stream().cuda_stream
    assert tk.lib().ntt_batch(d.data_ptr(), npoly, s) == 0
    fwd = _host(d)
    assert np.array_equal(fwd, canon(oracle.ntt(x)))
    assert tk.lib().invntt_batch(d.data_ptr(), npoly, s) == 0
    assert np.array_equal(_host(d), centred(oracle.invntt(fwd)))
    r = torch.empty_like(d)
    assert tk.lib().basemul_batch(r.data_ptr(), d.data_ptr(), d.data_ptr(), npoly, s) == 0
    h = _host(d)
    assert np.array_equal(_host(r), oracle.basemul(h, h))


@pytest.mark.parametrize("npoly", [32768 * 16 + 1, 32768 * 24 + 7, (1 << 20) - 3])
def test_mfma16_tile_swap_sizes(tk, gpu, oracle, npoly):
    """Batches around the size where the MFMA16 launcher switches to the tile
    swap (more tiles than the grid's wave slots: 256 CUs x 4 workgroups x 8 x
    4 waves = 32,768 tiles on MI355X), each ending in a partial tile: some
    waves swap into a second tile and others do not, in place and out of
    place, against the oracle on every poly."""
    import torch
    print(f"mfma16 tile swap npoly={npoly}", flush=True)
    x = _inputs(tk, "edge", npoly, 70 + npoly % 97)
    src = _dev(x, gpu)
    fwd = tk.ntt(src, out=torch.empty_like(src), method="mfma16", form="canonical")
    want_f = canon(oracle.ntt(x))
    assert np.array_equal(_host(fwd), want_f), "forward"
    inv = fwd.clone()
    tk.invntt(inv, out=inv, method="mfma16", form="centered")
    assert np.array_equal(_host(inv), centred(oracle.invntt(want_f))), "inverse in place"
    assert np.array_equal(_host(src), x), "input changed"
    torch.cuda.synchronize()


@pytest.mark.parametrize("method,npoly", [("mfma16", 65535), ("mfma16", 65536),
                                          ("mfma16", 524288), ("mfma32", 131071),
                                          ("mfma32", 131072), ("mfma32", 524288)])
def test_kept_output_policy_edges(tk, gpu, oracle, method, npoly):
    """Linear batches of 65,536 (MFMA16) or 131,072 (MFMA32) to 524,288 polys
    store their output with the default cache policy (it stays in the
    Infinity Cache for the next pass), others nontemporally: both sides of the
    lower edge and the upper edge (the next size up is covered above), out of
    place and in place, against the oracle on every poly."""
    import torch
    print(f"output policy edge {method} npoly={npoly}", flush=True)
    x = _inputs(tk, "edge", npoly, 80 + npoly % 89)
    src = _dev(x, gpu)
    fwd = tk.ntt(src, out=torch.empty_like(src), method=method, form="canonical")
    want_f = canon(oracle.ntt(x))
    assert np.array_equal(_host(fwd), want_f), "forward out of place"
    assert np.array_equal(_host(src), x), "input changed"
    tk.ntt(src, out=src, method=method, form="canonical")
    assert np.array_equal(_host(src), want_f), "forward in place"
    inv = tk.invntt(fwd, out=torch.empty_like(fwd), method=method, form="centered")
    assert np.array_equal(_host(inv), centred(oracle.invntt(want_f))), "inverse"
    torch.cuda.synchronize()
