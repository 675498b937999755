"""Python mirror without a GPU: every wrapper rejects an output of the wrong
size, or buffers on different devices, before anything reaches the C ABI.

The C ABI takes raw pointers and the size of the input only, so an `out`
smaller than the input would be written past its end on the device.  Here
the CUDA type check is relaxed so CPU and `meta` tensors reach the size and
device checks, and the library is replaced by one that fails the test if any
entry point is called."""
import pytest

torch = pytest.importorskip("torch")

N = 256


class _NoCalls:
    def __getattr__(self, name):
        raise AssertionError(f"{name} reached the C ABI")


@pytest.fixture
def tkw(monkeypatch):
    import tensorkyber_amd as tk

    def as_polys(t, name="polys"):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.int16:
            raise TypeError(name)
        if t.numel() % N:
            raise ValueError(name)
        return t

    monkeypatch.setattr(tk, "_as_polys", as_polys)
    monkeypatch.setattr(tk, "lib", lambda: _NoCalls())
    monkeypatch.setattr(tk, "_stream_ptr", lambda stream=None, device=None: None)
    return tk


def polys(n, device="cpu"):
    return torch.zeros((n, N), dtype=torch.int16, device=device)


@pytest.mark.parametrize("fn", ["ntt", "invntt"])
def test_transform_out_size_and_device(tkw, fn):
    f = getattr(tkw, fn)
    with pytest.raises(ValueError, match="elements"):
        f(polys(4), out=polys(3))
    with pytest.raises(ValueError, match="elements"):
        f(polys(4), out=polys(5))
    with pytest.raises(ValueError, match="different devices"):
        f(polys(4), out=polys(4, "meta"))


def test_basemul_out_size_and_devices(tkw):
    with pytest.raises(ValueError, match="differ in size"):
        tkw.basemul(polys(4), polys(3))
    with pytest.raises(ValueError, match="elements"):
        tkw.basemul(polys(4), polys(4), out=polys(2))
    with pytest.raises(ValueError, match="different devices"):
        tkw.basemul(polys(4), polys(4, "meta"))
    with pytest.raises(ValueError, match="different devices"):
        tkw.basemul(polys(4), polys(4), out=polys(4, "meta"))


def test_fused_wrappers_devices(tkw):
    k, n = 3, 2
    A, s = polys(n * k * k), polys(n * k)
    with pytest.raises(ValueError, match="ninst"):
        tkw.matvec(A, s, k, out=polys(n * k - 1))
    with pytest.raises(ValueError, match="different devices"):
        tkw.matvec(A, s, k, out=polys(n * k, "meta"))
    with pytest.raises(ValueError, match="different devices"):
        tkw.kpke_keygen(A, s, polys(n * k, "meta"), k)
    with pytest.raises(ValueError, match="different devices"):
        tkw.kpke_decrypt(s, s, polys(n), k, w=polys(n, "meta"))
    with pytest.raises(ValueError, match="different devices"):
        tkw.kpke_encrypt(A, s, s, s, polys(n), polys(n, "meta"), k)
