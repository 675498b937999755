"""ctypes bindings to the checker libraries built by oracle/Makefile.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, always as the checker and never
as the code that is measured or shipped.

* ``Oracle`` wraps ``oracle/liboracle.so``, the project's C restatement of
  ``ntt_bench/ref/ntt.c`` and ``ntt_bench/ref/reduce.c`` (see
  ``kyber_oracle.h`` for the file:line each function follows).
* ``Reference`` wraps ``oracle/_ref/libkyber_ref.so``, the reference's own C
  sources compiled unchanged; it exists only where the reference checkout was
  present when the checker was built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

Q = 3329
N = 256
_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libkyber_ref.so")

_i16p = ctypes.POINTER(ctypes.c_int16)


def _ptr(a: np.ndarray):
    assert a.dtype == np.int16 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i16p)


def canon(a) -> np.ndarray:
    """Canonical representatives in [0, q) as int16."""
    return np.mod(np.asarray(a, dtype=np.int32), Q).astype(np.int16)


def centred(a) -> np.ndarray:
    """Centred representatives in [-(q-1)/2, (q-1)/2] as int16."""
    c = np.mod(np.asarray(a, dtype=np.int32), Q)
    return np.where(c > Q // 2, c - Q, c).astype(np.int16)


class _Batch:
    """Shared batch wrappers; subclasses bind the symbol prefix."""

    _prefix = ""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        self.path = path
        self.lib = ctypes.CDLL(path)
        p = self._prefix
        for name in ("ntt_batch", "invntt_batch", "roundtrip_batch"):
            f = getattr(self.lib, p + name)
            f.argtypes = [_i16p, ctypes.c_size_t, ctypes.c_int]
            f.restype = None
        getattr(self.lib, p + "zetas").restype = _i16p
        getattr(self.lib, p + "montgomery_reduce").argtypes = [ctypes.c_int32]
        getattr(self.lib, p + "montgomery_reduce").restype = ctypes.c_int16
        getattr(self.lib, p + "barrett_reduce").argtypes = [ctypes.c_int16]
        getattr(self.lib, p + "barrett_reduce").restype = ctypes.c_int16
        getattr(self.lib, p + "max_threads").restype = ctypes.c_int
        sz, i = ctypes.c_size_t, ctypes.c_int
        getattr(self.lib, p + "kpke_keygen_batch").argtypes = [_i16p] * 5 + [i, sz, i]
        getattr(self.lib, p + "kpke_encrypt_batch").argtypes = [_i16p] * 7 + [ctypes.c_void_p,
                                                                            i, sz, i]
        getattr(self.lib, p + "kpke_decrypt_batch").argtypes = [_i16p] * 4 + [i, sz, i]

    # K-PKE arithmetic cores (pqcrystals ref/indcpa.c composition); inputs are
    # reshaped to [ninst, k(, k), 256], outputs returned the same way.
    def kpke_keygen(self, A_hat, s, e, k: int, nthreads: int = 0):
        s = np.ascontiguousarray(s, dtype=np.int16).reshape(-1, k, N)
        ninst = s.shape[0]
        A = np.ascontiguousarray(A_hat, dtype=np.int16).reshape(ninst, k, k, N)
        e = np.ascontiguousarray(e, dtype=np.int16).reshape(ninst, k, N)
        t_hat, s_hat = np.zeros_like(s), np.zeros_like(s)
        getattr(self.lib, self._prefix + "kpke_keygen_batch")(
            _ptr(t_hat), _ptr(s_hat), _ptr(A), _ptr(s), _ptr(e), k, ninst, nthreads)
        return t_hat, s_hat

    def kpke_encrypt(self, At_hat, t_hat, r, e1, e2, m, k: int, nthreads: int = 0):
        r = np.ascontiguousarray(r, dtype=np.int16).reshape(-1, k, N)
        ninst = r.shape[0]
        At = np.ascontiguousarray(At_hat, dtype=np.int16).reshape(ninst, k, k, N)
        t = np.ascontiguousarray(t_hat, dtype=np.int16).reshape(ninst, k, N)
        e1 = np.ascontiguousarray(e1, dtype=np.int16).reshape(ninst, k, N)
        e2 = np.ascontiguousarray(e2, dtype=np.int16).reshape(ninst, N)
        mp = None if m is None else np.ascontiguousarray(m, dtype=np.int16).reshape(ninst, N)
        u, v = np.zeros_like(r), np.zeros((ninst, N), dtype=np.int16)
        getattr(self.lib, self._prefix + "kpke_encrypt_batch")(
            _ptr(u), _ptr(v), _ptr(At), _ptr(t), _ptr(r), _ptr(e1), _ptr(e2),
            None if mp is None else mp.ctypes.data, k, ninst, nthreads)
        return u, v

    def kpke_decrypt(self, s_hat, u, v, k: int, nthreads: int = 0):
        u = np.ascontiguousarray(u, dtype=np.int16).reshape(-1, k, N)
        ninst = u.shape[0]
        sh = np.ascontiguousarray(s_hat, dtype=np.int16).reshape(ninst, k, N)
        v = np.ascontiguousarray(v, dtype=np.int16).reshape(ninst, N)
        w = np.zeros((ninst, N), dtype=np.int16)
        getattr(self.lib, self._prefix + "kpke_decrypt_batch")(
            _ptr(w), _ptr(sh), _ptr(u), _ptr(v), k, ninst, nthreads)
        return w

    def _call(self, name, polys, nthreads):
        a = np.ascontiguousarray(polys, dtype=np.int16).copy()
        assert a.size % N == 0
        getattr(self.lib, self._prefix + name)(_ptr(a), a.size // N, nthreads)
        return a

    def ntt(self, polys, nthreads: int = 0) -> np.ndarray:
        return self._call("ntt_batch", polys, nthreads)

    def invntt(self, polys, nthreads: int = 0) -> np.ndarray:
        return self._call("invntt_batch", polys, nthreads)

    def roundtrip(self, polys, nthreads: int = 0) -> np.ndarray:
        return self._call("roundtrip_batch", polys, nthreads)

    def ntt_inplace(self, a: np.ndarray, nthreads: int = 0) -> None:
        getattr(self.lib, self._prefix + "ntt_batch")(_ptr(a), a.size // N, nthreads)

    def invntt_inplace(self, a: np.ndarray, nthreads: int = 0) -> None:
        getattr(self.lib, self._prefix + "invntt_batch")(_ptr(a), a.size // N, nthreads)

    def roundtrip_inplace(self, a: np.ndarray, nthreads: int = 0) -> None:
        getattr(self.lib, self._prefix + "roundtrip_batch")(_ptr(a), a.size // N, nthreads)

    def zetas(self) -> np.ndarray:
        p = getattr(self.lib, self._prefix + "zetas")()
        return np.ctypeslib.as_array(p, shape=(128,)).copy()

    def montgomery_reduce(self, a: int) -> int:
        return int(getattr(self.lib, self._prefix + "montgomery_reduce")(a))

    def barrett_reduce(self, a: int) -> int:
        return int(getattr(self.lib, self._prefix + "barrett_reduce")(a))

    def max_threads(self) -> int:
        return int(getattr(self.lib, self._prefix + "max_threads")())


class Oracle(_Batch):
    _prefix = "oracle_"

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.oracle_basemul_batch.argtypes = [_i16p, _i16p, _i16p, ctypes.c_size_t, ctypes.c_int]
        L.oracle_basemul_batch.restype = None
        L.oracle_matvec_batch.argtypes = [_i16p, _i16p, _i16p, ctypes.c_int, ctypes.c_size_t,
                                          ctypes.c_int]
        L.oracle_matvec_batch.restype = None
        L.oracle_fqmul.argtypes = [ctypes.c_int16, ctypes.c_int16]
        L.oracle_packed_poly_bytes.argtypes = [ctypes.c_int]
        L.oracle_packed_poly_bytes.restype = ctypes.c_size_t
        L.oracle_pack_small.argtypes = [ctypes.c_void_p, _i16p, ctypes.c_size_t, ctypes.c_int]
        L.oracle_unpack_small.argtypes = [_i16p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.oracle_fqmul.restype = ctypes.c_int16
        L.oracle_check_reductions.restype = ctypes.c_longlong

    def packed_poly_bytes(self, bits: int) -> int:
        return int(self.lib.oracle_packed_poly_bytes(bits))

    def pack_small(self, polys, bits: int) -> np.ndarray:
        a = np.ascontiguousarray(polys, dtype=np.int16).reshape(-1, 256)
        ppb = self.packed_poly_bytes(bits)
        assert ppb > 0, f"unsupported width {bits}"
        out = np.zeros((a.shape[0], ppb), dtype=np.uint8)
        self.lib.oracle_pack_small(out.ctypes.data, _ptr(a), a.shape[0], bits)
        return out

    def unpack_small(self, packed, bits: int) -> np.ndarray:
        ppb = self.packed_poly_bytes(bits)
        assert ppb > 0, f"unsupported width {bits}"
        b = np.ascontiguousarray(packed, dtype=np.uint8).reshape(-1, ppb)
        out = np.zeros((b.shape[0], 256), dtype=np.int16)
        self.lib.oracle_unpack_small(_ptr(out), b.ctypes.data, b.shape[0], bits)
        return out

    def check_reductions(self) -> int:
        """Violations of the reductions' contracts over their whole domains."""
        return int(self.lib.oracle_check_reductions())

    def basemul(self, a, b, nthreads: int = 0) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.int16)
        b = np.ascontiguousarray(b, dtype=np.int16)
        assert a.shape == b.shape and a.size % N == 0
        r = np.empty_like(a)
        self.lib.oracle_basemul_batch(_ptr(r), _ptr(a), _ptr(b), a.size // N, nthreads)
        return r

    def matvec(self, A_hat, s, k: int, nthreads: int = 0) -> np.ndarray:
        A_hat = np.ascontiguousarray(A_hat, dtype=np.int16)
        s = np.ascontiguousarray(s, dtype=np.int16)
        ninst = s.size // (k * N)
        assert A_hat.size == ninst * k * k * N
        t = np.empty((ninst, k, N), dtype=np.int16)
        self.lib.oracle_matvec_batch(_ptr(t), _ptr(A_hat), _ptr(s), k, ninst, nthreads)
        return t

    def fqmul(self, a: int, b: int) -> int:
        return int(self.lib.oracle_fqmul(a, b))


class Reference(_Batch):
    _prefix = "refdrv_"

    def __init__(self, path: str = REF_SO):
        super().__init__(path)
        L = self.lib
        L.refdrv_basemul.argtypes = [_i16p, _i16p, _i16p, ctypes.c_int16]
        L.refdrv_basemul.restype = None
        L.refdrv_basemul_batch.argtypes = [_i16p, _i16p, _i16p, ctypes.c_size_t, ctypes.c_int]
        L.refdrv_basemul_batch.restype = None
        L.refdrv_matvec_batch.argtypes = [_i16p, _i16p, _i16p, ctypes.c_int, ctypes.c_size_t,
                                          ctypes.c_int]
        L.refdrv_matvec_batch.restype = None

    def basemul(self, a, b, nthreads: int = 0) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.int16)
        b = np.ascontiguousarray(b, dtype=np.int16)
        r = np.empty_like(a)
        self.lib.refdrv_basemul_batch(_ptr(r), _ptr(a), _ptr(b), a.size // N, nthreads)
        return r

    def matvec(self, A_hat, s, k: int, nthreads: int = 0) -> np.ndarray:
        A_hat = np.ascontiguousarray(A_hat, dtype=np.int16)
        s = np.ascontiguousarray(s, dtype=np.int16)
        ninst = s.size // (k * N)
        t = np.empty((ninst, k, N), dtype=np.int16)
        self.lib.refdrv_matvec_batch(_ptr(t), _ptr(A_hat), _ptr(s), k, ninst, nthreads)
        return t

    def basemul_pair(self, a2, b2, zeta: int) -> np.ndarray:
        a2 = np.ascontiguousarray(a2, dtype=np.int16)
        b2 = np.ascontiguousarray(b2, dtype=np.int16)
        r = np.empty(2, dtype=np.int16)
        self.lib.refdrv_basemul(_ptr(r), _ptr(a2), _ptr(b2), zeta)
        return r

    def poly_basemul(self, a, b) -> np.ndarray:
        """Whole-poly basemul built from the reference's pair basemul with the
        pairing of ntt()'s output order (zetas[64+i] / -zetas[64+i])."""
        a = np.ascontiguousarray(a, dtype=np.int16).reshape(-1, N)
        b = np.ascontiguousarray(b, dtype=np.int16).reshape(-1, N)
        z = self.zetas()
        r = np.empty_like(a)
        for p in range(a.shape[0]):
            for i in range(64):
                r[p, 4 * i:4 * i + 2] = self.basemul_pair(a[p, 4 * i:4 * i + 2],
                                                          b[p, 4 * i:4 * i + 2], int(z[64 + i]))
                r[p, 4 * i + 2:4 * i + 4] = self.basemul_pair(a[p, 4 * i + 2:4 * i + 4],
                                                              b[p, 4 * i + 2:4 * i + 4],
                                                              -int(z[64 + i]))
        return r


def reference_available() -> bool:
    return os.path.exists(REF_SO)
