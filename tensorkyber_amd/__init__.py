"""MI355X batched ML-KEM NTT engine: Python host side over the C ABI.

The product is the native library ``libtkyber.so`` (hand-written gfx950 HIP
kernels behind ``include/tkyber.h``).  This module loads it with ctypes and
mirrors the reference interface on device tensors:

=====================  ==================================================
reference (ntt_bench)  here
=====================  ==================================================
``ntt(r)``             ``ntt(polys)``      ref/ntt.h:10-11, ref/ntt.c:80-95
``invntt(r)``          ``invntt(polys)``   ref/ntt.h:13-14, ref/ntt.c:106-126
``basemul(r,a,b,z)``   ``basemul(a, b)``   ref/ntt.h:16-17, ref/ntt.c:139-146
``k_*(int16_t* d, int  the same calls on a [npoly, 256] int16 CUDA tensor
npoly)``               (ntt_bench.cu:61-62, 97-98, 181-182, 230-231)
=====================  ==================================================

PyTorch only provides device memory and the current stream; every transform
runs in the HIP kernels.  There is no CPU fallback: if the library is not
built, importing the product API raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

N = 256
Q = 3329

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libtkyber.so")

# tk_status
OK, EINVAL, ENOMEM, EDEVICE, ENOTSUP, ENODEV = 0, -22, -12, -5, -95, -19
# tk_method
METHOD_AUTO, METHOD_BUTTERFLY, METHOD_MFMA, METHOD_MFMA16, METHOD_MFMA32 = 0, 1, 2, 3, 4
# tk_form
FORM_DEFAULT, FORM_CANONICAL, FORM_CENTERED, FORM_REF = 0, 1, 2, 3

METHODS = {"auto": METHOD_AUTO, "butterfly": METHOD_BUTTERFLY, "mfma": METHOD_MFMA,
           "mfma16": METHOD_MFMA16, "mfma32": METHOD_MFMA32}
FORMS = {"default": FORM_DEFAULT, "canonical": FORM_CANONICAL, "centered": FORM_CENTERED,
         "ref": FORM_REF}


class TkError(RuntimeError):
    def __init__(self, status: int, what: str, hip: int = 0):
        self.status = status
        msg = f"{what}: {_LIB.tk_strerror(status).decode() if _LIB else status}"
        if status == EDEVICE:
            msg += f" (hipError {hip})"
        super().__init__(msg)


_i16p = ctypes.POINTER(ctypes.c_int16)
_LIB = None


class TkRequest(ctypes.Structure):
    _fields_ = [("base", ctypes.c_void_p), ("npoly", ctypes.c_uint32),
                ("k", ctypes.c_uint16), ("kind", ctypes.c_uint16)]


class TkFusedRequest(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("k", ctypes.c_int), ("ninst", ctypes.c_size_t),
                ("out0", ctypes.c_void_p), ("out1", ctypes.c_void_p), ("mat", ctypes.c_void_p),
                ("mat2", ctypes.c_void_p), ("x", ctypes.c_void_p), ("add0", ctypes.c_void_p),
                ("add1", ctypes.c_void_p), ("add2", ctypes.c_void_p)]


FUSED_MATVEC, FUSED_KEYGEN, FUSED_ENCRYPT, FUSED_DECRYPT = 0, 1, 2, 3


class TkWlRequest(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("npoly", ctypes.c_uint32),
                ("k", ctypes.c_uint16), ("kind", ctypes.c_uint16)]


def _declare(lib):
    vp, sz, i, u32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64
    sigs = {
        "tk_init": ([], i),
        "tk_strerror": ([i], ctypes.c_char_p),
        "tk_last_hip_error": ([], i),
        "tk_version": ([], ctypes.c_char_p),
        "ntt_batch": ([vp, sz, vp], i),
        "invntt_batch": ([vp, sz, vp], i),
        "basemul_batch": ([vp, vp, vp, sz, vp], i),
        "ntt_batch_ex": ([vp, vp, sz, i, i, vp], i),
        "invntt_batch_ex": ([vp, vp, sz, i, i, vp], i),
        "basemul_batch_ex": ([vp, vp, vp, sz, i, vp], i),
        "tk_host_tables": ([vp, vp, vp, vp, vp], i),
        "tk_host_mfma_operands": ([i, i, vp, vp], i),
        "tk_wl_draw": ([u64, u64, u64], u64),
        "tk_wl_below": ([u64, u64, u64, u32], u32),
        "tk_wl_uniform": ([vp, sz, u64, u64, i, i], None),
        "tk_wl_edge_mix": ([vp, sz, u64], None),
        "tk_wl_cbd2": ([vp, sz, u64, u64], None),
        "tk_wl_ragged_layout": ([vp, sz, sz, u64, ctypes.POINTER(sz)], sz),
        "tk_wl_canon_digest": ([vp, sz], u64),
        "tk_packed_poly_bytes": ([i], sz),
        "pack_small_batch": ([vp, vp, sz, i, vp], i),
        "ntt_batch_packed": ([vp, vp, sz, i, i, i, vp], i),
        "matvec_ntt_fused_packed": ([vp, vp, vp, i, sz, i, vp], i),
        "tk_fused_plan_create": ([ctypes.POINTER(vp), vp, sz], i),
        "tk_fused_plan_execute": ([vp, vp], i),
        "tk_fused_plan_launch": ([vp, vp], i),
        "tk_fused_plan_info": ([vp, ctypes.POINTER(i), ctypes.POINTER(sz)], i),
        "tk_fused_plan_destroy": ([vp], i),
        "tk_host_pipeline_create": ([ctypes.POINTER(vp), sz, i], i),
        "tk_host_pipeline_run": ([vp, vp, vp, sz, i, i, i], i),
        "tk_host_pipeline_destroy": ([vp], i),
        "ntt_batch_host": ([vp, sz], i),
        "invntt_batch_host": ([vp, sz], i),
        "kpke_keygen_core": ([vp, vp, vp, vp, vp, i, sz, vp], i),
        "kpke_encrypt_core": ([vp, vp, vp, vp, vp, vp, vp, vp, i, sz, vp], i),
        "kpke_decrypt_core": ([vp, vp, vp, vp, i, sz, vp], i),
        "tk_ragged_plan_create": ([ctypes.POINTER(vp), vp, sz, i, i, vp], i),
        "tk_ragged_plan_execute": ([vp, vp], i),
        "tk_ragged_plan_launch": ([vp, vp], i),
        "tk_ragged_plan_info": ([vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                                 ctypes.POINTER(sz)], i),
        "tk_ragged_plan_destroy": ([vp], i),
        "tk_ragged_plan_graph": ([vp, ctypes.POINTER(vp)], i),
        "tk_ragged_transform_noalloc": ([vp, sz, i, vp], i),
        "matvec_ntt_fused": ([vp, vp, vp, i, sz, vp], i),
        "tk_autotune": ([sz, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i), vp], i),
    }
    for name, (args, res) in sigs.items():
        f = getattr(lib, name)
        f.argtypes, f.restype = args, res
    return lib


def load_library(path: str = LIB_PATH):
    """Loads libtkyber.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m tensorkyber_amd.build`")
        # PyTorch-ROCm ships its own libamdhip64.so, which its libraries request
        # by the unversioned name.  Loaded first, it satisfies this library's
        # libamdhip64.so.7 too; loaded after this library, it would become a
        # second HIP runtime in the process and device initialisation fails.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _LIB = _declare(ctypes.CDLL(path))
    return _LIB


def lib():
    return load_library()


def _check(status: int, what: str):
    if status != OK:
        raise TkError(status, what, lib().tk_last_hip_error())


# ---------------------------------------------------------------------------
# device-tensor API (torch is plumbing: memory and the current stream)
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _stream_ptr(stream=None, device=None):
    """Raw handle of `stream`, or of the current stream of `device` (the
    tensors' device, which must be the current device: the C ABI launches on
    the current device).  torch.cuda.current_stream() builds a Stream object
    per call (about 2.7 us of host time, which a synchronised single call
    exposes), so the current stream's handle is read directly where torch
    offers that."""
    torch = _torch()
    cur = torch.cuda.current_device()
    if device is not None and device.index != cur:
        raise ValueError(f"tensors on {device}, but the current device is cuda:{cur}")
    if stream is not None:
        return stream.cuda_stream
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    return raw(cur) if raw is not None else torch.cuda.current_stream(cur).cuda_stream


def _as_polys(t, name="polys"):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.int16:
        raise TypeError(f"{name} must be an int16 CUDA tensor")
    if not t.is_contiguous() or t.numel() % N:
        raise ValueError(f"{name} must be contiguous with a multiple of {N} elements")
    return t


def _one_device(**tensors):
    """The one device of the given tensors (None skipped): the C ABI takes
    raw pointers and launches on one device."""
    dev = None
    for t in tensors.values():
        if t is None:
            continue
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError("tensors on different devices: " +
                             ", ".join(f"{n} on {t.device}" for n, t in tensors.items()
                                       if t is not None))
    return dev


def _same_size(out, ref, name="out", ref_name="the input"):
    """The C ABI takes the size from the input only, so an output that is
    smaller would be written past its end on the device."""
    if out.numel() != ref.numel():
        raise ValueError(f"{name} holds {out.numel()} elements, {ref_name} {ref.numel()}")


def _method(m):
    return METHODS[m] if isinstance(m, str) else int(m)


def _form(f):
    return FORMS[f] if isinstance(f, str) else int(f)


def ntt(polys, out=None, method="auto", form="default", stream=None):
    """Forward NTT of every poly (in place unless ``out`` is given)."""
    polys = _as_polys(polys)
    out = polys if out is None else _as_polys(out, "out")
    _same_size(out, polys)
    dev = _one_device(polys=polys, out=out)
    _check(lib().ntt_batch_ex(out.data_ptr(), polys.data_ptr(), polys.numel() // N,
                              _method(method), _form(form), _stream_ptr(stream, dev)), "ntt_batch")
    return out


def invntt(polys, out=None, method="auto", form="default", stream=None):
    """Inverse NTT (times 2^16, as invntt_tomont) of every poly."""
    polys = _as_polys(polys)
    out = polys if out is None else _as_polys(out, "out")
    _same_size(out, polys)
    dev = _one_device(polys=polys, out=out)
    _check(lib().invntt_batch_ex(out.data_ptr(), polys.data_ptr(), polys.numel() // N,
                                 _method(method), _form(form), _stream_ptr(stream, dev)),
           "invntt_batch")
    return out


def basemul(a, b, out=None, form="default", stream=None):
    """NTT-domain product of two poly batches (reference basemul per pair)."""
    a, b = _as_polys(a, "a"), _as_polys(b, "b")
    if a.numel() != b.numel():
        raise ValueError("a and b differ in size")
    torch = _torch()
    out = torch.empty_like(a) if out is None else _as_polys(out, "out")
    _same_size(out, a, ref_name="a")
    dev = _one_device(a=a, b=b, out=out)
    _check(lib().basemul_batch_ex(out.data_ptr(), a.data_ptr(), b.data_ptr(), a.numel() // N,
                                  _form(form), _stream_ptr(stream, dev)), "basemul_batch")
    return out


def matvec(A_hat, s, k: int, out=None, stream=None):
    """Fused t = invntt(A_hat o ntt(s)) for [ninst, k, k, 256] / [ninst, k, 256]."""
    A_hat, s = _as_polys(A_hat, "A_hat"), _as_polys(s, "s")
    ninst = s.numel() // (k * N)
    if A_hat.numel() != ninst * k * k * N:
        raise ValueError("A_hat and s shapes disagree")
    torch = _torch()
    out = torch.empty_like(s) if out is None else _as_polys(out, "out")
    if s.numel() != ninst * k * N or out.numel() != s.numel():
        raise ValueError("s and out must hold ninst*k polys")
    dev = _one_device(A_hat=A_hat, s=s, out=out)
    _check(lib().matvec_ntt_fused(out.data_ptr(), A_hat.data_ptr(), s.data_ptr(), k, ninst,
                                  _stream_ptr(stream, dev)), "matvec_ntt_fused")
    return out


def packed_poly_bytes(bits: int) -> int:
    """Bytes of one packed poly of `bits`-bit coefficients (0 if unsupported)."""
    return int(lib().tk_packed_poly_bytes(bits))


def _as_packed(t, bits: int, name="packed"):
    torch = _torch()
    ppb = packed_poly_bytes(bits)
    if ppb == 0:
        raise ValueError(f"unsupported coefficient width {bits}")
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.uint8:
        raise TypeError(f"{name} must be a uint8 CUDA tensor")
    if not t.is_contiguous() or t.numel() % ppb:
        raise ValueError(f"{name} must be contiguous with a multiple of {ppb} bytes")
    return t, ppb


def pack_small(polys, bits: int, out=None, stream=None):
    """Packs small-coefficient int16 polys into `bits`-bit fields (uint8)."""
    polys = _as_polys(polys)
    npoly = polys.numel() // N
    if out is None:
        ppb = packed_poly_bytes(bits)
        if ppb == 0:
            raise ValueError(f"unsupported coefficient width {bits}")
        out = _torch().empty(npoly * ppb, dtype=_torch().uint8, device=polys.device)
    out, ppb = _as_packed(out, bits, "out")
    if out.numel() != npoly * ppb:
        raise ValueError("out has the wrong size")
    dev = _one_device(polys=polys, out=out)
    _check(lib().pack_small_batch(out.data_ptr(), polys.data_ptr(), npoly, bits,
                                  _stream_ptr(stream, dev)), "pack_small_batch")
    return out


def ntt_packed(packed, bits: int, out=None, method="auto", form="default", stream=None):
    """Forward NTT of packed small-coefficient polys into int16 [npoly, 256]."""
    packed, ppb = _as_packed(packed, bits)
    npoly = packed.numel() // ppb
    torch = _torch()
    out = torch.empty((npoly, N), dtype=torch.int16, device=packed.device) if out is None \
        else _as_polys(out, "out")
    if out.numel() != npoly * N:
        raise ValueError("out has the wrong size")
    dev = _one_device(packed=packed, out=out)
    _check(lib().ntt_batch_packed(out.data_ptr(), packed.data_ptr(), npoly, bits, _method(method),
                                  _form(form), _stream_ptr(stream, dev)), "ntt_batch_packed")
    return out


def matvec_packed(A_hat, s_packed, k: int, bits: int, out=None, stream=None):
    """Fused t = invntt(A_hat o ntt(s)) with s packed: [ninst, k, packed bytes]."""
    A_hat = _as_polys(A_hat, "A_hat")
    s_packed, ppb = _as_packed(s_packed, bits, "s_packed")
    ninst = s_packed.numel() // (k * ppb)
    if s_packed.numel() != ninst * k * ppb or A_hat.numel() != ninst * k * k * N:
        raise ValueError("A_hat and s_packed shapes disagree")
    torch = _torch()
    out = torch.empty((ninst * k, N), dtype=torch.int16, device=A_hat.device) if out is None \
        else _as_polys(out, "out")
    if out.numel() != ninst * k * N:
        raise ValueError("out must hold ninst*k polys")
    dev = _one_device(A_hat=A_hat, s_packed=s_packed, out=out)
    _check(lib().matvec_ntt_fused_packed(out.data_ptr(), A_hat.data_ptr(), s_packed.data_ptr(), k,
                                         ninst, bits, _stream_ptr(stream, dev)),
           "matvec_ntt_fused_packed")
    return out


class HostPipeline:
    """Transforms host-resident poly batches (tk_host_pipeline_*).  Buffers:
    numpy int16 arrays or CPU torch tensors.  Pinned ones (torch.empty(...,
    pin_memory=True)) are read and written by the transform kernel over PCIe;
    others stream through the device in chunks of `chunk_polys`, `depth`
    chunks in flight."""

    def __init__(self, chunk_polys: int = 65536, depth: int = 2):
        self._p = ctypes.c_void_p()
        _check(lib().tk_host_pipeline_create(ctypes.byref(self._p), chunk_polys, depth),
               "tk_host_pipeline_create")

    @staticmethod
    def _host(a, name):
        torch = _torch()
        if isinstance(a, torch.Tensor):
            if a.is_cuda or a.dtype != torch.int16 or not a.is_contiguous() or a.numel() % N:
                raise ValueError(f"{name} must be a contiguous CPU int16 tensor of whole polys")
            return a.data_ptr(), a.numel() // N
        if not isinstance(a, np.ndarray) or a.dtype != np.int16 or not a.flags["C_CONTIGUOUS"] \
                or a.size % N:
            raise ValueError(f"{name} must be a contiguous int16 array of whole polys")
        return a.ctypes.data, a.size // N

    def run(self, inp, out=None, direction: int = 0, method="auto", form="default"):
        out = inp if out is None else out
        pi, n = self._host(inp, "in")
        po, n_out = self._host(out, "out")
        if n_out != n:
            raise ValueError("in and out differ in size")
        _check(lib().tk_host_pipeline_run(self._p, po, pi, n, direction, _method(method),
                                          _form(form)), "tk_host_pipeline_run")
        return out

    def close(self):
        if self._p:
            lib().tk_host_pipeline_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr_or_none(t):
    return None if t is None else t.data_ptr()


def kpke_keygen(A_hat, s, e, k: int, t_hat=None, s_hat=None, stream=None):
    """K-PKE key generation core: (t_hat, s_hat) in the NTT domain, centred."""
    A_hat, s, e = _as_polys(A_hat, "A_hat"), _as_polys(s, "s"), _as_polys(e, "e")
    ninst = s.numel() // (k * N)
    if s.numel() != ninst * k * N or e.numel() != s.numel() or A_hat.numel() != s.numel() * k:
        raise ValueError("A_hat, s and e shapes disagree")
    torch = _torch()
    t_hat = torch.empty_like(s) if t_hat is None else _as_polys(t_hat, "t_hat")
    s_hat = torch.empty_like(s) if s_hat is None else _as_polys(s_hat, "s_hat")
    if t_hat.numel() != s.numel() or s_hat.numel() != s.numel():
        raise ValueError("t_hat and s_hat must hold ninst*k polys")
    dev = _one_device(A_hat=A_hat, s=s, e=e, t_hat=t_hat, s_hat=s_hat)
    _check(lib().kpke_keygen_core(t_hat.data_ptr(), s_hat.data_ptr(), A_hat.data_ptr(),
                                  s.data_ptr(), e.data_ptr(), k, ninst, _stream_ptr(stream, dev)),
           "kpke_keygen_core")
    return t_hat, s_hat


def kpke_encrypt(At_hat, t_hat, r, e1, e2, m, k: int, u=None, v=None, stream=None):
    """K-PKE encryption core: (u [ninst*k, 256], v [ninst, 256]), centred."""
    At_hat, t_hat, r = _as_polys(At_hat, "At_hat"), _as_polys(t_hat, "t_hat"), _as_polys(r, "r")
    e1, e2 = _as_polys(e1, "e1"), _as_polys(e2, "e2")
    ninst = r.numel() // (k * N)
    if (r.numel() != ninst * k * N or t_hat.numel() != r.numel() or e1.numel() != r.numel()
            or At_hat.numel() != r.numel() * k or e2.numel() != ninst * N
            or (m is not None and _as_polys(m, "m").numel() != ninst * N)):
        raise ValueError("K-PKE encryption shapes disagree")
    torch = _torch()
    u = torch.empty((ninst * k, N), dtype=torch.int16, device=r.device) if u is None \
        else _as_polys(u, "u")
    v = torch.empty((ninst, N), dtype=torch.int16, device=r.device) if v is None \
        else _as_polys(v, "v")
    if u.numel() != ninst * k * N or v.numel() != ninst * N:
        raise ValueError("u must hold ninst*k polys and v ninst polys")
    dev = _one_device(At_hat=At_hat, t_hat=t_hat, r=r, e1=e1, e2=e2, m=m, u=u, v=v)
    _check(lib().kpke_encrypt_core(u.data_ptr(), v.data_ptr(), At_hat.data_ptr(), t_hat.data_ptr(),
                                   r.data_ptr(), e1.data_ptr(), e2.data_ptr(), _ptr_or_none(m), k,
                                   ninst, _stream_ptr(stream, dev)), "kpke_encrypt_core")
    return u, v


def kpke_decrypt(s_hat, u, v, k: int, w=None, stream=None):
    """K-PKE decryption core: w = reduce(v - invntt(s_hat . poly_ntt(u)))."""
    s_hat, u, v = _as_polys(s_hat, "s_hat"), _as_polys(u, "u"), _as_polys(v, "v")
    ninst = u.numel() // (k * N)
    if u.numel() != ninst * k * N or s_hat.numel() != u.numel() or v.numel() != ninst * N:
        raise ValueError("K-PKE decryption shapes disagree")
    torch = _torch()
    w = torch.empty((ninst, N), dtype=torch.int16, device=u.device) if w is None \
        else _as_polys(w, "w")
    if w.numel() != ninst * N:
        raise ValueError("w must hold ninst polys")
    dev = _one_device(s_hat=s_hat, u=u, v=v, w=w)
    _check(lib().kpke_decrypt_core(w.data_ptr(), s_hat.data_ptr(), u.data_ptr(), v.data_ptr(), k,
                                   ninst, _stream_ptr(stream, dev)), "kpke_decrypt_core")
    return w


# ---------------------------------------------------------------------------
# host introspection and synthetic workloads (no device needed)
# ---------------------------------------------------------------------------

def host_tables():
    L = lib()
    z = np.zeros(128, np.int16)
    f1 = np.zeros((16, 16), np.int16)
    f2 = np.zeros((16, 16, 16), np.int16)
    i2 = np.zeros((16, 16, 16), np.int16)
    i1 = np.zeros((16, 16), np.int16)
    _check(L.tk_host_tables(z.ctypes.data, f1.ctypes.data, f2.ctypes.data, i2.ctypes.data,
                            i1.ctypes.data), "tk_host_tables")
    return {"zetas": z, "fwd1": f1, "fwd2": f2, "inv2": i2, "inv1": i1}


def host_mfma_operands(shape: int, stage: int):
    nmat = 16 if stage in (1, 2) else 1
    if shape == 16:
        ops = np.zeros((nmat, 2, 64, 2), np.uint32)
    else:
        ops = np.zeros((nmat, 64, 4), np.uint32)
    bias = np.zeros((nmat, 16), np.int32)
    _check(lib().tk_host_mfma_operands(shape, stage, ops.ctypes.data, bias.ctypes.data),
           "tk_host_mfma_operands")
    return ops, bias


WL_COEF, WL_CAT, WL_POS, WL_REQ, WL_GAP, WL_MAT, WL_SEC = (
    0x434F4546, 0x434154, 0x504F53, 0x524551, 0x474150, 0x4D4154, 0x534543)


def wl_uniform(n: int, seed: int, lo: int = 0, hi: int = Q, stream: int = WL_COEF) -> np.ndarray:
    a = np.empty(n, np.int16)
    lib().tk_wl_uniform(a.ctypes.data, n, seed, stream, lo, hi)
    return a


def wl_edge_mix(npoly: int, seed: int) -> np.ndarray:
    a = np.empty(npoly * N, np.int16)
    lib().tk_wl_edge_mix(a.ctypes.data, npoly, seed)
    return a.reshape(npoly, N)


def wl_cbd2(n: int, seed: int, stream: int = WL_SEC) -> np.ndarray:
    a = np.empty(n, np.int16)
    lib().tk_wl_cbd2(a.ctypes.data, n, seed, stream)
    return a


def wl_ragged_layout(npoly_total: int, seed: int):
    L = lib()
    pool = ctypes.c_size_t(0)
    n = L.tk_wl_ragged_layout(None, 0, npoly_total, seed, ctypes.byref(pool))
    arr = (TkWlRequest * n)()
    L.tk_wl_ragged_layout(ctypes.addressof(arr), n, npoly_total, seed, ctypes.byref(pool))
    reqs = np.array([(r.offset, r.npoly, r.k, r.kind) for r in arr], dtype=np.int64).reshape(-1, 4)
    return reqs, int(pool.value)


def canon_digest(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a, dtype=np.int16)
    return int(lib().tk_wl_canon_digest(a.ctypes.data, a.size))


def workload(config: int, npoly: int | None = None):
    """Synthetic inputs of the benchmark configurations (int16 numpy arrays).

    1: uniform [0, q), seed 1, 65,536 polys; 2: uniform [0, q), seed 2,
    262,144 polys; 3: edge-value mix, seed 3, 2^20 polys.  ``npoly`` takes the
    first polys of the same stream (any prefix is identical to the full set).
    """
    sizes = {1: 65536, 2: 262144, 3: 1 << 20}
    n = sizes[config] if npoly is None else npoly
    if config in (1, 2):
        return wl_uniform(n * N, seed=config).reshape(n, N)
    if config == 3:
        return wl_edge_mix(n, seed=3)
    raise ValueError(config)


def matvec_workload(ninst: int = 65536, k: int = 3):
    """Configuration 5: A_hat uniform [0, q) (NTT domain) and CBD eta = 2
    secrets, seed 5; returns ([ninst, k, k, 256], [ninst, k, 256])."""
    A = wl_uniform(ninst * k * k * N, 5, 0, Q, stream=WL_MAT).reshape(ninst, k, k, N)
    s = wl_cbd2(ninst * k * N, 5, stream=WL_SEC).reshape(ninst, k, N)
    return A, s


def kpke_workload(ninst: int = 65536, k: int = 3, seed: int = 9) -> dict:
    """K-PKE inputs: A_hat uniform [0, q) (NTT domain, as matrix expansion
    yields it); s, e, r, e1, e2 CBD eta = 2 on consecutive seeds of the secret
    stream; m = (q+1)/2 times uniform message bits (poly_frommsg).  Arrays are
    [ninst, k, k, 256], [ninst, k, 256] or [ninst, 256]; "bits" holds the
    message bits."""
    A = wl_uniform(ninst * k * k * N, seed, 0, Q, stream=WL_MAT).reshape(ninst, k, k, N)
    vec = {name: wl_cbd2(ninst * k * N, seed + 1 + i).reshape(ninst, k, N)
           for i, name in enumerate(("s", "e", "r", "e1"))}
    e2 = wl_cbd2(ninst * N, seed + 5).reshape(ninst, N)
    bits = wl_uniform(ninst * N, seed + 6, 0, 2).reshape(ninst, N)
    m = (bits * ((Q + 1) // 2)).astype(np.int16)
    return {"A_hat": A, **vec, "e2": e2, "m": m, "bits": bits}


def packed_workload(bits: int, npoly: int = 1 << 20) -> np.ndarray:
    """Small-coefficient polys for the packed forward path: ternary (seed 6)
    for 2 bits, CBD eta = 2 (seed 7) for 3 bits, uniform [-8, 8) (seed 8) for
    4 bits."""
    if bits == 2:
        x = wl_uniform(npoly * N, 6, -1, 2)
    elif bits == 3:
        x = wl_cbd2(npoly * N, 7)
    elif bits == 4:
        x = wl_uniform(npoly * N, 8, -8, 8)
    else:
        raise ValueError(bits)
    return x.reshape(npoly, N)


__all__ = [
    "N", "Q", "TkError", "load_library", "lib", "ntt", "invntt", "basemul", "matvec",
    "packed_poly_bytes", "pack_small", "ntt_packed", "matvec_packed",
    "kpke_keygen", "kpke_encrypt", "kpke_decrypt", "HostPipeline", "TkFusedRequest",
    "host_tables", "host_mfma_operands", "workload", "matvec_workload", "packed_workload",
    "kpke_workload",
    "wl_uniform", "wl_edge_mix", "wl_cbd2", "wl_ragged_layout", "canon_digest", "TkRequest",
    "METHODS", "FORMS",
]
