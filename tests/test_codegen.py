"""Code-generation checks on the gfx950 kernels (CPU only: hipcc cross-compiles).

Each HIP source is compiled to assembly with the library's own flags and the
compiler's resource remarks.  Two properties that GPU tests only catch by
luck are checked on every kernel:

* no spills to memory, and the tile kernels at their occupancy targets (four
  waves per SIMD for the MFMA16 kernels, the queued packed MFMA16 and the
  wide decryption, three for the wide matvec, two for MFMA32): a spill or a
  lost wave costs 5-15% on the streaming kernels without changing any
  result;
* no store of more than 8 bytes whose data VGPRs the very next VALU
  instruction overwrites.  On gfx950 such a store writes the new values; the
  compiler only guards stores without a scalar offset register, so a buffer
  store with one can lose its data under load (seen as whole 16-byte chunks of
  garbage in a full-size run, never in small ones).
"""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess

import pytest

from tensorkyber_amd import build as tkbuild

HIPCC = tkbuild.HIPCC if os.path.exists(tkbuild.HIPCC) else shutil.which("hipcc")
pytestmark = pytest.mark.skipif(HIPCC is None, reason="hipcc not available")

KERNEL_SOURCES = sorted(f for f in os.listdir(tkbuild.CSRC) if f.endswith(".hip"))


def _compile(src, tmp):
    asm = os.path.join(tmp, src + ".s")
    flags = [f for f in tkbuild.HIP_FLAGS if f != "-fPIC"]
    cmd = [HIPCC, *flags, "--cuda-device-only", "-S", "-Rpass-analysis=kernel-resource-usage",
           os.path.join(tkbuild.CSRC, src), "-o", asm]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    with open(asm) as f:
        return src, f.read(), r.stderr


@pytest.fixture(scope="module")
def compiled(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("asm"))
    with cf.ThreadPoolExecutor(min(8, len(KERNEL_SOURCES))) as ex:
        return list(ex.map(lambda s: _compile(s, tmp), KERNEL_SOURCES))


def _resources(remarks):
    """{kernel: {field: value}} from the kernel-resource-usage remarks."""
    out, name = {}, None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z \[\]/]+): (\d+)", line)
        if m and name:
            out[name][m.group(1).strip()] = int(m.group(2))
    return out


def test_no_spills_and_tile_kernel_occupancy(compiled):
    seen = four = queued = wide_mv = 0
    for src, _, remarks in compiled:
        for name, r in _resources(remarks).items():
            seen += 1
            # SGPR spills go to VGPR lanes; VGPR spills and any scratch go to memory
            assert r.get("VGPRs Spill", 0) == 0, f"{src}: {name} spills VGPRs"
            assert r.get("ScratchSize [bytes/lane]", 0) == 0, f"{src}: {name} uses scratch"
            # (k_decrypt<K, TABLE, WIDE>, mangled ...Lb0ELb1E for the wide
            # variant: the table-driven one runs three)
            wide = "k_decrypt" in name and "Lb0ELb1E" in name
            # the queued packed instances, one 16-wave workgroup per CU: four
            # per SIMD (k_mfma16<FORM, INV, 16, PackedTiles<B>, true, true, false>)
            if "k_mfma16" in name and "ELi16ENS_11PackedTiles" in name:
                assert r["Occupancy [waves/SIMD]"] >= 4, f"{name}: occupancy below 4"
                queued += 1
            elif "k_mfma16" in name or wide:
                assert r["Occupancy [waves/SIMD]"] >= 4, f"{name}: occupancy below 4"
                four += 1
            if "k_mfma32" in name:
                assert r["Occupancy [waves/SIMD]"] >= 2, f"{name}: occupancy below 2"
            if "k_matvec_wide" in name:  # 12-wave workgroups, one per CU: three per SIMD
                assert r["Occupancy [waves/SIMD]"] >= 3, f"{name}: occupancy below 3"
                wide_mv += 1
    assert seen > 50, "kernel remarks not found"
    assert four >= 30, "four-wave kernels not found"
    assert queued == 6, "queued packed forwards: 2 output forms x 3 widths"
    assert wide_mv == 12, "wide matvec kernels for ranks 2, 3, 4 x s int16 or packed in 2, 3, 4 bits"


_WIDE_STORE = re.compile(r"^(buffer|global|flat)_store_dwordx(3|4)\b.*?v\[(\d+):(\d+)\]")
_VALU_DST = re.compile(r"^v_\S+\s+v\[?(\d+)(?::(\d+))?")


def _functions(asm):
    """(name, body) of every function in the assembly (data symbols, such as
    device arrays, have labels too but no .Lfunc_end)."""
    funcs = set(re.findall(r"^\s*\.type\s+(\w+),\s*@function", asm, re.M))
    for m in re.finditer(r"^(_Z\w+):", asm, re.M):
        if m.group(1) in funcs:
            yield m.group(1), asm[m.end():asm.index(".Lfunc_end", m.end())]


def _wide_store_hazards(asm):
    found, stores = [], 0
    for name, body in _functions(asm):
        ins = [l.strip() for l in body.split("\n")
               if l.startswith("\t") and not l.strip().startswith((".", ";"))]
        for j, line in enumerate(ins[:-1]):
            st = _WIDE_STORE.match(line)
            if not st:
                continue
            stores += 1
            lo, hi = int(st.group(3)), int(st.group(4))
            d = _VALU_DST.match(ins[j + 1])
            if d:
                a, b = int(d.group(1)), int(d.group(2) or d.group(1))
                if a <= hi and b >= lo:
                    found.append(f"{name}: {line} || {ins[j + 1]}")
    return found, stores


def test_wide_store_data_not_overwritten_next(compiled):
    total = 0
    for src, asm, _ in compiled:
        found, stores = _wide_store_hazards(asm)
        total += stores
        assert not found, f"{src}:\n" + "\n".join(found[:5])
    assert total > 100, "no wide stores found: the scan is not reading the assembly"


def test_scan_detects_the_pattern():
    """The scan itself flags a store followed at once by a write of its data."""
    asm = ("\t.type\t_Zfake,@function\n_Zfake:\n\tbuffer_store_dwordx4 v[8:11], v43, s[20:23], s65 offen nt\n"
           "\tv_bfe_i32 v9, v4, 0, 16\n\ts_endpgm\n.Lfunc_end0:\n")
    found, stores = _wide_store_hazards(asm)
    assert stores == 1 and len(found) == 1
    safe = asm.replace("\tv_bfe_i32 v9", "\ts_nop 0\n\tv_bfe_i32 v9")
    assert _wide_store_hazards(safe)[0] == []
