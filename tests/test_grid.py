"""Grid arithmetic of the streaming kernels (basemul, packer), on the CPU.

A dispatch packet holds the grid size in work-items as a 32-bit number, so a
launch of more than 2^32 - 1 lanes fails.  basemul gives each lane one
16-byte chunk (32 per poly) and the packer one output dword, so from 2^27
polys (basemul) or 2^28 (2-bit packer) the grid must stop at the largest one
and each lane stride over the rest; the API accepts up to 2^36 polys."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tensorkyber_amd", "csrc")
CXX = shutil.which("g++")

PROG = r"""
#include <stdio.h>
#include "tk_grid.h"
int main() {
  const size_t sizes[] = {1, 1000, (size_t)1 << 20, ((size_t)1 << 27) - 1, (size_t)1 << 27,
                          (size_t)1 << 28, (size_t)1 << 36};
  for (size_t npoly : sizes) {
    // basemul: 32 chunks per poly; packer: 16 (2 bits) or 32 (3, 4 bits) dwords per poly
    const size_t items[3] = {npoly * 32, npoly * 16, npoly * 32};
    for (size_t n : items)
      printf("%zu %zu %u\n", npoly, n, tk::stream_grid(n, 256));
  }
  return 0;
}
"""


@pytest.mark.skipif(CXX is None, reason="g++ not available")
def test_stream_grid_within_dispatch_limit(tmp_path):
    src, exe = tmp_path / "grid.cpp", tmp_path / "grid"
    src.write_text(PROG)
    subprocess.run([CXX, "-std=c++17", "-O1", f"-I{CSRC}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    rows = [tuple(int(v) for v in line.split()) for line in out.splitlines()]
    assert len(rows) == 21
    for npoly, items, grid in rows:
        lanes = grid * 256
        assert 1 <= grid and lanes <= 2**32 - 1, (npoly, items, grid)
        if items <= (2**32 - 1) // 256 * 256:
            assert lanes >= items and lanes - items < 256  # one item per lane
        else:
            assert grid == (2**32 - 1) // 256  # the largest grid, lanes stride
    # the sizes the old 2^30-workgroup cap launched with more than 2^32 lanes
    big = {(n, i): g for n, i, g in rows if n >= 2**27}
    assert all(g == 16777215 for (n, i), g in big.items() if i > 2**32)
