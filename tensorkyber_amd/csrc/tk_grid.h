// Grid sizing of the streaming kernels that give each lane one item of a
// batch and fall back to a grid-stride loop beyond the largest grid.  Host
// arithmetic only (no HIP types), so it is checked on the CPU.
#pragma once

#include <stddef.h>

namespace tk {

// A dispatch packet carries the grid size in work-items as a 32-bit number,
// so a grid holds at most 2^32 - 1 work-items: (2^32 - 1) / block workgroups.
constexpr size_t max_grid_blocks(unsigned block)
{
  return ((size_t(1) << 32) - 1) / block;
}

// Workgroups of `block` lanes for `items` items, one per lane, capped at the
// largest grid (each lane then strides over the rest).
constexpr unsigned stream_grid(size_t items, unsigned block)
{
  const size_t need = (items + block - 1) / block, cap = max_grid_blocks(block);
  return (unsigned)(need < cap ? need : cap);
}

}  // namespace tk
