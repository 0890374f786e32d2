// Work-item scheduling and tile-arena arithmetic of the item-persistent kernels, in host + device
// form: the kernels (mpm_kernels.hip) call these with gridDim / blockIdx and device atomics, and a
// CPU check (tests/item_loop_check.cpp) runs the same functions over a full scene's item list with
// bounds asserts.
#pragma once
#include <stdint.h>

#include "mpm_kernels.hpp"
#include "mpm_math.hpp"

namespace mpm {

// Work items of an item-persistent kernel. Workgroups are dealt to the 8 XCDs in turn (b % 8), and
// each XCD takes contiguous items so that neighbouring blocks share an L2.
//  - first_item_of: the first s = min(items, 8 floor(grid / 8)) items are assigned statically, XCD x
//    taking [s x / 8, s (x + 1) / 8) in workgroup order: no atomic for a workgroup's first item, and a
//    workgroup without one (fewer items than workgroups) learns it from one load.
//  - next_item_of: the items after s are split into kWorkSplits contiguous ranges, each with its own
//    counter (take(range) = that counter's atomic fetch-and-increment); a workgroup starts on range
//    blockIdx % kWorkSplits (its XCD) and moves on when its range is used up. Called by thread 0;
//    0xffffffff when every range is done (at once when no items are left after the static ones).
MPM_HD uint32_t static_items_of(uint32_t items, uint32_t grid) {
  const uint32_t s = grid / 8u * 8u;
  return items < s ? items : s;
}
MPM_HD uint32_t first_item_of(uint32_t items, uint32_t grid, uint32_t block) {
  const uint32_t s = static_items_of(items, grid), x = block % 8u, i = block / 8u;
  const uint32_t lo = (uint32_t)((uint64_t)s * x / 8u), hi = (uint32_t)((uint64_t)s * (x + 1u) / 8u);
  return lo + i < hi ? lo + i : 0xffffffffu;
}
template <class Take>
MPM_HD uint32_t next_item_of(Take&& take, uint32_t items, uint32_t grid, uint32_t& range, uint32_t& spent) {
  const uint32_t s = static_items_of(items, grid), rest = items - s;
  if (rest == 0) return 0xffffffffu;
  while (spent < (uint32_t)kWorkSplits) {
    const uint32_t lo = s + (uint32_t)((uint64_t)rest * range / kWorkSplits);
    const uint32_t hi = s + (uint32_t)((uint64_t)rest * (range + 1) / kWorkSplits);
    const uint32_t c = take(range);
    if (c < hi - lo) return lo + c;
    spent++;
    range = (range + 1) % kWorkSplits;
  }
  return 0xffffffffu;
}

// One axis of a Morton code (the inverse of spread3): block coordinates from a block key.
MPM_HD uint32_t compact3(uint32_t x) {
  x &= 0x09249249u;
  x = (x ^ (x >> 2)) & 0x030c30c3u;
  x = (x ^ (x >> 4)) & 0x0300f00fu;
  x = (x ^ (x >> 8)) & 0xff0000ffu;
  x = (x ^ (x >> 16)) & 0x000003ffu;
  return x;
}

// A stencil base (b0, b1, b2) is in the drift-D window of the work item whose lowest in-window base
// is (wx, wy, wz) when it lies within D cells of the item's 4^3-cell block on every axis.
template <int D = kDrift>
MPM_HD bool in_window(int b0, int b1, int b2, int wx, int wy, int wz) {
  return (unsigned)(b0 - wx) < (unsigned)(4 + 2 * D) && (unsigned)(b1 - wy) < (unsigned)(4 + 2 * D) &&
         (unsigned)(b2 - wz) < (unsigned)(4 + 2 * D);
}

// Tile storage order. A tile box (origin o, extents e per axis) is cut at node-block boundaries
// (multiples of 4) into segments per axis; the chunks (z segment, y segment, x segment) are stored one
// after another in that order, each x fastest, so the part of the tile one node block holds is one
// contiguous run of at most 64 nodes (the grid update reads it as whole lines).
// tile_segment: the segment of node g (o <= g < o + e) along one axis: its start in box coordinates,
// its width, and g's place in it.
MPM_HD void tile_segment(int o, int e, int g, int& start, int& width, int& at) {
  const int b = g & ~3, lo = b > o ? b : o, hi = b + 4 < o + e ? b + 4 : o + e;
  start = lo - o;
  width = hi - lo;
  at = g - lo;
}
// The chunk of node block (b0, b1, b2) in a box: its first node (offset) and widths, or false when the
// box does not reach the block.
MPM_HD bool tile_chunk(int ox, int oy, int oz, int ex, int ey, int ez, int b0, int b1, int b2, uint32_t& offset, int& wx,
                       int& wy, int& wz, int& lo0, int& lo1, int& lo2) {
  lo0 = 4 * b0 > ox ? 4 * b0 : ox;
  lo1 = 4 * b1 > oy ? 4 * b1 : oy;
  lo2 = 4 * b2 > oz ? 4 * b2 : oz;
  wx = (4 * b0 + 4 < ox + ex ? 4 * b0 + 4 : ox + ex) - lo0;
  wy = (4 * b1 + 4 < oy + ey ? 4 * b1 + 4 : oy + ey) - lo1;
  wz = (4 * b2 + 4 < oz + ez ? 4 * b2 + 4 : oz + ez) - lo2;
  if (wx <= 0 || wy <= 0 || wz <= 0) return false;
  offset = (uint32_t)((lo2 - oz) * ey * ex + wz * ((lo1 - oy) * ex + wy * (lo0 - ox)));
  return true;
}
// Storage index of box node (lx, ly, lz) (box coordinates) in a tile: a bijection of the box's nodes
// onto [0, ex ey ez).
MPM_HD uint32_t tile_index(int ox, int oy, int oz, int ex, int ey, int ez, int lx, int ly, int lz) {
  int sx, wx, px, sy, wy, py, sz, wz, pz;
  tile_segment(ox, ex, ox + lx, sx, wx, px);
  tile_segment(oy, ey, oy + ly, sy, wy, py);
  tile_segment(oz, ez, oz + lz, sz, wz, pz);
  return (uint32_t)(sz * ey * ex + wz * (sy * ex + wy * sx) + px + wx * (py + wy * pz));
}

// Arena offset of a tile of `pnodes` nodes for which region `region` (rsize nodes) returned o from
// its bump counter, or 0xffffffff when it does not fit (the tile then goes to A with atomics). A
// tile box never exceeds the window tile (pnodes <= kTileNodes, by the window test); the bound is
// checked here too, so no box, whatever its size, can take space past its region's end.
MPM_HD uint32_t arena_offset(uint32_t region, uint32_t rsize, uint32_t o, uint32_t pnodes) {
  const bool fits = pnodes <= (uint32_t)kTileNodes && pnodes <= rsize && o <= rsize - pnodes;
  return fits ? region * rsize + o : 0xffffffffu;
}

}  // namespace mpm
