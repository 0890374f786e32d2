// CPU check of the item-persistent kernels' work loop (shaderintel_amd/csrc/item_sched.hpp) over a
// full scene's item list, with bounds asserts. Input file (written by tests/test_item_loop.py):
//   u32 n, u32 items, i32 X, Y, Z, u32 key_bits, then u32 wl_start[items + 1], u32 wl_block[items]
//   (Morton block code, scene bits cleared), then float x[3 n] (positions, sorted order).
// For several grid sizes the workgroups take their items (first_item_of, then next_item_of on shared
// counters) in a random interleaving; every item must be taken exactly once. Each item's P2G is then
// followed through its window: every in-window stencil's LDS tile index, the tile box (pnodes <=
// kTileNodes, 8-bit extents), the tile store's indices and the arena offsets; and the same LDS
// bounds in the narrower drift-2 window (a 10^3-node tile).
// A second pass replays the loop as if a workgroup's item slot were overwritten with its next item
// before all of its waves had read it: waves 1-3 then work on the next item, so the box spans two
// windows. It counts the boxes that break the window bound and the arena offsets the unguarded check
// (o <= rsize - pnodes, without pnodes <= kTileNodes) would have given past the arena's end.
// The tile storage order (tile_index, tile_chunk) is checked over boxes of every extent up to 12.
// Prints one summary line; exit status 0 = the loop's bounds hold.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "item_sched.hpp"

using namespace mpm;

#define CHECK(c)                                                                  \
  do {                                                                            \
    if (!(c)) {                                                                   \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      std::exit(1);                                                               \
    }                                                                             \
  } while (0)

struct Box {
  int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  bool any() const { return lo[0] != INT32_MAX; }
  void add(const int b[3]) {
    for (int a = 0; a < 3; a++) {
      lo[a] = b[a] < lo[a] ? b[a] : lo[a];
      hi[a] = b[a] > hi[a] ? b[a] : hi[a];
    }
  }
  long extent(int a) const { return any() ? (long)hi[a] - lo[a] + 3 : 0; }
  long nodes() const { return extent(0) * extent(1) * extent(2); }
};

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  uint32_t n, items, key_bits;
  int32_t dims[3];
  CHECK(std::fread(&n, 4, 1, f) == 1 && std::fread(&items, 4, 1, f) == 1 && std::fread(dims, 4, 3, f) == 3 &&
        std::fread(&key_bits, 4, 1, f) == 1);
  std::vector<uint32_t> start(items + 1), block(items);
  std::vector<float> x(3 * (size_t)n);
  CHECK(std::fread(start.data(), 4, items + 1, f) == items + 1);
  CHECK(std::fread(block.data(), 4, items, f) == items);
  CHECK(std::fread(x.data(), 4, 3 * (size_t)n, f) == 3 * (size_t)n);
  std::fclose(f);
  CHECK(start[0] == 0 && start[items] == n);
  for (uint32_t k = 0; k < items; k++) CHECK(start[k] < start[k + 1]);

  // the pool's arena for a capacity of n particles (kArenaRegions regions)
  const uint64_t arena =
      std::min<uint64_t>(0x7fffffffu, std::max<uint64_t>((uint64_t)4096 * kTileNodes, 4ull * n)) / kArenaRegions *
      kArenaRegions;
  const uint32_t rsize = (uint32_t)(arena / kArenaRegions);
  CHECK(rsize > (uint32_t)kTileNodes);
  constexpr int S = kTileSpan;

  auto window = [&](uint32_t k, int w[3]) {
    const uint32_t mb = block[k];
    w[0] = 4 * (int)compact3(mb) - kDrift;
    w[1] = 4 * (int)compact3(mb >> 1) - kDrift;
    w[2] = 4 * (int)compact3(mb >> 2) - kDrift;
  };
  auto base = [&](uint32_t j, int b[3]) {
    for (int a = 0; a < 3; a++) b[a] = (int)floorf(x[3 * (size_t)j + a] - 0.5f);
  };

  // the tile box of item k (its in-window particles) with every LDS index checked
  uint64_t fallbacks = 0, inwin_total = 0;
  std::vector<Box> boxes(items);
  for (uint32_t k = 0; k < items; k++) {
    int w[3];
    window(k, w);
    Box bx;
    for (uint32_t j = start[k]; j < start[k + 1]; j++) {
      int b[3];
      base(j, b);
      if (!in_window(b[0], b[1], b[2], w[0], w[1], w[2])) {
        fallbacks++;
        continue;
      }
      inwin_total++;
      bx.add(b);
      const long tb = ((long)(b[2] - w[2]) * S + (b[1] - w[1])) * S + (b[0] - w[0]);
      CHECK(tb >= 0 && tb + (2L * S + 2) * S + 2 < kTileNodes);  // the 27 nodes of its stencil
    }
    if (bx.any()) {
      CHECK(bx.nodes() <= kTileNodes);
      for (int a = 0; a < 3; a++) CHECK(bx.extent(a) <= 255 && bx.lo[a] - w[a] >= 0 && bx.hi[a] + 2 - w[a] < S);
    }
    boxes[k] = bx;
  }
  // the narrower window (drift 2, a 10^3-node tile): the same bounds for its in-window particles
  uint64_t fallbacks2 = 0;
  {
    constexpr int D = 2, S2 = Window<D>::kSpan, N2 = Window<D>::kNodes;
    for (uint32_t k = 0; k < items; k++) {
      const uint32_t mb = block[k];
      const int w[3] = {4 * (int)compact3(mb) - D, 4 * (int)compact3(mb >> 1) - D, 4 * (int)compact3(mb >> 2) - D};
      Box bx;
      for (uint32_t j = start[k]; j < start[k + 1]; j++) {
        int b[3];
        base(j, b);
        if (!in_window<D>(b[0], b[1], b[2], w[0], w[1], w[2])) {
          fallbacks2++;
          continue;
        }
        bx.add(b);
        const long tb = ((long)(b[2] - w[2]) * S2 + (b[1] - w[1])) * S2 + (b[0] - w[0]);
        CHECK(tb >= 0 && tb + (2L * S2 + 2) * S2 + 2 < N2);
      }
      if (bx.any()) {
        CHECK(bx.nodes() <= N2);
        for (int a = 0; a < 3; a++) CHECK(bx.lo[a] - w[a] >= 0 && bx.hi[a] + 2 - w[a] < S2);
      }
    }
  }

  // the tile storage order: tile_index is a bijection of every box's nodes onto [0, nodes), and each
  // node block's chunk (tile_chunk, as the grid update reads it) holds exactly that block's nodes, in
  // one run at the chunk's offset
  for (int ox = 0; ox < 8; ox++)
    for (int ex = 1; ex <= 12; ex++)
      for (int ey = 1; ey <= 12; ey += 5)
        for (int ez = 1; ez <= 12; ez += 3) {
          const int oy = (ox * 3 + 1) % 8, oz = (ox * 5 + 2) % 8, nn = ex * ey * ez;
          std::vector<uint8_t> seen((size_t)nn, 0);
          for (int lz = 0; lz < ez; lz++)
            for (int ly = 0; ly < ey; ly++)
              for (int lx = 0; lx < ex; lx++) {
                const uint32_t t = tile_index(ox, oy, oz, ex, ey, ez, lx, ly, lz);
                CHECK(t < (uint32_t)nn && !seen[t]);
                seen[t] = 1;
                const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
                uint32_t off;
                int wx, wy, wz, l0, l1, l2;
                CHECK(tile_chunk(ox, oy, oz, ex, ey, ez, gx >> 2, gy >> 2, gz >> 2, off, wx, wy, wz, l0, l1, l2));
                CHECK(wx * wy * wz <= 64 && gx - l0 < wx && gy - l1 < wy && gz - l2 < wz);
                CHECK(off + (uint32_t)((gx - l0) + wx * ((gy - l1) + wy * (gz - l2))) == t);
              }
          for (int bz = 0; bz < 8; bz++)
            for (int by = 0; by < 8; by++)
              for (int bx = 0; bx < 8; bx++) {
                uint32_t off;
                int wx, wy, wz, l0, l1, l2;
                const bool any = tile_chunk(ox, oy, oz, ex, ey, ez, bx, by, bz, off, wx, wy, wz, l0, l1, l2);
                const bool overlaps = 4 * bx < ox + ex && 4 * bx + 4 > ox && 4 * by < oy + ey && 4 * by + 4 > oy &&
                                      4 * bz < oz + ez && 4 * bz + 4 > oz;
                CHECK(any == overlaps);
                if (any) CHECK(off + (uint32_t)(wx * wy * wz) <= (uint32_t)nn);
              }
        }

  std::mt19937_64 rng(12345);
  const uint32_t grids[] = {8, 64, 256, 2048, 2056, 4096};
  uint64_t mixed_over = 0, mixed_ext = 0, mixed_oob = 0, mixed_pairs = 0;
  long worst_end = 0;
  for (uint32_t grid : grids) {
    std::vector<uint32_t> ctr(kWorkSplits, 0), taken(items, 0), region_ctr(kArenaRegions, 0);
    std::vector<uint32_t> range(grid), spent(grid, 0), cur(grid);
    std::vector<uint8_t> full(grid, 0);
    std::vector<uint32_t> live;
    for (uint32_t b = 0; b < grid; b++) {
      range[b] = b % kWorkSplits;
      cur[b] = first_item_of(items, grid, b);
      if (cur[b] == 0xffffffffu)  // no static item: its first comes from the counters
        cur[b] = next_item_of([&](uint32_t r) { return ctr[r]++; }, items, grid, range[b], spent[b]);
      if (cur[b] != 0xffffffffu) live.push_back(b);
    }
    std::vector<uint32_t> region_ctr_old(kArenaRegions, 0);
    while (!live.empty()) {
      const size_t pick = (size_t)(rng() % live.size());
      const uint32_t b = live[pick];
      const uint32_t k = cur[b];
      CHECK(k < items);
      taken[k]++;
      // the committed store: arena space for the box, then the workgroup's next item
      const Box& bx = boxes[k];
      const uint32_t pn = (uint32_t)bx.nodes(), region = b % kArenaRegions;
      if (pn > 0 && !full[b]) {
        const uint32_t o = region_ctr[region];
        region_ctr[region] += pn;
        const uint32_t off = arena_offset(region, rsize, o, pn);
        if (off == 0xffffffffu) full[b] = 1;
        else CHECK((uint64_t)off + pn <= arena && off >= region * rsize && off + pn <= (region + 1) * rsize);
      }
      const uint32_t nk = next_item_of([&](uint32_t r) { return ctr[r]++; }, items, grid, range[b], spent[b]);
      // the overwritten-slot replay: waves 1-3 of this item read the slot after it held nk
      if (nk != 0xffffffffu && grid == 2048) {
        mixed_pairs++;
        int w0[3], w1[3];
        window(k, w0);
        window(nk, w1);
        Box mix;
        auto add_range = [&](uint32_t kk, const int* w, uint32_t from, uint32_t to) {
          for (uint32_t j = start[kk] + from; j < start[kk + 1] && j < start[kk] + to; j++) {
            int bb[3];
            base(j, bb);
            if (in_window(bb[0], bb[1], bb[2], w[0], w[1], w[2])) mix.add(bb);
          }
        };
        // a workgroup of 256 threads: wave 0 on lanes 0-63 of each pass of k, waves 1-3 on nk's
        for (uint32_t p = 0; p < 16; p++) {
          add_range(k, w0, 256 * p, 256 * p + 64);
          add_range(nk, w1, 256 * p + 64, 256 * p + 256);
        }
        if (mix.any()) {
          const long pn2 = mix.nodes();
          if (pn2 > kTileNodes) mixed_over++;
          for (int a = 0; a < 3; a++)
            if (mix.extent(a) > 255) {
              mixed_ext++;
              break;
            }
          // the unguarded arena check, with pnodes as the 32-bit product the kernel forms
          const uint32_t pn32 = (uint32_t)pn2;
          const uint32_t o = region_ctr_old[region];
          region_ctr_old[region] += pn32;
          if (o <= rsize - pn32) {  // wraps when pn32 > rsize
            const long end = (long)region * rsize + o + pn32;
            if (end > (long)arena) {
              mixed_oob++;
              worst_end = std::max(worst_end, end - (long)arena);
            }
          }
        }
      }
      if (nk == 0xffffffffu) {
        live[pick] = live.back();
        live.pop_back();
      } else {
        cur[b] = nk;
      }
    }
    for (uint32_t k = 0; k < items; k++) CHECK(taken[k] == 1);
  }
  std::printf(
      "ok items=%u particles=%u in_window=%llu fallback=%llu (drift 2: %llu) arena_nodes=%llu | overwritten-slot replay: %llu item "
      "pairs, %llu boxes over %d nodes, %llu with an extent over 255, %llu arena ends past the arena (worst by %ld "
      "nodes)\n",
      items, n, (unsigned long long)inwin_total, (unsigned long long)fallbacks, (unsigned long long)fallbacks2,
      (unsigned long long)arena,
      (unsigned long long)mixed_pairs, (unsigned long long)mixed_over, kTileNodes, (unsigned long long)mixed_ext,
      (unsigned long long)mixed_oob, worst_end);
  return 0;
}
