// Tile traversal orders from a GemmNode control tree, for the persistent and
// data-parallel GEMM kernels (host and device code, so a host build checks
// it: tests/native/tile_order_check.cpp).
//
// A control tree's loops over m and n (PartM / PartN, src/composables/part.rs:
// 7-79) decide which operand stays resident in which cache level: the
// reference's variant trees differ only in their order and block sizes
// (src/exper/l3_shapes.rs:69-100, src/exper/l4.rs:71-78). On the GPU the
// C tiles of one launch are the iteration space; the order in which the
// workgroups take them decides which A row-panels and B column-panels the
// tiles resident at one time share in an XCD's L2. TileOrder restates the
// tree's M / N loops, outermost first, in whole tiles: level i splits the
// current region of the tile grid along its dimension into blocks of size[i]
// tiles (the last may be smaller), the blocks taken in order; what the last
// level leaves is walked with m fastest. Loops over k do not reorder tiles:
// each tile keeps its whole k loop in one workgroup.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define MOMMS_TO_HD __host__ __device__
#else
#define MOMMS_TO_HD
#endif

namespace momms {

constexpr int kTileOrderLevels = 8;

struct TileOrder {
  int32_t nlev;
  int32_t dim[kTileOrderLevels];   // 0: split rows (M), 1: split columns (N)
  int64_t size[kTileOrderLevels];  // block size in tiles, >= 1
};

// (tm, tn) of logical tile t (0 <= t < TM * TN): a bijection onto the grid.
MOMMS_TO_HD inline void tile_order_decode(int64_t t, int64_t TM, int64_t TN, const TileOrder& o,
                                          int64_t& tm, int64_t& tn) {
  int64_t m0 = 0, m1 = TM, n0 = 0, n1 = TN;
  for (int i = 0; i < o.nlev && i < kTileOrderLevels; ++i) {
    const int64_t rows = m1 - m0, cols = n1 - n0;
    if (o.dim[i] == 0) {
      const int64_t s = o.size[i] < rows ? o.size[i] : rows;
      const int64_t b = t / (s * cols);  // every block before b is whole
      t -= b * s * cols;
      m0 += b * s;
      m1 = (m0 + s) < m1 ? m0 + s : m1;
    } else {
      const int64_t s = o.size[i] < cols ? o.size[i] : cols;
      const int64_t b = t / (s * rows);
      t -= b * s * rows;
      n0 += b * s;
      n1 = (n0 + s) < n1 ? n0 + s : n1;
    }
  }
  const int64_t rows = m1 - m0;
  tm = m0 + t % rows;
  tn = n0 + t / rows;
}

// TileOrder of a hierarchy description (kind 0 = M, 1 = N, 2 = K; block
// size in elements; innermost step first, as GemmNode::hierarchy_description
// returns it) for tiles of bm x bn: its M and N steps, outermost first, each
// rounded down to whole tiles (at least one). Returns false for more than
// kTileOrderLevels such steps or a non-positive block size.
inline bool tile_order_from_steps(const int32_t* kinds, const int64_t* bszs, int64_t nsteps,
                                  int64_t bm, int64_t bn, TileOrder& o) {
  o.nlev = 0;
  for (int64_t i = nsteps - 1; i >= 0; --i) {
    if (bszs[i] <= 0) return false;
    if (kinds[i] == 2) continue;
    if (kinds[i] != 0 && kinds[i] != 1) return false;
    if (o.nlev == kTileOrderLevels) return false;
    const int64_t tile = kinds[i] == 0 ? bm : bn;
    const int64_t s = bszs[i] / tile;
    o.dim[o.nlev] = kinds[i];
    o.size[o.nlev] = s < 1 ? 1 : s;
    ++o.nlev;
  }
  return true;
}

}  // namespace momms
