// Host-side dealing of independent tiles to the XCDs (plain C++, no device
// code): used by the batched min-plus products.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace mi355 {

// Workgroups are dealt to the 8 XCDs round-robin by workgroup id, so the
// entry at workgroup id j * 8 + x runs on XCD x.
constexpr int kXcds = 8;

// Tiles [first, first + count) of product `item`; each tile costs `k`.
struct TileRun {
  int item, first, count, k;
};
// One workgroup's work: tile `tile` of product `item`; item -1 is padding.
struct TileSlot {
  int item, tile;
};

// The runs are cut into pieces of at most 1/64 of all tiles, the pieces go
// longest-first (by k) to the XCD with the least work so far, and each XCD
// runs its pieces in that order: the XCDs finish together, long tiles start
// first, and a product's tiles share an XCD (and its L2).  out[j * kXcds + x]
// is XCD x's j-th tile; shorter XCD lists are padded with item -1.
inline void xcd_deal(const std::vector<TileRun>& runs, uint64_t total_tiles, std::vector<TileSlot>& out) {
  const int piece = (int)std::max<uint64_t>(1, (total_tiles + 63) / 64);
  std::vector<TileRun> pieces;
  for (const TileRun& r : runs)
    for (int f = 0; f < r.count; f += piece)
      pieces.push_back(TileRun{r.item, r.first + f, std::min(piece, r.count - f), r.k});
  std::stable_sort(pieces.begin(), pieces.end(), [](const TileRun& x, const TileRun& y) { return x.k > y.k; });
  std::vector<std::vector<TileSlot>> bins(kXcds);
  uint64_t work[kXcds] = {};
  for (const TileRun& r : pieces) {
    int b = 0;
    for (int x = 1; x < kXcds; ++x)
      if (work[x] < work[b]) b = x;
    work[b] += (uint64_t)r.count * (uint64_t)std::max(r.k, 1);
    for (int t = 0; t < r.count; ++t) bins[b].push_back(TileSlot{r.item, r.first + t});
  }
  size_t len = 0;
  for (auto& bn : bins) len = std::max(len, bn.size());
  out.assign(len * kXcds, TileSlot{-1, 0});
  for (int x = 0; x < kXcds; ++x)
    for (size_t j = 0; j < bins[x].size(); ++j) out[j * kXcds + x] = bins[x][j];
}

}  // namespace mi355
