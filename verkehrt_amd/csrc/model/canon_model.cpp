// Host-side model of the tiled exact canonical fold (canon_binade.h), for CPU verification of
// its exactness against the sequential definition. It runs the decisions of the device
// kernel one tile after another: an approximate start per chain picks the binade
// hypothesis, each run of 16 blocks is summarised, runs join in order into a tile summary,
// tiles with equal hypotheses join across up to `window` predecessors before the summary is
// applied to an exact start, and a failed check folds the tile sequentially.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../canon_binade.h"

using namespace vk;

extern "C" int vkm_canonical_fold(const double *vals, uint64_t blocks, double *acc, int window,
                                  uint64_t *stats) {
  const uint64_t kTileBlocks = 256, kRun = 16;
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  uint64_t fast = 0, slow = 0;
  for (int j = 0; j < 16; ++j) {
    double exact = acc[j];
    double approx = acc[j];
    std::vector<RunFn> pend;  // summaries of tiles since the last exact value
    std::vector<uint64_t> pend_tiles;
    Binade pend_h{0, 0, 0, 0};
    auto fold_tile_seq = [&](uint64_t t, double s) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      for (uint64_t b = b0; b < b1; ++b) s = s + vals[b * 16 + j];
      return s;
    };
    auto flush = [&]() {
      // apply the pending joined summary; on failure fold the pending tiles one by one
      if (pend.empty()) return;
      RunFn g = run_identity();
      for (auto &f : pend) g = run_join(g, f);
      double out;
      if (run_apply(g, pend_h, exact, &out)) {
        exact = out;
        fast += pend.size();
      } else {
        for (size_t i = 0; i < pend.size(); ++i) {
          double o;
          if (run_apply(pend[i], pend_h, exact, &o)) {
            exact = o;
            ++fast;
          } else {
            exact = fold_tile_seq(pend_tiles[i], exact);
            ++slow;
          }
        }
      }
      pend.clear();
      pend_tiles.clear();
    };
    for (uint64_t t = 0; t < tiles; ++t) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      double asum = 0.0;
      for (uint64_t b = b0; b < b1; ++b) asum += vals[b * 16 + j];
      const Binade h = binade_of(approx);
      approx = approx + asum;
      RunFn tile = run_identity();
      if (h.ok) {
        const BinadeRef r = binade_ref(h);
        for (uint64_t rb = b0; rb < b1; rb += kRun) {
          RunFn f = run_identity();
          for (uint64_t b = rb; b < std::min(b1, rb + kRun); ++b) run_push(f, r, vals[b * 16 + j]);
          tile = run_join(tile, f);
        }
      }
      if (!h.ok) {
        flush();
        exact = fold_tile_seq(t, exact);
        ++slow;
        continue;
      }
      if (!pend.empty() && !same_binade(pend_h, h)) flush();
      pend_h = h;
      pend.push_back(tile);
      pend_tiles.push_back(t);
      if (static_cast<int>(pend.size()) >= window) flush();
    }
    flush();
    acc[j] = exact;
  }
  if (stats) {
    stats[0] = fast;
    stats[1] = slow;
  }
  return 0;
}
