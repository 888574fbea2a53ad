// Host-side model of the tiled exact canonical fold (canon_binade.h, k_canon.hip), for CPU
// verification of its exactness against the sequential definition. It takes the decisions of
// the device kernel one tile after another, for one chain at a time:
//  - an approximate running sum per chain picks the tile's binade hypothesis;
//  - each run of 16 blocks is summarised under it, and the runs join into the tile summary;
//  - the look-back joins the summaries of preceding tiles back to a tile whose exact end is
//    published (here: every tile at least `lag` tiles back), to a tile with another
//    hypothesis (whose exact end it then waits for), or to the item's start;
//  - the joined summary applied to that exact value gives the tile's exact end; where the
//    check fails, the exact value at the tile's start (the joined predecessors applied, or
//    the previous tile's end) is carried through the tile by the segmented evaluation of the
//    kernel's fallback wave: runs summarised under the current binade and applied up to the
//    first run whose check fails, that run folded sequentially, and so on.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../canon_binade.h"

using namespace vk;

namespace {

constexpr uint64_t kTileBlocks = 1024, kRun = 16;

// The fallback wave of k_canon.hip: chain j's blocks [b0, b1) from the exact value s.
double eval_chain(const double *vals, uint64_t b0, uint64_t b1, int j, double s, uint64_t *rounds) {
  const uint64_t nruns = (b1 - b0 + kRun - 1) / kRun;
  double cur = s;
  for (uint64_t r0 = 0; r0 < nruns;) {
    ++*rounds;
    const Binade h = binade_of(cur);
    // inclusive scan of the run summaries r0..r and the check of each prefix
    RunFn p = run_identity();
    uint64_t m = nruns;
    double last = cur, before_m = cur;
    for (uint64_t r = r0; r < nruns; ++r) {
      RunFn f = run_identity();
      if (h.ok) {
        const BinadeRef ref = binade_ref(h);
        for (uint64_t b = b0 + r * kRun; b < std::min(b1, b0 + (r + 1) * kRun); ++b)
          run_push(f, ref, vals[b * 16 + j]);
      }
      p = run_join(p, f);
      double val;
      if (!run_apply(p, h, cur, &val)) {
        m = r;
        break;
      }
      last = val;
      before_m = val;
    }
    if (m == nruns) return last;
    double t = m > r0 ? before_m : cur;
    for (uint64_t b = b0 + m * kRun; b < std::min(b1, b0 + (m + 1) * kRun); ++b) t = t + vals[b * 16 + j];
    cur = t;
    r0 = m + 1;
  }
  return cur;
}

}  // namespace

// stats: [0] tiles applied from summaries, [1] tiles evaluated by the fallback, [2] its rounds,
// [3] look-back stops at a tile with another hypothesis.
extern "C" int vkm_canonical_fold(const double *vals, uint64_t blocks, double *acc, int lag,
                                  uint64_t *stats) {
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  uint64_t fast = 0, slow = 0, rounds = 0, mismatch = 0;
  if (lag < 1) lag = 1;
  for (int j = 0; j < 16; ++j) {
    double approx = acc[j];
    std::vector<RunFn> fn(tiles);
    std::vector<Binade> hyp(tiles);
    std::vector<double> end(tiles);
    for (uint64_t t = 0; t < tiles; ++t) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      double asum = 0.0;
      for (uint64_t b = b0; b < b1; ++b) asum += vals[b * 16 + j];
      const Binade h = binade_of(approx);
      approx = approx + asum;
      RunFn ft = run_identity();
      if (h.ok) {
        const BinadeRef ref = binade_ref(h);
        for (uint64_t rb = b0; rb < b1; rb += kRun) {
          RunFn f = run_identity();
          for (uint64_t b = rb; b < std::min(b1, rb + kRun); ++b) run_push(f, ref, vals[b * 16 + j]);
          ft = run_join(ft, f);
        }
      }
      hyp[t] = h;
      fn[t] = ft;
      // look-back
      RunFn g = run_identity();
      double base = acc[j];
      int64_t base_tile = -1;
      for (int64_t k = int64_t(t) - 1; k >= 0; --k) {
        if (k <= int64_t(t) - lag) {
          base = end[k];
          base_tile = k;
          break;
        }
        if (!same_binade(hyp[k], h)) {
          ++mismatch;
          base = end[k];
          base_tile = k;
          break;
        }
        g = run_join(fn[k], g);
      }
      double ex;
      if (run_apply(run_join(g, ft), h, base, &ex)) {
        ++fast;
      } else {
        if (base_tile != int64_t(t) - 1) {
          double start;
          base = run_apply(g, h, base, &start) ? start : end[t - 1];
        }
        ex = eval_chain(vals, b0, b1, j, base, &rounds);
        ++slow;
      }
      end[t] = ex;
    }
    if (tiles) acc[j] = end[tiles - 1];
  }
  if (stats) {
    stats[0] = fast;
    stats[1] = slow;
    stats[2] = rounds;
    stats[3] = mismatch;
  }
  return 0;
}
