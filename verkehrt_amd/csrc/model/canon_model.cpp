// Host-side model of the tiled exact canonical fold (canon_binade.h, k_canon.hip), for CPU
// verification of its exactness against the sequential definition. It takes the decisions of
// the device kernel one tile after another, for one chain at a time:
//  - an approximate running sum per chain picks the tile's binade hypothesis;
//  - each run of 16 blocks is summarised under it, and the runs join into the tile summary;
//  - a tile whose summary fails at its own approximate start publishes it as opaque;
//  - the look-back joins the summaries of preceding tiles back to a tile whose exact end is
//    published (here: every tile at least `lag` tiles back), to a tile with another
//    hypothesis or an opaque summary (whose exact end it then waits for), or to the start;
//  - the joined summary applied to that exact value gives the tile's exact end; where the
//    check fails, the exact value at the tile's start (the joined predecessors applied, or
//    the previous tile's end) is carried through the tile by the kernel's fallback wave: each
//    run summarised under the binade of its own approximate start and applied in order, a run
//    whose check fails folded element by element.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../canon_binade.h"

using namespace vk;

namespace {

constexpr uint64_t kTileBlocks = 1024, kRun = 16;

// The fallback wave of k_canon.hip: chain j's blocks [b0, b1) from the exact value s. With at
// most two changes of the per-run binade hypotheses (each run's approximate start is s plus
// the approximate sums of the runs before it, summed in the device's scan order), rounds:
// summaries under the binade of the current value, applied up to the first run whose check
// fails, that run folded element by element. Otherwise each run is summarised under its own
// hypothesis and applied in order, a run whose check fails folded element by element.
double eval_chain(const double *vals, uint64_t b0, uint64_t b1, int j, double s, uint64_t *rounds) {
  const uint64_t nruns = (b1 - b0 + kRun - 1) / kRun;
  auto run_of = [&](uint64_t r, const Binade &h) {
    RunFn f = run_identity();
    if (h.ok) {
      const BinadeRef ref = binade_ref(h);
      for (uint64_t b = b0 + r * kRun; b < std::min(b1, b0 + (r + 1) * kRun); ++b) run_push(f, ref, vals[b * 16 + j]);
    }
    return f;
  };
  auto fold_run = [&](uint64_t r, double t) {
    for (uint64_t b = b0 + r * kRun; b < std::min(b1, b0 + (r + 1) * kRun); ++b) t = t + vals[b * 16 + j];
    return t;
  };
  std::vector<double> asum(64, 0.0);
  for (uint64_t r = 0; r < nruns; ++r)
    for (uint64_t b = b0 + r * kRun; b < std::min(b1, b0 + (r + 1) * kRun); ++b) asum[r] += vals[b * 16 + j];
  // the device's inclusive scan: Hillis-Steele over 64 lanes
  std::vector<double> inc(asum);
  for (int d = 1; d < 64; d <<= 1) {
    std::vector<double> nxt(inc);
    for (int l = d; l < 64; ++l) nxt[l] = inc[l] + inc[l - d];
    inc = nxt;
  }
  std::vector<Binade> ha(64);
  int changes = 0;
  for (uint64_t r = 0; r < 64; ++r) {
    ha[r] = binade_of(s + (inc[r] - asum[r]));
    if (r > 0 && r < nruns && !same_binade(ha[r], ha[r - 1])) ++changes;
  }
  double cur = s;
  if (changes <= 2) {
    for (uint64_t r0 = 0; r0 < nruns;) {
      const Binade h = binade_of(cur);
      RunFn p = run_identity();
      uint64_t m = nruns;
      double last = cur;
      for (uint64_t r = r0; r < nruns; ++r) {
        p = run_join(p, run_of(r, h));
        double val;
        if (!run_apply(p, h, cur, &val)) {
          m = r;
          break;
        }
        last = val;
      }
      if (m == nruns) return last;
      ++*rounds;
      cur = fold_run(m, last);
      r0 = m + 1;
    }
    return cur;
  }
  for (uint64_t r = 0; r < nruns; ++r) {
    double out;
    if (run_apply(run_of(r, ha[r]), ha[r], cur, &out)) {
      cur = out;
    } else {
      ++*rounds;
      cur = fold_run(r, cur);
    }
  }
  return cur;
}

}  // namespace

// stats: [0] tiles applied from summaries, [1] tiles evaluated by the fallback, [2] runs it
// folded element by element, [3] look-back stops at a tile with another hypothesis or opaque.
extern "C" int vkm_canonical_fold(const double *vals, uint64_t blocks, double *acc, int lag,
                                  uint64_t *stats) {
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  uint64_t fast = 0, slow = 0, rounds = 0, mismatch = 0;
  if (lag < 1) lag = 1;
  for (int j = 0; j < 16; ++j) {
    double approx = acc[j];
    std::vector<RunFn> fn(tiles);
    std::vector<Binade> hyp(tiles);
    std::vector<char> opaque(tiles);
    std::vector<double> end(tiles);
    for (uint64_t t = 0; t < tiles; ++t) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      double asum = 0.0;
      for (uint64_t b = b0; b < b1; ++b) asum += vals[b * 16 + j];
      const Binade h = binade_of(approx);
      const double approx_start = approx;
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
      double probe;
      opaque[t] = !run_apply(ft, h, approx_start, &probe);
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
        if (!same_binade(hyp[k], h) || opaque[k]) {
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
