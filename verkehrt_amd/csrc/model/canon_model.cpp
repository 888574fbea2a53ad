// Host-side models of the tiled exact canonical fold (canon_binade.h, k_canon.hip), for CPU
// verification of their exactness against the sequential definition.
//
// vkm_canonical_fold follows the single-pass kernel, one tile after another, one chain at a time:
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

// The summary of chain j's blocks [b, min(b + 16, b1)) through the device's run_of.
RunFn run_at(const double *vals, uint64_t b, uint64_t b1, int j, const BinadeRef &ref) {
  double v[kRun];
  const int cnt = static_cast<int>(std::min<uint64_t>(kRun, b1 - b));
  for (int k = 0; k < static_cast<int>(kRun); ++k) v[k] = k < cnt ? vals[(b + k) * 16 + j] : 0.0;
  return run_of(v, cnt, ref);
}

// The fallback wave of k_canon.hip: chain j's blocks [b0, b1) from the exact value s. With at
// most two changes of the per-run binade hypotheses (each run's approximate start is s plus
// the approximate sums of the runs before it, summed in the device's scan order), rounds:
// summaries under the binade of the current value, applied up to the first run whose check
// fails, that run folded element by element. Otherwise the elements are folded one by one.
double eval_chain(const double *vals, uint64_t b0, uint64_t b1, int j, double s, uint64_t *rounds) {
  const uint64_t nruns = (b1 - b0 + kRun - 1) / kRun;
  auto run_of = [&](uint64_t r, const Binade &h) {
    RunFn f = run_identity();
    if (h.ok) f = run_at(vals, b0 + r * kRun, b1, j, binade_ref(h));
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
  for (uint64_t r = 0; r < nruns; ++r) cur = fold_run(r, cur);
  return cur;
}

// The wide table (canon_binade.h) of chain j over blocks [b0, b1): 16 residues simulated from
// B = approx less its residue modulo 16 u0, then the entries the greatest rise M needs.
void wide_table(const double *vals, uint64_t b0, uint64_t b1, int j, double approx, WideHead *hd, WideEnt *ents) {
  const Binade h = binade_of(approx);
  *hd = WideHead{0.0, 0, 0};
  if (!h.ok) return;
  const double b = wide_base(approx, h), u0 = pow2(h.e - 52);
  int m = -2000;
  for (int r = 0; r < kWideRes; ++r) {
    // as the device: 16 elements at a time (zeros past b1)
    WideAcc a = wide_start(b + r * u0);
    for (uint64_t k0 = b0; k0 < b1; k0 += 16) {
      double v[16];
      for (uint64_t m = 0; m < 16; ++m) v[m] = k0 + m < b1 ? vals[(k0 + m) * 16 + j] : 0.0;
      wide_push_n(a, v);
    }
    ents[r] = wide_end(a);
    if (ents[r].lo <= ents[r].hi) m = std::max(m, a.emax - h.e);
  }
  *hd = WideHead{b, h.e, wide_nres(m)};
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
          ft = run_join(ft, run_at(vals, rb, b1, j, ref));
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

// The streamed form of k_canon.hip for long inputs, one chain at a time (its predicted first
// hypotheses only decide which tiles it summarises twice, so they are left out): approximate
// tile sums (summed in the device's order), their exclusive prefix from the start state (in 1024 parts,
// as the prefix kernel does), tile summaries under the binade of that prefix, usable
// where they apply at it; groups of `group_tiles` tiles are clean when every tile is usable
// under one hypothesis; the walk takes 64 groups at a time, joins consecutive clean groups of
// one hypothesis and applies the joined summary to the exact value (group by group where that
// fails), and splits other groups into segments of usable same-hypothesis tiles, each applied
// once or tile by tile, a tile that fails evaluated through its wide table (`records` bit 1),
// then its run records (bit 0) or by the fallback wave.
// stats: [0] joined clean spans applied, [1] spans that failed, [2] segments applied, [3] tiles
// whose summary did not apply, [4] run segments applied, [5] runs folded element by element,
// [6] tiles applied from their wide table, [7] wide tables that did not apply.
extern "C" int vkm_canonical_fold_streamed(const double *vals, uint64_t blocks, double *acc,
                                           int group_tiles, int records, uint64_t *stats) {
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  if (group_tiles < 1) group_tiles = 1;
  const uint64_t gt = static_cast<uint64_t>(group_tiles);
  const uint64_t groups = (tiles + gt - 1) / gt;
  uint64_t st[8] = {0, 0, 0, 0, 0, 0, 0, 0}, rounds = 0;
  for (int j = 0; j < 16; ++j) {
    // 1. approximate tile sums as the first summary pass forms them: each run's elements in
    //    order (from 0.0, missing elements as 0.0), runs in fours, fours in fours, and so on
    std::vector<double> asum(tiles);
    for (uint64_t t = 0; t < tiles; ++t) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      double part[64];
      for (uint64_t r = 0; r < 64; ++r) {
        double a = 0.0;
        for (uint64_t k = 0; k < kRun; ++k) {
          const uint64_t b = b0 + r * kRun + k;
          a = a + (b < b1 ? vals[b * 16 + j] : 0.0);
        }
        part[r] = a;
      }
      double w[16];
      for (int i = 0; i < 16; ++i) w[i] = (part[4 * i] + part[4 * i + 1]) + (part[4 * i + 2] + part[4 * i + 3]);
      double q[4];
      for (int i = 0; i < 4; ++i) q[i] = (w[4 * i] + w[4 * i + 1]) + (w[4 * i + 2] + w[4 * i + 3]);
      asum[t] = (q[0] + q[1]) + (q[2] + q[3]);
    }
    // 2. exclusive prefix in 1024 parts, the part sums scanned as the device does: Hillis-Steele
    //    within each wave of 64 parts, then over the 16 wave totals, which are added after
    std::vector<double> pre(tiles);
    const uint64_t parts = 1024, per = (tiles + parts - 1) / parts;
    std::vector<double> psum(parts, 0.0);
    for (uint64_t p = 0; p < parts; ++p)
      for (uint64_t k = std::min(tiles, p * per); k < std::min(tiles, (p + 1) * per); ++k) psum[p] = psum[p] + asum[k];
    auto hs = [](double *v, int n) {
      for (int d = 1; d < n; d <<= 1) {
        std::vector<double> nxt(v, v + n);
        for (int i = d; i < n; ++i) nxt[i] = v[i - d] + v[i];
        std::copy(nxt.begin(), nxt.end(), v);
      }
    };
    double tot[16];
    for (int w = 0; w < 16; ++w) {
      hs(&psum[64 * w], 64);
      tot[w] = psum[64 * w + 63];
    }
    hs(tot, 16);
    for (uint64_t p = 64; p < parts; ++p) psum[p] = tot[p / 64 - 1] + psum[p];
    for (uint64_t p = 0; p < parts; ++p) {
      double run = acc[j] + (p > 0 ? psum[p - 1] : 0.0);
      for (uint64_t k = std::min(tiles, p * per); k < std::min(tiles, (p + 1) * per); ++k) {
        pre[k] = run;
        run = run + asum[k];
      }
    }
    // 3. tile summaries
    std::vector<RunFn> fn(tiles);
    std::vector<Binade> hyp(tiles);
    std::vector<char> usable(tiles);
    for (uint64_t t = 0; t < tiles; ++t) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      const Binade h = binade_of(pre[t]);
      RunFn ft = run_identity();
      if (h.ok) {
        const BinadeRef ref = binade_ref(h);
        for (uint64_t rb = b0; rb < b1; rb += kRun) {
          ft = run_join(ft, run_at(vals, rb, b1, j, ref));
        }
      }
      double probe;
      hyp[t] = h;
      fn[t] = ft;
      usable[t] = run_apply(ft, h, pre[t], &probe);
    }
    // a tile whose summary does not apply: by its run records (each run summarised under the
    // binade of its own approximate start, the tile's plus the run sums before it; segments of
    // usable runs under one hypothesis applied at once, then run by run, then element by
    // element), or, without records, by the fallback wave
    auto by_runs = [&](uint64_t t, double cur) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      const uint64_t nruns = (b1 - b0 + kRun - 1) / kRun;
      std::vector<RunFn> rf(nruns);
      std::vector<Binade> rh(nruns);
      std::vector<char> ru(nruns);
      double rs = pre[t];
      for (uint64_t r = 0; r < nruns; ++r) {
        const uint64_t rb = b0 + r * kRun;
        rh[r] = binade_of(rs);
        rf[r] = rh[r].ok ? run_at(vals, rb, b1, j, binade_ref(rh[r])) : run_identity();
        double probe;
        ru[r] = run_apply(rf[r], rh[r], rs, &probe);
        for (uint64_t b = rb; b < std::min(b1, rb + kRun); ++b) rs = rs + vals[b * 16 + j];
      }
      auto fold = [&](uint64_t r, double c) {
        ++st[5];
        for (uint64_t b = b0 + r * kRun; b < std::min(b1, b0 + (r + 1) * kRun); ++b) c = c + vals[b * 16 + j];
        return c;
      };
      for (uint64_t s0 = 0; s0 < nruns;) {
        uint64_t e = s0 + 1;
        if (ru[s0])
          while (e < nruns && ru[e] && same_binade(rh[e], rh[s0])) ++e;
        RunFn f = run_identity();
        for (uint64_t r = s0; r < e; ++r) f = run_join(f, rf[r]);
        double out;
        if (ru[s0] && run_apply(f, rh[s0], cur, &out)) {
          ++st[4];
          cur = out;
        } else {
          for (uint64_t r = s0; r < e; ++r) cur = ru[r] && run_apply(rf[r], rh[r], cur, &out) ? out : fold(r, cur);
        }
        s0 = e;
      }
      return cur;
    };
    // the tile's wide table (canon_binade.h) built from its approximate start
    auto by_wide = [&](uint64_t t, double cur, double *out) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      WideHead hd;
      WideEnt ents[kWideRes];
      wide_table(vals, b0, b1, j, pre[t], &hd, ents);
      return wide_apply(hd, ents, cur, out);
    };
    auto one = [&](uint64_t t, double cur) {
      double out;
      if (usable[t] && run_apply(fn[t], hyp[t], cur, &out)) return out;
      ++st[3];
      if (records & 2) {
        if (by_wide(t, cur, &out)) {
          ++st[6];
          return out;
        }
        ++st[7];
      }
      if (records & 1) return by_runs(t, cur);
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      return eval_chain(vals, b0, b1, j, cur, &rounds);
    };
    auto walk_group = [&](uint64_t g, double cur, bool tiles_only) {
      const uint64_t t0 = g * gt, t1 = std::min(tiles, t0 + gt);
      for (uint64_t s = t0; s < t1;) {
        uint64_t e = s + 1;
        if (!tiles_only && usable[s])
          while (e < t1 && usable[e] && same_binade(hyp[e], hyp[s])) ++e;
        RunFn f = run_identity();
        for (uint64_t t = s; t < e; ++t) f = run_join(f, fn[t]);
        double out;
        if (usable[s] && run_apply(f, hyp[s], cur, &out)) {
          ++st[2];
          cur = out;
        } else {
          for (uint64_t t = s; t < e; ++t) cur = one(t, cur);
        }
        s = e;
      }
      return cur;
    };
    // 4-5. clean groups and the walk: 64 groups at a time, segments of consecutive clean groups
    //      of one hypothesis joined and applied, a failing segment group by group, a group that
    //      is not clean or fails through walk_group
    double cur = acc[j];
    std::vector<RunFn> gfn(groups);
    std::vector<char> gclean(groups);
    for (uint64_t g = 0; g < groups; ++g) {
      const uint64_t t0 = g * gt, t1 = std::min(tiles, t0 + gt);
      bool clean = true;
      RunFn gf = run_identity();
      for (uint64_t t = t0; t < t1; ++t) {
        clean = clean && usable[t] && same_binade(hyp[t], hyp[t0]);
        gf = run_join(gf, fn[t]);
      }
      gfn[g] = gf;
      gclean[g] = clean;
    }
    for (uint64_t b0 = 0; b0 < groups; b0 += 64) {
      const uint64_t b1 = std::min(groups, b0 + 64);
      for (uint64_t s0 = b0; s0 < b1;) {
        uint64_t e = s0 + 1;
        if (gclean[s0])
          while (e < b1 && gclean[e] && same_binade(hyp[e * gt], hyp[s0 * gt])) ++e;
        RunFn f = run_identity();
        for (uint64_t g = s0; g < e; ++g) f = run_join(f, gfn[g]);
        double out;
        if (gclean[s0] && run_apply(f, hyp[s0 * gt], cur, &out)) {
          ++st[0];
          cur = out;
        } else {
          if (gclean[s0]) ++st[1];
          for (uint64_t g = s0; g < e; ++g)
            cur = gclean[g] && run_apply(gfn[g], hyp[g * gt], cur, &out) ? out : walk_group(g, cur, false);
        }
        s0 = e;
      }
    }
    acc[j] = cur;
  }
  if (stats)
    for (int k = 0; k < 8; ++k) stats[k] = st[k];
  return 0;
}

// The wide table of chain j over blocks [b0, b1) from `approx`, as k_canon_wide builds it.
extern "C" int vkm_wide_table_apply(const double *chain, uint64_t n, double approx, const double *starts,
                                    uint64_t nstarts, double *out, uint8_t *applied, int32_t *nres) {
  WideHead hd;
  WideEnt ents[kWideRes];
  std::vector<double> vals(n * 16, 0.0);
  for (uint64_t i = 0; i < n; ++i) vals[i * 16] = chain[i];
  wide_table(vals.data(), 0, n, 0, approx, &hd, ents);
  *nres = hd.nres;
  for (uint64_t i = 0; i < nstarts; ++i) {
    double o = 0.0;
    applied[i] = wide_apply(hd, ents, starts[i], &o) ? 1 : 0;
    out[i] = o;
  }
  return 0;
}

// ------------------------------------------------------------------ f32 chains
// The streamed form of k_canon32.hip, chain by chain: tiles of 1024 blocks; the approximate
// start of every tile (the start plus the approximate sums of the tiles before) picks the
// tile's hypothesis at f32 precision; the tile summary joins its runs' summaries in order; the
// walk applies a tile's summary when its hypothesis is the running value's binade and the check
// holds, and otherwise folds the tile's rows in f32 arithmetic as the kernel's wave does, after
// the tile's wide table when its summary failed at its approximate start. vals are f32 values
// held in doubles; every fold step rounds to f32. stats: [0] tiles applied, [1] tiles evaluated,
// [2] unused (0), [3] runs of 16 rows folded element by element, [4] wide tables applied.
namespace {

double eval_tile32_model(const double *vals, uint64_t b0, uint64_t b1, int j, double s, uint64_t *st) {
  st[3] += (b1 - b0 + kRun - 1) / kRun;
  for (uint64_t b = b0; b < b1; ++b) s = PrecF32::rnd(s + vals[b * 16 + j]);
  return s;
}

// The f32 wide table of chain j over blocks [b0, b1) from `approx`, as k_c32_wide builds it.
void wide_table32(const double *vals, uint64_t b0, uint64_t b1, int j, double approx, WideHead *hd, WideEnt *ents) {
  const Binade h = binade_of<PrecF32>(approx);
  *hd = WideHead{0.0, 0, 0};
  if (!h.ok) return;
  const double b = wide_base<PrecF32>(approx, h), u0 = pow2(h.e - PrecF32::kMant);
  int m = -2000;
  for (int r = 0; r < kWideRes; ++r) {
    WideAcc a = wide_start(b + r * u0);
    for (uint64_t k0 = b0; k0 < b1; k0 += 16) {
      double v[16];
      for (uint64_t q = 0; q < 16; ++q) v[q] = k0 + q < b1 ? vals[(k0 + q) * 16 + j] : 0.0;
      wide_push_n<16, PrecF32>(a, v);
    }
    ents[r] = wide_end<PrecF32>(a);
    if (ents[r].lo <= ents[r].hi) m = std::max(m, a.emax - h.e);
  }
  *hd = WideHead{b, h.e, wide_nres(m)};
}

}  // namespace

extern "C" int vkm_c32_fold(const double *vals, uint64_t blocks, double *acc, uint64_t *stats) {
  uint64_t st[5] = {0, 0, 0, 0, 0};
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  for (int j = 0; j < 16; ++j) {
    std::vector<double> pre(tiles);
    double run = acc[j];
    for (uint64_t t = 0; t < tiles; ++t) {
      pre[t] = run;
      double a = 0.0;
      for (uint64_t b = t * kTileBlocks; b < std::min(blocks, (t + 1) * kTileBlocks); ++b) a += vals[b * 16 + j];
      run += a;
    }
    double s = acc[j];
    for (uint64_t t = 0; t < tiles; ++t) {
      const uint64_t b0 = t * kTileBlocks, b1 = std::min(blocks, b0 + kTileBlocks);
      const Binade h = binade_of<PrecF32>(pre[t]);
      RunFn f = run_identity();
      if (h.ok)
        for (uint64_t b = b0; b < b1; b += kRun) {
          double v[kRun];
          const int cnt = static_cast<int>(std::min<uint64_t>(kRun, b1 - b));
          for (int k = 0; k < static_cast<int>(kRun); ++k) v[k] = k < cnt ? vals[(b + k) * 16 + j] : 0.0;
          f = run_join(f, run_of<kRun, PrecF32>(v, cnt, binade_ref<PrecF32>(h)));
        }
      const Binade hs = binade_of<PrecF32>(s);
      double out = s, probe;
      if (h.ok && same_binade(h, hs) && run_apply<PrecF32>(f, hs, s, &out)) {
        s = out;
        st[0]++;
        continue;
      }
      // the summary does not apply: the tile's wide table when its summary failed at its
      // approximate start (the device builds one then), else the rows
      if (h.ok && !run_apply<PrecF32>(f, h, pre[t], &probe)) {
        WideHead hd;
        WideEnt ents[kWideRes];
        wide_table32(vals, b0, b1, j, pre[t], &hd, ents);
        if (wide_apply<PrecF32>(hd, ents, s, &out)) {
          s = out;
          st[4]++;
          continue;
        }
      }
      s = eval_tile32_model(vals, b0, b1, j, s, st);
      st[1]++;
    }
    acc[j] = s;
  }
  if (stats)
    for (int k = 0; k < 5; ++k) stats[k] = st[k];
  return 0;
}
