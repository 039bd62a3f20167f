// CPU check of the stream-K work division (GemmPlan::use_streamk /
// streamk_schedule, mps_amd/csrc/planner.cpp), built and run by
// tests/test_streamk_cpu.py.  For seeded random grouped-GEMM plans and the
// H_eff step-4 problems of the 6x24 cylinder at chi = 2048 it checks, for several
// tile shapes, kernel modes and work-group counts:
//   * every tile's K chunks are covered exactly once, by pieces in K order;
//   * no piece is shorter than kStreamKMinPiece chunks unless it is its tile's
//     only piece;
//   * cut tiles are listed once with their piece count, their partial tiles are
//     disjoint and inside the plan's partial buffer (part_elems);
//   * every work group's work (chunks + kStreamKOverhead per piece) is within
//     kStreamKMinPiece + 2 * kStreamKOverhead of its XCD's mean.
// For the same plans it checks the per-tile records the DMA kernel reads
// (GemmPlan::tile_records) of the stream-K pieces and of split-K tile lists
// (tiles() and schedule()) against the plan, the first chunk's segment and k
// offset found by enumerating the problem's chunks one by one.
// Prints one line per case and exits non-zero on the first violation.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <tuple>
#include <vector>

#include "../mps_amd/csrc/planner.hpp"

using namespace mpsh;

static int fail(const char* what, int c) {
  std::printf("FAIL case %d: %s\n", c, what);
  return 1;
}

static int64_t rec64(const int32_t* w, int i) {
  return (int64_t)(((uint64_t)(uint32_t)w[i + 1] << 32) | (uint64_t)(uint32_t)w[i]);
}

static int check_records(const GemmPlan& g, const std::vector<GemmTile>& t, int BK, int c) {
  const std::vector<int32_t> r = g.tile_records(t, BK);
  if (r.size() != t.size() * (size_t)kTileRecWords) return fail("record count", c);
  for (size_t i = 0; i < t.size(); ++i) {
    const GemmTile& x = t[i];
    const GemmProb& p = g.probs[x.prob];
    const int32_t* w = r.data() + i * kTileRecWords;
    // (segment, k offset) of every K chunk of the problem, in K order
    std::vector<std::pair<int, int64_t>> ch;
    for (int s = 0; s < p.nseg; ++s)
      for (int64_t k = 0; k < g.segs[p.seg0 + s].K; k += BK) ch.push_back({s, k});
    const int64_t n = (int64_t)ch.size();
    int64_t cb = x.kc0, ce = x.kc1;
    if (ce < 0) {
      cb = n * x.split / p.nsplit;
      ce = n * (x.split + 1) / p.nsplit;
    }
    if (w[0] != x.m0 || w[1] != x.n0 || w[2] != p.M || w[3] != p.N || w[4] != p.seg0 || w[5] != p.nseg)
      return fail("record tile / problem fields", c);
    if (w[6] != ce - cb) return fail("record chunk count", c);
    if (w[9] != x.split || w[10] != p.nsplit || w[11] != p.pld || w[12] != x.pld)
      return fail("record split fields", c);
    if (rec64(w, 14) != p.crow0 || rec64(w, 16) != p.c_off || rec64(w, 18) != p.ldc || rec64(w, 20) != p.part_off ||
        rec64(w, 22) != x.poff)
      return fail("record output fields", c);
    if (ce > cb) {
      const int s = ch[cb].first;
      const GemmSeg& sg = g.segs[p.seg0 + s];
      if (w[7] != s || w[8] != ch[cb].second) return fail("record first chunk", c);
      if (rec64(w, 24) != sg.K || rec64(w, 26) != sg.atab0 || rec64(w, 28) != sg.btab0 || rec64(w, 30) != sg.a_off ||
          rec64(w, 32) != sg.lda || rec64(w, 34) != sg.b_off || rec64(w, 36) != sg.ldb)
        return fail("record segment descriptor", c);
    }
  }
  return 0;
}

// records of the split-K tile lists of a plan (splits for `mode` at 2048 work items)
static int check_split_records(const GemmPlan& g0, int mode, int slots, int c) {
  GemmPlan g = g0;
  set_splits(g, 2048, mode);
  const int BK = chunk_k(mode, 0);
  std::vector<int32_t> qoff, woff;
  if (check_records(g, g.tiles(&qoff), BK, c)) return 1;
  return check_records(g, g.schedule(slots, BK, &woff), BK, c);
}

static int check(const GemmPlan& g0, int mode, int slots, int c) {
  GemmPlan g = g0;
  g.use_streamk(mode);
  std::vector<int32_t> woff;
  std::vector<TileRed> reds;
  const int BK = chunk_k(mode, 0);
  std::vector<GemmTile> t = g.streamk_schedule(slots, BK, &woff, &reds);
  if ((int)woff.size() != NXCD * slots + 1 || woff.back() != (int)t.size()) return fail("woff size", c);
  if (check_records(g, t, BK, c) || check_split_records(g0, mode, slots, c)) return 1;
  std::vector<int64_t> ctot(g.probs.size(), 0);
  for (size_t p = 0; p < g.probs.size(); ++p)
    for (int s = 0; s < g.probs[p].nseg; ++s) ctot[p] += (g.segs[g.probs[p].seg0 + s].K + BK - 1) / BK;
  // coverage
  std::map<std::tuple<int, int, int>, std::vector<const GemmTile*>> by;
  for (const GemmTile& x : t) by[{x.prob, x.m0, x.n0}].push_back(&x);
  int64_t ntiles = 0;
  for (const GemmProb& p : g.probs) ntiles += (int64_t)((p.M + g.BM - 1) / g.BM) * ((p.N + g.BN - 1) / g.BN);
  if ((int64_t)by.size() != ntiles) return fail("tile count", c);
  std::map<std::tuple<int, int, int>, const TileRed*> rr;
  for (const TileRed& r : reds) {
    if (rr.count({r.prob, r.m0, r.n0})) return fail("tile reduced twice", c);
    rr[{r.prob, r.m0, r.n0}] = &r;
  }
  std::vector<char> blk_used((size_t)(g.part_elems / ((int64_t)g.BM * g.BN)) + 1, 0);
  for (auto& kv : by) {
    auto ps = kv.second;
    std::sort(ps.begin(), ps.end(), [](const GemmTile* a, const GemmTile* b) { return a->kc0 < b->kc0; });
    const int64_t c0 = ctot[std::get<0>(kv.first)];
    int64_t k = 0;
    for (const GemmTile* x : ps) {
      if (x->kc0 != k || x->kc1 < x->kc0) return fail("chunk coverage", c);
      if (ps.size() > 1 && x->kc1 - x->kc0 < kStreamKMinPiece) return fail("short piece", c);
      k = x->kc1;
    }
    if (k != c0) return fail("chunk total", c);
    const bool cut = ps.size() > 1;
    auto it = rr.find(kv.first);
    if (cut != (it != rr.end())) return fail("reduction list", c);
    if (cut) {
      if (it->second->np != (int)ps.size()) return fail("piece count", c);
      for (size_t i = 0; i < ps.size(); ++i) {
        if (ps[i]->pld != g.BM) return fail("partial leading dimension", c);
        const int64_t b = ps[i]->poff / ((int64_t)g.BM * g.BN);
        if (ps[i]->poff != b * g.BM * g.BN || b != it->second->blk0 + (int64_t)i) return fail("partial block", c);
        if ((b + 1) * g.BM * g.BN > g.part_elems) return fail("partial buffer bound", c);
        if (blk_used[b]++) return fail("partial block reused", c);
      }
    } else if (ps[0]->pld != 0) {
      return fail("whole tile writes a partial", c);
    }
  }
  // balance per XCD
  double worst = 0;
  for (int x = 0; x < NXCD; ++x) {
    std::vector<int64_t> w(slots, 0);
    int64_t tot = 0;
    for (int j = 0; j < slots; ++j) {
      for (int i = woff[x * slots + j]; i < woff[x * slots + j + 1]; ++i) w[j] += t[i].kc1 - t[i].kc0 + kStreamKOverhead;
      tot += w[j];
    }
    const double mean = (double)tot / slots;
    const int64_t mx = *std::max_element(w.begin(), w.end());
    worst = std::max(worst, mx - mean);
    if (mx > mean + kStreamKMinPiece + 2 * kStreamKOverhead + 1) {
      std::printf("  xcd %d: max %lld mean %.1f\n", x, (long long)mx, mean);
      return fail("work-group balance", c);
    }
  }
  std::printf("case %d: mode %d tile %dx%d slots %d: %zu pieces, %zu cut tiles, worst excess %.1f chunks\n", c, mode,
              g.BM, g.BN, slots, t.size(), reds.size(), worst);
  return 0;
}

int main() {
  int c = 0;
  // the step-4 problems of the 6x24 cylinder at chi = 2048 (60 sectors, MPO bond {-1:6, 0:8, +1:6})
  {
    const int dims[60] = {1, 1, 1,  1,   1,   1,   1,   1,   1,   1,   1,   1,  1,  1,  1,  1,  1,  2,  3,  7,
                          12, 21, 35, 54, 79, 107, 137, 165, 187, 198, 198, 187, 165, 137, 107, 79, 54, 35, 21, 12,
                          7,  3,  2,  1,  1,   1,   1,   1,   1,   1,   1,   1,  1,  1,  1,  1,  1,  1,  1,  1};
    std::map<int, int> dl, dr;
    for (int i = 0; i < 60; ++i) {
      dl[i] = dims[i];
      dr[i + 1] = dims[i];
    }
    const std::map<int, int> dw = {{-1, 6}, {0, 8}, {1, 6}};
    GemmPlan g;
    for (auto& [br, dbr] : dr) {
      int64_t M = 0;
      for (int p1 = 0; p1 < 2; ++p1)
        for (int p2 = 0; p2 < 2; ++p2) {
          auto it = dl.find(br - p1 - p2);
          if (it != dl.end()) M += it->second;
        }
      if (!M) continue;
      g.add_prob(M, dbr);
      g.add_crows_affine(0, M, M);
      for (auto& [w, n] : dw) {
        auto it = dr.find(br - w);
        if (it != dr.end()) g.add_seg((int64_t)it->second * n, 0, 8, 0, 8);
      }
    }
    for (int mode : {1, 0, 4})
      for (auto bmbn : {std::make_pair(64, 64), std::make_pair(64, 128), std::make_pair(128, 64)})
        for (int slots : {128, 96, 37, 1}) {
          GemmPlan h = g;
          h.BM = bmbn.first;
          h.BN = bmbn.second;
          h.KS = 0;
          if (slots > sk_max_slots(mode, h.BM, h.BN)) continue;
          if (check(h, mode, slots, ++c)) return 1;
        }
  }
  // seeded random plans: uneven problems, zero problems, several K segments
  std::mt19937_64 rng(1911);
  for (int rep = 0; rep < 40; ++rep) {
    GemmPlan g;
    const int np = 1 + (int)(rng() % 40);
    for (int p = 0; p < np; ++p) {
      const int64_t M = 1 + (int64_t)(rng() % 900), N = 1 + (int64_t)(rng() % 300);
      g.add_prob(M, N);
      g.add_crows_affine(0, M, M);
      const int ns = (int)(rng() % 5);  // 0: a zero problem
      for (int s = 0; s < ns; ++s) g.add_seg(1 + (int64_t)(rng() % 700), 0, 8, 0, 8);
    }
    g.BM = 64;
    g.BN = (rng() % 2) ? 64 : 128;
    g.KS = 0;
    const int mode = (int)(rng() % 2);
    const int slots = 1 + (int)(rng() % sk_max_slots(mode, g.BM, g.BN));
    if (check(g, mode, slots, ++c)) return 1;
  }
  std::printf("OK %d cases\n", c);
  return 0;
}
