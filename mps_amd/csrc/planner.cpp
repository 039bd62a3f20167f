// Host-side integer bookkeeping (see planner.hpp).  Each routine cites the
// reference function whose sector logic it restates.
#include "planner.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <map>
#include <numeric>
#include <set>
#include <tuple>

namespace mpsh {

static int64_t prod(const int64_t* p, int n) {
  int64_t r = 1;
  for (int i = 0; i < n; ++i) r *= p[i];
  return r;
}

// ---------------------------------------------------------------------------
// Legs (src/vspace.jl)
// ---------------------------------------------------------------------------
int64_t Leg::dim_of(int64_t c) const {
  auto it = std::lower_bound(charges.begin(), charges.end(), c);
  if (it == charges.end() || *it != c) return 0;
  return dims[it - charges.begin()];
}

int64_t Leg::total() const { return std::accumulate(dims.begin(), dims.end(), (int64_t)0); }

Leg make_leg(bool isdual, const std::vector<std::pair<int64_t, int64_t>>& pairs) {
  std::map<int64_t, int64_t> m;
  for (auto& p : pairs) {
    if (m.count(p.first)) throw Error(ERR_INVALID, "duplicate charges!");
    if (p.second <= 0) throw Error(ERR_INVALID, "space dimension can only be positive!");
    m[p.first] = p.second;
  }
  Leg l;
  l.isdual = isdual;
  for (auto& kv : m) {
    l.charges.push_back(kv.first);
    l.dims.push_back(kv.second);
  }
  return l;
}

bool isdual_pair(const Leg& a, const Leg& b) { return a.same_map(b) && a.isdual != b.isdual; }

// fuse(_dual, Vs): right-to-left recursion; the fused dims are the charge-flow
// convolution of the constituent legs (src/vspace.jl:112-138).
Leg fuse(bool dual, const std::vector<Leg>& Vs) {
  const size_t N = Vs.size();
  if (N == 0) throw Error(ERR_INVALID, "fuse of no legs");
  if (N == 1) {
    if (dual == Vs[0].isdual) return Vs[0];
    std::vector<std::pair<int64_t, int64_t>> p;
    for (size_t i = 0; i < Vs[0].charges.size(); ++i) p.push_back({-Vs[0].charges[i], Vs[0].dims[i]});
    return make_leg(!Vs[0].isdual, p);
  }
  if (N > 2) {
    std::vector<Leg> head(Vs.begin(), Vs.end() - 2);
    head.push_back(fuse(dual, {Vs[N - 2], Vs[N - 1]}));
    return fuse(dual, head);
  }
  std::map<int64_t, int64_t> acc;
  const Leg& V1 = Vs[0];
  const Leg& V2 = Vs[1];
  for (size_t j = 0; j < V2.charges.size(); ++j) {
    int64_t c2 = V2.isdual ? -V2.charges[j] : V2.charges[j];
    for (size_t i = 0; i < V1.charges.size(); ++i) {
      int64_t c1 = V1.isdual ? -V1.charges[i] : V1.charges[i];
      int64_t c = dual ? -(c1 + c2) : c1 + c2;
      acc[c] += V1.dims[i] * V2.dims[j];
    }
  }
  Leg l;
  l.isdual = dual;
  for (auto& kv : acc) {
    l.charges.push_back(kv.first);
    l.dims.push_back(kv.second);
  }
  return l;
}

// ---------------------------------------------------------------------------
// Sector enumeration (src/sector.jl:69-90): last leg outermost, ascending raw
// labels, recursion on the head; output is column-major sorted by construction.
// ---------------------------------------------------------------------------
static void enum_rec(int64_t charge, const std::vector<Leg>& legs, int N, std::vector<int64_t>& cur,
                     std::vector<int64_t>& curz, std::vector<std::vector<int64_t>>& outs,
                     std::vector<std::vector<int64_t>>& outz) {
  if (N == 1) {
    const Leg& V = legs[0];
    int64_t c = V.isdual ? -charge : charge;
    int64_t d = V.dim_of(c);
    if (d > 0) {
      cur[0] = c;
      curz[0] = d;
      outs.push_back(cur);
      outz.push_back(curz);
    }
    return;
  }
  const Leg& V = legs[N - 1];
  for (size_t i = 0; i < V.charges.size(); ++i) {
    int64_t c = V.charges[i];
    int64_t cc = V.isdual ? -c : c;
    cur[N - 1] = c;
    curz[N - 1] = V.dims[i];
    enum_rec(charge - cc, legs, N - 1, cur, curz, outs, outz);
  }
}

void all_sectors(int64_t charge, const std::vector<Leg>& legs, std::vector<int64_t>& sects,
                 std::vector<int64_t>& sizes) {
  const int N = (int)legs.size();
  sects.clear();
  sizes.clear();
  if (N == 0) return;
  std::vector<int64_t> cur(N), curz(N);
  std::vector<std::vector<int64_t>> outs, outz;
  enum_rec(charge, legs, N, cur, curz, outs, outz);
  for (size_t i = 0; i < outs.size(); ++i) {
    sects.insert(sects.end(), outs[i].begin(), outs[i].end());
    sizes.insert(sizes.end(), outz[i].begin(), outz[i].end());
  }
}

int64_t Struct::block_numel(int64_t b) const { return prod(size(b), rank()); }

int64_t Struct::find(const std::vector<int64_t>& s) const {
  auto it = index.find(s);
  return it == index.end() ? -1 : it->second;
}

Struct make_struct(int64_t charge, const std::vector<Leg>& legs) {
  Struct S;
  S.charge = charge;
  S.legs = legs;
  if (legs.empty()) throw Error(ERR_INVALID, "rank-0 structure");
  all_sectors(charge, legs, S.sectors, S.sizes);
  const int r = (int)legs.size();
  S.nblocks = (int64_t)S.sectors.size() / r;
  S.offsets.assign(S.nblocks + 1, 0);
  for (int64_t b = 0; b < S.nblocks; ++b) {
    S.offsets[b + 1] = S.offsets[b] + prod(&S.sizes[b * r], r);
    S.index.emplace(std::vector<int64_t>(S.sectors.begin() + b * r, S.sectors.begin() + (b + 1) * r), b);
  }
  return S;
}

bool same_struct(const Struct& a, const Struct& b) { return a.charge == b.charge && a.legs == b.legs; }

static std::vector<int64_t> colmajor_strides(const int64_t* dims, int n) {
  std::vector<int64_t> s(n);
  int64_t acc = 1;
  for (int i = 0; i < n; ++i) {
    s[i] = acc;
    acc *= dims[i];
  }
  return s;
}

// ---------------------------------------------------------------------------
// Copy descriptors
// ---------------------------------------------------------------------------
void CopyPlan::add(int ndim, const int64_t* dims, const int64_t* sst, const int64_t* dst, int64_t soff,
                   int64_t doff) {
  struct D {
    int64_t n, s, d;
  };
  std::vector<D> v;
  for (int i = 0; i < ndim; ++i)
    if (dims[i] > 1) v.push_back({dims[i], sst[i], dst[i]});
  // destination-contiguous first, then merge runs contiguous on both sides
  std::stable_sort(v.begin(), v.end(), [](const D& a, const D& b) { return a.d < b.d; });
  std::vector<D> m;
  for (auto& x : v) {
    if (!m.empty() && x.d == m.back().d * m.back().n && x.s == m.back().s * m.back().n)
      m.back().n *= x.n;
    else
      m.push_back(x);
  }
  if (m.empty()) m.push_back({1, 1, 1});
  int tr = 0;
  if (m[0].s != 1) {
    for (size_t i = 1; i < m.size(); ++i)
      if (m[i].s == 1) {
        D t = m[i];
        m.erase(m.begin() + i);
        m.insert(m.begin() + 1, t);
        tr = 1;
        break;
      }
  }
  if ((int)m.size() > MAXD) throw Error(ERR_UNSUPPORTED, "copy with more than MAXD non-mergeable dims");
  CopyDesc c{};
  c.ndim = (int32_t)m.size();
  c.transpose = tr;
  for (size_t i = 0; i < m.size(); ++i) {
    if (m[i].n > INT32_MAX) throw Error(ERR_UNSUPPORTED, "dimension too large");
    c.dims[i] = (int32_t)m[i].n;
    c.src_st[i] = m[i].s;
    c.dst_st[i] = m[i].d;
  }
  c.src_off = soff;
  c.dst_off = doff;
  descs.push_back(c);
  int64_t e = 1;
  for (auto& x : m) e *= x.n;
  elems += e;
}

void CopyPlan::finalize() {
  tile_start.assign(descs.size() + 1, 0);
  for (size_t i = 0; i < descs.size(); ++i) {
    const CopyDesc& c = descs[i];
    int64_t t;
    int64_t d0 = c.dims[0], d1 = c.ndim > 1 ? c.dims[1] : 1;
    int64_t z = 1;
    for (int k = 2; k < c.ndim; ++k) z *= c.dims[k];
    if (c.transpose)
      t = ((d0 + 63) / 64) * ((d1 + 63) / 64) * z;
    else
      t = ((d0 * d1 + 4095) / 4096) * z;
    tile_start[i + 1] = tile_start[i] + t;
  }
  ntiles = tile_start.back();
  if (ntiles > INT32_MAX) throw Error(ERR_UNSUPPORTED, "copy with more than 2^31 tiles");
  tile_map.resize(2 * ntiles);
  for (size_t i = 0; i < descs.size(); ++i)
    for (int64_t t = tile_start[i]; t < tile_start[i + 1]; ++t) {
      tile_map[2 * t] = (int32_t)i;
      tile_map[2 * t + 1] = (int32_t)(t - tile_start[i]);
    }
}

void GemmPlan::add_seg_tab(int64_t K, const std::vector<int64_t>& arows, const std::vector<int64_t>& brows) {
  GemmProb& p = probs.back();
  if ((int64_t)arows.size() != p.M || (int64_t)brows.size() != p.N)
    throw Error(ERR_INVALID, "gemm segment row tables do not match the problem");
  segs.push_back(GemmSeg{K, (int64_t)atab.size(), (int64_t)btab.size(), 0, 0, 0, 0});
  atab.insert(atab.end(), arows.begin(), arows.end());
  btab.insert(btab.end(), brows.begin(), brows.end());
  p.nseg++;
  flops += 2 * (int64_t)p.M * p.N * K;
}

void GemmPlan::add_seg(int64_t K, int64_t a_off, int64_t lda, int64_t b_off, int64_t ldb) {
  GemmProb& p = probs.back();
  segs.push_back(GemmSeg{K, -1, -1, a_off, lda, b_off, ldb});
  p.nseg++;
  flops += 2 * (int64_t)p.M * p.N * K;
}

static int64_t prob_chunks(const GemmPlan& g, const GemmProb& p, int BK) {
  int64_t c = 0;
  for (int s = 0; s < p.nseg; ++s) c += (g.segs[p.seg0 + s].K + BK - 1) / BK;
  return c;
}

void GemmPlan::choose_splits(int64_t target, int BK, int64_t min_chunks_per_split) {
  streamk = false;
  sk_mode = -1;
  int64_t tiles = 0;
  for (auto& p : probs) tiles += ((p.M + BM - 1) / BM) * ((p.N + BN - 1) / BN);
  part_elems = 0;
  for (auto& p : probs) {
    p.nsplit = 1;
    p.pld = (p.M + 3) & ~3;
    p.part_off = 0;
  }
  if (tiles == 0 || tiles >= target) return;
  const int64_t want = (target + tiles - 1) / tiles;
  for (auto& p : probs) {
    const int64_t ch = prob_chunks(*this, p, BK);
    int64_t s = std::min<int64_t>(want, ch / std::max<int64_t>(1, min_chunks_per_split));
    if (s > 1 && (int64_t)p.pld * p.N >= ((int64_t)1 << 31)) s = 1;  // the reduction indexes M*N in 32 bits
    if (s > 1) {
      p.nsplit = (int32_t)s;
      p.part_off = part_elems;
      part_elems += s * (int64_t)p.pld * p.N;
    }
  }
}

void GemmPlan::use_streamk(int mode) {
  choose_splits(0, 16, 32);  // no per-problem split
  streamk = true;
  sk_mode = mode;
  // partial tiles: at most two per work group, and a tile of c chunks is cut
  // into at most c / kStreamKMinPiece pieces (only tiles of >= 2 minimum pieces)
  const int BK = chunk_k(mode, 0);
  int64_t cut = 0;
  for (const GemmProb& p : probs) {
    const int64_t c = prob_chunks(*this, p, BK);
    if (c >= 2 * kStreamKMinPiece)
      cut += (int64_t)((p.M + BM - 1) / BM) * ((p.N + BN - 1) / BN) * (c / kStreamKMinPiece);
  }
  part_elems = std::min<int64_t>(2 * (int64_t)NXCD * sk_max_slots(mode, BM, BN), cut) * BM * BN;
}

// split target < 0: stream-K (DMA-streamed variants; others keep split-K at 1024 items)
void set_splits(GemmPlan& g, int64_t split_target, int mode) {
  if (split_target < 0 && g.KS == 0)
    g.use_streamk(mode);
  else
    g.choose_splits(split_target < 0 ? 1024 : split_target, chunk_k(mode, g.KS), 32);  // target 0: no split-K
}

GemmPlan gemm_variant(const GemmPlan& base, const TileCfg& c, int64_t split_target, int mode) {
  GemmPlan g = base;
  g.BM = c.BM;
  g.BN = c.BN;
  g.KS = c.KS;
  set_splits(g, split_target, mode);
  return g;
}

std::vector<GemmTile> GemmPlan::tiles(std::vector<int32_t>* qoff) const {
  // Panel pieces: the tiles of one (problem, split) over a contiguous range of
  // m0 rows (all n0).  Tiles of a piece share A rows (same m0) and B rows (same
  // n0), so a piece is placed whole on one XCD; large (problem, split) units are
  // cut by m0 rows into pieces of at most a quarter of one XCD's share.  Inside
  // a piece the tiles run along the larger operand's rows in the outer loop.
  struct Piece {
    int64_t work;
    int32_t prob, split, mr0, mr1;
  };
  std::vector<std::pair<int32_t, int64_t>> units;  // (prob, chunk K of one split)
  int64_t total = 0;
  for (size_t pi = 0; pi < probs.size(); ++pi) {
    const GemmProb& q = probs[pi];
    int64_t K = 0;
    for (int s = 0; s < q.nseg; ++s) K += segs[q.seg0 + s].K;
    const int64_t Ks = std::max<int64_t>(1, K / std::max(1, q.nsplit));
    const int64_t nm = (q.M + BM - 1) / BM, nn = (q.N + BN - 1) / BN;
    total += nm * nn * q.nsplit * Ks;
    units.push_back({(int32_t)pi, Ks});
  }
  const int64_t cap = std::max<int64_t>(1, total / (NXCD * 4));
  std::vector<Piece> pieces;
  for (auto& u : units) {
    const GemmProb& q = probs[u.first];
    const int64_t nm = (q.M + BM - 1) / BM, nn = (q.N + BN - 1) / BN;
    const int64_t row_work = nn * u.second;
    const int64_t rows_per = std::max<int64_t>(1, cap / std::max<int64_t>(1, row_work));
    for (int sp = 0; sp < q.nsplit; ++sp)
      for (int64_t r0 = 0; r0 < nm; r0 += rows_per) {
        const int64_t r1 = std::min(nm, r0 + rows_per);
        pieces.push_back(Piece{(r1 - r0) * row_work, u.first, sp, (int32_t)r0, (int32_t)r1});
      }
  }
  std::stable_sort(pieces.begin(), pieces.end(), [](const Piece& a, const Piece& b) { return a.work > b.work; });
  std::vector<std::vector<GemmTile>> queue(NXCD);
  std::vector<int64_t> load(NXCD, 0);
  for (const Piece& pc : pieces) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += pc.work;
    const GemmProb& q = probs[pc.prob];
    // the larger operand is walked once in the outer loop, the smaller one stays
    // in L2: column panels (n0 outer) when the piece's B rows outnumber its A rows
    const int64_t arows = (int64_t)(pc.mr1 - pc.mr0) * BM;
    if ((int64_t)q.N > arows) {
      for (int n0 = 0; n0 < q.N; n0 += BN)
        for (int32_t r = pc.mr0; r < pc.mr1; ++r) queue[x].push_back(GemmTile{pc.prob, r * BM, n0, pc.split});
    } else {
      for (int32_t r = pc.mr0; r < pc.mr1; ++r)
        for (int n0 = 0; n0 < q.N; n0 += BN) queue[x].push_back(GemmTile{pc.prob, r * BM, n0, pc.split});
    }
  }
  std::vector<GemmTile> out;
  if (qoff) qoff->assign(NXCD + 1, 0);
  for (int x = 0; x < NXCD; ++x) {
    out.insert(out.end(), queue[x].begin(), queue[x].end());
    if (qoff) (*qoff)[x + 1] = (int32_t)out.size();
  }
  return out;
}

std::vector<GemmTile> GemmPlan::schedule(int slots, int BK, std::vector<int32_t>* woff) const {
  if (slots < 1) throw Error(ERR_INVALID, "schedule: slots < 1");
  std::vector<int32_t> qoff;
  const std::vector<GemmTile> q = tiles(&qoff);
  // K chunks of each problem (the split ranges follow from them)
  std::vector<int64_t> ctot(probs.size(), 0);
  for (size_t pi = 0; pi < probs.size(); ++pi)
    for (int s2 = 0; s2 < probs[pi].nseg; ++s2) ctot[pi] += (segs[probs[pi].seg0 + s2].K + BK - 1) / BK;
  // a tile's cost: its K chunks plus a fixed cost for its set-up and its stores
  constexpr int64_t kTileCost = 3;
  auto cost = [&](const GemmTile& t) {
    const int64_t c = ctot[t.prob], ns = probs[t.prob].nsplit;
    return c * (t.split + 1) / ns - c * t.split / ns + kTileCost;
  };
  std::vector<GemmTile> out;
  out.reserve(q.size());
  woff->assign((size_t)NXCD * slots + 1, 0);
  for (int x = 0; x < NXCD; ++x) {
    const int32_t b = qoff[x], e = qoff[x + 1];
    std::vector<int32_t> order(e - b);
    std::iota(order.begin(), order.end(), b);
    std::stable_sort(order.begin(), order.end(), [&](int32_t i, int32_t j) { return cost(q[i]) > cost(q[j]); });
    std::vector<std::vector<int32_t>> lists(slots);
    std::vector<int64_t> load(slots, 0);
    for (int32_t i : order) {
      const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[w] += cost(q[i]);
      lists[w].push_back(i);
    }
    for (int w = 0; w < slots; ++w) {
      std::sort(lists[w].begin(), lists[w].end());  // queue order: the panel walk of tiles()
      for (int32_t i : lists[w]) out.push_back(q[i]);
      (*woff)[(size_t)x * slots + w + 1] = (int32_t)out.size();
    }
  }
  return out;
}

std::vector<GemmTile> GemmPlan::streamk_schedule(int slots, int BK, std::vector<int32_t>* woff,
                                                 std::vector<TileRed>* reds) const {
  if (slots < 1) throw Error(ERR_INVALID, "streamk_schedule: slots < 1");
  std::vector<int32_t> qoff;
  const std::vector<GemmTile> q = tiles(&qoff);  // every problem unsplit (use_streamk)
  std::vector<int64_t> ctot(probs.size(), 0);
  for (size_t pi = 0; pi < probs.size(); ++pi)
    for (int s2 = 0; s2 < probs[pi].nseg; ++s2) ctot[pi] += (segs[probs[pi].seg0 + s2].K + BK - 1) / BK;
  struct Piece {
    int32_t tile;
    int64_t k0, k1;
  };
  std::vector<GemmTile> out;
  woff->assign((size_t)NXCD * slots + 1, 0);
  reds->clear();
  int64_t blk = 0;
  const int64_t ovh = kStreamKOverhead, minp = kStreamKMinPiece;
  for (int x = 0; x < NXCD; ++x) {
    const int32_t b = qoff[x], e = qoff[x + 1];
    // work of the queue: chunks plus one overhead per piece (each work-group
    // boundary adds at most one piece)
    int64_t total = ovh * (int64_t)(slots - 1);
    for (int32_t t = b; t < e; ++t) total += ctot[q[t].prob] + ovh;
    std::vector<std::vector<Piece>> lists(slots);
    int j = 0;
    double cum = 0;  // work placed so far on this XCD
    // work group j's share ends at `end`: the work left spread evenly over the
    // work groups left, so rounding and short-piece rules do not pile up
    double end = (double)total / slots;
    auto next = [&]() {
      ++j;
      end = cum + ((double)total - cum) / (slots - j);
    };
    for (int32_t t = b; t < e; ++t) {
      const int64_t c = ctot[q[t].prob];
      int64_t k = 0;
      for (;;) {
        int64_t take = c - k;
        if (j < slots - 1) {
          const int64_t room = (int64_t)std::floor(end - cum + 0.5) - ovh;
          if (room < take) {
            if (room < minp && !lists[j].empty()) {  // too little room left: the next work group starts it
              next();
              continue;
            }
            take = std::max(room, std::min(take, minp));
            if (c - k - take < minp) take = c - k;  // no short remainder
          }
        }
        lists[j].push_back(Piece{t, k, k + take});
        cum += (double)(ovh + take);
        k += take;
        if (j < slots - 1 && cum >= end - 0.5) next();
        if (k >= c) break;
      }
    }
    // pieces of each tile, in K order (= work-group order)
    std::vector<std::vector<std::pair<int, int>>> where(e - b);  // (list, index)
    for (int w = 0; w < slots; ++w)
      for (int i = 0; i < (int)lists[w].size(); ++i) where[lists[w][i].tile - b].push_back({w, i});
    std::vector<std::vector<GemmTile>> lt(slots);
    for (int w = 0; w < slots; ++w) lt[w].resize(lists[w].size());
    for (int32_t t = b; t < e; ++t) {
      const auto& ps = where[t - b];
      const GemmTile& tl = q[t];
      const bool cut = ps.size() > 1;
      if (cut) reds->push_back(TileRed{tl.prob, tl.m0, tl.n0, (int32_t)ps.size(), blk});
      for (size_t i = 0; i < ps.size(); ++i) {
        const Piece& pc = lists[ps[i].first][ps[i].second];
        GemmTile g = tl;
        g.split = 0;
        g.kc0 = (int32_t)pc.k0;
        g.kc1 = (int32_t)pc.k1;
        if (cut) {
          g.pld = BM;
          g.poff = (blk + (int64_t)i) * BM * BN;
        }
        lt[ps[i].first][ps[i].second] = g;
      }
      if (cut) blk += (int64_t)ps.size();
    }
    for (int w = 0; w < slots; ++w) {
      out.insert(out.end(), lt[w].begin(), lt[w].end());
      (*woff)[(size_t)x * slots + w + 1] = (int32_t)out.size();
    }
  }
  if (blk * BM * BN > part_elems)
    throw Error(ERR_INVALID, "stream-K schedule needs more partial tiles than its plan holds");
  return out;
}

std::vector<int32_t> GemmPlan::tile_records(const std::vector<GemmTile>& tl, int BK) const {
  std::vector<int32_t> r(tl.size() * kTileRecWords, 0);
  auto put64 = [](int32_t* w, int i, int64_t v) {
    w[i] = (int32_t)(uint32_t)((uint64_t)v & 0xffffffffu);
    w[i + 1] = (int32_t)(uint32_t)((uint64_t)v >> 32);
  };
  for (size_t t = 0; t < tl.size(); ++t) {
    const GemmTile& x = tl[t];
    const GemmProb& p = probs[x.prob];
    int64_t ctot = 0;
    for (int s = 0; s < p.nseg; ++s) ctot += (segs[p.seg0 + s].K + BK - 1) / BK;
    int64_t cb = x.kc0, ce = x.kc1;
    if (ce < 0) {
      cb = ctot * x.split / p.nsplit;
      ce = ctot * (x.split + 1) / p.nsplit;
    }
    int seg = 0;
    int64_t k0 = 0, cpos = 0;
    for (; seg < p.nseg; ++seg) {
      const int64_t nc = (segs[p.seg0 + seg].K + BK - 1) / BK;
      if (cpos + nc > cb) {
        k0 = (cb - cpos) * BK;
        break;
      }
      cpos += nc;
    }
    int32_t* w = r.data() + t * kTileRecWords;
    w[0] = x.m0;
    w[1] = x.n0;
    w[2] = p.M;
    w[3] = p.N;
    w[4] = p.seg0;
    w[5] = p.nseg;
    w[6] = (int32_t)(ce - cb);
    w[7] = seg;
    w[8] = (int32_t)k0;
    w[9] = x.split;
    w[10] = p.nsplit;
    w[11] = p.pld;
    w[12] = x.pld;
    put64(w, 14, p.crow0);
    put64(w, 16, p.c_off);
    put64(w, 18, p.ldc);
    put64(w, 20, p.part_off);
    put64(w, 22, x.poff);
    if (ce > cb && seg < p.nseg) {
      const GemmSeg& g = segs[p.seg0 + seg];
      put64(w, 24, g.K);
      put64(w, 26, g.atab0);
      put64(w, 28, g.btab0);
      put64(w, 30, g.a_off);
      put64(w, 32, g.lda);
      put64(w, 34, g.b_off);
      put64(w, 36, g.ldb);
    }
  }
  return r;
}

std::vector<RedDesc> GemmPlan::reductions() const {
  std::vector<RedDesc> r;
  int64_t start = 0;
  for (size_t pi = 0; pi < probs.size(); ++pi) {
    const GemmProb& q = probs[pi];
    if (q.nsplit <= 1) continue;
    r.push_back(RedDesc{(int32_t)pi, 0, start});
    start += ((int64_t)q.M * q.N + 1023) / 1024;
  }
  r.push_back(RedDesc{-1, 0, start});  // sentinel: total work groups
  return r;
}

// ---------------------------------------------------------------------------
// contract index bookkeeping (src/contract.jl:82-99)
// ---------------------------------------------------------------------------
IndexPerm contract_index_perm(const std::vector<int>& idx) {
  std::vector<int> rems, cons, cons_order, tof;
  for (size_t i = 0; i < idx.size(); ++i) {
    if (idx[i] < 0) {
      cons.push_back((int)i + 1);
      cons_order.push_back(idx[i]);
    } else {
      rems.push_back((int)i + 1);
      tof.push_back(idx[i]);
    }
  }
  std::vector<size_t> p(tof.size()), cp(cons_order.size());
  std::iota(p.begin(), p.end(), 0);
  std::iota(cp.begin(), cp.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](size_t a, size_t b) { return tof[a] < tof[b]; });
  std::stable_sort(cp.begin(), cp.end(),
                   [&](size_t a, size_t b) { return std::abs(cons_order[a]) < std::abs(cons_order[b]); });
  IndexPerm r;
  for (size_t j : p) {
    r.rems.push_back(rems[j]);
    r.tofinals.push_back(tof[j]);
  }
  for (size_t j : cp) r.cons.push_back(cons[j]);
  return r;
}

static int64_t flow(const int64_t* ch, const std::vector<bool>& dual) {
  int64_t r = 0;
  for (size_t i = 0; i < dual.size(); ++i) r += dual[i] ? -ch[i] : ch[i];
  return r;
}

static bool rev_less(const std::vector<int64_t>& a, const std::vector<int64_t>& b) {
  for (size_t i = a.size(); i-- > 0;) {
    if (a[i] < b[i]) return true;
    if (a[i] > b[i]) return false;
  }
  return false;
}

// ---------------------------------------------------------------------------
// SymMatrix(A, rows, cols) (src/contract.jl:112-171): blocks are consumed in
// the order (permuted key, then fused key), both stable; inside each fused
// block, column slabs outer and row slabs inner.
// ---------------------------------------------------------------------------
FusedMatrix symmatrix_plan(const Struct& A, const std::vector<int>& rows, const std::vector<int>& cols,
                           bool transposed, int64_t dst_base, int kpad) {
  const int N = A.rank();
  std::vector<int> idxperm(rows);
  idxperm.insert(idxperm.end(), cols.begin(), cols.end());
  {
    std::vector<int> s(idxperm);
    std::sort(s.begin(), s.end());
    for (int i = 0; i < (int)s.size(); ++i)
      if ((int)s.size() != N || s[i] != i + 1)
        throw Error(ERR_INVALID, "Incorrect index set for conversion to SymMatrix!");
  }
  if (rows.empty()) throw Error(ERR_INVALID, "empty row for matrix not allowed!");
  if (cols.empty()) throw Error(ERR_INVALID, "empty col for matrix not allowed!");
  const int n = (int)rows.size();
  std::vector<bool> rowdual, coldual;
  std::vector<Leg> rlegs, clegs;
  for (int i : rows) {
    rowdual.push_back(A.legs[i - 1].isdual);
    rlegs.push_back(A.legs[i - 1]);
  }
  for (int i : cols) {
    coldual.push_back(A.legs[i - 1].isdual);
    clegs.push_back(A.legs[i - 1]);
  }
  const int64_t nb = A.nblocks;
  std::vector<std::vector<int64_t>> cs(nb);
  for (int64_t b = 0; b < nb; ++b) {
    cs[b].resize(N);
    for (int t = 0; t < N; ++t) cs[b][t] = A.sector(b)[idxperm[t] - 1];
  }
  std::vector<int64_t> sperm1(nb);
  std::iota(sperm1.begin(), sperm1.end(), 0);
  std::stable_sort(sperm1.begin(), sperm1.end(), [&](int64_t a, int64_t b) { return rev_less(cs[a], cs[b]); });
  std::vector<std::vector<int64_t>> fs(nb);
  for (int64_t i = 0; i < nb; ++i) {
    const auto& sct = cs[sperm1[i]];
    fs[i] = {flow(sct.data(), rowdual), -flow(sct.data() + n, coldual)};
  }
  std::vector<int64_t> sperm2(nb);
  std::iota(sperm2.begin(), sperm2.end(), 0);
  std::stable_sort(sperm2.begin(), sperm2.end(), [&](int64_t a, int64_t b) { return rev_less(fs[a], fs[b]); });
  std::vector<int64_t> semits(nb);
  for (int64_t j = 0; j < nb; ++j) semits[j] = sperm1[sperm2[j]];

  FusedMatrix FM;
  FM.F = make_struct(A.charge, {fuse(false, rlegs), fuse(true, clegs)});
  FM.k_contig_rows = transposed;
  FM.ld.resize(FM.F.nblocks);
  FM.poff.assign(FM.F.nblocks + 1, 0);
  for (int64_t f = 0; f < FM.F.nblocks; ++f) {
    const int64_t r = FM.F.size(f)[0], c = FM.F.size(f)[1];
    // transposed: rows of length c (K) stored with stride ld; else column-major, ld >= r (K)
    FM.ld[f] = transposed ? (c + kpad - 1) / kpad * kpad : (r + kpad - 1) / kpad * kpad;
    FM.poff[f + 1] = FM.poff[f] + (transposed ? r : c) * FM.ld[f];
  }
  int64_t ptr = 0;
  std::vector<int64_t> pd(N), sst(N), dst(N);
  for (int64_t f = 0; f < FM.F.nblocks; ++f) {
    const int64_t rlim = FM.F.size(f)[0], clim = FM.F.size(f)[1];
    const int64_t foff = FM.poff[f], ldf = FM.ld[f];
    int64_t pc = 0;
    while (pc < clim) {
      int64_t pr = 0, fdc = 0;
      while (pr < rlim) {
        if (ptr >= nb) throw Error(ERR_SECTOR_MISMATCH, "SymMatrix: blocks exhausted");
        const int64_t b = semits[ptr];
        const int64_t* bz = A.size(b);
        std::vector<int64_t> st = colmajor_strides(bz, N);
        for (int t = 0; t < N; ++t) {
          pd[t] = bz[idxperm[t] - 1];
          sst[t] = st[idxperm[t] - 1];
        }
        const int64_t fdr = prod(pd.data(), n);
        fdc = prod(pd.data() + n, N - n);
        if (pr + fdr > rlim || pc + fdc > clim) throw Error(ERR_SECTOR_MISMATCH, "SymMatrix: slab overflow");
        int64_t acc = 1;
        for (int t = 0; t < n; ++t) {
          dst[t] = transposed ? acc * ldf : acc;
          acc *= pd[t];
        }
        acc = 1;
        for (int t = n; t < N; ++t) {
          dst[t] = transposed ? acc : acc * ldf;
          acc *= pd[t];
        }
        const int64_t doff = dst_base + foff + (transposed ? pr * ldf + pc : pr + pc * ldf);
        FM.gather.add(N, pd.data(), sst.data(), dst.data(), A.offsets[b], doff);
        pr += fdr;
        ++ptr;
      }
      pc += fdc;
    }
  }
  if (ptr != nb) throw Error(ERR_SECTOR_MISMATCH, "SymMatrix: blocks left over");
  FM.gather.finalize();
  return FM;
}

static std::vector<int> invperm(const std::vector<int>& p) {
  std::vector<int> r(p.size());
  for (size_t i = 0; i < p.size(); ++i) r[p[i] - 1] = (int)i + 1;
  return r;
}

// ---------------------------------------------------------------------------
// contract (src/contract.jl:11-47 with :234-298 and :187-224)
// ---------------------------------------------------------------------------
ContractPlan contract_plan(const Struct& A, const std::vector<int>& idxA, const Struct& B,
                           const std::vector<int>& idxB) {
  if ((int)idxA.size() != A.rank() || (int)idxB.size() != B.rank())
    throw Error(ERR_INVALID, "index set length differs from tensor rank");
  IndexPerm pa = contract_index_perm(idxA), pb = contract_index_perm(idxB);
  if (pa.cons.size() != pb.cons.size()) throw Error(ERR_INVALID, "contracted label counts differ");
  if (pa.rems.empty() || pb.rems.empty())
    throw Error(ERR_UNSUPPORTED, "rank-1 output branches of contract (src/contract.jl:24-42)");
  ContractPlan P;
  P.A = A;
  P.B = B;
  P.fa_off = 0;
  P.FA = symmatrix_plan(A, pa.rems, pa.cons, true, P.fa_off, 4);
  P.fb_off = P.FA.numel_padded();
  P.FB = symmatrix_plan(B, pb.cons, pb.rems, false, P.fb_off, 4);
  if (!isdual_pair(P.FA.F.legs[1], P.FB.F.legs[0])) throw Error(ERR_SPACE_MISMATCH, "not contractible!");
  const int64_t charge = A.charge + B.charge;
  P.FC = make_struct(charge, {P.FA.F.legs[0], P.FB.F.legs[1]});
  P.fc_off = P.fb_off + P.FB.numel_padded();
  P.fc_off = (P.fc_off + 3) / 4 * 4;
  P.ws_elems = P.fc_off + P.FC.numel();
  const Leg& a0 = P.FA.F.legs[0];
  const Leg& a1 = P.FA.F.legs[1];
  const Leg& b1 = P.FB.F.legs[1];
  for (int64_t f = 0; f < P.FC.nblocks; ++f) {
    const int64_t c1 = P.FC.sector(f)[0], c2 = P.FC.sector(f)[1];
    const int64_t M = P.FC.size(f)[0], Nn = P.FC.size(f)[1];
    const int64_t _c1 = a0.isdual ? -c1 : c1;
    const int64_t _c2 = b1.isdual ? -c2 : c2;
    std::vector<int64_t> sA, sB;
    if (a1.isdual) {
      sA = {c1, -A.charge + _c1};
      sB = {B.charge - _c2, c2};
    } else {
      sA = {c1, A.charge - _c1};
      sB = {-B.charge + _c2, c2};
    }
    int64_t ia = P.FA.F.find(sA), ib = P.FB.F.find(sB);
    P.gemm.add_prob(M, Nn);
    P.gemm.add_crows_affine(P.FC.offsets[f], M, M);
    if (ia >= 0 && ib >= 0) {
      const int64_t K = P.FA.F.size(ia)[1];
      if (P.FB.F.size(ib)[0] != K) throw Error(ERR_SECTOR_MISMATCH, "fused contraction dims differ");
      P.gemm.add_seg(K, P.FA.poff[ia], P.FA.ld[ia], P.FB.poff[ib], P.FB.ld[ib]);
    }
  }
  // split + final permutation in one scatter
  std::vector<Leg> rlegs, clegs, pre;
  for (int i : pa.rems) rlegs.push_back(A.legs[i - 1]);
  for (int i : pb.rems) clegs.push_back(B.legs[i - 1]);
  pre = rlegs;
  pre.insert(pre.end(), clegs.begin(), clegs.end());
  std::vector<int> tof(pa.tofinals);
  tof.insert(tof.end(), pb.tofinals.begin(), pb.tofinals.end());
  const int R = (int)pre.size();
  {
    std::vector<int> s(tof);
    std::sort(s.begin(), s.end());
    for (int i = 0; i < R; ++i)
      if (s[i] != i + 1) throw Error(ERR_INVALID, "output labels must be 1..K");
  }
  std::vector<int> Pm = invperm(tof);  // final leg k = pre leg Pm[k]
  std::vector<Leg> fin;
  for (int k = 0; k < R; ++k) fin.push_back(pre[Pm[k] - 1]);
  P.C = make_struct(charge, fin);
  const int nA = (int)rlegs.size();
  std::vector<char> covered(P.C.nblocks, 0);
  std::vector<int64_t> rs, rz, cs_, cz;
  std::vector<int64_t> dims(R), sst(R), dstv(R), prez(R), key(R);
  for (int64_t f = 0; f < P.FC.nblocks; ++f) {
    int64_t c1 = P.FC.sector(f)[0], c2 = P.FC.sector(f)[1];
    c1 = P.FC.legs[0].isdual ? -c1 : c1;
    c2 = P.FC.legs[1].isdual ? -c2 : c2;
    const int64_t Mf = P.FC.size(f)[0];
    all_sectors(c1, rlegs, rs, rz);
    all_sectors(c2, clegs, cs_, cz);
    const int nr = nA, nc = R - nA;
    const int64_t npr = nr ? (int64_t)rs.size() / nr : 0, npc = nc ? (int64_t)cs_.size() / nc : 0;
    int64_t pc = 0;
    for (int64_t j = 0; j < npc; ++j) {
      const int64_t csl = prod(&cz[j * nc], nc);
      int64_t pr = 0;
      for (int64_t i = 0; i < npr; ++i) {
        const int64_t rsl = prod(&rz[i * nr], nr);
        for (int t = 0; t < nr; ++t) prez[t] = rz[i * nr + t];
        for (int t = 0; t < nc; ++t) prez[nr + t] = cz[j * nc + t];
        std::vector<int64_t> prekey(R);
        for (int t = 0; t < nr; ++t) prekey[t] = rs[i * nr + t];
        for (int t = 0; t < nc; ++t) prekey[nr + t] = cs_[j * nc + t];
        for (int k = 0; k < R; ++k) key[k] = prekey[Pm[k] - 1];
        const int64_t cb = P.C.find(key);
        if (cb < 0) throw Error(ERR_SECTOR_MISMATCH, "split produced a sector outside the output space");
        if (covered[cb]) throw Error(ERR_SECTOR_MISMATCH, "split produced a sector twice");
        covered[cb] = 1;
        // pre-permutation strides inside the fused C block (column-major, ld Mf)
        std::vector<int64_t> pst(R);
        int64_t acc = 1;
        for (int t = 0; t < nr; ++t) {
          pst[t] = acc;
          acc *= prez[t];
        }
        acc = Mf;
        for (int t = nr; t < R; ++t) {
          pst[t] = acc;
          acc *= prez[t];
        }
        acc = 1;
        for (int k = 0; k < R; ++k) {
          dims[k] = prez[Pm[k] - 1];
          sst[k] = pst[Pm[k] - 1];
          dstv[k] = acc;
          acc *= dims[k];
        }
        P.scatter.add(R, dims.data(), sst.data(), dstv.data(), P.fc_off + P.FC.offsets[f] + pr + pc * Mf,
                      P.C.offsets[cb]);
        pr += rsl;
      }
      pc += csl;
    }
  }
  for (int64_t b = 0; b < P.C.nblocks; ++b)
    if (!covered[b]) throw Error(ERR_SECTOR_MISMATCH, "output sector not produced by the split");
  P.scatter.finalize();
  return P;
}

ContractCost contract_cost(const Struct& A, const std::vector<int>& idxA, const Struct& B,
                           const std::vector<int>& idxB) {
  if ((int)idxA.size() != A.rank() || (int)idxB.size() != B.rank())
    throw Error(ERR_INVALID, "index set length differs from tensor rank");
  IndexPerm pa = contract_index_perm(idxA), pb = contract_index_perm(idxB);
  if (pa.cons.size() != pb.cons.size()) throw Error(ERR_INVALID, "contracted label counts differ");
  if (pa.rems.empty() || pb.rems.empty())
    throw Error(ERR_UNSUPPORTED, "rank-1 output branches of contract (src/contract.jl:24-42)");
  std::vector<Leg> ar, ac, bc, br;
  for (int i : pa.rems) ar.push_back(A.legs[i - 1]);
  for (int i : pa.cons) ac.push_back(A.legs[i - 1]);
  for (int i : pb.cons) bc.push_back(B.legs[i - 1]);
  for (int i : pb.rems) br.push_back(B.legs[i - 1]);
  // the fused legs of SymMatrix(A, rems, cons) and SymMatrix(B, cons, rems)
  const Leg a0 = fuse(false, ar), a1 = fuse(true, ac), b0 = fuse(false, bc), b1 = fuse(true, br);
  if (!isdual_pair(a1, b0)) throw Error(ERR_SPACE_MISMATCH, "not contractible!");
  const int64_t charge = A.charge + B.charge;
  ContractCost R;
  const Struct FC = make_struct(charge, {a0, b1});
  for (int64_t f = 0; f < FC.nblocks; ++f) {
    const int64_t c1 = FC.sector(f)[0], c2 = FC.sector(f)[1];
    const int64_t _c1 = a0.isdual ? -c1 : c1;
    const int64_t _c2 = b1.isdual ? -c2 : c2;
    const int64_t ka = a1.isdual ? -A.charge + _c1 : A.charge - _c1;
    const int64_t kb = a1.isdual ? B.charge - _c2 : -B.charge + _c2;
    const int64_t K = a1.dim_of(ka);
    if (K > 0 && b0.dim_of(kb) > 0) {
      if (b0.dim_of(kb) != K) throw Error(ERR_SECTOR_MISMATCH, "fused contraction dims differ");
      R.flops += 2 * FC.size(f)[0] * FC.size(f)[1] * K;
    }
  }
  std::vector<Leg> pre = ar;
  pre.insert(pre.end(), br.begin(), br.end());
  std::vector<int> tof(pa.tofinals);
  tof.insert(tof.end(), pb.tofinals.begin(), pb.tofinals.end());
  std::vector<int> Pm = invperm(tof);
  std::vector<Leg> fin;
  for (size_t k = 0; k < pre.size(); ++k) fin.push_back(pre[Pm[k] - 1]);
  R.C = make_struct(charge, fin);
  return R;
}

// ---------------------------------------------------------------------------
// Re-legging (src/symreleg.jl)
// ---------------------------------------------------------------------------
RelegPlan permutelegs_plan(const Struct& A, const std::vector<int>& perm) {
  const int N = A.rank();
  {
    std::vector<int> s(perm);
    std::sort(s.begin(), s.end());
    if ((int)s.size() != N) throw Error(ERR_INVALID, "Not a valid permutation!");
    for (int i = 0; i < N; ++i)
      if (s[i] != i + 1) throw Error(ERR_INVALID, "Not a valid permutation!");
  }
  RelegPlan P;
  P.in = A;
  std::vector<Leg> legs;
  for (int k = 0; k < N; ++k) legs.push_back(A.legs[perm[k] - 1]);
  P.out = make_struct(A.charge, legs);
  std::vector<int64_t> dims(N), sst(N), dst(N), key(N);
  std::vector<char> cov(P.out.nblocks, 0);
  for (int64_t b = 0; b < A.nblocks; ++b) {
    std::vector<int64_t> st = colmajor_strides(A.size(b), N);
    int64_t acc = 1;
    for (int k = 0; k < N; ++k) {
      key[k] = A.sector(b)[perm[k] - 1];
      dims[k] = A.size(b)[perm[k] - 1];
      sst[k] = st[perm[k] - 1];
      dst[k] = acc;
      acc *= dims[k];
    }
    int64_t ob = P.out.find(key);
    if (ob < 0 || cov[ob]) throw Error(ERR_SECTOR_MISMATCH, "permutelegs: sector mismatch");
    cov[ob] = 1;
    P.copy.add(N, dims.data(), sst.data(), dst.data(), A.offsets[b], P.out.offsets[ob]);
  }
  for (char c : cov)
    if (!c) throw Error(ERR_SECTOR_MISMATCH, "permutelegs: sector mismatch");
  P.copy.finalize();
  return P;
}

RelegPlan fuselegs_plan(const Struct& A, int l, int n, bool dual) {
  const int N = A.rank();
  if (!(0 < l && l < N + 2 - n) || n < 1) throw Error(ERR_INVALID, "fuselegs: leg range out of bounds");
  const int lo = l - 1, hi = l - 1 + n;
  std::vector<bool> dualinfo;
  std::vector<Leg> fl;
  for (int t = lo; t < hi; ++t) {
    dualinfo.push_back(A.legs[t].isdual);
    fl.push_back(A.legs[t]);
  }
  const int64_t nb = A.nblocks;
  std::vector<std::vector<int64_t>> fsects(nb);
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t* s = A.sector(b);
    int64_t fc = flow(s + lo, dualinfo);
    if (dual) fc = -fc;
    std::vector<int64_t> k(s, s + lo);
    k.push_back(fc);
    k.insert(k.end(), s + hi, s + N);
    fsects[b] = k;
  }
  std::vector<int64_t> sperm(nb);
  std::iota(sperm.begin(), sperm.end(), 0);
  std::stable_sort(sperm.begin(), sperm.end(), [&](int64_t a, int64_t b) { return rev_less(fsects[a], fsects[b]); });
  std::vector<Leg> legs(A.legs.begin(), A.legs.begin() + lo);
  legs.push_back(fuse(dual, fl));
  legs.insert(legs.end(), A.legs.begin() + hi, A.legs.end());
  RelegPlan P;
  P.in = A;
  P.out = make_struct(A.charge, legs);
  const int No = P.out.rank();
  int64_t ptr = 0;
  std::vector<int64_t> sst(N), dst(N);
  for (int64_t ob = 0; ob < P.out.nblocks; ++ob) {
    std::vector<int64_t> ost = colmajor_strides(P.out.size(ob), No);
    const int64_t sizel = P.out.size(ob)[lo];
    int64_t p = 0;
    while (p < sizel) {
      if (ptr >= nb) throw Error(ERR_SECTOR_MISMATCH, "fuselegs: blocks exhausted");
      const int64_t b = sperm[ptr];
      const int64_t* bz = A.size(b);
      const int64_t fd = prod(bz + lo, n);
      std::vector<int64_t> st = colmajor_strides(bz, N);
      for (int t = 0; t < N; ++t) {
        sst[t] = st[t];
        if (t < lo)
          dst[t] = ost[t];
        else if (t >= hi)
          dst[t] = ost[t - n + 1];
        else {
          int64_t acc = ost[lo];
          for (int u = lo; u < t; ++u) acc *= bz[u];
          dst[t] = acc;
        }
      }
      P.copy.add(N, bz, sst.data(), dst.data(), A.offsets[b], P.out.offsets[ob] + p * ost[lo]);
      p += fd;
      ++ptr;
    }
  }
  if (ptr != nb) throw Error(ERR_SECTOR_MISMATCH, "fuselegs: blocks left over");
  P.copy.finalize();
  return P;
}

RelegPlan splitleg_plan(const Struct& A, int l, const std::vector<Leg>& spaces) {
  const int N = A.rank();
  if (!(0 < l && l <= N)) throw Error(ERR_INVALID, "splitleg: leg out of range");
  const int M = (int)spaces.size();
  if (M == 0) throw Error(ERR_INVALID, "splitleg: no spaces");
  std::vector<Leg> legs(A.legs.begin(), A.legs.begin() + (l - 1));
  legs.insert(legs.end(), spaces.begin(), spaces.end());
  legs.insert(legs.end(), A.legs.begin() + l, A.legs.end());
  RelegPlan P;
  P.in = A;
  if (M == 1 && A.legs[l - 1] != spaces[0]) throw Error(ERR_SPACE_MISMATCH, "SpaceMismatch in splitleg");
  P.out = make_struct(A.charge, legs);
  const int No = P.out.rank();
  std::vector<char> cov(P.out.nblocks, 0);
  std::map<int64_t, std::pair<std::vector<int64_t>, std::vector<int64_t>>> cache;
  std::vector<int64_t> dims(No), sst(No), dst(No), key(No);
  for (int64_t b = 0; b < A.nblocks; ++b) {
    const int64_t* s = A.sector(b);
    const int64_t* bz = A.size(b);
    const int64_t charge = A.legs[l - 1].isdual ? -s[l - 1] : s[l - 1];
    if (!cache.count(charge)) {
      std::vector<int64_t> ps, pz;
      all_sectors(charge, spaces, ps, pz);
      cache[charge] = {ps, pz};
    }
    const auto& ps = cache[charge].first;
    const auto& pz = cache[charge].second;
    std::vector<int64_t> st = colmajor_strides(bz, N);
    int64_t pivot = 0;
    const int64_t np = (int64_t)ps.size() / M;
    for (int64_t j = 0; j < np; ++j) {
      const int64_t sl = prod(&pz[j * M], M);
      if (pivot + sl > bz[l - 1]) throw Error(ERR_SECTOR_MISMATCH, "splitleg: patterns exceed the fused dim");
      int k = 0;
      for (int t = 0; t < l - 1; ++t, ++k) {
        key[k] = s[t];
        dims[k] = bz[t];
        sst[k] = st[t];
      }
      int64_t acc = st[l - 1];
      for (int u = 0; u < M; ++u, ++k) {
        key[k] = ps[j * M + u];
        dims[k] = pz[j * M + u];
        sst[k] = acc;
        acc *= pz[j * M + u];
      }
      for (int t = l; t < N; ++t, ++k) {
        key[k] = s[t];
        dims[k] = bz[t];
        sst[k] = st[t];
      }
      int64_t ob = P.out.find(key);
      if (ob < 0 || cov[ob]) throw Error(ERR_SECTOR_MISMATCH, "splitleg: sector mismatch");
      cov[ob] = 1;
      acc = 1;
      for (int t = 0; t < No; ++t) {
        dst[t] = acc;
        acc *= dims[t];
      }
      P.copy.add(No, dims.data(), sst.data(), dst.data(), A.offsets[b] + pivot * st[l - 1], P.out.offsets[ob]);
      pivot += sl;
    }
  }
  for (char c : cov)
    if (!c) throw Error(ERR_SECTOR_MISMATCH, "splitleg: output sector not produced");
  P.copy.finalize();
  return P;
}

RelegPlan symmatrix_releg(const Struct& A, const std::vector<int>& rows, const std::vector<int>& cols) {
  RelegPlan P;
  P.in = A;
  FusedMatrix F = symmatrix_plan(A, rows, cols, false, 0);
  P.out = F.F;
  P.copy = F.gather;
  return P;
}

// ---------------------------------------------------------------------------
// Fused H_eff plan and the environment plans built on its first stages
// ---------------------------------------------------------------------------
typedef std::vector<int64_t> Key;

static void need_pair(const Leg& a, const Leg& b, const char* what) {
  if (!isdual_pair(a, b)) throw Error(ERR_SPACE_MISMATCH, std::string("not contractible! ") + what);
}
static void need_same(const Leg& a, const Leg& b, const char* what) {
  if (a != b) throw Error(ERR_SPACE_MISMATCH, std::string("not contractible! ") + what);
}

// K-contiguous operand rows are padded to multiples of 8 elements: 16-byte
// aligned rows for the DMA-streamed GEMMs in every storage type (BF16 included)
static int64_t up8(int64_t x) { return (x + 7) / 8 * 8; }
static int64_t up64(int64_t x) { return (x + 63) / 64 * 64; }  // 256-byte (FP32) region starts

namespace {
// One stored block pair of the step-1 GEMM output T1_k = v_k^T envL_k: key (b, w, s1, s2, r).
struct T1B {
  Key key;
  int64_t base, rowoff, coloff, Mf, db, Dw, d1, d2, dR, dRp;  // dRp: dR rounded up to the plan's row alignment
};
// One T3 cell (b, s1o, s2o, x, r): the MPO pass writes output row rr = jx + Dx*(i1o + d1o*i2o)
// of pixel (m, n) (m the bra_l index, n the ket_r index) at
//   base + jx*sx + (i1o + d1o*i2o)*so + n*sn + m*sm.
struct Cell {
  int64_t base, sx, so, sn, sm, db, d1o, d2o, Dx, dR;
};
}  // namespace

// ---- step 1: one fused GEMM per ket charge k (the reference's fused sizes),
// T1_k = [v blocks (k,.)]^T x [envL blocks (k,.)], stored column-major (ld M_k
// rounded up to 32).  Operand rows (v block columns, envL block columns) are
// packed into K-contiguous rows of stride Kp = roundup(dk, 4): v per matvec,
// envL per bond.  padr > 1: each v block's rows hold the right-bond index in
// segments of dRp = roundup(dR, padr) rows (rows past dR are not written and
// their T1 rows never read), so every MPO-pass channel row starts on a
// multiple of padr.  Fills H.g1, H.pack_v, H.pack_envL, H.t1_elems, H.vp_elems,
// H.el_elems, H.padr and the T1 block list (sorted by (b, r, w, s1, s2)).
static void step1_layout(Step13& H, const Struct& v, const Struct& envL, std::vector<T1B>& t1, int padr) {
  H.padr = padr;
  auto padded = [&](int64_t dR) { return (dR + padr - 1) / padr * padr; };
  std::map<int64_t, std::vector<int64_t>> vby, eby;
  for (int64_t x = 0; x < v.nblocks; ++x) vby[v.sector(x)[0]].push_back(x);
  for (int64_t e = 0; e < envL.nblocks; ++e) eby[envL.sector(e)[0]].push_back(e);
  int64_t off = 0;
  H.vp_elems = H.el_elems = 0;
  for (auto& kv : eby) {
    auto itv = vby.find(kv.first);
    if (itv == vby.end()) continue;
    const std::vector<int64_t>& X = itv->second;
    const std::vector<int64_t>& E = kv.second;
    const int64_t dk = envL.size(E[0])[0], Kp = up8(dk);
    std::vector<int64_t> rowoff(X.size()), coloff(E.size());
    int64_t Mf = 0, Nf = 0;
    for (size_t i = 0; i < X.size(); ++i) {
      const int64_t* z = v.size(X[i]);
      rowoff[i] = Mf;
      Mf += z[1] * z[2] * padded(z[3]);
    }
    for (size_t j = 0; j < E.size(); ++j) {
      const int64_t* z = envL.size(E[j]);
      coloff[j] = Nf;
      Nf += z[1] * z[2];
    }
    const int64_t vbase = H.vp_elems, ebase = H.el_elems;
    H.vp_elems += Mf * Kp;
    H.el_elems += Nf * Kp;
    // v block (dk, d1, d2, dR) -> packed rows ordered ir + dR*(i1 + d1*i2): the
    // right-bond index runs fastest, so the MPO pass reads T1 with unit stride
    for (size_t i = 0; i < X.size(); ++i) {
      const int64_t* z = v.size(X[i]);
      const int64_t d1 = z[1], d2 = z[2], dR = z[3], dRp = padded(dR);
      const int64_t dims[4] = {dk, d1, d2, dR}, sst[4] = {1, dk, dk * d1, dk * d1 * d2},
                    dst[4] = {1, dRp * Kp, dRp * d1 * Kp, Kp};
      H.pack_v.add(4, dims, sst, dst, v.offsets[X[i]], vbase + rowoff[i] * Kp);
    }
    for (size_t j = 0; j < E.size(); ++j) {
      const int64_t* z = envL.size(E[j]);
      const int64_t dims[2] = {dk, z[1] * z[2]}, sst[2] = {1, dk}, dst[2] = {1, Kp};
      H.pack_envL.add(2, dims, sst, dst, envL.offsets[E[j]], ebase + coloff[j] * Kp);
    }
    // columns of T1_k start on 128-byte boundaries (ld = Mf rounded up to 32)
    const int64_t Mld = (Mf + 31) / 32 * 32;
    H.g1.add_prob(Mf, Nf);
    H.g1.add_crows_affine(off, Mld, Mf);
    H.g1.add_seg(dk, vbase, Kp, ebase, Kp);
    for (size_t i = 0; i < X.size(); ++i)
      for (size_t j = 0; j < E.size(); ++j) {
        const int64_t* xs = v.sector(X[i]);
        const int64_t* es = envL.sector(E[j]);
        T1B t;
        t.key = {es[2], es[1], xs[1], xs[2], xs[3]};
        t.base = off;
        t.rowoff = rowoff[i];
        t.coloff = coloff[j];
        t.Mf = Mld;
        t.db = envL.size(E[j])[2];
        t.Dw = envL.size(E[j])[1];
        t.d1 = v.size(X[i])[1];
        t.d2 = v.size(X[i])[2];
        t.dR = v.size(X[i])[3];
        t.dRp = padded(t.dR);
        t1.push_back(t);
      }
    off += Mld * Nf;
  }
  H.t1_elems = off;
  std::stable_sort(t1.begin(), t1.end(), [](const T1B& a, const T1B& b) {
    return std::make_tuple(a.key[0], a.key[4], a.key[1], a.key[2], a.key[3]) <
           std::make_tuple(b.key[0], b.key[4], b.key[1], b.key[2], b.key[3]);
  });
}

// ---- steps 2+3: pixel groups (b, r); their out types are the T3 cells of (b, r).
// Per group the combined two-site MPO W_group = WL (x) WR over the middle MPO bond
// maps the C input channels (w, s1, s2) of T1 to the R output rows (s1o, s2o, x)
// of the cells; groups with the same channel signature share one W_group.
static void mpo_layout(Step13& H, const Struct& WL, const Struct& WR, const std::vector<T1B>& t1,
                       const std::map<Key, Cell>& cells) {
  auto wl_by = [&]() {
    std::map<std::pair<int64_t, int64_t>, std::vector<int64_t>> m;  // (w, s1) -> WL blocks
    for (int64_t b = 0; b < WL.nblocks; ++b) m[{WL.sector(b)[0], WL.sector(b)[3]}].push_back(b);
    return m;
  }();
  auto wr_by = [&]() {
    std::map<std::pair<int64_t, int64_t>, std::vector<int64_t>> m;  // (m, s2) -> WR blocks
    for (int64_t b = 0; b < WR.nblocks; ++b) m[{WR.sector(b)[0], WR.sector(b)[3]}].push_back(b);
    return m;
  }();
  std::map<std::pair<int64_t, int64_t>, std::vector<Key>> cells_by_group;  // (b, r) -> out types
  for (auto& c : cells) cells_by_group[{c.first[0], c.first[4]}].push_back({c.first[1], c.first[2], c.first[3]});
  std::map<std::pair<int64_t, int64_t>, std::pair<size_t, size_t>> t1_range;
  for (size_t i = 0; i < t1.size();) {
    size_t j = i;
    while (j < t1.size() && t1[j].key[0] == t1[i].key[0] && t1[j].key[4] == t1[i].key[4]) ++j;
    t1_range[{t1[i].key[0], t1[i].key[4]}] = {i, j};
    i = j;
  }
  std::map<std::vector<int64_t>, int32_t> sigindex;
  int64_t w_off = 0;
  for (auto& gc : cells_by_group) {
    const int64_t b = gc.first.first, r = gc.first.second;
    const std::vector<Key>& outtypes = gc.second;  // sorted (map order)
    size_t gi0 = 0, gi1 = 0;
    auto itr = t1_range.find(gc.first);
    if (itr != t1_range.end()) {
      gi0 = itr->second.first;
      gi1 = itr->second.second;
    }
    auto cell = [&](const Key& ot) -> const Cell& { return cells.at({b, ot[0], ot[1], ot[2], r}); };
    const Cell& c0 = cell(outtypes[0]);
    const int64_t db = c0.db, dR = c0.dR;
    Key sigkey;
    for (size_t i = gi0; i < gi1; ++i) sigkey.insert(sigkey.end(), {t1[i].key[1], t1[i].key[2], t1[i].key[3]});
    sigkey.push_back(INT64_MIN);
    for (auto& ot : outtypes) sigkey.insert(sigkey.end(), ot.begin(), ot.end());
    std::vector<int64_t> col0s, row0s;
    int64_t C = 0, R = 0;
    for (size_t i = gi0; i < gi1; ++i) {
      col0s.push_back(C);
      C += t1[i].Dw * t1[i].d1 * t1[i].d2;
    }
    for (auto& ot : outtypes) {
      row0s.push_back(R);
      const Cell& cc = cell(ot);
      R += cc.Dx * cc.d1o * cc.d2o;
    }
    int32_t sig;
    auto its = sigindex.find(sigkey);
    if (its == sigindex.end()) {
      sig = (int32_t)H.sigs.size();
      sigindex[sigkey] = sig;
      H.sigs.push_back(MpoSig{(int32_t)R, (int32_t)C, w_off});
      w_off += R * C;
      H.maxR = std::max<int>(H.maxR, (int)R);
      H.maxC = std::max<int>(H.maxC, (int)C);
      for (size_t oi = 0; oi < outtypes.size(); ++oi) {
        const Key& ot = outtypes[oi];
        const Cell& cc = cell(ot);
        for (size_t i = gi0; i < gi1; ++i) {
          const Key& k = t1[i].key;
          MpoTerm tm{};
          tm.sig = sig;
          tm.row0 = (int32_t)row0s[oi];
          tm.col0 = (int32_t)col0s[i - gi0];
          tm.Dx = (int32_t)cc.Dx;
          tm.d1o = (int32_t)cc.d1o;
          tm.d2o = (int32_t)cc.d2o;
          tm.Dw = (int32_t)t1[i].Dw;
          tm.d1i = (int32_t)t1[i].d1;
          tm.d2i = (int32_t)t1[i].d2;
          tm.R = tm.Dx * tm.d1o * tm.d2o;
          tm.C = tm.Dw * tm.d1i * tm.d2i;
          tm.pair0 = (int32_t)H.pairs.size();
          tm.npair = 0;
          auto itl = wl_by.find({k[1], k[2]});
          if (itl != wl_by.end())
            for (int64_t wl : itl->second) {
              if (WL.sector(wl)[1] != ot[0]) continue;
              const int64_t m = WL.sector(wl)[2];
              auto itw = wr_by.find({m, k[3]});
              if (itw == wr_by.end()) continue;
              for (int64_t wr : itw->second) {
                if (WR.sector(wr)[1] != ot[1] || WR.sector(wr)[2] != ot[2]) continue;
                H.pairs.push_back(MpoPair{WL.offsets[wl], WR.offsets[wr], (int32_t)WL.size(wl)[2], 0});
                tm.npair++;
              }
            }
          H.terms.push_back(tm);
        }
      }
    } else {
      sig = its->second;
    }
    const int64_t dRp = (dR + H.padr - 1) / H.padr * H.padr;
    MpoGroup g{};
    g.sig = sig;
    g.dR = (int32_t)dR;
    g.db = (int32_t)db;
    g.dRp = (int32_t)dRp;
    g.in0 = (int32_t)H.inrows.size();
    g.out0 = (int32_t)H.outrows.size();
    g.npix = db * dR;
    for (size_t i = gi0; i < gi1; ++i) {
      const T1B& t = t1[i];
      if (t.db != db || t.dR != dR || t.dRp != dRp)
        throw Error(ERR_SECTOR_MISMATCH, "mpo pass: pixel group dims differ");
      // column order c = jw + Dw*(i1 + d1*i2); T1 element (n + dRp (i1 + d1 i2), jw + Dw m)
      for (int64_t c = 0; c < t.Dw * t.d1 * t.d2; ++c) {
        const int64_t jw = c % t.Dw, i1 = (c / t.Dw) % t.d1, i2 = c / (t.Dw * t.d1);
        H.inrows.push_back(MpoInRow{t.base + t.rowoff + t.dRp * (i1 + t.d1 * i2) + t.Mf * (t.coloff + jw), 1,
                                    t.Mf * t.Dw});
      }
    }
    for (auto& ot : outtypes) {
      const Cell& cc = cell(ot);
      if (cc.db != db || cc.dR != dR) throw Error(ERR_SECTOR_MISMATCH, "mpo pass: cell dims differ");
      // row order rr = jx + Dx*(i1o + d1o*i2o)
      for (int64_t rr = 0; rr < cc.Dx * cc.d1o * cc.d2o; ++rr) {
        const int64_t jx = rr % cc.Dx, io = rr / cc.Dx;
        H.outrows.push_back(MpoOutRow{cc.base + cc.sx * jx + cc.so * io, cc.sn, cc.sm});
      }
    }
    const int32_t gidx = (int32_t)H.groups.size();
    H.groups.push_back(g);
    // runs cover the group's padded range dRp (== dR in unpadded plans)
    const bool pad = H.padr > 1;
    for (int i = 0; i < 5; ++i) {
      if (i == 4 && !pad) break;
      const int P = i < 3 ? kMpoRuns[i]
                          : (pad ? mpo_vec_run((int)C, i == 3 ? 4 : 8) : mpo_mixed_run((int)C));
      const int64_t runs = db * ((dRp + P - 1) / P);
      for (int64_t r0 = 0; r0 < runs; r0 += 256) H.work[i].push_back(MpoWork{gidx, (int32_t)r0});
    }
    H.mpo_flops += 2 * R * C * g.npix;
  }
  H.w_elems = w_off;
}

static void check_two_site(const Struct& v, const Struct& envL, const Struct& WL, const Struct& WR) {
  if (v.rank() != 4 || envL.rank() != 3 || WL.rank() != 4 || WR.rank() != 4)
    throw Error(ERR_INVALID, "ranks must be v 4, env 3, WL 4, WR 4");
  need_pair(envL.legs[0], v.legs[0], "(envL ket, v left)");
  need_pair(envL.legs[1], WL.legs[0], "(envL mpo, WL left)");
  need_pair(v.legs[1], WL.legs[3], "(v p1, WL p_in)");
  need_pair(WL.legs[2], WR.legs[0], "(WL right, WR left)");
  need_pair(v.legs[2], WR.legs[3], "(v p2, WR p_in)");
}

HeffPlan heff_plan(const Struct& v, const Struct& envL, const Struct& envR, const Struct& WL, const Struct& WR,
                   const HeffTiles& tiles) {
  if (envR.rank() != 3) throw Error(ERR_INVALID, "heff: ranks must be v 4, envL 3, envR 3, WL 4, WR 4");
  check_two_site(v, envL, WL, WR);
  need_pair(WR.legs[2], envR.legs[1], "(WR right, envR mpo)");
  need_pair(v.legs[3], envR.legs[0], "(v right, envR ket)");
  HeffPlan H;
  H.v = v;
  H.envL = envL;
  H.envR = envR;
  H.WL = WL;
  H.WR = WR;
  const int64_t charge = envL.charge + v.charge + WL.charge + WR.charge + envR.charge;
  H.out = make_struct(charge, {envL.legs[2], WL.legs[1], WR.legs[1], envR.legs[2]});

  // reference-order cost from the reference's own four contracts
  {
    ContractCost P1 = contract_cost(envL, {-1, 2, 1}, v, {-1, 3, 4, 5});
    ContractCost P2 = contract_cost(P1.C, {1, -1, -2, 4, 5}, WL, {-1, 2, 3, -2});
    ContractCost P3 = contract_cost(P2.C, {1, 2, -1, -2, 5}, WR, {-1, 3, 4, -2});
    ContractCost P4 = contract_cost(P3.C, {1, 2, 3, -1, -2}, envR, {-2, -1, 4});
    const ContractCost* Ps[4] = {&P1, &P2, &P3, &P4};
    const Struct* ins[4][2] = {{&envL, &v}, {&P1.C, &WL}, {&P2.C, &WR}, {&P3.C, &envR}};
    for (int s = 0; s < 4; ++s) {
      H.step_flops[s] = Ps[s]->flops;
      H.step_elems[s] = ins[s][0]->numel() + ins[s][1]->numel() + Ps[s]->C.numel();
      H.ref_flops += H.step_flops[s];
      H.ref_elems += H.step_elems[s];
    }
    if (!same_struct(P4.C, H.out) || P4.C.sectors != H.out.sectors)
      throw Error(ERR_SECTOR_MISMATCH, "heff: output structure differs from the contraction chain");
  }

  // T1 / T3 rows aligned to kMpoPadR (MPSH_HEFF_PADR=0: unpadded), see step1_layout
  const char* pad_env = std::getenv("MPSH_HEFF_PADR");
  const int padr = pad_env && std::atoi(pad_env) == 0 ? 1 : kMpoPadR;
  std::vector<T1B> t1;
  step1_layout(H, v, envL, t1, padr);

  // ---- step 4 layout: one fused GEMM per output ket charge r'; rows = the out
  // blocks (b, s1o, s2o, r'), K segments = envR blocks (r, x, r').  T3 region of
  // (r', f) is K_f x M_f, K-contiguous (row stride Kfp = roundup(K_f, 8), K_f =
  // dRp Dx); cell (o, f) is T3 block (b, s1o, s2o, x, r): element
  // (n + dRp jx, rowoff + m + db (i1o + d1o i2o)).
  int64_t er_elems = 0;
  std::map<Key, Cell> cells;  // key (b, s1o, s2o, x, r)
  {
    std::map<int64_t, std::vector<int64_t>> oby, fby;
    for (int64_t o = 0; o < H.out.nblocks; ++o) oby[H.out.sector(o)[3]].push_back(o);
    for (int64_t f = 0; f < envR.nblocks; ++f) fby[envR.sector(f)[2]].push_back(f);
    int64_t off = 0;
    for (auto& kv : oby) {
      const std::vector<int64_t>& O = kv.second;
      std::vector<int64_t> rowoff(O.size());
      int64_t Mf = 0;
      for (size_t i = 0; i < O.size(); ++i) {
        rowoff[i] = Mf;
        const int64_t* z = H.out.size(O[i]);
        Mf += z[0] * z[1] * z[2];
      }
      const int64_t N = H.out.size(O[0])[3];
      H.g4.add_prob(Mf, N);
      for (size_t i = 0; i < O.size(); ++i) {
        const int64_t* z = H.out.size(O[i]);
        const int64_t Mo = z[0] * z[1] * z[2];
        for (int64_t m = 0; m < Mo; ++m) H.g4.add_crow(H.out.offsets[O[i]] + m, Mo);
      }
      auto itf = fby.find(kv.first);
      if (itf == fby.end()) continue;
      for (int64_t f : itf->second) {
        const int64_t* fs = envR.sector(f);
        // K index n + dRp jx: segments of dRp per MPO index, zeros past dR (the
        // MPO pass writes them into T3, the prepare zero-fills envR_p first)
        const int64_t dR = envR.size(f)[0], Dx = envR.size(f)[1], dRp = (dR + padr - 1) / padr * padr;
        const int64_t Kf = dRp * Dx, Kfp = up8(Kf);
        const int64_t rbase = er_elems;
        er_elems += Kfp * N;
        {
          const int64_t dims[3] = {dR, Dx, N}, sst[3] = {1, dR, dR * Dx}, dst[3] = {1, dRp, Kfp};
          H.pack_envR.add(3, dims, sst, dst, envR.offsets[f], rbase);
        }
        H.g4.add_seg(Kf, off, Kfp, rbase, Kfp);
        for (size_t i = 0; i < O.size(); ++i) {
          const int64_t* os = H.out.sector(O[i]);
          const int64_t* z = H.out.size(O[i]);
          cells[{os[0], os[1], os[2], fs[1], fs[0]}] =
              Cell{off + Kfp * rowoff[i], dRp, Kfp * z[0], 1, Kfp, z[0], z[1], z[2], Dx, dR};
        }
        off += Kfp * Mf;
      }
    }
    H.t3_elems = off;
  }
  mpo_layout(H, WL, WR, t1, cells);

  // kernel variants and split-K choices (step 1 is never split: K = one bond charge)
  auto cfg = [](GemmPlan& g, const TileCfg& c) {
    g.BM = c.BM;
    g.BN = c.BN;
    g.KS = c.KS;
  };
  H.g1d = H.g1h = H.g1;
  H.g4d = H.g4h = H.g4;
  cfg(H.g1, tiles.s1);
  cfg(H.g4, tiles.s4);
  cfg(H.g1h, tiles.s1h);
  cfg(H.g4h, tiles.s4h);
  cfg(H.g1d, tiles.s1d);
  cfg(H.g4d, tiles.s4d);
  // K chunk of each kernel (chunk_k); BF16 runs the BF16-storage kernels (modes 3, 4)
  set_splits(H.g4, tiles.split_target, 1);
  set_splits(H.g4h, tiles.split_target_h, 4);
  set_splits(H.g4d, tiles.split_target_d, 0);
  H.pack_v.finalize();
  H.pack_envL.finalize();
  H.pack_envR.finalize();
  // prepared operator: [envL_p | envR_p | W]
  H.el_off = 0;
  H.er_off = up8(H.el_elems);
  H.pw_off = up8(H.er_off + er_elems);
  H.prep_elems = H.pw_off + H.w_elems;
  // per-matvec workspace: [v_p | T1 | T3 | partials]
  H.vp_off = 0;
  H.t1_off = up64(H.vp_elems);
  H.t3_off = up64(H.t1_off + H.t1_elems);
  H.part_off = up64(H.t3_off + H.t3_elems);
  // partials sized for the largest split-K layout any tuning candidate may pick
  int64_t part = std::max({H.g4.part_elems, H.g4h.part_elems, H.g4d.part_elems});
  for (int mode : {0, 1, 4})  // the step-4 kernel modes of FP64, FP32, BF16
    for (const TileCfg& c : kTuneTiles) {
      if (!tune_tile_ok(mode, c)) continue;
      for (int64_t sp : kTuneSplits) part = std::max(part, gemm_variant(H.g4, c, sp, mode).part_elems);
    }
  H.ws_elems = H.part_off + part;
  return H;
}

// ---------------------------------------------------------------------------
// Environment plans: out[x, y, x'] = sum env[a, m, b] K[a, p1, p2, x] WL WR K[b, q1, q2, x']
// (the left update of src/apply.jl:49-58 over one or two sites), computed as the
// H_eff's step 1 and MPO pass followed by one grouped GEMM over the bra side.
// ---------------------------------------------------------------------------
Struct dual_struct(const Struct& A) {
  std::vector<Leg> l;
  for (const Leg& x : A.legs) l.push_back(x.dual());
  return make_struct(-A.charge, l);
}

static Leg trivial_leg(bool dual) { return make_leg(dual, {{0, 1}}); }

// A rank-3 site tensor (l, p, r) as a rank-4 one (l, p, t, r) with a trivial leg t
// (charge 0, dim 1): the sectors keep their order and the blocks their layout.
static Struct with_trivial_leg(const Struct& A) {
  Struct S = make_struct(A.charge, {A.legs[0], A.legs[1], trivial_leg(false), A.legs[2]});
  if (S.nblocks != A.nblocks || S.offsets != A.offsets)
    throw Error(ERR_SECTOR_MISMATCH, "trivial leg changed the block layout");
  return S;
}

// identity MPO tensor (W.legs[2]^dual, t, W.legs[2], t^dual) continuing a one-site MPO W
static Struct identity_mpo(const Struct& W, std::vector<double>& data) {
  Struct I = make_struct(0, {W.legs[2].dual(), trivial_leg(false), W.legs[2], trivial_leg(true)});
  data.assign(I.numel(), 0.0);
  for (int64_t b = 0; b < I.nblocks; ++b) {
    const int64_t* z = I.size(b);
    if (I.sector(b)[0] != I.sector(b)[2] || z[0] != z[2]) throw Error(ERR_SECTOR_MISMATCH, "identity MPO block");
    for (int64_t i = 0; i < z[0]; ++i) data[I.offsets[b] + i + z[0] * i] = 1.0;
  }
  return I;
}

static EnvCore env_core(const Struct& env, const Struct& ket, const Struct& WL, const Struct& WR) {
  check_two_site(ket, env, WL, WR);
  // the bra is the ket with dual legs: its left, p1 and p2 legs must carry the
  // env's third leg and the MPO output legs (src/apply.jl:56, `dual(mat)`)
  need_same(env.legs[2], ket.legs[0], "(env bra, bra left)");
  need_same(WL.legs[1], ket.legs[1], "(WL p_out, bra p1)");
  need_same(WR.legs[1], ket.legs[2], "(WR p_out, bra p2)");
  EnvCore H;
  H.env = env;
  H.ket = ket;
  H.WL = WL;
  H.WR = WR;
  H.out = make_struct(env.charge + WL.charge + WR.charge, {ket.legs[3], WR.legs[2], ket.legs[3].dual()});
  std::vector<T1B> t1;
  step1_layout(H, ket, env, t1, 1);  // T3 rows run along the bra side here: no padding

  // ---- final GEMM: one problem per output block (r, x, y): rows m = ir + dR*ix,
  // columns the bra's right index iy, K = the bra-left side (b, q1, q2) of every
  // bra block (b, q1, q2, y), concatenated in sector order.  T3 region of charge y
  // is M_y x K_y with K contiguous (stride Kp_y = roundup(K_y, 4)); the bra is
  // packed into rows of the same K order.
  std::map<int64_t, std::vector<int64_t>> oby, kby;  // by y charge: out blocks, bra (= ket) blocks
  for (int64_t o = 0; o < H.out.nblocks; ++o) oby[H.out.sector(o)[2]].push_back(o);
  for (int64_t x = 0; x < ket.nblocks; ++x) kby[ket.sector(x)[3]].push_back(x);
  std::map<Key, Cell> cells;
  int64_t t3 = 0, bp = 0;
  for (auto& kv : oby) {
    const int64_t y = kv.first;
    const std::vector<int64_t>& O = kv.second;
    auto itk = kby.find(y);
    std::vector<int64_t> K = itk == kby.end() ? std::vector<int64_t>() : itk->second;
    std::vector<int64_t> koff(K.size());
    int64_t Ky = 0;
    for (size_t i = 0; i < K.size(); ++i) {
      const int64_t* z = ket.size(K[i]);
      koff[i] = Ky;
      Ky += z[0] * z[1] * z[2];
    }
    const int64_t Kp = up8(Ky), dY = H.out.size(O[0])[2];
    const int64_t breg = bp;
    if (Ky > 0) {
      for (size_t i = 0; i < K.size(); ++i) {
        const int64_t* z = ket.size(K[i]);
        const int64_t nk = z[0] * z[1] * z[2];
        const int64_t dims[2] = {nk, z[3]}, sst[2] = {1, nk}, dst[2] = {1, Kp};
        H.pack_bra.add(2, dims, sst, dst, ket.offsets[K[i]], breg + koff[i]);
      }
      bp += up8(Kp * dY);
    }
    const int64_t reg = t3;
    int64_t rowoff = 0;
    for (int64_t o : O) {
      const int64_t* os = H.out.sector(o);
      const int64_t* z = H.out.size(o);
      const int64_t dR = z[0], Dx = z[1];
      H.gF.add_prob(dR * Dx, dY);
      H.gF.add_crows_affine(H.out.offsets[o], dR * Dx, dR * Dx);
      if (Ky > 0) {
        H.gF.add_seg(Ky, reg + Kp * rowoff, Kp, breg, Kp);
        for (size_t i = 0; i < K.size(); ++i) {
          const int64_t* ks = ket.sector(K[i]);
          const int64_t* kz = ket.size(K[i]);
          cells[{ks[0], ks[1], ks[2], os[1], os[0]}] =
              Cell{reg + koff[i] + Kp * rowoff, Kp * dR, kz[0], Kp, 1, kz[0], kz[1], kz[2], Dx, dR};
        }
      }
      rowoff += dR * Dx;
    }
    if (Ky > 0) t3 += up8(Kp * rowoff);
  }
  H.t3_elems = t3;
  H.bp_elems = bp;
  mpo_layout(H, WL, WR, t1, cells);
  H.g1.BM = 64;
  H.g1.BN = 128;
  H.g1.KS = 0;
  H.gF.BM = H.gF.BN = 64;
  H.gF.KS = 0;
  H.pack_v.finalize();
  H.pack_envL.finalize();
  H.pack_bra.finalize();
  // workspace: [ket_p | env_p | bra_p | W | T1 | T3 | partials]
  H.kp_off = 0;
  H.ep_off = up64(H.vp_elems);
  H.bp_off = up64(H.ep_off + H.el_elems);
  H.w_off = up64(H.bp_off + H.bp_elems);
  H.t1_off = up64(H.w_off + H.w_elems);
  H.t3_off = up64(H.t1_off + H.t1_elems);
  H.part_off = up64(H.t3_off + H.t3_elems);
  return H;
}

static void env_core_splits(EnvCore& H, int mode) {
  H.gF.choose_splits(mode == 0 ? 2048 : 1024, chunk_k(mode, 0), 32);
  H.ws_elems = H.part_off + H.gF.part_elems;
}

static void finish_env_plan(EnvPlan& P) {
  P.ws_elems = 0;
  for (int c = 0; c < P.ncore; ++c) {
    P.core_f64[c] = P.core[c];
    env_core_splits(P.core[c], 1);
    env_core_splits(P.core_f64[c], 0);
    P.ws_elems = std::max({P.ws_elems, P.core[c].ws_elems, P.core_f64[c].ws_elems});
  }
  // pre-permuted operands live after the cores' workspace
  int64_t off = up64(P.ws_elems);
  for (int c = 0; c < P.ncore; ++c)
    for (int k = 0; k < 3; ++k) {
      P.perm_off[c][k] = off;
      if (P.has_perm[c][k]) off = up64(off + P.perm[c][k].out.numel());
    }
  P.ws_elems = off;
}

EnvPlan envupdate_plan(int side, const Struct& env, const Struct& A, const Struct& W) {
  if (env.rank() != 3 || A.rank() != 3 || W.rank() != 4) throw Error(ERR_INVALID, "ranks must be env 3, A 3, W 4");
  EnvPlan P;
  P.kind = side;
  P.ncore = 1;
  P.ins = {env, A, W};
  const Struct dA = dual_struct(A);
  ContractCost P1, P2, P3;
  if (side == 0) {  // _mpsupdateleft (src/apply.jl:49-58)
    P1 = contract_cost(env, {-1, 3, 4}, A, {-1, 2, 1});
    P2 = contract_cost(P1.C, {1, -1, -2, 4}, W, {-2, 2, 3, -1});
    P3 = contract_cost(P2.C, {1, -1, 2, -2}, dA, {-2, -1, 3});
  } else if (side == 1) {  // _mpsupdateright (src/apply.jl:17-25)
    P1 = contract_cost(env, {-1, 3, 4}, A, {1, 2, -1});
    P2 = contract_cost(P1.C, {1, -1, -2, 4}, W, {3, 2, -2, -1});
    P3 = contract_cost(P2.C, {1, -1, 2, -2}, dA, {3, -1, -2});
  } else {
    throw Error(ERR_INVALID, "side must be 0 (left) or 1 (right)");
  }
  P.ref_flops = P1.flops + P2.flops + P3.flops;
  P.ref_elems = env.numel() + A.numel() + P1.C.numel() + P1.C.numel() + W.numel() + P2.C.numel() + P2.C.numel() +
                dA.numel() + P3.C.numel();
  Struct ket = A, WL = W;
  if (side == 1) {  // the mirror image: A^M = A(r, p, l), W^M = W(w_r, p_out, w_l, p_in)
    P.perm[0][0] = permutelegs_plan(A, {3, 2, 1});
    P.perm[0][1] = permutelegs_plan(W, {3, 2, 1, 4});
    P.has_perm[0][0] = P.has_perm[0][1] = true;
    ket = P.perm[0][0].out;
    WL = P.perm[0][1].out;
  }
  const Struct WR = identity_mpo(WL, P.wr_identity);
  P.wr_is_identity = true;
  P.core[0] = env_core(env, with_trivial_leg(ket), WL, WR);
  if (!same_struct(P.core[0].out, P3.C) || P.core[0].out.sectors != P3.C.sectors)
    throw Error(ERR_SECTOR_MISMATCH, "envupdate: output structure differs from the contraction chain");
  P.outs[0] = P3.C;
  finish_env_plan(P);
  return P;
}

EnvPlan heff_adjoint_plan(const Struct& v, const Struct& envL, const Struct& envR, const Struct& WL,
                          const Struct& WR) {
  // validates the operands like the forward H_eff
  HeffPlan fwd_check = heff_plan(v, envL, envR, WL, WR);
  (void)fwd_check;
  EnvPlan P;
  P.kind = 2;
  P.ncore = 2;
  P.ins = {v, envL, envR, WL, WR};
  const Struct bra = dual_struct(v);
  // d<v|H v>/d envR[c,k,e] = sum envL[a,m,b] v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] v[b,q,r,e]
  ContractCost Y1 = contract_cost(envL, {-1, 2, 1}, v, {-1, 3, 4, 5});
  ContractCost Y2 = contract_cost(Y1.C, {1, -1, -2, 4, 5}, WL, {-1, 2, 3, -2});
  ContractCost Y3 = contract_cost(Y2.C, {1, 2, -1, -2, 5}, WR, {-1, 3, 4, -2});
  ContractCost Y4 = contract_cost(Y3.C, {-1, -2, -3, 2, 1}, bra, {-1, -2, -3, 3});
  // d<v|H v>/d envL[a,m,b] = sum v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] envR[c,k,e] v[b,q,r,e]
  ContractCost X1 = contract_cost(v, {1, 2, 3, -1}, envR, {-1, 4, 5});
  ContractCost X2 = contract_cost(X1.C, {1, 2, -1, -2, 3}, WR, {4, 5, -2, -1});
  ContractCost X3 = contract_cost(X2.C, {1, -1, 2, -2, 3}, WL, {4, 5, -2, -1});
  ContractCost X4 = contract_cost(X3.C, {1, -1, -2, 2, -3}, bra, {3, -3, -2, -1});
  P.ref_flops = Y1.flops + Y2.flops + Y3.flops + Y4.flops + X1.flops + X2.flops + X3.flops + X4.flops;
  P.ref_elems = 0;
  for (const ContractCost* c : {&Y1, &Y2, &Y3, &Y4, &X1, &X2, &X3, &X4}) P.ref_elems += 2 * c->C.numel();
  // core 0: dE/denvR directly; core 1: dE/denvL as the mirror image (v^M = v(r, p2, p1, l),
  // the MPO sites swapped and each mirrored: W^M = W(w_r, p_out, w_l, p_in))
  P.core[0] = env_core(envL, v, WL, WR);
  P.perm[1][0] = permutelegs_plan(v, {4, 3, 2, 1});
  P.perm[1][1] = permutelegs_plan(WR, {3, 2, 1, 4});
  P.perm[1][2] = permutelegs_plan(WL, {3, 2, 1, 4});
  P.has_perm[1][0] = P.has_perm[1][1] = P.has_perm[1][2] = true;
  P.core[1] = env_core(envR, P.perm[1][0].out, P.perm[1][1].out, P.perm[1][2].out);
  if (!same_struct(P.core[0].out, Y4.C) || P.core[0].out.sectors != Y4.C.sectors)
    throw Error(ERR_SECTOR_MISMATCH, "heff adjoint: dE/denvR structure differs from the contraction chain");
  if (!same_struct(P.core[1].out, X4.C) || P.core[1].out.sectors != X4.C.sectors)
    throw Error(ERR_SECTOR_MISMATCH, "heff adjoint: dE/denvL structure differs from the contraction chain");
  P.outs[0] = X4.C;  // gL
  P.outs[1] = Y4.C;  // gR
  finish_env_plan(P);
  return P;
}

}  // namespace mpsh
