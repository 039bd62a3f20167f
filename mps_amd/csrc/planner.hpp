// Host-side integer bookkeeping for the U(1) block-sparse contraction engine.
//
// Everything here is exact integer work done once per tensor structure and
// cached by the caller: sector enumeration in column-major order, fused
// spaces, block-pair matching, and the descriptor tables the HIP kernels
// consume (gather/scatter copies, grouped-GEMM problems, MPO groups).
// No device calls live in this file; it builds with plain g++.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace mpsh {

enum Status : int {
  OK = 0,
  ERR_INVALID = 1,
  ERR_SPACE_MISMATCH = 2,
  ERR_SECTOR_MISMATCH = 3,
  ERR_NOT_SIMILAR = 4,
  ERR_NONFINITE = 5,
  ERR_DEVICE = 6,
  ERR_OOM = 7,
  ERR_UNSUPPORTED = 8,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// One U(1) leg: sorted charge -> dim map plus the dual flag
// (reference src/vspace.jl:5-18).
struct Leg {
  bool isdual = false;
  std::vector<int64_t> charges;
  std::vector<int64_t> dims;
  int64_t dim_of(int64_t c) const;  // 0 when absent
  int64_t total() const;
  bool same_map(const Leg& o) const { return charges == o.charges && dims == o.dims; }
  bool operator==(const Leg& o) const { return isdual == o.isdual && same_map(o); }
  bool operator!=(const Leg& o) const { return !(*this == o); }
  Leg dual() const {
    Leg l = *this;
    l.isdual = !isdual;
    return l;
  }
};

Leg make_leg(bool isdual, const std::vector<std::pair<int64_t, int64_t>>& pairs);
bool isdual_pair(const Leg& a, const Leg& b);  // reference src/vspace.jl:24-25
Leg fuse(bool dual, const std::vector<Leg>& Vs);  // reference src/vspace.jl:112-138

struct KeyHash {
  size_t operator()(const std::vector<int64_t>& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int64_t x : k) {
      h ^= (uint64_t)x + 0x9e3779b97f4a7c15ull;
      h *= 1099511628211ull;
    }
    return (size_t)h;
  }
};

// A tensor structure: charge + legs + every allowed sector in column-major
// order with its block size and packed element offset (reference
// src/symtensor.jl:12-30, src/sector.jl:69-90).
struct Struct {
  int64_t charge = 0;
  std::vector<Leg> legs;
  int64_t nblocks = 0;
  std::vector<int64_t> sectors;  // nblocks * rank
  std::vector<int64_t> sizes;    // nblocks * rank
  std::vector<int64_t> offsets;  // nblocks + 1
  std::unordered_map<std::vector<int64_t>, int64_t, KeyHash> index;

  int rank() const { return (int)legs.size(); }
  int64_t numel() const { return offsets.empty() ? 0 : offsets.back(); }
  const int64_t* sector(int64_t b) const { return &sectors[(size_t)b * rank()]; }
  const int64_t* size(int64_t b) const { return &sizes[(size_t)b * rank()]; }
  int64_t block_numel(int64_t b) const;
  int64_t find(const std::vector<int64_t>& s) const;
};

void all_sectors(int64_t charge, const std::vector<Leg>& legs, std::vector<int64_t>& sects,
                 std::vector<int64_t>& sizes);
Struct make_struct(int64_t charge, const std::vector<Leg>& legs);
bool same_struct(const Struct& a, const Struct& b);  // same charge and legs

// ---------------------------------------------------------------------------
// Device descriptor records (plain data, copied verbatim to the GPU)
// ---------------------------------------------------------------------------
constexpr int MAXD = 10;

// Strided N-d copy: dst[dst_off + sum i_k*dst_st[k]] = src[src_off + sum i_k*src_st[k]].
// dims[0] is the destination-contiguous dimension where one exists and
// dims[1] the source-contiguous one when it differs ("transpose").
struct CopyDesc {
  int32_t ndim;
  int32_t transpose;
  int32_t dims[MAXD];
  int64_t src_st[MAXD];
  int64_t dst_st[MAXD];
  int64_t src_off;
  int64_t dst_off;
};

struct CopyPlan {
  std::vector<CopyDesc> descs;
  std::vector<int64_t> tile_start;  // ndesc + 1, prefix of 64x64 tiles
  std::vector<int32_t> tile_map;    // per tile: (descriptor, tile within it)
  int64_t ntiles = 0;
  int64_t elems = 0;
  void add(int ndim, const int64_t* dims, const int64_t* sst, const int64_t* dst, int64_t soff,
           int64_t doff);
  void finalize();
};

// One grouped-GEMM problem: C[M x N] = sum over its K segments of A_s^T B_s.
// Operand rows are gathered through base-address tables, so a problem may
// span several stored blocks without copies:
//   A_s(k, m) = A[atab[s.atab0 + m] + k],   B_s(k, n) = B[btab[s.btab0 + n] + k],
//   C(m, n)   = C[crow[p.crow0 + m] + n * cld[p.crow0 + m]].
// Affine rows need no table: atab0 < 0 means A row m at a_off + m * lda (the
// same for B), crow0 < 0 means C(m, n) = C[c_off + m + n * ldc].
// nseg == 0 means "write zeros" (reference src/contract.jl:292-294).  With
// nsplit > 1 the K chunks are split over work items that write dense
// partials at part_off (+ split * pld * N, column-major with leading dimension
// pld = M rounded up to 4, so every partial column starts 16-byte aligned),
// summed by a reduction.
struct GemmProb {
  int32_t M, N, seg0, nseg;
  int64_t crow0;
  int32_t nsplit, pld;
  int64_t part_off;
  int64_t c_off, ldc;
};
struct GemmSeg {
  int64_t K, atab0, btab0;
  int64_t a_off, lda, b_off, ldb;
};
constexpr int NXCD = 8;  // MI355X: 8 accelerator complex dies, one L2 each

struct GemmTile {
  int32_t prob, m0, n0, split;
  // explicit K-chunk range [kc0, kc1) of the problem's flattened segment chunks
  // (kc1 < 0: the split's share, chunks * split / nsplit .. chunks * (split+1) / nsplit);
  // pld > 0: a piece of a stream-K tile, writing its partial tile (m, n) to
  // P[poff + (m - m0) + (n - n0) * pld] (poff = block * BM * BN, pld = BM)
  int32_t kc0 = 0, kc1 = -1, pld = 0, pad = 0;
  int64_t poff = 0;
};
// Per-tile record of a DMA-streamed GEMM launch in one kernel mode (K chunk
// fixed): everything the kernel reads at a tile's start, so a work group loads
// one record (prefetched a tile ahead) instead of the tile -> problem ->
// segment chain.  Words (int32; int64 fields low word first):
//   0 m0, 1 n0, 2 M, 3 N, 4 seg0, 5 nseg, 6 chunks of this tile, 7 start
//   segment (relative to seg0), 8 start k (elements), 9 split, 10 nsplit,
//   11 pld, 12 piece pld (stream-K, else 0), 13 unused, 14 crow0, 16 c_off,
//   18 ldc, 20 part_off, 22 piece poff, 24.. the start segment: 24 K, 26 atab0,
//   28 btab0, 30 a_off, 32 lda, 34 b_off, 36 ldb
constexpr int kTileRecWords = 40;
// One stream-K tile cut into np > 1 pieces: its partial tiles are the BM x BN
// blocks blk0 .. blk0 + np - 1 of P, summed in that order by the reduction.
struct TileRed {
  int32_t prob, m0, n0, np;
  int64_t blk0;
};
struct RedDesc {  // one split problem of the reduction pass
  int32_t prob, pad;
  int64_t start;  // prefix over ceil(M*N/1024) work groups
};

struct GemmPlan {
  std::vector<GemmProb> probs;
  std::vector<GemmSeg> segs;
  std::vector<int64_t> atab, btab, crow, cld;
  int64_t flops = 0;  // 2*M*N*K summed (algorithmic)
  int BM = 64, BN = 64, KS = 1;  // tile shape and K-chunk multiplier of the kernel variant
  int64_t part_elems = 0;
  // stream-K (use_streamk): no per-problem split; every work group of the launch
  // gets a contiguous range of equal K-chunk work across the tiles of its XCD's
  // queue (streamk_schedule), tiles cut between work groups write partial tiles
  bool streamk = false;
  int sk_mode = -1;  // kernel mode the stream-K bound was sized for
  // open a problem; its C rows follow with add_crow()
  void add_prob(int64_t M, int64_t N) {
    probs.push_back(GemmProb{(int32_t)M, (int32_t)N, (int32_t)segs.size(), 0, (int64_t)crow.size(), 1, 0, 0, 0, 0});
  }
  void add_crow(int64_t base, int64_t ld) {
    crow.push_back(base);
    cld.push_back(ld);
  }
  // affine C rows of the open problem: C(m, n) = C[base + m + n * ld]
  void add_crows_affine(int64_t base, int64_t ld, int64_t M) {
    GemmProb& p = probs.back();
    if (M != p.M || (int64_t)crow.size() != p.crow0) throw Error(ERR_INVALID, "affine C rows: bad problem");
    p.crow0 = -1;
    p.c_off = base;
    p.ldc = ld;
  }
  // a K segment with explicit row tables (M entries for A, N for B)
  void add_seg_tab(int64_t K, const std::vector<int64_t>& arows, const std::vector<int64_t>& brows);
  // a K segment with affine rows: A rows a_off + m*lda, B rows b_off + n*ldb
  void add_seg(int64_t K, int64_t a_off, int64_t lda, int64_t b_off, int64_t ldb);
  // split long-K problems so that the number of work items reaches `target`
  void choose_splits(int64_t target, int BK, int64_t min_chunks_per_split);
  // Tiles in per-XCD queues: queue x = tiles [qoff[x], qoff[x+1]) runs on the
  // work groups with blockIdx % NXCD == x.  Work is balanced over the queues
  // (LPT over panel groups) and the tiles that share operand panels stay in
  // one queue, so each XCD's L2 fetches its panels once.
  std::vector<GemmTile> tiles(std::vector<int32_t>* qoff = nullptr) const;
  // Per-work-group tile lists for `slots` resident work groups per XCD: each
  // XCD's queue (tiles()) is dealt to its work groups by LPT on the tiles' K
  // chunks (chunk BK elements) plus a fixed per-tile cost, each list kept in
  // queue order.  Returns the lists concatenated; woff (NXCD*slots + 1 entries)
  // bounds list j of XCD x at [woff[x*slots + j], woff[x*slots + j + 1]).
  std::vector<GemmTile> schedule(int slots, int BK, std::vector<int32_t>* woff) const;
  std::vector<RedDesc> reductions() const;
  // Stream-K instead of split-K for a DMA-streamed kernel variant (KS = 0) of
  // `mode`: resets every split and sizes part_elems for the largest partial
  // layout streamk_schedule can produce (two partial tiles per work group, at
  // most sk_max_slots(mode) work groups per XCD).
  void use_streamk(int mode);
  // Per-work-group piece lists for `slots` work groups per XCD (slots <=
  // sk_max_slots(mode)): XCD x's queue (tiles()) is cut into `slots` contiguous
  // ranges of equal work (K chunks of BK elements plus kStreamKOverhead chunks
  // per piece); a tile that straddles ranges becomes one piece per range (no
  // piece shorter than kStreamKMinPiece chunks unless its tile is).  woff as in
  // schedule(); reds lists the cut tiles.
  std::vector<GemmTile> streamk_schedule(int slots, int BK, std::vector<int32_t>* woff,
                                         std::vector<TileRed>* reds) const;
  // the kTileRecWords-word records of `tiles` (in that order) for K chunks of BK
  std::vector<int32_t> tile_records(const std::vector<GemmTile>& tiles, int BK) const;
};
constexpr int kStreamKOverhead = 4;  // chunks: a piece's set-up, pipeline fill and stores
constexpr int kStreamKMinPiece = 8;  // chunks
// upper bound of the resident work groups per XCD of a DMA-streamed GEMM variant
// (32 CUs per XCD, at most 8 work groups of 256 threads per CU, 160 KiB of LDS
// per CU over the variant's LDS ring: 3 stages for FP64, 4 otherwise, of
// (BM + BN) rows of 64 bytes)
inline int sk_max_slots(int mode, int BM, int BN) {
  const int ns = mode == 0 ? 3 : 4;
  const int lds = ns * (BM + BN) * 64;
  int occ = (160 * 1024) / lds;
  if (occ > 8) occ = 8;
  if (occ < 1) occ = 1;
  return 32 * occ;
}

// ---------------------------------------------------------------------------
// contract(A, idxA, B, idxB): the reference's fuse -> GEMM -> split -> permute
// (src/contract.jl:11-47, :112-171, :187-224, :234-298).
// ---------------------------------------------------------------------------
struct IndexPerm {
  std::vector<int> rems, cons, tofinals;  // 1-based, as in the reference
};
IndexPerm contract_index_perm(const std::vector<int>& idx);  // src/contract.jl:82-99

struct FusedMatrix {
  Struct F;          // the fused rank-2 structure
  CopyPlan gather;   // A blocks -> fused blocks
  bool k_contig_rows;  // storage: true = rows K-contiguous (A^T), false = column-major
  // storage of each fused block with its K-contiguous dimension padded to a
  // multiple of kpad (16-byte aligned rows for vector loads): leading dim and offsets
  std::vector<int64_t> ld, poff;  // nblocks, nblocks + 1
  int64_t numel_padded() const { return poff.empty() ? 0 : poff.back(); }
};

// SymMatrix(A, rows, cols) (src/contract.jl:112-171).  `transposed` stores
// each fused block row-major (element (r,c) at r*ncols + c) instead of
// column-major, which is the K-contiguous layout of a left GEMM operand.
FusedMatrix symmatrix_plan(const Struct& A, const std::vector<int>& rows, const std::vector<int>& cols,
                           bool transposed, int64_t dst_base, int kpad = 1);

struct ContractPlan {
  Struct A, B, C;
  FusedMatrix FA, FB;
  Struct FC;
  GemmPlan gemm;      // A base = ws, B base = ws + FA.numel, C base = ws + FA+FB
  CopyPlan scatter;   // fused C (ws) -> C
  int64_t ws_elems = 0;
  int64_t fa_off = 0, fb_off = 0, fc_off = 0;
};
ContractPlan contract_plan(const Struct& A, const std::vector<int>& idxA, const Struct& B,
                           const std::vector<int>& idxB);

// Cost of contract(A, idxA, B, idxB) without building its copy/GEMM tables:
// the reference's fused-charge GEMM FLOPs (sum 2*M*N*K over matched fused
// charges) and the output structure C.
struct ContractCost {
  int64_t flops = 0;
  Struct C;
};
ContractCost contract_cost(const Struct& A, const std::vector<int>& idxA, const Struct& B,
                           const std::vector<int>& idxB);

// Re-legging plans: pure data movement (src/symreleg.jl:34-214).
struct RelegPlan {
  Struct in, out;
  CopyPlan copy;
};
RelegPlan permutelegs_plan(const Struct& A, const std::vector<int>& perm);
RelegPlan fuselegs_plan(const Struct& A, int l, int n, bool dual);
RelegPlan splitleg_plan(const Struct& A, int l, const std::vector<Leg>& spaces);
RelegPlan symmatrix_releg(const Struct& A, const std::vector<int>& rows, const std::vector<int>& cols);

// ---------------------------------------------------------------------------
// Fused two-site H_eff (src/apply.jl:93-105) computed as
//   step 1: T1 = v^T . envL per matched block pair (grouped TN GEMM),
//   step 2+3: T3 = W_group . T1 per (bra_l, ket_r) pixel group, where
//             W_group = WL (x) WR combined over the middle MPO bond,
//   step 4: out = T3^T . envR, K segments over (mpo_r, ket_r).
// ---------------------------------------------------------------------------
struct MpoSig {
  int32_t R, C;
  int64_t w_off;  // row-major R x C matrix in the W workspace
};
struct MpoTerm {   // one (out type, in type) sub-matrix of a signature's W
  int32_t sig, row0, col0, R, C;  // R = Dx*d1o*d2o rows, C = Dw*d1*d2 cols
  int32_t Dw, d1i, d2i, Dx, d1o, d2o;
  int32_t pair0, npair;
};
struct MpoPair {   // WL block (w, s1o, m, s1) x WR block (m, s2o, x, s2)
  int64_t wl_off, wr_off;
  int32_t Dm, pad;
};
struct MpoGroup {
  // dRp: the ket_r range the pixel runs cover, dR rounded up to the plan's
  // row alignment (T1 / T3 rows of padded plans hold zeros past dR)
  int32_t sig, dR, db, dRp;
  int32_t in0, out0;  // first row in the in/out row tables
  int64_t npix;
};
struct MpoInRow {
  int64_t base, nstride, mstride;
};
struct MpoOutRow {  // T3 address = base + n*nstride + m*mstride
  int64_t base, nstride, mstride;
};
struct MpoWork {  // one work group of the MPO pass: runs [run0, run0 + 256) of a pixel group
  int32_t group, run0;
};
// The MPO pass gives each thread a run of P consecutive pixels along the ket_r
// index n of one bra_l row m; P = 1, 2 or 4 (by storage type and channel count)
constexpr int kMpoRuns[3] = {1, 2, 4};
// the mixed work list: runs of 4 pixels in groups of at most 8 channels (little
// work per pixel: fewer work groups), one pixel per thread otherwise
constexpr int mpo_mixed_run(int C) { return C <= 8 ? 4 : 1; }
// the work lists of padded plans (every channel row and T3 row segment starts on a
// multiple of 4 elements): runs of P pixels moved by one aligned vector access
// per row, P as large as the P x C inputs a thread holds allow (esize: bytes of
// the arithmetic type)
constexpr int mpo_vec_run(int C, int esize) {
  return esize == 8 ? (C <= 16 ? 2 : 1) : (C <= 16 ? 4 : (C <= 32 ? 2 : 1));
}
constexpr int kMpoPadR = 4;  // row alignment (elements) of padded plans

// Kernel variant of one grouped GEMM: tile BM x BN, K chunk multiplier KS.
struct TileCfg {
  int BM = 64, BN = 64, KS = 1;
};
struct HeffTiles {
  TileCfg s1{64, 128, 0}, s4{64, 64, 0};     // FP32 (DMA-streamed operands)
  TileCfg s1h{64, 128, 0}, s4h{64, 64, 0};   // BF16 storage (DMA-streamed BF16 operands)
  TileCfg s1d{64, 128, 0}, s4d{64, 64, 0};   // FP64 (DMA-streamed operands)
  // step-4 GEMM work division: split-K to at least this many work items, or
  // stream-K (< 0).  Measured on the H_eff shapes (cyl6x24_chi2048, MI355X):
  // split-K at 2048 items beats 1024 (FP32 3911 vs 3696 matvecs/s); stream-K runs
  // within 1 % of it in FP32 and FP64 (8 % slower in BF16) but fetches 3x the HBM
  // bytes per launch (FP32 474 vs 153 MB: the pieces of one K range of an A panel's
  // tiles run at different times, so the panel is fetched once per tile)
  int64_t split_target = 2048;               // FP32
  int64_t split_target_h = 2048;             // BF16 storage
  int64_t split_target_d = 2048;             // FP64
};

// Autotuner candidates for the two H_eff GEMMs (tile BM x BN x KS; step-4
// split-K work-item targets).  FP64 uses the KS = 1 tiles up to 128 wide and
// KS = 0; KS = 0 (DMA-streamed operands) is FP32 / FP64 only.
constexpr TileCfg kTuneTiles[] = {{64, 64, 1}, {64, 128, 1}, {128, 64, 1}, {64, 64, 2}, {128, 128, 1},
                                  {128, 64, 2}, {64, 64, 0},  {128, 64, 0}, {64, 128, 0}};
constexpr int64_t kTuneSplits[] = {512, 1024, 2048};
// KS = 0: the kernel with DMA-streamed operands (FP32 / FP64 / BF16 storage), tiles
// up to 128 x 64.  Modes: 0 FP64, 1 FP32, 2 BF16 operands from FP32 storage
// (register-staged only), 3 / 4 BF16 storage (DMA only; 3 writes BF16, 4 FP32).
inline bool tune_tile_ok(int mode, const TileCfg& c) {
  if (mode >= 3) return c.KS == 0 && c.BM * c.BN <= 128 * 64;
  if (c.KS == 0) return mode != 2 && c.BM * c.BN <= 128 * 64;
  return mode != 0 || (c.KS == 1 && c.BM * c.BN <= 128 * 64);
}
// K chunk of a GEMM mode (elements): 16 * KS for FP64 / FP32, 32 * KS for BF16
// operands; KS = 0 (one 64-byte row per operand row): 8 doubles, 16 floats, 32 bf16
inline int chunk_k(int mode, int KS) {
  if (KS == 0) return mode == 0 ? 8 : (mode >= 3 ? 32 : 16);
  return (mode == 2 ? 32 : 16) * KS;
}
// the step-4 GEMM of a plan with another tile and split target
GemmPlan gemm_variant(const GemmPlan& base, const TileCfg& c, int64_t split_target, int mode);
// split-K (target >= 0 work items, 0: none) or stream-K (target < 0, DMA variants;
// register-staged variants fall back to split-K at 1024 items) for a kernel mode
void set_splits(GemmPlan& g, int64_t split_target, int mode);

// Step 1 (T1 = v^T envL) and the combined-MPO pass (T3 = W_group T1): the first
// stages of the fused H_eff, shared with the environment plans.
struct Step13 {
  GemmPlan g1;  // A = packed v, B = packed envL, C = T1
  int64_t t1_elems = 0;
  std::vector<MpoSig> sigs;
  std::vector<MpoTerm> terms;
  std::vector<MpoPair> pairs;
  std::vector<MpoGroup> groups;
  std::vector<MpoInRow> inrows;
  std::vector<MpoOutRow> outrows;
  // work lists for runs of kMpoRuns[i] pixels; [3]: mixed runs (mpo_mixed_run, or
  // mpo_vec_run for 4- and 2-byte types in padded plans); [4]: mpo_vec_run of
  // 8-byte types (padded plans only)
  std::vector<MpoWork> work[5];
  int32_t padr = 1;  // ket_r row alignment of T1 and T3 (kMpoPadR: padded plan)
  int64_t w_elems = 0;
  int64_t t3_elems = 0;
  int maxR = 0, maxC = 0;
  int64_t mpo_flops = 0;  // dense W_group application flops of this engine
  CopyPlan pack_v, pack_envL;  // v and envL into 16-byte aligned K-contiguous rows
  int64_t vp_elems = 0, el_elems = 0;
};

struct HeffPlan : Step13 {
  Struct v, envL, envR, WL, WR, out;
  GemmPlan g1h, g4h;  // the same problems with the BF16 tile variants
  GemmPlan g1d, g4d;  // the same problems with the FP64 tile variants
  GemmPlan g4;  // A = T3 (ws), B = envR_p (prepared), C = out
  // Prepared operator (per bond): envL and envR repacked into 16-byte aligned
  // K-contiguous rows (K padded to a multiple of 4) and the combined-MPO matrices.
  CopyPlan pack_envR;
  int64_t el_off = 0, er_off = 0, pw_off = 0, prep_elems = 0;
  // Per-matvec workspace: [v_p | T1 | T3 | split-K partials]; v is packed the same way.
  int64_t vp_off = 0, t1_off = 0, t3_off = 0, part_off = 0, ws_elems = 0;
  // reference-order algorithmic cost (src/apply.jl:99-103 via src/contract.jl:290-291)
  int64_t ref_flops = 0;
  int64_t ref_elems = 0;  // sum over the 4 steps of numel(A)+numel(B)+numel(C)
  int64_t step_flops[4] = {0, 0, 0, 0};
  int64_t step_elems[4] = {0, 0, 0, 0};
};
HeffPlan heff_plan(const Struct& v, const Struct& envL, const Struct& envR, const Struct& WL,
                   const Struct& WR, const HeffTiles& tiles = HeffTiles());

// ---------------------------------------------------------------------------
// Environment plans.  One core computes, for a ket K (l, p1, p2, r) whose bra is
// K with dual legs,
//   out[x, y, x'] = sum env[a, m, b] K[a, p1, p2, x] WL[m, q1, n, p1] WR[n, q2, y, p2] K[b, q1, q2, x']
// as the H_eff's step 1 and MPO pass (T3 written with the bra side (b, q1, q2)
// K-contiguous) followed by one grouped GEMM out = T3^T bra per output block.
// _mpsupdateleft (src/apply.jl:49-58) is one core over one site (K = A with a
// trivial second physical leg, WR = identity); _mpsupdateright (:17-25) is the
// left update of the mirror image A(r, p, l), W(w_r, p_out, w_l, p_in); the
// adjoint of <v|H_eff v> (src/apply.jl:93-105) w.r.t. envR is one two-site
// core and w.r.t. envL the two-site core of the mirror image.
// ---------------------------------------------------------------------------
struct EnvCore : Step13 {
  Struct env, ket, WL, WR, out;
  GemmPlan gF;        // A = T3 (ws), B = packed bra (ws), C = out
  CopyPlan pack_bra;  // the bra into rows of its K order
  int64_t bp_elems = 0;
  // workspace: [ket_p | env_p | bra_p | W | T1 | T3 | split-K partials]
  int64_t kp_off = 0, ep_off = 0, bp_off = 0, w_off = 0, t1_off = 0, t3_off = 0, part_off = 0, ws_elems = 0;
};

struct EnvPlan {
  int kind = 0;  // 0 left update, 1 right update, 2 H_eff adjoint
  int ncore = 1;
  EnvCore core[2];      // FP32 split-K layout
  EnvCore core_f64[2];  // FP64 split-K layout (same tables otherwise)
  // operands permuted before a core runs: [core][ket, WL, WR]
  RelegPlan perm[2][3];
  bool has_perm[2][3] = {{false, false, false}, {false, false, false}};
  int64_t perm_off[2][3] = {{0, 0, 0}, {0, 0, 0}};
  bool wr_is_identity = false;  // one-site plans: core 0's WR is the identity below
  std::vector<double> wr_identity;
  std::vector<Struct> ins;  // the operand structures, in argument order
  Struct outs[2];           // left/right update: outs[0]; adjoint: gL, gR
  int64_t ws_elems = 0;
  int64_t ref_flops = 0, ref_elems = 0;  // the reference contraction chain(s)
};
Struct dual_struct(const Struct& A);  // src/symtensor.jl:201-212
EnvPlan envupdate_plan(int side, const Struct& env, const Struct& A, const Struct& W);
EnvPlan heff_adjoint_plan(const Struct& v, const Struct& envL, const Struct& envR, const Struct& WL,
                          const Struct& WR);

}  // namespace mpsh
