// Device-side views of the planner tables and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "planner.hpp"

namespace mpsh {

struct DevCopy {
  const CopyDesc* descs = nullptr;
  const int32_t* tile_map = nullptr;  // per tile: (descriptor, tile within it)
  int ndesc = 0;
  int64_t ntiles = 0;
};

// thread-local message returned by mpsh_last_error (defined in api.hip)
void set_last_error(const std::string& m);
// synchronous host-to-device copy of plan tables that does not wait for other streams' work (api.hip)
hipError_t upload_tables(void* dst, const void* src, size_t bytes);

// Tile queues of a persistent GEMM launch.  With woff == nullptr, work group b
// walks queue b % NXCD ([off[x], off[x+1])) with stride gridDim.x / NXCD; with
// woff, the grid has NXCD * slots work groups and work group b runs the list
// [woff[x*slots + j], woff[x*slots + j + 1]), x = b % NXCD, j = b / NXCD.
struct XcdQueues {
  int32_t off[NXCD + 1];
  int32_t slots;
  const int32_t* woff;
};

struct DevGemm {
  const GemmProb* probs = nullptr;
  const GemmSeg* segs = nullptr;
  const int64_t *atab = nullptr, *btab = nullptr, *crow = nullptr, *cld = nullptr;
  const GemmTile* tiles = nullptr;
  const RedDesc* reds = nullptr;
  const int32_t* red_map = nullptr;  // per reduction work group: (problem, group within it)
  int ntiles = 0, nred = 0;
  int64_t red_groups = 0;
  const TileRed* treds = nullptr;  // stream-K: cut tiles, one reduction work group each
  int ntred = 0;
  const int32_t* recs = nullptr;   // per-tile records (DMA variants) for kernel mode rec_mode, or none
  int rec_mode = -1;
  int BM = 64, BN = 64, KS = 1;  // tile and K-chunk multiplier
  XcdQueues q{};                 // tiles [q.off[x], q.off[x+1]) run on XCD x
};

struct DevMpo {
  const MpoSig* sigs = nullptr;
  const MpoTerm* terms = nullptr;
  const MpoPair* pairs = nullptr;
  const MpoGroup* groups = nullptr;
  const MpoInRow* inrows = nullptr;
  const MpoOutRow* outrows = nullptr;
  // runs of kMpoRuns[i] pixels; [3] mixed; [4] mixed of 8-byte types in padded plans
  const MpoWork* work[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int nwork[5] = {0, 0, 0, 0, 0};
  bool vec = false;  // padded plan: T1 / T3 row segments aligned, pixel runs by vector access
  int nterms = 0, cmax = 32, rmax = 32;
  int64_t t1_elems = 0, t3_elems = 0;  // T1 / T3 sizes (32-bit byte offsets when both fit)
  int run = 0;                          // pixels per thread (1, 2, 4; FP64 and CMAX 64 at most 2), 0: mixed
  bool off64 = false;                   // force 64-bit offsets
};
// the MPO pass options of newly uploaded plans: MPSH_MPO_RUN (1, 2, 4) and MPSH_MPO_OFF64 (0 / 1)
void mpo_options(DevMpo& d);

hipError_t launch_copy(const DevCopy& c, int elem_size, const void* src, void* dst, hipStream_t s);
// dst[0 .. bytes) = 0; dst 16-byte aligned, bytes a multiple of 16 (a kernel, not a
// memset node, so the call may be captured in a HIP graph)
hipError_t launch_zero(void* dst, size_t bytes, hipStream_t s);
// the same copy converting FP32 to BF16 (packs of the BF16-storage H_eff)
hipError_t launch_copy_to_bf16(const DevCopy& c, const float* src, void* dst, hipStream_t s);
// resident work groups per XCD of a GEMM kernel variant (CUs per XCD x occupancy)
int gemm_slots_per_xcd(int mode, int BM, int BN, int KS);
// mode 0 = f64, 1 = f32, 2 = bf16 operands from f32 storage, 3 / 4 = bf16 storage
// (DMA-streamed; 3 writes bf16, 4 fp32; split-K partials fp32)
// P: split-K partial buffer (unused when no problem is split)
hipError_t launch_gemm(const DevGemm& g, int mode, const void* A, const void* B, void* C, void* P, hipStream_t s);
hipError_t launch_mpo_combine(const DevMpo& d, int elem_size, const void* WL, const void* WR, void* W,
                              hipStream_t s);
// elem_size 8 / 4: T1, T3, W of that type; 2: BF16 T1 and T3, FP32 W and arithmetic
hipError_t launch_mpo_apply(const DevMpo& d, int elem_size, const void* W, const void* T1, void* T3, hipStream_t s);
hipError_t launch_dot(int elem_size, int64_t n, const void* x, const void* y, double* out, double* partials,
                      int max_partials, hipStream_t s);
hipError_t launch_axpby(int elem_size, int64_t n, double a, const void* x, double b, void* y, hipStream_t s);
hipError_t launch_scal(int elem_size, int64_t n, double a, void* x, hipStream_t s);
hipError_t launch_convert(int src_size, const void* src, int dst_size, void* dst, int64_t n, hipStream_t s);
constexpr int GATHER_FIELDS = 8;
hipError_t launch_gather_blocks(int nblk, int64_t ntiles, const int64_t* table, const double* src,
                                const double* scale, double* dst, hipStream_t s);
hipError_t launch_lz_dots(int k, int64_t n, const double* V, int64_t ldv, const double* w, double* c, double* part,
                          hipStream_t s);
hipError_t launch_lz_update(int k, int64_t n, const double* V, int64_t ldv, const double* c, double sign, double* w,
                            hipStream_t s);
hipError_t launch_lz_normalize(int64_t n, const double* w, double* out, double* beta, double* part, hipStream_t s);

}  // namespace mpsh
