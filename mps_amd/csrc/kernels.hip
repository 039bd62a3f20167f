// HIP kernels for gfx950 (MI355X / CDNA4) and their launchers.
//
//  * copy_kernel     -- batched strided N-d copy; 64x64 LDS-tiled transpose when
//                       the source- and destination-contiguous dims differ.
//                       Implements fuse / split / permute (SymMatrix,
//                       SymTensor(::SymMatrix), permutelegs, fuselegs, splitleg).
//  * gemm_grouped    -- persistent grouped GEMM over uneven block problems with
//                       K segments; MFMA 16x16x4 f64 / 16x16x4 f32 / 16x16x32 bf16.
//  * mpo_combine     -- builds the combined two-site MPO matrices W_group.
//  * mpo_apply       -- T3 = W_group . T1 per (bra_l, ket_r) pixel group.
//  * dot/axpby/scal/convert -- block-wise vector ops on packed buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

namespace mpsh {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// copy
// ---------------------------------------------------------------------------
// One work group per tile; the plan's tile map names the tile's descriptor and
// its index within it, so a work group starts with one table load.
// TD: destination type (a conversion on the way, e.g. FP32 -> BF16 packs)
template <typename T, typename TD = T>
__global__ __launch_bounds__(256) void copy_kernel(const CopyDesc* __restrict__ descs,
                                                   const int32_t* __restrict__ tmap,
                                                   const T* __restrict__ src, TD* __restrict__ dst) {
  const int2 tm = ((const int2*)tmap)[blockIdx.x];
  const CopyDesc* d = descs + tm.x;
  int64_t t = tm.y;
  const int ndim = d->ndim;
  const int64_t d0 = d->dims[0], d1 = ndim > 1 ? d->dims[1] : 1;
  const int64_t s0 = d->src_st[0], s1 = ndim > 1 ? d->src_st[1] : 0;
  const int64_t w0 = d->dst_st[0], w1 = ndim > 1 ? d->dst_st[1] : 0;
  int64_t sb = d->src_off, db = d->dst_off;
  if (d->transpose) {
    __shared__ T tile[64][65];
    const int64_t nx = (d0 + 63) >> 6, ny = (d1 + 63) >> 6;
    const int64_t tx = t % nx;
    t /= nx;
    const int64_t ty = t % ny;
    t /= ny;
    for (int k = 2; k < ndim; ++k) {
      const int64_t i = t % d->dims[k];
      t /= d->dims[k];
      sb += i * d->src_st[k];
      db += i * d->dst_st[k];
    }
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int i = ly; i < 64; i += 4) {
      const int64_t x = tx * 64 + i, y = ty * 64 + lx;
      if (x < d0 && y < d1) tile[i][lx] = src[sb + x * s0 + y * s1];
    }
    __syncthreads();
    for (int j = ly; j < 64; j += 4) {
      const int64_t x = tx * 64 + lx, y = ty * 64 + j;
      if (x < d0 && y < d1) dst[db + x * w0 + y * w1] = (TD)tile[lx][j];
    }
  } else {
    const int64_t n01 = d0 * d1;
    const int64_t nt = (n01 + 4095) >> 12;
    const int64_t tt = t % nt;
    t /= nt;
    for (int k = 2; k < ndim; ++k) {
      const int64_t i = t % d->dims[k];
      t /= d->dims[k];
      sb += i * d->src_st[k];
      db += i * d->dst_st[k];
    }
    // all sixteen loads of a thread are issued before its first store, so a
    // tile is one memory round trip rather than sixteen
    const uint32_t ud0 = (uint32_t)d0;
    const int64_t e0 = tt * 4096 + threadIdx.x;
    T val[16];
    int64_t to[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int64_t e = e0 + 256 * u;
      if (e < n01) {
        const uint32_t ue = (uint32_t)e;
        const uint32_t y = ue / ud0, x = ue - y * ud0;
        val[u] = src[sb + x * s0 + y * s1];
        to[u] = db + x * w0 + y * w1;
      }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (e0 + 256 * u < n01) dst[to[u]] = (TD)val[u];
  }
}

__global__ __launch_bounds__(256) void zero_kernel(floatx4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    dst[i] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
}

hipError_t launch_zero(void* dst, size_t bytes, hipStream_t s) {
  const int64_t n = (int64_t)(bytes / 16);
  if (n == 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, s, (floatx4*)dst, n);
  return hipGetLastError();
}

hipError_t launch_copy_to_bf16(const DevCopy& c, const float* src, void* dst, hipStream_t s) {
  if (c.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL((copy_kernel<float, __bf16>), dim3((unsigned)c.ntiles), dim3(256), 0, s, c.descs, c.tile_map,
                     src, (__bf16*)dst);
  return hipGetLastError();
}

hipError_t launch_copy(const DevCopy& c, int elem_size, const void* src, void* dst, hipStream_t s) {
  if (c.ntiles == 0) return hipSuccess;
  dim3 grid((unsigned)c.ntiles), block(256);
  if (elem_size == 8)
    hipLaunchKernelGGL(copy_kernel<double>, grid, block, 0, s, c.descs, c.tile_map,
                       (const double*)src, (double*)dst);
  else
    hipLaunchKernelGGL(copy_kernel<float>, grid, block, 0, s, c.descs, c.tile_map,
                       (const float*)src, (float*)dst);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// block gather (truncated SVD factors): table = nblk records of GATHER_FIELDS
// int64 {src_off, rs, cs, rows, cols, dst_off, sc_off, smode}, then one int64
// per 4096-element tile: (block << 32) | tile within the block.
//   dst[dst_off + i + rows*j] = src[src_off + i*rs + j*cs] * f,
//   f = 1 (smode 0), scale[sc_off + i] (smode 1), scale[sc_off + j] (smode 2)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gather_blocks_kernel(int nblk, const int64_t* __restrict__ table,
                                                            const double* __restrict__ src,
                                                            const double* __restrict__ scale,
                                                            double* __restrict__ dst) {
  const int64_t tm = table[(int64_t)nblk * GATHER_FIELDS + blockIdx.x];
  const int64_t* d = table + (tm >> 32) * GATHER_FIELDS;
  const int64_t so = d[0], rs = d[1], cs = d[2], rows = d[3], cols = d[4], doff = d[5], sco = d[6];
  const int smode = (int)d[7];
  const int64_t n = rows * cols;
  const int64_t e0 = (tm & 0xffffffffLL) * 4096 + threadIdx.x;
  double val[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int64_t e = e0 + 256 * u;
    if (e < n) {
      const int64_t j = e / rows, i = e - j * rows;
      double f = 1.0;
      if (smode == 1) f = scale[sco + i];
      if (smode == 2) f = scale[sco + j];
      val[u] = src[so + i * rs + j * cs] * f;
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u)
    if (e0 + 256 * u < n) dst[doff + e0 + 256 * u] = val[u];
}

hipError_t launch_gather_blocks(int nblk, int64_t ntiles, const int64_t* table, const double* src,
                                const double* scale, double* dst, hipStream_t s) {
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_blocks_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, nblk, table, src, scale, dst);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// grouped GEMM
//   C[M x N] (col-major, ldc) = sum_s A_s^T B_s, A_s(k,m) = A[a_off + k + m*lda],
//   B_s(k,n) = B[b_off + k + n*ldb].  Tile 64x64, 4 waves in 2x2, each wave
//   2x2 MFMA 16x16 sub-tiles.  Operands are staged K-contiguous in LDS; each
//   lane reads a contiguous K-run and feeds successive MFMAs (the K order
//   inside a chunk is permuted identically for A and B).
// ---------------------------------------------------------------------------
template <int MODE>
struct GT;
// T: storage type, L: LDS/MFMA operand type, V: 16-byte global load vector (VEC elements)
template <>
struct GT<0> {  // FP64
  typedef double T;
  typedef double L;
  typedef doublex4 Acc;
  typedef double2 V;
  static constexpr int BK = 16, PAD = 2, VEC = 2;
};
template <>
struct GT<1> {  // FP32
  typedef float T;
  typedef float L;
  typedef floatx4 Acc;
  typedef floatx4 V;
  static constexpr int BK = 16, PAD = 4, VEC = 4;
};
template <>
struct GT<2> {  // BF16 operands, FP32 accumulate/storage
  typedef float T;
  typedef __bf16 L;
  typedef floatx4 Acc;
  typedef floatx4 V;
  static constexpr int BK = 32, PAD = 8, VEC = 4;
};

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// keep the first `rem` lanes of a loaded vector (rem <= 0: all zero)
template <int MODE>
__device__ __forceinline__ typename GT<MODE>::V vmask(typename GT<MODE>::V x, int64_t rem) {
  if constexpr (MODE == 0) {
    if (rem < 2) x.y = 0;
    if (rem < 1) x.x = 0;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (e >= rem) x[e] = 0;
  }
  return x;
}

template <int MODE>
__device__ __forceinline__ typename GT<MODE>::V vzero() {
  typename GT<MODE>::V x;
  if constexpr (MODE == 0) {
    x.x = 0;
    x.y = 0;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = 0;
  }
  return x;
}

template <int MODE>
__device__ __forceinline__ void vstore(typename GT<MODE>::L* dst, const typename GT<MODE>::V& x) {
  if constexpr (MODE == 2) {
    bf16x4 y;
#pragma unroll
    for (int e = 0; e < 4; ++e) y[e] = (__bf16)x[e];
    *(bf16x4*)dst = y;
  } else {
    *(typename GT<MODE>::V*)dst = x;
  }
}

// One lane's K-run of an operand row in LDS and the MFMAs that consume it:
// FP64 16x16x4 (4 k per lane = 4 MFMAs), FP32 16x16x4 (4 k = 4 MFMAs),
// BF16 16x16x32 (8 k = 1 MFMA).  mfma(c, b, a, kk) issues k-slice kk.
template <int MODE>
struct Frag;
template <>
struct Frag<0> {
  struct F {
    double2 lo, hi;
  };
  static constexpr int NK = 4;
  static __device__ __forceinline__ F load(const double* p) { return F{*(const double2*)p, *(const double2*)(p + 2)}; }
  static __device__ __forceinline__ void mfma(doublex4& c, const F& b, const F& a, int kk) {
    const double bv = kk == 0 ? b.lo.x : kk == 1 ? b.lo.y : kk == 2 ? b.hi.x : b.hi.y;
    const double av = kk == 0 ? a.lo.x : kk == 1 ? a.lo.y : kk == 2 ? a.hi.x : a.hi.y;
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, av, c, 0, 0, 0);
  }
};
template <>
struct Frag<1> {
  typedef floatx4 F;
  static constexpr int NK = 4;
  static __device__ __forceinline__ F load(const float* p) { return *(const floatx4*)p; }
  static __device__ __forceinline__ void mfma(floatx4& c, const F& b, const F& a, int kk) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(b[kk], a[kk], c, 0, 0, 0);
  }
};
template <>
struct Frag<2> {
  typedef bf16x8 F;
  static constexpr int NK = 1;
  static __device__ __forceinline__ F load(const __bf16* p) { return *(const bf16x8*)p; }
  static __device__ __forceinline__ void mfma(floatx4& c, const F& b, const F& a, int) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a, c, 0, 0, 0);
  }
};

// This work group's tiles: its own list (balanced schedule) or every
// (gridDim.x / NXCD)-th tile of its XCD's queue.  False: nothing to do.
__device__ __forceinline__ bool tile_range(const XcdQueues& xq, int& t0, int& tend, int& tstep) {
  const int xcd = blockIdx.x % NXCD, j = blockIdx.x / NXCD;
  if (xq.woff) {
    if (j >= xq.slots) return false;
    t0 = xq.woff[xcd * xq.slots + j];
    tend = xq.woff[xcd * xq.slots + j + 1];
    tstep = 1;
  } else {
    t0 = xq.off[xcd] + j;
    tend = xq.off[xcd + 1];
    tstep = gridDim.x / NXCD;
  }
  return true;
}

// Tile BM x BN, 4 waves in 2x2, each wave (BM/2) x (BN/2) = SM x SN MFMA 16x16
// sub-tiles; K chunk = KS k-slices of one MFMA group each.  Operand rows come
// from base-address tables and are 16-byte aligned and K-contiguous, so each
// thread moves 16-byte vectors (masked past K); the K loop walks the problem's
// segments chunk by chunk (split-K: a contiguous chunk range).  Global loads
// run DEPTH chunks ahead in registers; LDS is double buffered (one barrier per chunk).
template <int MODE, int BM, int BN, int KS>
__global__ __launch_bounds__(256) void gemm_grouped(
    const GemmProb* __restrict__ probs, const GemmSeg* __restrict__ segs, const int64_t* __restrict__ atab,
    const int64_t* __restrict__ btab, const int64_t* __restrict__ crow, const int64_t* __restrict__ cld,
    const GemmTile* __restrict__ tiles, XcdQueues xq, const typename GT<MODE>::T* __restrict__ A,
    const typename GT<MODE>::T* __restrict__ B, typename GT<MODE>::T* __restrict__ C,
    typename GT<MODE>::T* __restrict__ P) {
  typedef typename GT<MODE>::T T;
  typedef typename GT<MODE>::L L;
  typedef typename GT<MODE>::Acc Acc;
  typedef typename GT<MODE>::V V;
  constexpr int KL = (MODE == 2) ? 8 : 4;
  constexpr int BK = 4 * KL * KS, LD = BK + GT<MODE>::PAD;
  constexpr int VEC = GT<MODE>::VEC;
  constexpr int TPR = BK / VEC;   // threads per operand row
  constexpr int RP = 256 / TPR;   // rows per pass
  constexpr int NA = BM / RP, NB = BN / RP;
  static_assert(NA >= 1 && NB >= 1, "tile too small for the vector loader");
  constexpr int SM = BM / 32, SN = BN / 32;
  constexpr int DEPTH = (NA + NB) <= 4 ? 2 : 1;
  __shared__ __attribute__((aligned(16))) L As[2][BM][LD];
  __shared__ __attribute__((aligned(16))) L Bs[2][BN][LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int q = tid % TPR, row = tid / TPR;
  const int g = lane >> 4, l16 = lane & 15;

  // work group b runs on XCD b % NXCD (round-robin dispatch) and walks that
  // XCD's tile queue; gridDim.x is a multiple of NXCD
  int t0, tend, tstep;
  if (!tile_range(xq, t0, tend, tstep)) return;
  for (int t = t0; t < tend; t += tstep) {
    const GemmTile tile = tiles[t];
    const GemmProb p = probs[tile.prob];
    Acc acc[SM][SN];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;

    // chunk range [cb, ce) of this split over the flattened segment chunks
    int64_t ctot = 0;
    for (int s = 0; s < p.nseg; ++s) ctot += (segs[p.seg0 + s].K + BK - 1) / BK;
    const int64_t cb = ctot * tile.split / p.nsplit, ce = ctot * (tile.split + 1) / p.nsplit;
    const int64_t nch = ce - cb;
    // loader state: segment, k offset, row bases
    int seg = 0;
    int64_t k0 = 0, K = 0, cpos = 0;
    for (; seg < p.nseg; ++seg) {
      const int64_t nc = (segs[p.seg0 + seg].K + BK - 1) / BK;
      if (cpos + nc > cb) {
        k0 = (cb - cpos) * BK;
        break;
      }
      cpos += nc;
    }
    // Row bases for the current segment.  Out-of-range rows read row 0 (a valid
    // address) and are zeroed at the LDS store: loads are never predicated, so a
    // load in flight is only waited for where its data is consumed.
    int64_t ab[NA], bb[NB];
    bool av[NA], bv[NB];
#pragma unroll
    for (int i = 0; i < NA; ++i) av[i] = tile.m0 + row + i * RP < p.M;
#pragma unroll
    for (int i = 0; i < NB; ++i) bv[i] = tile.n0 + row + i * RP < p.N;
    auto bases = [&]() {
      const GemmSeg sg = segs[p.seg0 + seg];
      K = sg.K;
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int64_t r = av[i] ? tile.m0 + row + i * RP : 0;
        ab[i] = sg.atab0 >= 0 ? atab[sg.atab0 + r] : sg.a_off + r * sg.lda;
      }
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int64_t r = bv[i] ? tile.n0 + row + i * RP : 0;
        bb[i] = sg.btab0 >= 0 ? btab[sg.btab0 + r] : sg.b_off + r * sg.ldb;
      }
    };
    if (nch > 0) bases();
    V ra[DEPTH][NA], rb[DEPTH][NB];
    int64_t krem[DEPTH];
    bool full[DEPTH];  // chunk needs no K or row masking (uniform over the work group)
    const bool rows_full = tile.m0 + BM <= p.M && tile.n0 + BN <= p.N;
    // 16x16 sub-tiles of this wave that touch the problem (uniform over the wave)
    const int wr0 = tile.m0 + wm * (BM / 2), wc0 = tile.n0 + wn * (BN / 2);
    const int wvi = wr0 >= p.M ? 0 : min(SM, (p.M - wr0 + 15) / 16);
    const int wvj = wc0 >= p.N ? 0 : min(SN, (p.N - wc0 + 15) / 16);
    const bool wave_full = wvi == SM && wvj == SN;
    auto load = [&](auto D) {
      constexpr int d = decltype(D)::value;
      const int64_t k = k0 + VEC * q;
      const int64_t rem = K - k;
      const int64_t kk_ = rem > 0 ? k : 0;
      krem[d] = rem;
      full[d] = rows_full && k0 + BK <= K;
#pragma unroll
      for (int i = 0; i < NA; ++i) ra[d][i] = *(const V*)(A + ab[i] + kk_);
#pragma unroll
      for (int i = 0; i < NB; ++i) rb[d][i] = *(const V*)(B + bb[i] + kk_);
      // advance the loader by one chunk
      k0 += BK;
      if (k0 >= K && seg + 1 < p.nseg) {
        ++seg;
        k0 = 0;
        bases();
      }
    };
    auto step = [&](auto D, int64_t it) {
      constexpr int d = decltype(D)::value;
      // with two register stages chunk `it` always sits in stage it & 1, so the
      // LDS buffer index is a compile-time constant (immediate LDS offsets)
      const int lb = DEPTH == 2 ? d : (int)(it & 1);
      if (full[d]) {
#pragma unroll
        for (int i = 0; i < NA; ++i) vstore<MODE>(&As[lb][row + i * RP][VEC * q], ra[d][i]);
#pragma unroll
        for (int i = 0; i < NB; ++i) vstore<MODE>(&Bs[lb][row + i * RP][VEC * q], rb[d][i]);
      } else {
        const int64_t rem = krem[d];
#pragma unroll
        for (int i = 0; i < NA; ++i)
          vstore<MODE>(&As[lb][row + i * RP][VEC * q], vmask<MODE>(ra[d][i], av[i] ? rem : 0));
#pragma unroll
        for (int i = 0; i < NB; ++i)
          vstore<MODE>(&Bs[lb][row + i * RP][VEC * q], vmask<MODE>(rb[d][i], bv[i] ? rem : 0));
      }
      __syncthreads();
      if (it + DEPTH < nch) load(D);
      if (wave_full) {
#pragma unroll
        for (int kc = 0; kc < BK; kc += 4 * KL) {
          // operands of all sub-tiles first, then the MFMAs k-slice-outer so that
          // consecutive MFMAs accumulate into different sub-tiles (no dependent chain)
          typename Frag<MODE>::F fa[SM], fb[SN];
#pragma unroll
          for (int j = 0; j < SN; ++j) fb[j] = Frag<MODE>::load(&Bs[lb][wn * (BN / 2) + j * 16 + l16][kc + KL * g]);
#pragma unroll
          for (int i = 0; i < SM; ++i) fa[i] = Frag<MODE>::load(&As[lb][wm * (BM / 2) + i * 16 + l16][kc + KL * g]);
#pragma unroll
          for (int kk = 0; kk < Frag<MODE>::NK; ++kk)
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
              for (int j = 0; j < SN; ++j) Frag<MODE>::mfma(acc[i][j], fb[j], fa[i], kk);
        }
      } else {
        // edge tile: skip the 16x16 sub-tiles that lie wholly outside the problem
#pragma unroll
        for (int kc = 0; kc < BK; kc += 4 * KL) {
          typename Frag<MODE>::F fa[SM], fb[SN];
#pragma unroll
          for (int j = 0; j < SN; ++j)
            if (j < wvj) fb[j] = Frag<MODE>::load(&Bs[lb][wn * (BN / 2) + j * 16 + l16][kc + KL * g]);
#pragma unroll
          for (int i = 0; i < SM; ++i)
            if (i < wvi) fa[i] = Frag<MODE>::load(&As[lb][wm * (BM / 2) + i * 16 + l16][kc + KL * g]);
#pragma unroll
          for (int kk = 0; kk < Frag<MODE>::NK; ++kk)
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
              for (int j = 0; j < SN; ++j)
                if (i < wvi && j < wvj) Frag<MODE>::mfma(acc[i][j], fb[j], fa[i], kk);
        }
      }
    };
    typedef std::integral_constant<int, 0> D0;
    typedef std::integral_constant<int, DEPTH - 1> D1;
    if (nch > 0) load(D0{});
    if (DEPTH > 1 && nch > 1) load(D1{});
    for (int64_t it = 0; it < nch; it += DEPTH) {
      step(D0{}, it);
      if (DEPTH > 1 && it + 1 < nch) step(D1{}, it + 1);
    }
    // epilogue: the MFMA is fed (B, A), so sub-tile (i, j) holds C^T: lane l16
    // indexes the row m (contiguous in C) and accumulator register r the column
    // n = g + 4r (FP64 16x16x4 form) or 4g + r (FP32/BF16 forms).  Each store
    // instruction writes 16 consecutive rows of 4 columns.
    const bool split = p.nsplit > 1;
    T* Pp = P + (split ? p.part_off + (int64_t)tile.split * p.pld * p.N : 0);
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      const int m = tile.m0 + wm * (BM / 2) + i * 16 + l16;
      if (m >= p.M) continue;
      int64_t cb_ = m, cl_ = p.pld;
      if (!split) {
        if (p.crow0 >= 0) {
          cb_ = crow[p.crow0 + m];
          cl_ = cld[p.crow0 + m];
        } else {
          cb_ = p.c_off + m;
          cl_ = p.ldc;
        }
      }
      T* Cb = split ? Pp : C;
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = tile.n0 + wn * (BN / 2) + j * 16 + (MODE == 0 ? g + 4 * r : 4 * g + r);
          if (n < p.N) Cb[cb_ + (int64_t)n * cl_] = (T)acc[i][j][r];
        }
    }
    __syncthreads();  // LDS reuse by the next tile
  }
}

// ---------------------------------------------------------------------------
// FP32 / FP64 grouped GEMM with operands streamed to LDS by DMA
// (global_load_lds dwordx4: no register staging), an LDS ring of 4 (FP32) or 3
// (FP64) stages (loads run 3 or 2 K chunks ahead) and one barrier per chunk.  Same problem / segment /
// tile / split-K / output conventions as gemm_grouped (tile variant KS = 0).
// In LDS a stage holds BM + BN operand rows of 64 bytes (a K chunk of 16 floats
// or 8 doubles) in four 16-byte K groups; lane l of wave w fills the slot
// (row, q) = (w*16 + l/4 + 64 i, l%4) with the K group q ^ dma_swz(row), so
// the MFMA operand reads (16 rows at one K group) hit distinct banks.  A lane
// of MFMA lane group g reads K group g of its row and feeds one MFMA per
// element (FP32 16x16x4: 4, FP64 16x16x4: 2), the same K permutation for A
// and B.  K groups past a segment's K are read from the row start and zeroed in
// LDS (only in a segment's last chunk), as are the entries past K of the last
// group.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N < 64, "vmcnt");
  __builtin_amdgcn_s_waitcnt((N & 15) | (((N >> 4) & 3) << 14) | (7 << 4) | (15 << 8));
}
__device__ __forceinline__ void wait_lgkmcnt0() { __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4)); }
// a barrier without the compiler's release fence (which would wait for every
// outstanding vector memory operation, the DMA loads of later chunks included);
// the callers wait for the counters they need first
__device__ __forceinline__ void barrier_raw() { asm volatile("s_barrier" ::: "memory"); }

typedef __attribute__((address_space(3))) void* lds_void_ptr;

// K-group swizzle of LDS row r (rows of 16 floats = 64 bytes): 8 consecutive
// rows read at one K group hit 8 distinct 16-byte bank groups of a 128-byte
// bank line, and 16 consecutive rows 16 distinct groups of a 256-byte line
__device__ __forceinline__ int dma_swz(int r) { return ((r >> 2) ^ (r & 2)) & 3; }

// a work-group-uniform value in scalar registers
__device__ __forceinline__ int32_t uniform_i32(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uniform_i64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(x & 0xffffffffu));
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <int MODE>
struct DmaT;
template <>
struct DmaT<1> {  // FP32: 4 floats per K group, MFMA 16x16x4 f32 (output column 4g + r)
  typedef float T;
  typedef float TC;  // output and split-K partial types
  typedef float TP;
  typedef floatx4 F;
  typedef floatx4 Acc;
  static constexpr int EG = 4, NK = 4;  // elements per 16-byte K group, MFMAs per group
  static __device__ __forceinline__ void mfma(Acc& c, const F& b, const F& a, int kk) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(b[kk], a[kk], c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int g, int r) { return 4 * g + r; }
};
template <>
struct DmaT<0> {  // FP64: 2 doubles per K group, MFMA 16x16x4 f64 (output column g + 4r)
  typedef double T;
  typedef double TC;
  typedef double TP;
  typedef double2 F;
  typedef doublex4 Acc;
  static constexpr int EG = 2, NK = 2;
  static __device__ __forceinline__ void mfma(Acc& c, const F& b, const F& a, int kk) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(kk == 0 ? b.x : b.y, kk == 0 ? a.x : a.y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int g, int r) { return g + 4 * r; }
};

// BF16 storage: 8 bf16 per K group, one MFMA 16x16x32 bf16 per group (FP32
// accumulation, output column 4g + r); mode 3 writes BF16, mode 4 FP32
template <>
struct DmaT<3> {
  typedef __bf16 T;
  typedef __bf16 TC;
  typedef float TP;
  typedef bf16x8 F;
  typedef floatx4 Acc;
  static constexpr int EG = 8, NK = 1;
  static __device__ __forceinline__ void mfma(Acc& c, const F& b, const F& a, int) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int g, int r) { return 4 * g + r; }
};
template <>
struct DmaT<4> : DmaT<3> {
  typedef float TC;
};

// Write-back of one wave's staged (MS x NW) FP32 accumulator sub-tile E (column
// n at E + n * (MS + 4); the pad of 4 spreads the four MFMA lane groups' columns
// over distinct banks) to out[m + n * ld], m = m0 + 0 .. MS-1, n = n0 + 0 .. NW-1:
// 16-byte vectors of VW outputs along m (8 lanes cover 32 FP32 rows of one
// column), element stores where a vector crosses M or is not 16-byte aligned.
template <typename O, int MS, int NW>
__device__ __forceinline__ void staged_store_impl(const float* E, O* out, int64_t ld, int m0, int n0, int M, int N,
                                                  int lane) {
  constexpr int LDE = MS + 4, VW = 16 / (int)sizeof(O), MQ = MS / VW, NV = MS * NW / VW / 64;
  static_assert(NV * VW * 64 == MS * NW, "whole vectors per lane");
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const int v = lane + 64 * q, ml = (v % MQ) * VW, nl = v / MQ;
    const int m = m0 + ml, n = n0 + nl;
    if (n >= N || m >= M) continue;
    float x[VW];
#pragma unroll
    for (int e = 0; e < VW; e += 4) {
      const floatx4 f = *(const floatx4*)(E + nl * LDE + ml + e);
      x[e] = f[0];
      x[e + 1] = f[1];
      x[e + 2] = f[2];
      x[e + 3] = f[3];
    }
    O* dst = out + m + (int64_t)n * ld;
    if (m + VW <= M && ((uintptr_t)dst & 15) == 0) {
      if constexpr (sizeof(O) == 4) {
        *(floatx4*)dst = floatx4{x[0], x[1], x[2], x[3]};
      } else {
        bf16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = (__bf16)x[e];
        *(bf16x8*)dst = b;
      }
    } else {
#pragma unroll
      for (int e = 0; e < VW; ++e)
        if (m + e < M) dst[e] = (O)x[e];
    }
  }
}

// NS: LDS ring depth (loads run NS - 1 chunks ahead).  Measured on the H_eff
// GEMMs (autotuned tiles): FP32 is faster with 4 stages (3547 vs 3491
// matvecs/s), FP64 with 3 (2120 vs 2050: a smaller ring leaves LDS for one
// more resident work group per CU).
// Waves per SIMD the register allocation must allow: 64x64 FP32/BF16 tiles with
// a 4-stage ring take 32 KiB of LDS, five work groups per CU, which needs at
// most 96 registers per lane (unconstrained the variant took 82 VGPRs + 16
// AGPRs, allocated as 104: four waves per SIMD, the fifth work group of each
// CU then ran as a second wave).  Measured 1-2 % faster matvecs.
template <int MODE, int BM, int BN, int NS>
constexpr int dma_waves_per_eu() {
  return MODE != 0 && BM == 64 && BN == 64 && NS == 4 ? 5 : 1;
}
template <int MODE, int BM, int BN, int NS = (MODE == 0 ? 3 : 4)>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(dma_waves_per_eu<MODE, BM, BN, NS>()))) void gemm_dma(
    const GemmProb* __restrict__ probs, const GemmSeg* __restrict__ segs, const int64_t* __restrict__ atab,
    const int64_t* __restrict__ btab, const int64_t* __restrict__ crow, const int64_t* __restrict__ cld,
    const GemmTile* __restrict__ tiles, XcdQueues xq, const typename DmaT<MODE>::T* __restrict__ A,
    const typename DmaT<MODE>::T* __restrict__ B, typename DmaT<MODE>::TC* __restrict__ C,
    typename DmaT<MODE>::TP* __restrict__ P, const int32_t* __restrict__ recs) {
  typedef typename DmaT<MODE>::T T;
  typedef typename DmaT<MODE>::TC TC;
  typedef typename DmaT<MODE>::TP TP;
  typedef typename DmaT<MODE>::F F;
  typedef typename DmaT<MODE>::Acc Acc;
  constexpr int EG = DmaT<MODE>::EG;  // elements per 16-byte K group
  constexpr int BK = 4 * EG, NA = BM / 64, NB = BN / 64, SM = BM / 32, SN = BN / 32;
  constexpr int STG = (BM + BN) * BK;  // elements per stage
  // the four stages are distinct LDS variables so the compiler can tell a DMA
  // into one stage from the operand reads of another
  static_assert(NS == 3 || NS == 4, "LDS ring of 3 or 4 stages");
  __shared__ __attribute__((aligned(16))) T S0[STG], S1[STG], S2[STG], S3[NS == 4 ? STG : 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int g = lane >> 4, l16 = lane & 15;
  const int lrow = wave * 16 + (lane >> 2), lq = lane & 3;  // + 64 i for the i-th slot of a lane
  const int lkl = lq ^ dma_swz(lrow);  // dma_swz(64 i + lrow) does not depend on i
  int offa[SM], offb[SN];  // operand read offsets (elements) within a stage
#pragma unroll
  for (int i = 0; i < SM; ++i) {
    const int r = wm * (BM / 2) + i * 16 + l16;
    offa[i] = (r * 4 + (g ^ dma_swz(r))) * EG;
  }
#pragma unroll
  for (int j = 0; j < SN; ++j) {
    const int r = wn * (BN / 2) + j * 16 + l16;
    offb[j] = BM * BK + (r * 4 + (g ^ dma_swz(r))) * EG;
  }
  typedef std::integral_constant<int, 0> I0;
  typedef std::integral_constant<int, 1> I1;
  typedef std::integral_constant<int, 2> I2;
  typedef std::integral_constant<int, 3> I3;
  auto stage = [&](auto ST) -> T* {
    constexpr int st = decltype(ST)::value;
    return st == 0 ? S0 : st == 1 ? S1 : st == 2 ? S2 : S3;
  };
  // LDS-staged epilogue: FP32 accumulators of FP32 / BF16 modes, one wave's
  // sub-tile per ring stage (FP64 keeps the element-wise epilogue: its 3-stage
  // ring has no stage per wave)
  constexpr int LDE = BM / 2 + 4;
  constexpr bool kStaged = MODE != 0 && NS == 4 && (BN / 2) * LDE * 4 <= STG * (int)sizeof(T);
  auto staged_store = [&](const float* E, auto* out, int64_t ld, int m0, int n0, int M, int N, int ln) {
    typedef typename std::remove_pointer<decltype(out)>::type O;
    staged_store_impl<O, BM / 2, BN / 2>(E, out, ld, m0, n0, M, N, ln);
  };

  int t0, tend, tstep;
  if (!tile_range(xq, t0, tend, tstep)) return;
  // with per-tile records (recs): lane i < kTileRecWords holds word i of the next
  // tile's record, loaded one tile ahead; its words move to scalar registers at
  // the tile's start (readlane)
  int32_t rnext = 0;
  if (recs && t0 < tend && lane < kTileRecWords) rnext = recs[(int64_t)t0 * kTileRecWords + lane];
  for (int t = t0; t < tend; t += tstep) {
    GemmTile tile;
    GemmProb p;
    int nch = 0, seg = 0;
    int64_t k0 = 0, K = 0;
    GemmSeg sg0;
    if (recs) {
      const int32_t rc = rnext;
      if (t + tstep < tend && lane < kTileRecWords) rnext = recs[(int64_t)(t + tstep) * kTileRecWords + lane];
      auto w = [&](int i) { return (int32_t)__builtin_amdgcn_readlane(rc, i); };
      auto w64 = [&](int i) { return (int64_t)(((uint64_t)(uint32_t)w(i + 1) << 32) | (uint64_t)(uint32_t)w(i)); };
      tile.m0 = w(0);
      tile.n0 = w(1);
      p.M = w(2);
      p.N = w(3);
      p.seg0 = w(4);
      p.nseg = w(5);
      nch = w(6);
      seg = w(7);
      k0 = w(8);
      tile.split = w(9);
      p.nsplit = w(10);
      p.pld = w(11);
      tile.pld = w(12);
      p.crow0 = w64(14);
      p.c_off = w64(16);
      p.ldc = w64(18);
      p.part_off = w64(20);
      tile.poff = w64(22);
      sg0 = GemmSeg{w64(24), w64(26), w64(28), w64(30), w64(32), w64(34), w64(36)};
    } else {
    // tile and problem fields moved to scalar registers here: the loop below then
    // uses no register written by a pending vector load, and the compiler places
    // no full vmcnt wait (which would also drain the DMA of later chunks) in it
    tile = tiles[t];
    tile.prob = uniform_i32(tile.prob);
    tile.m0 = uniform_i32(tile.m0);
    tile.n0 = uniform_i32(tile.n0);
    tile.split = uniform_i32(tile.split);
    tile.kc0 = uniform_i32(tile.kc0);
    tile.kc1 = uniform_i32(tile.kc1);
    tile.pld = uniform_i32(tile.pld);
    tile.poff = uniform_i64(tile.poff);
    p = probs[tile.prob];
    p.M = uniform_i32(p.M);
    p.N = uniform_i32(p.N);
    p.seg0 = uniform_i32(p.seg0);
    p.nseg = uniform_i32(p.nseg);
    p.nsplit = uniform_i32(p.nsplit);
    p.crow0 = uniform_i64(p.crow0);
    p.part_off = uniform_i64(p.part_off);
    p.c_off = uniform_i64(p.c_off);
    p.ldc = uniform_i64(p.ldc);
    int64_t cb = tile.kc0, ce = tile.kc1;  // an explicit chunk range (stream-K pieces)
    if (ce < 0) {                          // else this split's share of the problem's chunks
      int64_t ctot = 0;
      for (int s = 0; s < p.nseg; ++s) ctot += (uniform_i64(segs[p.seg0 + s].K) + BK - 1) / BK;
      cb = ctot * tile.split / p.nsplit;
      ce = ctot * (tile.split + 1) / p.nsplit;
    }
    nch = (int)(ce - cb);
    // issue side: segment, k offset, this lane's row bases
    int64_t cpos = 0;
    for (; seg < p.nseg; ++seg) {
      const int64_t nc = (uniform_i64(segs[p.seg0 + seg].K) + BK - 1) / BK;
      if (cpos + nc > cb) {
        k0 = (cb - cpos) * BK;
        break;
      }
      cpos += nc;
    }
    if (nch > 0) sg0 = segs[p.seg0 + seg];
    }
    Acc acc[SM][SN];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;
    int64_t ab[NA], bb[NB];
    auto bases_from = [&](GemmSeg sg) {
      sg.K = uniform_i64(sg.K);
      sg.atab0 = uniform_i64(sg.atab0);
      sg.btab0 = uniform_i64(sg.btab0);
      sg.a_off = uniform_i64(sg.a_off);
      sg.lda = uniform_i64(sg.lda);
      sg.b_off = uniform_i64(sg.b_off);
      sg.ldb = uniform_i64(sg.ldb);
      K = sg.K;
      // rows from tables are waited for here, inside their branch, so no row
      // base is a pending load in the chunk loop
      if (sg.atab0 >= 0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
          const int64_t r = tile.m0 + lrow + 64 * i < p.M ? tile.m0 + lrow + 64 * i : 0;
          ab[i] = atab[sg.atab0 + r];
        }
        wait_vmcnt<0>();
      } else {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
          const int64_t r = tile.m0 + lrow + 64 * i < p.M ? tile.m0 + lrow + 64 * i : 0;
          ab[i] = sg.a_off + r * sg.lda;
        }
      }
      if (sg.btab0 >= 0) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          const int64_t r = tile.n0 + lrow + 64 * i < p.N ? tile.n0 + lrow + 64 * i : 0;
          bb[i] = btab[sg.btab0 + r];
        }
        wait_vmcnt<0>();
      } else {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          const int64_t r = tile.n0 + lrow + 64 * i < p.N ? tile.n0 + lrow + 64 * i : 0;
          bb[i] = sg.b_off + r * sg.ldb;
        }
      }
    };
    auto bases = [&]() { bases_from(segs[p.seg0 + seg]); };
    if (nch > 0) bases_from(sg0);
    int krem[4];  // K entries left at the start of the chunk held by each stage (>= BK: no tail)
    auto issue = [&](auto ST) {
      constexpr int st = decltype(ST)::value;
      T* sp = stage(ST);
      const int64_t kg = k0 + EG * lkl;
      const int64_t ko = kg < K ? kg : 0;  // groups past K: read the row start, zeroed later
      krem[st] = (int)std::min<int64_t>(K - k0, BK);
#pragma unroll
      for (int i = 0; i < NA; ++i)
        __builtin_amdgcn_global_load_lds((const void*)(A + ab[i] + ko),
                                         (lds_void_ptr)(sp + (i * 256 + wave * 64) * EG), 16, 0, 0);
#pragma unroll
      for (int i = 0; i < NB; ++i)
        __builtin_amdgcn_global_load_lds((const void*)(B + bb[i] + ko),
                                         (lds_void_ptr)(sp + BM * BK + (i * 256 + wave * 64) * EG), 16, 0, 0);
      k0 += BK;
      if (k0 >= K && seg + 1 < p.nseg) {
        ++seg;
        k0 = 0;
        bases();
      }
    };
    // FP64 64x128 tiles: a wave whose whole (BM/2) x (BN/2) sub-tile lies past M
    // or N skips its MFMAs (one scalar test per tile; it still issues its DMA
    // slots and meets every barrier), leaving its SIMD's MFMA pipe to the other
    // resident waves.  Measured on the step-1 GEMM: 3.5 % faster; on the 64x64
    // step-4 tiles it raised the launch's HBM fetch by 35 % for no gain, and in
    // FP32 / BF16 it gained nothing, so those keep the unconditional loop.
    constexpr bool kSkipIdle = MODE == 0 && BN == 128;
    const int wave_s = __builtin_amdgcn_readfirstlane(wave);
    const bool wact = !kSkipIdle || (tile.m0 + (wave_s >> 1) * (BM / 2) < p.M &&
                                     tile.n0 + (wave_s & 1) * (BN / 2) < p.N);
    auto compute = [&](auto ST) {
      constexpr int st = decltype(ST)::value;
      const T* sp = stage(ST);
      if (krem[st] < BK) {
        // segment tail: zero this lane's own slots past K, then make them visible
        T* wp = stage(ST);
        const int kr = krem[st];
#pragma unroll
        for (int i = 0; i < NA + NB; ++i) {
          T* slot = wp + (i < NA ? 0 : BM * BK) + ((i < NA ? i : i - NA) * 256 + wave * 64 + lane) * EG;
#pragma unroll
          for (int e = 0; e < EG; ++e)
            if (EG * lkl + e >= kr) slot[e] = T(0.0f);
        }
        wait_lgkmcnt0();
        barrier_raw();
      }
      if (!wact) return;
      F fa[SM], fb[SN];
#pragma unroll
      for (int j = 0; j < SN; ++j) fb[j] = *(const F*)(sp + offb[j]);
#pragma unroll
      for (int i = 0; i < SM; ++i) fa[i] = *(const F*)(sp + offa[i]);
#pragma unroll
      for (int kk = 0; kk < DmaT<MODE>::NK; ++kk)
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
          for (int j = 0; j < SN; ++j) DmaT<MODE>::mfma(acc[i][j], fb[j], fa[i], kk);
    };
    // chunk it: wait for its DMA (the chunks after it may stay in flight), barrier
    // (every wave's part landed; every wave is done with chunk it-1, whose stage
    // (it+3)%4 is refilled now), issue chunk it+3, compute chunk it
    auto iter = [&](auto ST, int it) {
      constexpr int st = decltype(ST)::value;
      if constexpr (NS == 4) {
        if (nch - 1 - it >= 2)
          wait_vmcnt<2 * (NA + NB)>();
        else if (nch - 1 - it == 1)
          wait_vmcnt<NA + NB>();
        else
          wait_vmcnt<0>();
      } else {
        if (nch - 1 - it >= 1)
          wait_vmcnt<NA + NB>();
        else
          wait_vmcnt<0>();
      }
      barrier_raw();
      if (it + NS - 1 < nch) issue(std::integral_constant<int, (st + NS - 1) % NS>{});
      compute(ST);
    };
    if (0 < nch) issue(I0{});
    if (1 < nch) issue(I1{});
    if (NS == 4 && 2 < nch) issue(I2{});
    for (int it = 0; it < nch; it += NS) {
      iter(I0{}, it);
      if (it + 1 < nch) iter(I1{}, it + 1);
      if (it + 2 < nch) iter(I2{}, it + 2);
      if (NS == 4 && it + 3 < nch) iter(I3{}, it + 3);
    }
    // epilogue (split-K partials and stream-K partial tiles in TP, outputs in TC).
    // Partial element (m, n) at Pp[(m - pm0) + (n - pn0) * pl]: the split's dense
    // M x N partial (pm0 = pn0 = 0, pl = pld), or the piece's BM x BN partial tile
    const bool piece = tile.pld > 0;
    const bool split = p.nsplit > 1 || piece;
    TP* Pp = P + (piece ? tile.poff : (p.nsplit > 1 ? p.part_off + (int64_t)tile.split * p.pld * p.N : 0));
    const int pm0 = piece ? tile.m0 : 0, pn0 = piece ? tile.n0 : 0;
    const int64_t pl = piece ? (int64_t)tile.pld : (int64_t)p.pld;
    if constexpr (kStaged) {
      if (split || p.crow0 < 0) {
        // outputs along m contiguous (split-K partials, or affine rows): each wave
        // stages its (BM/2) x (BN/2) accumulators in its own ring stage as FP32,
        // column n at E + n * LDE, and writes them back in 16-byte vectors along m
        barrier_raw();  // every wave is done reading the ring
        float* E = reinterpret_cast<float*>(wave == 0 ? S0 : wave == 1 ? S1 : wave == 2 ? S2 : S3);
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
          for (int j = 0; j < SN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) E[(j * 16 + DmaT<MODE>::col(g, r)) * LDE + i * 16 + l16] = acc[i][j][r];
        const int mw = tile.m0 + wm * (BM / 2), nw = tile.n0 + wn * (BN / 2);
        // the write-back's lane offsets are formed here, not hoisted out of the
        // tile loop into registers that the K loop would hold
        int ln = lane;
        asm volatile("" : "+v"(ln));
        if (split)
          staged_store(E, Pp, pl, mw - pm0, nw - pn0, p.M - pm0, p.N - pn0, ln);
        else
          staged_store(E, C + p.c_off, p.ldc, mw, nw, p.M, p.N, ln);
        barrier_raw();  // the next tile refills the stages
        continue;
      }
    }
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      const int m = tile.m0 + wm * (BM / 2) + i * 16 + l16;
      if (m >= p.M) continue;
      int64_t cb_ = m - pm0 - (int64_t)pn0 * pl, cl_ = pl;
      if (!split) {
        if (p.crow0 >= 0) {
          cb_ = crow[p.crow0 + m];
          cl_ = cld[p.crow0 + m];
        } else {
          cb_ = p.c_off + m;
          cl_ = p.ldc;
        }
      }
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = tile.n0 + wn * (BN / 2) + j * 16 + DmaT<MODE>::col(g, r);
          if (n < p.N) {
            if (split)
              Pp[cb_ + (int64_t)n * cl_] = acc[i][j][r];
            else
              C[cb_ + (int64_t)n * cl_] = (TC)acc[i][j][r];
          }
        }
    }
    barrier_raw();  // the next tile refills the stages
  }
}

// C(m, n) = sum_s P[part_off + s*pld*N + m + pld*n] for every split problem
template <typename T, typename TC = T>
// One work group per 1024 outputs of a split problem, named by the plan's
// reduction map.  Each output sums its partials in split order.  A thread's four
// outputs are summed side by side, up to eight splits' partials of them loaded
// before they are added (one memory round trip per eight splits), and the
// output row table is read before the partials, so its latency overlaps theirs.
__global__ __launch_bounds__(256) void splitk_reduce(const int32_t* __restrict__ rmap,
                                                     const GemmProb* __restrict__ probs,
                                                     const int64_t* __restrict__ crow,
                                                     const int64_t* __restrict__ cld, const T* __restrict__ P,
                                                     TC* __restrict__ C) {
  constexpr int SB = 8;  // splits loaded per round trip
  const int2 rm = ((const int2*)rmap)[blockIdx.x];
  const GemmProb p = probs[rm.x];
  const int64_t MN = (int64_t)p.M * p.N, PN = (int64_t)p.pld * p.N;
  const int64_t e0 = (int64_t)rm.y * 1024 + threadIdx.x;
  // output e = m + M * n; pld * N < 2^31 for every split problem (checked by the
  // planner): 32-bit division
  int64_t off[4], dst[4];  // partial offset and output offset of the thread's four outputs
  bool ok[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    ok[u] = e0 + 256 * u < MN;
    const int64_t e = ok[u] ? e0 + 256 * u : 0;  // past MN: output 0 is read, nothing stored
    const uint32_t ue = (uint32_t)e, uM = (uint32_t)p.M;
    const uint32_t un = ue / uM;
    const int64_t m = ue - un * uM, n = un;
    off[u] = m + n * p.pld;
    if (p.crow0 >= 0)
      dst[u] = crow[p.crow0 + m] + n * cld[p.crow0 + m];
    else
      dst[u] = p.c_off + m + n * p.ldc;
  }
  T acc[4] = {0, 0, 0, 0};
  const T* pp = P + p.part_off;
  for (int s0 = 0; s0 < p.nsplit; s0 += SB) {
    T v[SB][4];
#pragma unroll
    for (int j = 0; j < SB; ++j)
#pragma unroll
      for (int u = 0; u < 4; ++u) v[j][u] = s0 + j < p.nsplit ? pp[(int64_t)(s0 + j) * PN + off[u]] : T(0);
#pragma unroll
    for (int j = 0; j < SB; ++j)
      if (s0 + j < p.nsplit)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += v[j][u];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (ok[u]) C[dst[u]] = (TC)acc[u];
}

// Stream-K reduction: one work group per cut tile r sums the tile's np partial
// tiles (BM x BN, column-major, leading dimension BM) in piece order and writes
// the tile's part of C; each thread sums four consecutive rows of one column.
template <typename T, typename TC>
__global__ __launch_bounds__(256) void tile_reduce(const TileRed* __restrict__ reds,
                                                   const GemmProb* __restrict__ probs,
                                                   const int64_t* __restrict__ crow,
                                                   const int64_t* __restrict__ cld, const T* __restrict__ P,
                                                   TC* __restrict__ C, int bm, int bn) {
  const TileRed r = reds[blockIdx.x];
  const GemmProb p = probs[r.prob];
  const int64_t bs = (int64_t)bm * bn;
  const T* src = P + r.blk0 * bs;
  for (int e = 4 * (int)threadIdx.x; e < bm * bn; e += 1024) {
    const int ml = e % bm, nl = e / bm;  // bm is a multiple of 4: the four rows share a column
    const int m = r.m0 + ml, n = r.n0 + nl;
    if (m >= p.M || n >= p.N) continue;
    T acc[4] = {0, 0, 0, 0};
    for (int i = 0; i < r.np; ++i) {
      const T* q = src + i * bs + e;
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] += q[u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (m + u >= p.M) break;
      if (p.crow0 >= 0)
        C[crow[p.crow0 + m + u] + (int64_t)n * cld[p.crow0 + m + u]] = (TC)acc[u];
      else
        C[p.c_off + m + u + (int64_t)n * p.ldc] = (TC)acc[u];
    }
  }
}

static int g_num_cu = 0;
static int num_cu() {
  if (g_num_cu == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      g_num_cu = n;
    else
      g_num_cu = 256;
  }
  return g_num_cu;
}

// The occupancy query's count, lowered to what the CU's LDS holds (a persistent
// grid sized past it would run its last work groups as a second wave).
static int lds_fit(const void* fn, int occ) {
  hipFuncAttributes a;
  int dev = 0, lds = 0;
  if (hipFuncGetAttributes(&a, fn) != hipSuccess || a.sharedSizeBytes == 0) return occ;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || lds <= 0)
    return occ;
  const int fit = (int)((size_t)lds / a.sharedSizeBytes);
  return std::max(1, std::min(occ, fit));
}

// work groups of a variant one CU holds at once (queried once per variant)
template <int MODE, int BM, int BN, int KS>
static int occ_grouped() {
  static int occ = 0;
  if (occ == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gemm_grouped<MODE, BM, BN, KS>, 256, 0) != hipSuccess ||
        nb <= 0)
      nb = 2;
    occ = lds_fit((const void*)gemm_grouped<MODE, BM, BN, KS>, nb);
  }
  return occ;
}
template <int MODE, int BM, int BN>
static int occ_dma() {
  static int occ = 0;
  if (occ == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gemm_dma<MODE, BM, BN>, 256, 0) != hipSuccess || nb <= 0)
      nb = 2;
    occ = lds_fit((const void*)gemm_dma<MODE, BM, BN>, nb);
  }
  return occ;
}

template <int MODE, int BM, int BN, int KS>
static void launch_variant(const DevGemm& gm, const void* A, const void* B, void* C, void* P, hipStream_t s) {
  typedef typename GT<MODE>::T T;
  // persistent grid: per XCD as many work groups as its CUs hold at once
  // (occupancy of this variant), no more than its queue holds
  const int occ = occ_grouped<MODE, BM, BN, KS>();
  int qmax = 0;
  for (int x = 0; x < NXCD; ++x) qmax = std::max(qmax, gm.q.off[x + 1] - gm.q.off[x]);
  const int per = gm.q.woff ? gm.q.slots : std::min(std::max(1, num_cu() / NXCD) * occ, qmax);
  dim3 grid((unsigned)(per * NXCD)), block(256);
  hipLaunchKernelGGL((gemm_grouped<MODE, BM, BN, KS>), grid, block, 0, s, gm.probs, gm.segs, gm.atab, gm.btab,
                     gm.crow, gm.cld, gm.tiles, gm.q, (const T*)A, (const T*)B, (T*)C, (T*)P);
}

template <int MODE, int BM, int BN>
static void launch_variant_dma(const DevGemm& gm, const void* A, const void* B, void* C, void* P, hipStream_t s) {
  typedef typename DmaT<MODE>::T T;
  typedef typename DmaT<MODE>::TC TC;
  typedef typename DmaT<MODE>::TP TP;
  const int occ = occ_dma<MODE, BM, BN>();
  int qmax = 0;
  for (int x = 0; x < NXCD; ++x) qmax = std::max(qmax, gm.q.off[x + 1] - gm.q.off[x]);
  const int per = gm.q.woff ? gm.q.slots : std::min(std::max(1, num_cu() / NXCD) * occ, qmax);
  hipLaunchKernelGGL((gemm_dma<MODE, BM, BN>), dim3((unsigned)(per * NXCD)), dim3(256), 0, s, gm.probs, gm.segs,
                     gm.atab, gm.btab, gm.crow, gm.cld, gm.tiles, gm.q, (const T*)A, (const T*)B, (TC*)C, (TP*)P,
                     gm.rec_mode == MODE ? gm.recs : nullptr);
}

// Compiled tile variants: (BM, BN) in {64x64, 128x64, 64x128, 128x128} with K
// chunk multiplier KS in {1, 2} (FP64: KS = 1 only); FP32 and FP64 also KS = 0,
// the DMA-streamed kernel (K chunk 16 floats / 8 doubles) for 64x64, 128x64 and
// 64x128.
template <int MODE>
static hipError_t launch_mode(const DevGemm& gm, const void* A, const void* B, void* C, void* P, hipStream_t s) {
  const int key = gm.BM * 10000 + gm.BN * 10 + gm.KS;
#define MPSH_V(bm, bn, ks)                                  \
  case bm * 10000 + bn * 10 + ks:                          \
    launch_variant<MODE, bm, bn, ks>(gm, A, B, C, P, s); \
    break;
  if constexpr (MODE >= 3) {  // BF16 storage: DMA-streamed operands only
    switch (key) {
      case 64 * 10000 + 64 * 10:
        launch_variant_dma<MODE, 64, 64>(gm, A, B, C, P, s);
        break;
      case 128 * 10000 + 64 * 10:
        launch_variant_dma<MODE, 128, 64>(gm, A, B, C, P, s);
        break;
      case 64 * 10000 + 128 * 10:
        launch_variant_dma<MODE, 64, 128>(gm, A, B, C, P, s);
        break;
      default:
        return hipErrorInvalidValue;
    }
  } else if constexpr (MODE == 0) {
    switch (key) {
      MPSH_V(64, 64, 1)
      MPSH_V(128, 64, 1)
      MPSH_V(64, 128, 1)
      case 64 * 10000 + 64 * 10:
        launch_variant_dma<0, 64, 64>(gm, A, B, C, P, s);
        break;
      case 128 * 10000 + 64 * 10:
        launch_variant_dma<0, 128, 64>(gm, A, B, C, P, s);
        break;
      case 64 * 10000 + 128 * 10:
        launch_variant_dma<0, 64, 128>(gm, A, B, C, P, s);
        break;
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (key) {
      MPSH_V(64, 64, 1)
      MPSH_V(128, 64, 1)
      MPSH_V(64, 128, 1)
      MPSH_V(128, 128, 1)
      MPSH_V(64, 64, 2)
      MPSH_V(128, 64, 2)
      MPSH_V(64, 128, 2)
      MPSH_V(128, 128, 2)
      case 64 * 10000 + 64 * 10:
        if (MODE != 1) return hipErrorInvalidValue;
        launch_variant_dma<1, 64, 64>(gm, A, B, C, P, s);
        break;
      case 128 * 10000 + 64 * 10:
        if (MODE != 1) return hipErrorInvalidValue;
        launch_variant_dma<1, 128, 64>(gm, A, B, C, P, s);
        break;
      case 64 * 10000 + 128 * 10:
        if (MODE != 1) return hipErrorInvalidValue;
        launch_variant_dma<1, 64, 128>(gm, A, B, C, P, s);
        break;
      default:
        return hipErrorInvalidValue;
    }
  }
#undef MPSH_V
  return hipGetLastError();
}

int gemm_slots_per_xcd(int mode, int BM, int BN, int KS) {
  const int key = BM * 10000 + BN * 10 + KS;
  int occ = 0;
#define MPSH_O(m, bm, bn, ks) \
  if (mode == m && key == bm * 10000 + bn * 10 + ks) occ = occ_grouped<m, bm, bn, ks>();
#define MPSH_OD(m, bm, bn) \
  if (mode == m && key == bm * 10000 + bn * 10) occ = occ_dma<m, bm, bn>();
  MPSH_O(0, 64, 64, 1) MPSH_O(0, 128, 64, 1) MPSH_O(0, 64, 128, 1)
  MPSH_OD(0, 64, 64) MPSH_OD(0, 128, 64) MPSH_OD(0, 64, 128)
  MPSH_OD(1, 64, 64) MPSH_OD(1, 128, 64) MPSH_OD(1, 64, 128)
  MPSH_O(1, 64, 64, 1) MPSH_O(1, 128, 64, 1) MPSH_O(1, 64, 128, 1) MPSH_O(1, 128, 128, 1)
  MPSH_O(1, 64, 64, 2) MPSH_O(1, 128, 64, 2) MPSH_O(1, 64, 128, 2) MPSH_O(1, 128, 128, 2)
  MPSH_O(2, 64, 64, 1) MPSH_O(2, 128, 64, 1) MPSH_O(2, 64, 128, 1) MPSH_O(2, 128, 128, 1)
  MPSH_O(2, 64, 64, 2) MPSH_O(2, 128, 64, 2) MPSH_O(2, 64, 128, 2) MPSH_O(2, 128, 128, 2)
  MPSH_OD(3, 64, 64) MPSH_OD(3, 128, 64) MPSH_OD(3, 64, 128)
  MPSH_OD(4, 64, 64) MPSH_OD(4, 128, 64) MPSH_OD(4, 64, 128)
#undef MPSH_O
#undef MPSH_OD
  if (occ <= 0) return 0;  // not a compiled variant
  return std::max(1, num_cu() / NXCD) * occ;
}

hipError_t launch_gemm(const DevGemm& gm, int mode, const void* A, const void* B, void* C, void* P,
                       hipStream_t s) {
  if (gm.ntiles == 0) return hipSuccess;
  if ((gm.nred > 0 || gm.ntred > 0) && P == nullptr) return hipErrorInvalidValue;
  hipError_t e;
  if (mode == 0)
    e = launch_mode<0>(gm, A, B, C, P, s);
  else if (mode == 1)
    e = launch_mode<1>(gm, A, B, C, P, s);
  else if (mode == 2)
    e = launch_mode<2>(gm, A, B, C, P, s);
  else if (mode == 3)
    e = launch_mode<3>(gm, A, B, C, P, s);
  else if (mode == 4)
    e = launch_mode<4>(gm, A, B, C, P, s);
  else
    return hipErrorInvalidValue;
  if (e != hipSuccess) return e;
  if (gm.ntred > 0) {
    const dim3 tg((unsigned)gm.ntred), tb(256);
    if (mode == 0)
      hipLaunchKernelGGL((tile_reduce<double, double>), tg, tb, 0, s, gm.treds, gm.probs, gm.crow, gm.cld,
                         (const double*)P, (double*)C, gm.BM, gm.BN);
    else if (mode == 3)
      hipLaunchKernelGGL((tile_reduce<float, __bf16>), tg, tb, 0, s, gm.treds, gm.probs, gm.crow, gm.cld,
                         (const float*)P, (__bf16*)C, gm.BM, gm.BN);
    else
      hipLaunchKernelGGL((tile_reduce<float, float>), tg, tb, 0, s, gm.treds, gm.probs, gm.crow, gm.cld,
                         (const float*)P, (float*)C, gm.BM, gm.BN);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (gm.nred == 0) return hipSuccess;
  const dim3 rg((unsigned)gm.red_groups), rb(256);
  if (mode == 0)
    hipLaunchKernelGGL((splitk_reduce<double, double>), rg, rb, 0, s, gm.red_map, gm.probs, gm.crow, gm.cld,
                       (const double*)P, (double*)C);
  else if (mode == 3)
    hipLaunchKernelGGL((splitk_reduce<float, __bf16>), rg, rb, 0, s, gm.red_map, gm.probs, gm.crow, gm.cld,
                       (const float*)P, (__bf16*)C);
  else
    hipLaunchKernelGGL((splitk_reduce<float, float>), rg, rb, 0, s, gm.red_map, gm.probs, gm.crow, gm.cld,
                       (const float*)P, (float*)C);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// combined two-site MPO
// W_sig[(row0+r)*C + col0 + c] = sum_pairs sum_jm WL[jw,i1o,jm,i1] * WR[jm,i2o,jx,i2]
//   r = jx + Dx*(i1o + d1o*i2o),  c = jw + Dw*(i1 + d1i*i2)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void mpo_combine(const MpoSig* __restrict__ sigs,
                                                   const MpoTerm* __restrict__ terms,
                                                   const MpoPair* __restrict__ pairs,
                                                   const T* __restrict__ WL, const T* __restrict__ WR,
                                                   T* __restrict__ W) {
  const MpoTerm tm = terms[blockIdx.x];
  const MpoSig sg = sigs[tm.sig];
  const int n = tm.R * tm.C;
  for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
    const int r = idx / tm.C, c = idx - r * tm.C;
    const int jx = r % tm.Dx, i1o = (r / tm.Dx) % tm.d1o, i2o = r / (tm.Dx * tm.d1o);
    const int jw = c % tm.Dw, i1 = (c / tm.Dw) % tm.d1i, i2 = c / (tm.Dw * tm.d1i);
    T acc = 0;
    for (int q = 0; q < tm.npair; ++q) {
      const MpoPair pr = pairs[tm.pair0 + q];
      const int Dm = pr.Dm;
      const int64_t lbase = pr.wl_off + jw + (int64_t)tm.Dw * (i1o + (int64_t)tm.d1o * ((int64_t)Dm * i1));
      const int64_t lstep = (int64_t)tm.Dw * tm.d1o;
      const int64_t rbase = pr.wr_off + (int64_t)Dm * (i2o + (int64_t)tm.d2o * (jx + (int64_t)tm.Dx * i2));
      for (int jm = 0; jm < Dm; ++jm) acc += WL[lbase + jm * lstep] * WR[rbase + jm];
    }
    W[sg.w_off + (int64_t)(tm.row0 + r) * sg.C + tm.col0 + c] = acc;
  }
}

// Address of element offset `off` of a T1 / T3 buffer.  IX = uint32_t: `off` is
// a byte offset (zero-extended onto the buffer's base, the global-load form with
// a scalar base and a 32-bit vector offset); IX = int64_t: an element offset.
template <typename IX, typename E>
__device__ __forceinline__ E* mpo_at(E* base, IX off) {
  if constexpr (sizeof(IX) == 4)
    return reinterpret_cast<E*>(reinterpret_cast<typename std::conditional<std::is_const<E>::value, const char,
                                                                              char>::type*>(base) + off);
  else
    return base + off;
}

// One run of P pixels (m, n0 .. n0+P-1): out[u][r] = sum_c W[r][c] in[u][c], the
// C x P inputs in registers.  Each channel's address is formed once per run and
// each W value read once for the P pixels.  CM >= C is the group's channel
// class; channels past C read row C-1 (a real row) and meet zero W columns, and
// pixels past the row end re-read its last pixel, so every load is
// unconditional; only the stores are masked.  ib/ins/im and ob/on/om are in
// the units of IX (bytes for uint32_t, elements for int64_t).
// VEC (padded plans): channel rows and T3 rows have unit pixel stride and start
// on multiples of 4 elements, and the run lies inside the group's padded range,
// so a run's P pixels of one row are one aligned vector load / store; pixels
// past dR compute on the padding and store zeros (T3's padding enters the
// step-4 GEMM).
template <typename T, typename TI, typename TO, typename IX, int CM, int CMAX, int P, bool VEC>
__device__ __forceinline__ void mpo_run(const T* Ws, const IX* ib, const IX* ins, const IX* im, const IX* ob,
                                        const IX* on, const IX* om, int R, IX m, IX n0, IX dR, bool live,
                                        const TI* __restrict__ T1, TO* __restrict__ T3) {
  T v[P][CM];
  if constexpr (VEC) {
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      const TI* a = mpo_at(T1, (IX)(ib[c] + m * im[c] + n0 * ins[c]));
      if constexpr (P == 1) {
        v[0][c] = (T)*a;
      } else {
        typedef TI VI __attribute__((ext_vector_type(P)));
        const VI x = *(const VI*)a;
#pragma unroll
        for (int u = 0; u < P; ++u) v[u][c] = (T)x[u];
      }
    }
  } else {
    IX nu[P];
#pragma unroll
    for (int u = 0; u < P; ++u) nu[u] = n0 + u < dR ? n0 + u : dR - 1;
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      const IX o = ib[c] + m * im[c], st = ins[c];
#pragma unroll
      for (int u = 0; u < P; ++u) v[u][c] = (T)*mpo_at(T1, (IX)(o + nu[u] * st));
    }
  }
  if (!live) return;
  for (int r = 0; r < R; ++r) {
    const T* wr = Ws + r * CMAX;
    T acc[P];
#pragma unroll
    for (int u = 0; u < P; ++u) acc[u] = 0;
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      const T w = wr[c];
#pragma unroll
      for (int u = 0; u < P; ++u) acc[u] += w * v[u][c];
    }
    const IX o = ob[r] + m * om[r], st = on[r];
    if constexpr (VEC) {
      TO* d = mpo_at(T3, (IX)(o + n0 * st));
      if constexpr (P == 1) {
        *d = n0 < dR ? (TO)acc[0] : TO(0.0f);
      } else {
        typedef TO VO __attribute__((ext_vector_type(P)));
        VO y;
#pragma unroll
        for (int u = 0; u < P; ++u) y[u] = n0 + u < dR ? (TO)acc[u] : TO(0.0f);
        *(VO*)d = y;
      }
    } else {
#pragma unroll
      for (int u = 0; u < P; ++u)
        if (n0 + u < dR) *mpo_at(T3, (IX)(o + (n0 + u) * st)) = (TO)acc[u];
    }
  }
}

// out[r][pix] = sum_c W[r][c] * in[c][pix] over the pixels (m, n) of one
// (bra_l, ket_r) group (output row r of pixel (m, n) at ob[r] + n*on[r] + m*om[r]);
// each thread takes a run of P consecutive n of one row m, a work group 256
// runs.  CMAX bounds the plan's channel count (LDS layout); each group runs the
// smallest channel class (multiple of 4) that holds its own C.
// T: arithmetic and W type; TI / TO: T1 / T3 storage (BF16 storage computes in FP32);
// IX: offset arithmetic (uint32_t byte offsets when T1 and T3 are below 4 GiB)
template <typename T, int CMAX, int P, typename TI = T, typename TO = T, typename IX = int64_t, bool VEC = false>
__global__ __launch_bounds__(256) void mpo_apply(const MpoSig* __restrict__ sigs,
                                                 const MpoGroup* __restrict__ groups,
                                                 const MpoInRow* __restrict__ inrows,
                                                 const MpoOutRow* __restrict__ outrows,
                                                 const MpoWork* __restrict__ work, const T* __restrict__ W,
                                                 const TI* __restrict__ T1, TO* __restrict__ T3) {
  constexpr IX si = sizeof(IX) == 4 ? (IX)sizeof(TI) : 1, so = sizeof(IX) == 4 ? (IX)sizeof(TO) : 1;
  __shared__ __attribute__((aligned(16))) T Ws[64 * CMAX];
  __shared__ IX ib[CMAX], ins[CMAX], im[CMAX];
  __shared__ IX ob[64], on[64], om[64];
  const MpoWork wk = work[blockIdx.x];
  const MpoGroup g = groups[wk.group];
  const MpoSig sg = sigs[g.sig];
  const int R = sg.R, C = sg.C;
  if (R <= 0) return;  // uniform: the whole block leaves before any barrier
  for (int i = threadIdx.x; i < R * CMAX; i += blockDim.x) {
    const int r = i / CMAX, c = i - r * CMAX;
    Ws[i] = c < C ? W[sg.w_off + (int64_t)r * C + c] : T(0);
  }
  if (C > 0)
    for (int i = threadIdx.x; i < CMAX; i += blockDim.x) {
      const MpoInRow q = inrows[g.in0 + (i < C ? i : C - 1)];
      ib[i] = (IX)q.base * si;
      ins[i] = (IX)q.nstride * si;
      im[i] = (IX)q.mstride * si;
    }
  for (int i = threadIdx.x; i < R; i += blockDim.x) {
    const MpoOutRow q = outrows[g.out0 + i];
    ob[i] = (IX)q.base * so;
    on[i] = (IX)q.nstride * so;
    om[i] = (IX)q.mstride * so;
  }
  __syncthreads();
  // run q of the group: row m = q / nr, first pixel n0 = Pg * (q % nr), nr runs per
  // row over the padded range dRp (== dR unless padded); P = 0: the mixed list,
  // runs of mpo_mixed_run(C) pixels (mpo_vec_run(C, sizeof(T)) in padded plans)
  const int Pg = P > 0 ? P : (VEC ? mpo_vec_run(C, (int)sizeof(T)) : mpo_mixed_run(C));
  const uint32_t dR = (uint32_t)g.dR, dRr = (uint32_t)g.dRp, nr = (dRr + Pg - 1) / Pg;
  const int64_t nruns = (int64_t)g.db * nr;
  const int64_t q0 = (int64_t)wk.run0 + threadIdx.x;
  const bool live = q0 < nruns;
  const uint32_t q = live ? (uint32_t)q0 : 0u;
  const uint32_t mq = q / nr;
  const IX m = (IX)mq, n0 = (IX)((q - mq * nr) * Pg);
  if (C <= 0) {  // a group no input type reaches: its T3 cells (and their padding) are zero
    if (live)
      for (int r = 0; r < R; ++r)
        for (int u = 0; u < Pg; ++u)
          if (n0 + u < (IX)dRr) *mpo_at(T3, (IX)(ob[r] + (n0 + u) * on[r] + m * om[r])) = TO(0.0f);
    return;
  }
  const int cls = (C + 3) / 4;  // uniform over the block
  // P = 0: classes 1 and 2 (C <= 8) run 4 pixels per thread, the others 1; padded
  // plans mpo_vec_run(C, sizeof(T)) pixels (C <= 4 K, the same thresholds)
#define MPSH_PIX(K)                                                                                       \
  case K:                                                                                                 \
    if constexpr (4 * K <= CMAX)                                                                          \
      mpo_run<T, TI, TO, IX, 4 * K, CMAX,                                                                 \
              (P > 0 ? P : (VEC ? mpo_vec_run(4 * K, (int)sizeof(T)) : (K <= 2 ? 4 : 1))), VEC>(         \
          Ws, ib, ins, im, ob, on, om, R, m, n0, (IX)dR, live, T1, T3);                                   \
    break;
  switch (cls) {
    MPSH_PIX(1)
    MPSH_PIX(2)
    MPSH_PIX(3)
    MPSH_PIX(4)
    MPSH_PIX(5)
    MPSH_PIX(6)
    MPSH_PIX(7)
    MPSH_PIX(8)
    default:
      if constexpr (CMAX == 64)
        mpo_run<T, TI, TO, IX, 64, CMAX, (P > 0 ? P : (VEC ? mpo_vec_run(64, (int)sizeof(T)) : 1)), VEC>(
            Ws, ib, ins, im, ob, on, om, R, m, n0, (IX)dR, live, T1, T3);
      break;
  }
#undef MPSH_PIX
}

hipError_t launch_mpo_combine(const DevMpo& d, int elem_size, const void* WL, const void* WR, void* W,
                              hipStream_t s) {
  if (d.nterms == 0) return hipSuccess;
  if (elem_size == 8)
    hipLaunchKernelGGL(mpo_combine<double>, dim3(d.nterms), dim3(256), 0, s, d.sigs, d.terms, d.pairs,
                       (const double*)WL, (const double*)WR, (double*)W);
  else
    hipLaunchKernelGGL(mpo_combine<float>, dim3(d.nterms), dim3(256), 0, s, d.sigs, d.terms, d.pairs,
                       (const float*)WL, (const float*)WR, (float*)W);
  return hipGetLastError();
}

void mpo_options(DevMpo& d) {
  const char* e = std::getenv("MPSH_MPO_RUN");
  const int v = e ? std::atoi(e) : 0;
  d.run = (v == 1 || v == 2 || v == 4) ? v : 0;
  const char* f = std::getenv("MPSH_MPO_OFF64");
  d.off64 = f && std::atoi(f) != 0;
}

hipError_t launch_mpo_apply(const DevMpo& d, int elem_size, const void* W, const void* T1, void* T3, hipStream_t s) {
  // 32-bit byte offsets when every T1 / T3 byte offset fits
  const int64_t bytes = std::max(d.t1_elems, d.t3_elems) * elem_size;
  const bool off32 = !d.off64 && bytes < ((int64_t)1 << 32);
  if (d.vec) {
    // padded plans: the vector-access runs of mpo_vec_run (the list of the
    // arithmetic type's size; MPSH_MPO_RUN does not apply)
    const int pi = elem_size == 8 ? 4 : 3;
    if (d.nwork[pi] == 0) return hipSuccess;
    dim3 grid(d.nwork[pi]), block(256);
#define MPSH_MPOV(T, CM, TI, IX)                                                                           \
  hipLaunchKernelGGL((mpo_apply<T, CM, 0, TI, TI, IX, true>), grid, block, 0, s, d.sigs, d.groups, d.inrows, \
                     d.outrows, d.work[pi], (const T*)W, (const TI*)T1, (TI*)T3)
#define MPSH_MPOV_T(IX)                                                                       \
  if (elem_size == 8) {                                                                       \
    if (d.cmax <= 32) MPSH_MPOV(double, 32, double, IX); else MPSH_MPOV(double, 64, double, IX); \
  } else if (elem_size == 4) {                                                                \
    if (d.cmax <= 32) MPSH_MPOV(float, 32, float, IX); else MPSH_MPOV(float, 64, float, IX);    \
  } else {                                                                                    \
    if (d.cmax <= 32) MPSH_MPOV(float, 32, __bf16, IX); else MPSH_MPOV(float, 64, __bf16, IX);  \
  }
    if (off32) {
      MPSH_MPOV_T(uint32_t)
    } else {
      MPSH_MPOV_T(int64_t)
    }
#undef MPSH_MPOV_T
#undef MPSH_MPOV
    return hipGetLastError();
  }
  // d.run pixels per thread (0: the mixed list; FP64 then one pixel per thread,
  // measured 2 % faster than mixed there, while FP32 / BF16 gain 0-4 % from it);
  // wide channel classes (CMAX 64) and FP64 at most 2
  const int pw = d.run == 0 ? (elem_size == 8 ? 1 : 0) : (elem_size == 8 ? std::min(2, d.run) : d.run);
  const int P = pw == 0 ? 0 : (d.cmax > 32 ? std::max(1, pw / 2) : pw);
  const int pi = P == 0 ? 3 : (P == 4 ? 2 : (P == 2 ? 1 : 0));  // work list
  if (d.nwork[pi] == 0) return hipSuccess;
  dim3 grid(d.nwork[pi]), block(256);
#define MPSH_MPO(T, CM, PP, TI, IX)                                                                                \
  hipLaunchKernelGGL((mpo_apply<T, CM, PP, TI, TI, IX>), grid, block, 0, s, d.sigs, d.groups, d.inrows, d.outrows, \
                     d.work[pi], (const T*)W, (const TI*)T1, (TI*)T3)
#define MPSH_MPO_P(T, TI, IX)                   \
  if (d.cmax <= 32) {                           \
    if (P == 0) MPSH_MPO(T, 32, 0, TI, IX);     \
    else if (P == 4) MPSH_MPO(T, 32, 4, TI, IX);\
    else if (P == 2) MPSH_MPO(T, 32, 2, TI, IX);\
    else MPSH_MPO(T, 32, 1, TI, IX);            \
  } else {                                      \
    if (P == 0) MPSH_MPO(T, 64, 0, TI, IX);     \
    else if (P == 2) MPSH_MPO(T, 64, 2, TI, IX);\
    else MPSH_MPO(T, 64, 1, TI, IX);            \
  }
#define MPSH_MPO_T(IX)                                            \
  if (elem_size == 8) {                                           \
    if (d.cmax <= 32) {                                           \
      if (P == 0) MPSH_MPO(double, 32, 0, double, IX);            \
      else if (P == 2) MPSH_MPO(double, 32, 2, double, IX);       \
      else MPSH_MPO(double, 32, 1, double, IX);                   \
    } else {                                                      \
      if (P == 0) MPSH_MPO(double, 64, 0, double, IX);            \
      else MPSH_MPO(double, 64, 1, double, IX);                   \
    }                                                             \
  } else if (elem_size == 4) {                                    \
    MPSH_MPO_P(float, float, IX)                                  \
  } else { /* BF16 T1 / T3, FP32 W and arithmetic */              \
    MPSH_MPO_P(float, __bf16, IX)                                 \
  }
  if (off32) {
    MPSH_MPO_T(uint32_t)
  } else {
    MPSH_MPO_T(int64_t)
  }
#undef MPSH_MPO_T
#undef MPSH_MPO_P
#undef MPSH_MPO
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// vector ops
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void dot_partial(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                   double* __restrict__ partials) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    acc += (double)x[i] * (double)y[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void dot_final(int np, const double* __restrict__ partials,
                                                 double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) acc += partials[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

hipError_t launch_dot(int elem_size, int64_t n, const void* x, const void* y, double* out, double* partials,
                      int max_partials, hipStream_t s) {
  int64_t nb = (n + 2047) / 2048;
  if (nb < 1) nb = 1;
  if (nb > max_partials) nb = max_partials;
  if (elem_size == 8)
    hipLaunchKernelGGL(dot_partial<double>, dim3((unsigned)nb), dim3(256), 0, s, n, (const double*)x,
                       (const double*)y, partials);
  else
    hipLaunchKernelGGL(dot_partial<float>, dim3((unsigned)nb), dim3(256), 0, s, n, (const float*)x,
                       (const float*)y, partials);
  hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, s, (int)nb, partials, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Lanczos basis operations (FP64, scalars stay on the device)
//   dots:      c[i] = <V_i, w>, i < k       (block partials, fixed-order reduction)
//   update:    w += sign * sum_i c[i] V_i   (c read from device memory)
//   normalize: beta = ||w||, out = w / beta
// V_i = V + i * ldv.  Partials: k * LZ_PARTS doubles.
// ---------------------------------------------------------------------------
constexpr int LZ_PARTS = 256;

__global__ __launch_bounds__(256) void lz_dots_partial(int k, int64_t n, const double* __restrict__ V, int64_t ldv,
                                                       const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[8][256];
  // block b handles a contiguous slice of n for up to 8 basis vectors at a time
  const int64_t per = (n + LZ_PARTS - 1) / LZ_PARTS;
  const int64_t e0 = (int64_t)blockIdx.x * per, e1 = e0 + per < n ? e0 + per : n;
  // one work group per CU holds only four waves, so each thread issues the
  // loads of four elements of every basis vector before its first FMA
  for (int i0 = 0; i0 < k; i0 += 8) {
    const int ni = k - i0 < 8 ? k - i0 : 8;
    double acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 1024) {
      double we[4], vv[8][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) we[u] = e + 256 * u < e1 ? w[e + 256 * u] : 0.0;
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          vv[t][u] = t < ni && e + 256 * u < e1 ? V[(int64_t)(i0 + t) * ldv + e + 256 * u] : 0.0;
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[t] += vv[t][u] * we[u];
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) sh[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h)
#pragma unroll
        for (int t = 0; t < 8; ++t) sh[t][threadIdx.x] += sh[t][threadIdx.x + h];
      __syncthreads();
    }
    if ((int)threadIdx.x < ni) part[(int64_t)(i0 + threadIdx.x) * LZ_PARTS + blockIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
  }
}

// one block per basis vector: c[i] = sum of its LZ_PARTS partials (fixed order)
__global__ __launch_bounds__(256) void lz_dots_final(const double* __restrict__ part, double* __restrict__ c,
                                                     int sqrt_out) {
  __shared__ double sh[256];
  sh[threadIdx.x] = part[(int64_t)blockIdx.x * LZ_PARTS + threadIdx.x];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) c[blockIdx.x] = sqrt_out ? sqrt(sh[0]) : sh[0];
}

__global__ __launch_bounds__(256) void lz_update(int k, int64_t n, const double* __restrict__ V, int64_t ldv,
                                                 const double* __restrict__ c, double sign, double* __restrict__ w) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += c[i] * V[(int64_t)i * ldv + e];
    w[e] += sign * acc;
  }
}

__global__ __launch_bounds__(256) void lz_scale(int64_t n, const double* __restrict__ w, const double* __restrict__ beta,
                                                double* __restrict__ out) {
  const double b = beta[0];
  const double r = b > 0.0 ? 1.0 / b : 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) out[e] = w[e] * r;
}

hipError_t launch_lz_dots(int k, int64_t n, const double* V, int64_t ldv, const double* w, double* c, double* part,
                          hipStream_t s) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(lz_dots_partial, dim3(LZ_PARTS), dim3(256), 0, s, k, n, V, ldv, w, part);
  hipLaunchKernelGGL(lz_dots_final, dim3(k), dim3(256), 0, s, part, c, 0);
  return hipGetLastError();
}

hipError_t launch_lz_update(int k, int64_t n, const double* V, int64_t ldv, const double* c, double sign, double* w,
                            hipStream_t s) {
  if (k <= 0 || n <= 0) return hipSuccess;
  int64_t nb = (n + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(lz_update, dim3((unsigned)nb), dim3(256), 0, s, k, n, V, ldv, c, sign, w);
  return hipGetLastError();
}

hipError_t launch_lz_normalize(int64_t n, const double* w, double* out, double* beta, double* part, hipStream_t s) {
  hipLaunchKernelGGL(lz_dots_partial, dim3(LZ_PARTS), dim3(256), 0, s, 1, n, w, (int64_t)0, w, part);
  hipLaunchKernelGGL(lz_dots_final, dim3(1), dim3(256), 0, s, part, beta, 1);
  int64_t nb = (n + 255) / 256;
  if (nb > 4096) nb = 4096;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(lz_scale, dim3((unsigned)nb), dim3(256), 0, s, n, w, beta, out);
  return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void axpby_kernel(int64_t n, double a, const T* __restrict__ x, double b,
                                                    T* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    y[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void scal_kernel(int64_t n, double a, T* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    x[i] = (T)(a * (double)x[i]);
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void convert_kernel(int64_t n, const S* __restrict__ x, D* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    y[i] = (D)x[i];
}

static unsigned elt_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (unsigned)g;
}

hipError_t launch_axpby(int elem_size, int64_t n, double a, const void* x, double b, void* y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (elem_size == 8)
    hipLaunchKernelGGL(axpby_kernel<double>, dim3(elt_grid(n)), dim3(256), 0, s, n, a, (const double*)x, b,
                       (double*)y);
  else
    hipLaunchKernelGGL(axpby_kernel<float>, dim3(elt_grid(n)), dim3(256), 0, s, n, a, (const float*)x, b,
                       (float*)y);
  return hipGetLastError();
}

hipError_t launch_scal(int elem_size, int64_t n, double a, void* x, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (elem_size == 8)
    hipLaunchKernelGGL(scal_kernel<double>, dim3(elt_grid(n)), dim3(256), 0, s, n, a, (double*)x);
  else
    hipLaunchKernelGGL(scal_kernel<float>, dim3(elt_grid(n)), dim3(256), 0, s, n, a, (float*)x);
  return hipGetLastError();
}

hipError_t launch_convert(int src_size, const void* src, int dst_size, void* dst, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (src_size == 8 && dst_size == 4)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(elt_grid(n)), dim3(256), 0, s, n,
                       (const double*)src, (float*)dst);
  else if (src_size == 4 && dst_size == 8)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(elt_grid(n)), dim3(256), 0, s, n,
                       (const float*)src, (double*)dst);
  else if (src_size == 8)
    hipLaunchKernelGGL((convert_kernel<double, double>), dim3(elt_grid(n)), dim3(256), 0, s, n,
                       (const double*)src, (double*)dst);
  else
    hipLaunchKernelGGL((convert_kernel<float, float>), dim3(elt_grid(n)), dim3(256), 0, s, n,
                       (const float*)src, (float*)dst);
  return hipGetLastError();
}

}  // namespace mpsh
