// Batched per-block FP64 SVD on the device: the per-sector SVD of
// svdtrunc / svd (reference src/mpsutils.jl:116-179, src/symtools.jl:1-35).
//
// Default path: one-sided block Jacobi (Hestenes) on all blocks at once.  A
// block A (m x n) is worked as W = A (m >= n) or W = A^T (m < n), r x c with
// r >= c, columns split into blocks of w columns.  One launch per round of a
// round-robin tournament over the column blocks: every work group loads one
// pair of column blocks (r x 2w) into LDS and runs one cyclic sweep of plane
// rotations over its 2w columns, each rotation computed from the current
// columns' dot products (Rutishauser's formula), then writes the columns back
// and applies the accumulated 2w x 2w rotation to the c x c matrix Q.  A
// matrix is done after a sweep in which no pair had
// |x_i.x_j| > sqrt(r) eps |x_i| |x_j|; then sigma_j = |W_j|, the left vectors
// are W_j / sigma_j and the right vectors the columns of Q, sorted by
// descending sigma.  Matrices with at most 2w columns are solved in one work
// group.  Rows up to 2048 fit LDS; larger blocks use the rocSOLVER path.
// Matrices of 96 or more columns are first rotated by the eigenvectors of
// their Gram matrix (rocSOLVER syevd), which cuts the sweeps on graded DMRG
// spectra from ~30 to ~7 (details at JAC_PRE_MIN).
//
// rocSOLVER path (MPSH_SVD_ALGO=gesdd, or rows > 2048): blocks are zero-padded
// into square shape buckets and each bucket is one gesdd strided-batched call,
// the divide-and-conquer SVD the reference gets from LAPACK gesdd.  Zero
// padding adds only zero singular values; the leading min(m, n) singular
// triplets of a block are read back from its padded slot.  A bucket whose
// gesdd reports non-convergence is redone with gesvd (QR iteration), as the
// reference retries gesdd failures with gesvd (src/symtools.jl:14-19).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpsh.h"
#include "kernels.hpp"
#include "planner.hpp"

using namespace mpsh;

namespace {

// Square shape buckets.  A rocSOLVER batched call costs a fixed ~4 ms plus
// roughly c * P^3 per padded P x P matrix (measured on MI355X: 60 matrices of
// 100 / 200 / 400 take 4.7 / 12.4 / 34.7 ms), so the buckets are chosen by a
// dynamic program over the blocks sorted by size that minimises
// sum_buckets (fixed + c * count * P_max^3); the fixed cost equals about 766^3
// of padded work.
constexpr double kBucketFixed = 766.0 * 766.0 * 766.0;

int64_t pad16(int64_t x) { return (x + 15) / 16 * 16; }

// bucket edges over the ascending padded sizes P[0..n): returns the end index of each bucket
std::vector<int64_t> choose_buckets(const std::vector<int64_t>& P) {
  const int64_t n = (int64_t)P.size();
  std::vector<double> best(n + 1, 0.0);
  std::vector<int64_t> from(n + 1, 0);
  for (int64_t j = 1; j <= n; ++j) {
    best[j] = 1e300;
    const double p3 = (double)P[j - 1] * P[j - 1] * P[j - 1];
    for (int64_t i = 0; i < j; ++i) {
      const double c = best[i] + kBucketFixed + (double)(j - i) * p3;
      if (c < best[j]) {
        best[j] = c;
        from[j] = i;
      }
    }
  }
  std::vector<int64_t> ends;
  for (int64_t j = n; j > 0; j = from[j]) ends.push_back(j);
  std::reverse(ends.begin(), ends.end());
  return ends;
}

struct Bucket {
  int64_t Mp, Np, kp, count;
  int64_t a_off, s_off, u_off, v_off, info_off;  // workspace offsets (doubles; info in int32 units)
};

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

void ck(hipError_t e, const char* w) {
  if (e != hipSuccess) throw Error(ERR_DEVICE, std::string(w) + ": " + hipGetErrorString(e));
}
void ckb(rocblas_status e, const char* w) {
  if (e != rocblas_status_success) throw Error(ERR_DEVICE, std::string(w) + ": rocblas status " + std::to_string(e));
}

// ---------------------------------------------------------------------------
// one-sided block Jacobi
// ---------------------------------------------------------------------------
constexpr int JT = 256;                 // threads per work group
constexpr int64_t JAC_MAX_ROWS = 2048;  // r limit of the LDS-resident column pairs
constexpr int JAC_LDS = 17472;          // doubles: 2w * ldx + 2w * (2w + 1) for the largest (w, r) class
constexpr int JAC_MAX_SWEEPS = 60;
constexpr double JAC_SETTLED = 1e-10;  // a sweep whose rotations all stay below this ends the matrix

struct JacMat {
  int64_t r, c;          // work matrix W (r x c, ld r), r >= c >= 1
  int64_t w_off, q_off;  // workspace offsets (doubles): W and Q (c x c, ld c)
  int64_t w0_off;        // the scaled block before preconditioning (== w_off when not preconditioned)
  int64_t g_off, gld;    // preconditioned: Gram / eigenvector slot (gld x gld) and its leading dimension
  int64_t m, n;          // the block's shape
  int64_t u_off, s_off, v_off;  // packed outputs
  double ctol;           // sqrt(r) * eps
  int32_t w, nb;         // column-block width and block count (even, >= 2)
  int32_t trans, pre;    // trans: W = A^T; pre: preconditioned by the Gram eigenvectors
  int32_t g_info, pad;   // preconditioned: index of the syevd info entry
  int64_t sg_off, rk_off;  // sigma (c doubles) and rank (c int32) of the finish
};
struct JacTask {
  int32_t mat, k;  // matrix (index into the JacMat table) and pair index within a round
};

// columns and width of a block width: w in {16, 8, 4} so that 2w * (r | 1) doubles fit the LDS budget
int jac_width(int64_t r) { return r <= 512 ? 16 : (r <= 1024 ? 8 : 4); }

// Preconditioning.  Cyclic one-sided Jacobi converges only linearly for a
// long while on the graded spectra of DMRG two-site blocks (27-30 sweeps for
// 900 x 800 blocks whose singular values span 1 .. 1e-9, against 13 for
// Gaussian blocks).  Large matrices are first rotated by the eigenvectors V of
// their Gram matrix W^T W (rocSOLVER syevd, FP64, batched over padded size
// buckets): the columns of W V are orthogonal up to eps |W|^2 / (sigma_i
// sigma_j) relative, and the Jacobi sweeps that follow (started from Q = V)
// restore full relative orthogonality in 5-7 sweeps.  V enters only as an
// orthogonal change of basis (syevd's vectors are orthonormal to rounding), so
// the singular values and vectors keep the accuracy of the Jacobi sweeps.  A
// matrix whose syevd reports failure keeps Q = I and is swept
// unpreconditioned.  Every matrix of at least JAC_PRE_MIN columns is
// preconditioned: leaving the smaller ones plain (their ~30 sweeps run after
// the large matrices are done, each sweep a host synchronisation and a launch
// per round) measured slower than their syevd buckets, which run on the side
// streams next to the largest one.
constexpr int64_t JAC_PRE_MIN = 96;
// syevd cost model (MI355X, rocSOLVER 3.30, ms): a bucket of `count` padded
// P x P matrices costs ~ 0.022 P (the sequential Householder and merge steps)
// + 1.64e-9 count P^3 (16 matrices of 800: 31 ms, one: 18 ms; 16 of 400: 12 ms)
std::vector<int64_t> choose_gram_buckets(const std::vector<int64_t>& P) {
  const int64_t n = (int64_t)P.size();
  std::vector<double> best(n + 1, 0.0);
  std::vector<int64_t> from(n + 1, 0);
  for (int64_t j = 1; j <= n; ++j) {
    best[j] = 1e300;
    const double p = (double)P[j - 1];
    for (int64_t i = 0; i < j; ++i) {
      const double c = best[i] + 0.022 * p + 1.64e-9 * (double)(j - i) * p * p * p;
      if (c < best[j]) {
        best[j] = c;
        from[j] = i;
      }
    }
  }
  std::vector<int64_t> ends;
  for (int64_t j = n; j > 0; j = from[j]) ends.push_back(j);
  std::reverse(ends.begin(), ends.end());
  return ends;
}
struct GramBucket {
  int64_t P, count;       // padded order and number of matrices
  int64_t g_off, d_off, e_off;  // (jacobi region, doubles) slots P x P, eigenvalues, syevd scratch
  int32_t mat_start, info_start;  // first entry in the bucket-ordered matrix list / info array
};

__device__ __forceinline__ void jac_pair_of(int n, int t, int k, int& p, int& q) {
  // round t of the circle method over n (even) players: player n-1 fixed
  if (k == 0) {
    p = t % (n - 1);
    q = n - 1;
  } else {
    p = (t + k) % (n - 1);
    q = (t - k + (n - 1)) % (n - 1);
  }
}

// sum over each row of 16 lanes, returned in every lane of the row: DPP quad
// swaps, then rotations by 4 and 8 within the row, then lane 0's sum broadcast
// to the row (the rotation sums are added in lane-dependent order, and every
// lane of a rotation must use bitwise the same value)
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_row_sum(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x128>(v);  // row_ror:8
  return dpp_mov<0x150>(v);  // row_newbcast:0
}

// One pair of column blocks (p, q) per work group, X = [W_p W_q] (r x 2w) in
// LDS.  A sweep over the pair: (a) the pairs inside p and inside q (w - 1
// steps, circle order; only in round 0 of a sweep, where every block appears
// once), (b) the w x w cross pairs in w steps: group k keeps column k of p in
// registers and meets column (k + s) mod w of q at step s.  Over the nb - 1
// rounds of a sweep every column pair of the matrix is rotated once.
// iterate = 1 (matrices of two blocks, the pair holds every column): sweeps
// until converged.  GSZ lanes per rotation (w GSZ threads per work group); a
// lane holds RPL rows of a column, r <= GSZ RPL.  Rows r .. GSZ RPL of every
// column are zero, so loads and stores run unmasked.  The 2w x 2w rotation R is kept
// transposed (Rt[row * (2w + 1) + col]).
// Squared column norms live in LDS (nrm): computed from the columns at the
// start of every sweep of the work group (every round, for the round
// launches), then carried through each rotation by |x_i'|^2 = |x_i|^2 - t g,
// |x_j'|^2 = |x_j|^2 + t g (t the rotation's tangent, g = x_i.x_j), so a
// rotation needs one dot product instead of three; a norm that falls below
// 1/8 of its old value (cancellation) is recomputed from the column.
// offmax[mat] is set to 1 when a pair of the sweep was rotated at a relative
// product |x_i.x_j| / (|x_i| |x_j|) above JAC_SETTLED.  A sweep whose rotations
// all stayed below that ends the matrix: a later rotation (i, k) by an angle
// <= 1e-10 changes x_i by at most 1e-10 |x_k| in the direction of a column k
// whose product with x_j is itself <= 1e-10 relative, so every pair ends the
// sweep with a relative product of order c 1e-20 (below ctol = sqrt(r) eps for
// c up to 2048), and the sweep that would only confirm this is not run.
__device__ __forceinline__ double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
// sum over the 64 lanes of a wave, the same value in every lane: row sums by
// DPP, then the four rows' sums read from lanes 0, 16, 32, 48
__device__ __forceinline__ double wave_sum(double v) {
  v = wave_row_sum(v);
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// 1/x and 1/sqrt(x): the hardware estimates refined by two Newton steps (each
// squares the relative error).  The rotation is orthogonal to rounding because
// cs is the refined 1/sqrt(1 + t^2) and sn = cs t.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = fma(y, fma(-h * y, y, 0.5), y);
  return fma(y, fma(-h * y, y, 0.5), y);
}

// squared norm of a column held as RPL rows per lane (one wave), in every lane
template <int RPL>
__device__ __forceinline__ double jac_sumsq(const double (&x)[RPL]) {
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int u = 0; u < RPL; u += 2) {
    s0 = fma(x[u], x[u], s0);
    s1 = fma(x[u + 1], x[u + 1], s1);
  }
  return wave_sum(s0 + s1);
}

// One rotation of columns (i, j) held in registers by one wave (rows lane +
// 64 u); nrm: the columns' squared norms (LDS), Rt: the transposed accumulated
// rotation (LDS, row stride RTS, NC rows).  True when rotated.
template <int RPL, int NC, int RTS>
__device__ __forceinline__ bool jac_rotate(double (&xi)[RPL], double (&xj)[RPL], int i, int j, double* nrm,
                                           double* Rt, int lane, double ctol, bool& unsettled) {
  double g0 = 0.0, g1 = 0.0;
#pragma unroll
  for (int u = 0; u < RPL; u += 2) {
    g0 = fma(xi[u], xj[u], g0);
    g1 = fma(xi[u + 1], xj[u + 1], g1);
  }
  const double a = nrm[i], b = nrm[j], g = wave_sum(g0 + g1);
  // rotate when |g| > ctol |x_i| |x_j|; columns of norm^2 <= 1e-280 (the
  // matrix is scaled to entries < 1) count as zero, and so does NaN
  // (the threshold test needs no full precision: raw hardware estimates)
  const double off = fabs(g) * __builtin_amdgcn_rsq(a) * __builtin_amdgcn_rsq(b);
  const bool rot = a > 1e-280 && b > 1e-280 && off > ctol;
  if (rot && off > JAC_SETTLED) unsettled = true;
  if (!__any(rot)) return false;
  // t = tan of the angle that zeroes the pair's product (Rutishauser):
  // t = sign(z) / (|z| + sqrt(1 + z^2)), z = (b - a) / (2g); for |z| > 1e150
  // sqrt(1 + z^2) is taken as |z| (z^2 would overflow), t = 1 / (2z).
  const double z = (b - a) * rcp_nr(2.0 * (rot ? g : 1.0));
  const double az = fabs(z);
  const double w1 = fma(az, az, 1.0);
  const double sq = az > 1e150 ? az : w1 * rsq_nr(w1);
  const double tt = copysign(rcp_nr(az + sq), z);
  const double c0 = rsq_nr(fma(tt, tt, 1.0));
  const double cs = rot ? c0 : 1.0, sn = rot ? c0 * tt : 0.0;
#pragma unroll
  for (int u = 0; u < RPL; ++u) {
    const double x = xi[u], y = xj[u];
    xi[u] = cs * x - sn * y;
    xj[u] = sn * x + cs * y;
  }
  if (rot && lane < NC) {
    const double x = Rt[lane * RTS + i], y = Rt[lane * RTS + j];
    Rt[lane * RTS + i] = cs * x - sn * y;
    Rt[lane * RTS + j] = sn * x + cs * y;
  }
  // the norms after the rotation (uniform over the wave, as a, b, g, tt are)
  double na = rot ? fma(-tt, g, a) : a, nb = rot ? fma(tt, g, b) : b;
  const bool ra = rot && na < 0.125 * a, rb = rot && nb < 0.125 * b;
  if (__any(ra)) {
    const double e = jac_sumsq<RPL>(xi);
    if (ra) na = e;
  }
  if (__any(rb)) {
    const double e = jac_sumsq<RPL>(xj);
    if (rb) nb = e;
  }
  if (rot && lane == 0) {
    nrm[i] = na;
    nrm[j] = nb;
  }
  return rot;
}

template <int WB, int GSZ>
__global__ __launch_bounds__(WB * GSZ) void jac_pairs(const JacMat* __restrict__ mats,
                                                      const JacTask* __restrict__ tasks, int t, int iterate,
                                                      double* __restrict__ ws,
                                                      unsigned long long* __restrict__ offmax,
                                                      int* __restrict__ done, int* __restrict__ flags) {
  // one wave (GSZ = 64 lanes) per rotation -- two or four rotations per wave
  // sharing its scalar instructions measured no faster; a lane holds RPL rows
  // of a column
  constexpr int NT = WB * GSZ, NC = 2 * WB, RPL = (WB == 16 ? 512 : (WB == 8 ? 1024 : 2048)) / GSZ,
                LDX = GSZ * RPL + 1, RTS = NC + 1;
  static_assert(NC <= GSZ, "one Rt row per lane of a group");
  static_assert(GSZ == 64, "one wave per rotation");
  static_assert(NC * LDX + NC * RTS <= JAC_LDS, "LDS budget");
  __shared__ __align__(16) double lds[JAC_LDS];
  __shared__ double nrm[NC];
  __shared__ int rot_any, touched;
  const JacTask task = tasks[blockIdx.x];
  const JacMat M = mats[task.mat];
  if (done[task.mat]) return;
  if (!iterate && t >= M.nb - 1) return;
  const int r = (int)M.r, c = (int)M.c;
  double* X = lds;
  double* Rt = lds + NC * LDX;
  int p, q;
  jac_pair_of(M.nb, iterate ? 0 : t, task.k, p, q);
  auto gcol = [&](int lc) -> int {
    const int g = lc < WB ? p * WB + lc : q * WB + (lc - WB);
    return g < c ? g : -1;
  };
  double* W = ws + M.w_off;
  double* Q = ws + M.q_off;
  const int tid = threadIdx.x;
  // staging, write-back and Q loops run column by column: wave k takes local
  // columns k and k + WB, lane l rows l + 64 u (coalesced, no per-element
  // index division); loads are unpredicated from clamped addresses with
  // out-of-range values zeroed after, all of a column in flight at once
  const int grp = tid / GSZ, lane = tid % GSZ;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lc = grp + h * WB, g = gcol(lc);
    const double* src = W + (int64_t)(g >= 0 ? g : 0) * r;
    double v[RPL];
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int row = lane + GSZ * u;
      const bool in = g >= 0 && row < r;
      const double x = src[in ? row : 0];
      v[u] = in ? x : 0.0;
    }
#pragma unroll
    for (int u = 0; u < RPL; ++u) X[lc * LDX + lane + GSZ * u] = v[u];
  }
  for (int e = tid; e < NC * RTS; e += NT) Rt[e] = (e / RTS == e % RTS) ? 1.0 : 0.0;
  if (tid == 0) {
    rot_any = 0;
    touched = 0;
  }
  __syncthreads();
  const double ctol = M.ctol;
  const int max_sweeps = iterate ? JAC_MAX_SWEEPS : 1;
  bool unsettled = false;

  auto sumsq = [&](const double (&x)[RPL]) -> double { return jac_sumsq<RPL>(x); };
  // rotate columns (i, j) held in registers (rows lane + 64 u); true when rotated
  auto rotate = [&](double (&xi)[RPL], double (&xj)[RPL], int i, int j) {
    return jac_rotate<RPL, NC, RTS>(xi, xj, i, j, nrm, Rt, lane, ctol, unsettled);
  };
  auto load = [&](double (&x)[RPL], int col) {
    const double* src = X + col * LDX + lane;
#pragma unroll
    for (int u = 0; u < RPL; ++u) x[u] = src[u * GSZ];
  };
  auto store = [&](const double (&x)[RPL], int col) {
    double* dst = X + col * LDX + lane;
#pragma unroll
    for (int u = 0; u < RPL; ++u) dst[u * GSZ] = x[u];
  };

  int sweep = 0;
  int conv = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double xi[RPL], xj[RPL];
    // exact squared norms of the 2w columns: group k takes columns k and k + WB
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      load(xi, grp + h * WB);
      const double e = sumsq(xi);
      if (lane == 0) nrm[grp + h * WB] = e;
    }
    __syncthreads();
    if (iterate || t == 0) {
      // (a) pairs inside p (groups 0 .. WB/2-1) and inside q (groups WB/2 ..)
      const int half = WB / 2, base = grp < half ? 0 : WB, k = grp < half ? grp : grp - half;
#pragma unroll 1
      for (int s2 = 0; s2 < WB - 1; ++s2) {
        int i, j;
        jac_pair_of(WB, s2, k, i, j);
        i += base;
        j += base;
        load(xi, i);
        load(xj, j);
        if (rotate(xi, xj, i, j)) {
          store(xi, i);
          store(xj, j);
          touched = 1;
        }
        __syncthreads();
      }
    }
    // (b) cross pairs: group k keeps column k of p
    load(xi, grp);
    bool moved = false;
#pragma unroll 1
    for (int s2 = 0; s2 < WB; ++s2) {
      const int j = WB + (grp + s2) % WB;
      load(xj, j);
      if (rotate(xi, xj, grp, j)) {
        store(xj, j);
        moved = true;
        touched = 1;
      }
      __syncthreads();
    }
    if (moved) store(xi, grp);
    if (unsettled) rot_any = 1;
    __syncthreads();
    conv = !rot_any;
    __syncthreads();
    if (tid == 0) rot_any = 0;
    unsettled = false;
    if (!iterate && !conv && tid == 0) offmax[task.mat] = 1ull;
    if (conv) break;
  }
  if (iterate && tid == 0) {
    if (conv)
      done[task.mat] = 1;
    else
      flags[1] = 1;  // not converged
  }
  // nothing rotated: W and Q are unchanged (the last barrier of the sweep
  // loop orders every write of `touched` before this read)
  if (!touched) return;
  // columns back to W
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lc = grp + h * WB, g = gcol(lc);
    if (g >= 0) {
      double* dst = W + (int64_t)g * r;
#pragma unroll
      for (int u = 0; u < RPL; ++u) {
        const int row = lane + GSZ * u;
        if (row < r) dst[row] = X[lc * LDX + row];
      }
    }
  }
  __syncthreads();
  // Q[:, cols] <- Q[:, cols] R, staged through the X region (c <= r rows);
  // a work item is one row times 8 columns
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lc = grp + h * WB, g = gcol(lc);
    const double* src = Q + (int64_t)(g >= 0 ? g : 0) * c;
    double v[RPL];
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int row = lane + GSZ * u;
      const bool in = g >= 0 && row < c;
      const double x = src[in ? row : 0];
      v[u] = in ? x : 0.0;
    }
#pragma unroll
    for (int u = 0; u < RPL; ++u) X[lc * LDX + lane + GSZ * u] = v[u];
  }
  __syncthreads();
  for (int cg = 0; cg < NC / 8; ++cg)
    for (int row = tid; row < c; row += NT) {
      double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
      for (int li = 0; li < NC; ++li) {
        const double x = X[li * LDX + row];
        const double* rr = Rt + li * RTS + cg * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = fma(x, rr[k], acc[k]);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int g = gcol(cg * 8 + k);
        if (g >= 0) Q[(int64_t)g * c + row] = acc[k];
      }
    }
}

// Scale every work matrix by a power of two so its largest |entry| lies in
// [0.5, 1): products of columns then neither underflow (blocks of 1e-120 occur
// in DMRG) nor overflow.  Two passes over JSPLIT slices per matrix: the largest
// |entry| (atomic max of the non-negative doubles' bit patterns, which order
// like the values) and a flag for non-finite entries, then the scaling.
// sexp[mat] keeps the exponent for the finish; NaN / inf leave the matrix
// unscaled (reported by the finish).
constexpr int JSPLIT = 32;
__device__ __forceinline__ void slice_of(int64_t n, int64_t& b0, int64_t& b1) {
  const int64_t per = (n + JSPLIT - 1) / JSPLIT;
  b0 = (int64_t)blockIdx.y * per;
  b1 = b0 + per < n ? b0 + per : n;
}
__global__ __launch_bounds__(JT) void jac_absmax(const JacMat* __restrict__ mats, const double* __restrict__ jw,
                                                 unsigned long long* __restrict__ amax, int* __restrict__ bad) {
  __shared__ double red[JT / 64];
  __shared__ int nf_any;
  const JacMat M = mats[blockIdx.x];
  const double* W = jw + M.w0_off;
  int64_t b0, b1;
  slice_of(M.r * M.c, b0, b1);
  if (threadIdx.x == 0) nf_any = 0;
  __syncthreads();
  double mx = 0.0;
  int nf = 0;
  for (int64_t e = b0 + threadIdx.x; e < b1; e += JT) {
    const double x = W[e];
    mx = fmax(mx, fabs(x));
    nf |= !(x == x) || fabs(x) > 1.7976931348623157e308;
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = mx;
  if (nf) nf_any = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int k = 0; k < JT / 64; ++k) m = fmax(m, red[k]);
    if (m > 0.0) atomicMax(&amax[blockIdx.x], (unsigned long long)__double_as_longlong(m));
    if (nf_any) atomicOr(&bad[blockIdx.x], 1);
  }
}
__global__ __launch_bounds__(JT) void jac_rescale(const JacMat* __restrict__ mats, double* __restrict__ jw,
                                                  const unsigned long long* __restrict__ amax,
                                                  int* __restrict__ sexp) {
  const JacMat M = mats[blockIdx.x];
  const double m = __longlong_as_double((long long)amax[blockIdx.x]);
  int ex = 0;
  if (m > 0.0 && m <= 1.7976931348623157e308) frexp(m, &ex);
  if (blockIdx.y == 0 && threadIdx.x == 0) sexp[blockIdx.x] = ex;
  if (ex == 0) return;
  double* W = jw + M.w0_off;
  int64_t b0, b1;
  slice_of(M.r * M.c, b0, b1);
  for (int64_t e = b0 + threadIdx.x; e < b1; e += JT) W[e] = ldexp(W[e], -ex);
}

// After a sweep of rounds: matrices whose largest relative off-diagonal stayed
// below tolerance are done; count the rest (left[0]) and the rounds a sweep of
// the rest takes (left[3]: the largest nb - 1).  One thread (few matrices).
__global__ void jac_check(const JacMat* __restrict__ mats, const int32_t* __restrict__ big, int nbig,
                          unsigned long long* __restrict__ offmax, int* __restrict__ done, int* __restrict__ left) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int n = 0, rounds = 0;
  for (int i = 0; i < nbig; ++i) {
    const int b = big[i];
    if (!done[b]) {
      if (offmax[b] == 0ull) {  // a sweep without rotations
        done[b] = 1;
      } else {
        ++n;
        rounds = max(rounds, mats[b].nb - 1);
      }
    }
    offmax[b] = 0ull;
  }
  left[0] = n;
  left[3] = rounds;
}

// Gram slots of one bucket (after the syrk wrote the upper triangles of the
// c x c Gram matrices): padded rows / columns c .. P get -1 on the diagonal
// and 0 elsewhere, so the padding's eigenpairs (-1, unit vectors) lie below
// every eigenvalue of the Gram matrix (>= 0) and stay apart from them.  The
// slot of a matrix with a non-finite entry is set to the identity (the sweeps
// then report the input as not finite).  JSPLIT slices per slot.
__global__ __launch_bounds__(JT) void jac_gram_pad(const JacMat* __restrict__ mats, const int32_t* __restrict__ list,
                                                   const int* __restrict__ bad, double* __restrict__ jw, int64_t P) {
  const int mi = list[blockIdx.x];
  const JacMat M = mats[mi];
  double* G = jw + M.g_off;
  const int64_t c = M.c;
  int64_t b0, b1;
  if (bad[mi]) {
    slice_of(P * P, b0, b1);
    for (int64_t e = b0 + threadIdx.x; e < b1; e += JT) G[e] = (e / P == e % P) ? 1.0 : 0.0;
    return;
  }
  // columns c .. P-1 (all rows), then rows c .. P-1 of columns 0 .. c-1
  const int64_t n1 = (P - c) * P, n2 = (P - c) * c;
  slice_of(n1 + n2, b0, b1);
  for (int64_t e = b0 + threadIdx.x; e < b1; e += JT) {
    int64_t i, j;
    if (e < n1) {
      j = c + e / P;
      i = e - (j - c) * P;
    } else {
      j = (e - n1) / (P - c);
      i = c + (e - n1) - j * (P - c);
    }
    G[i + j * P] = (i == j) ? -1.0 : 0.0;
  }
}

// Q <- the eigenvectors of the Gram matrix by descending eigenvalue (syevd
// returns them ascending, the padding's -1 first): Q(:, j) = V(0:c, P-1-j).
// A matrix whose syevd failed keeps Q = I.  JSPLIT slices per matrix.
__global__ __launch_bounds__(JT) void jac_take_v(const JacMat* __restrict__ mats, const int32_t* __restrict__ list,
                                                 double* __restrict__ jw, const rocblas_int* __restrict__ info,
                                                 int64_t P) {
  const JacMat M = mats[list[blockIdx.x]];
  if (info[M.g_info] != 0) return;
  const double* G = jw + M.g_off;
  double* Q = jw + M.q_off;
  const int64_t c = M.c;
  int64_t b0, b1;
  slice_of(c * c, b0, b1);
  for (int64_t e = b0 + threadIdx.x; e < b1; e += JT) {
    const int64_t j = e / c, i = e - j * c;
    Q[e] = G[i + (P - 1 - j) * P];
  }
}

__global__ __launch_bounds__(JT) void jac_init_q(const JacMat* __restrict__ mats, double* __restrict__ ws) {
  const JacMat M = mats[blockIdx.x];
  double* Q = ws + M.q_off;
  const int64_t cc = M.c * M.c;
  for (int64_t e = threadIdx.x; e < cc; e += JT) Q[e] = (e / M.c == e % M.c) ? 1.0 : 0.0;
}

// The finish, in four launches: sigma_j = |W_j| (one wave per column), the
// rank of each column by descending sigma (ties by column index) with S and the
// count of zero columns, the left / right vectors in rank order (one wave per
// column), and the completion of zero columns' left vectors to an orthonormal
// set by Gram-Schmidt on unit vectors (one work group per matrix that has any).
constexpr int JF_MAX = 2048;
constexpr int JF_COLS = JT / 64;  // columns per work group (one per wave)
__global__ __launch_bounds__(JT) void jac_sigma(const JacMat* __restrict__ mats, double* __restrict__ jw) {
  const JacMat M = mats[blockIdx.x];
  const int64_t j = (int64_t)blockIdx.y * JF_COLS + threadIdx.x / 64;
  if (j >= M.c) return;
  const int ln = threadIdx.x % 64;
  const double* W = jw + M.w_off + j * M.r;
  double a = 0.0;
  for (int64_t row = ln; row < M.r; row += 64) a = fma(W[row], W[row], a);
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if (ln == 0) jw[M.sg_off + j] = sqrt(a);
}
__global__ __launch_bounds__(JT) void jac_rank(const JacMat* __restrict__ mats, double* __restrict__ jw,
                                               const int* __restrict__ sexp, double* __restrict__ S,
                                               int* __restrict__ nzero, int* __restrict__ flags) {
  __shared__ double sig[JF_MAX];
  const JacMat M = mats[blockIdx.x];
  const int64_t c = M.c;
  if ((int64_t)blockIdx.y * JT >= c) return;
  for (int64_t i = threadIdx.x; i < c; i += JT) sig[i] = jw[M.sg_off + i];
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.y * JT + threadIdx.x;
  if (j >= c) return;
  const double sj = sig[j];
  if (!(sj == sj) || sj > 1.7e308) flags[0] = 1;
  int n = 0;
  for (int64_t i = 0; i < c; ++i) n += (sig[i] > sj) || (sig[i] == sj && i < j);
  ((int*)(jw + M.rk_off))[j] = n;
  S[M.s_off + n] = ldexp(sj, sexp[blockIdx.x]);
  if (!(sj > 1e-140)) atomicAdd(&nzero[blockIdx.x], 1);
}
// left vector (length r) of rank q: U column (trans 0) or Vt row (trans 1); right vector likewise
__global__ __launch_bounds__(JT) void jac_vectors(const JacMat* __restrict__ mats, const double* __restrict__ jw,
                                                  double* __restrict__ U, double* __restrict__ Vt) {
  const JacMat M = mats[blockIdx.x];
  const int64_t j = (int64_t)blockIdx.y * JF_COLS + threadIdx.x / 64;
  if (j >= M.c) return;
  const int ln = threadIdx.x % 64;
  const int64_t r = M.r, c = M.c, k = c;
  const double* W = jw + M.w_off;
  const double* Q = jw + M.q_off;
  const int64_t ldl = M.trans ? k : 1, stl = M.trans ? 1 : M.m;
  double* Lb = M.trans ? Vt + M.v_off : U + M.u_off;
  double* Rb = M.trans ? U + M.u_off : Vt + M.v_off;
  const int64_t ldr = M.trans ? 1 : k, str = M.trans ? M.m : 1;
  const double sj = jw[M.sg_off + j];
  const int64_t q = ((const int*)(jw + M.rk_off))[j];
  const bool live = sj > 1e-140;  // columns the rotations treat as zero get completed below
  const double inv = live ? 1.0 / sj : 0.0;
  for (int64_t row = ln; row < r; row += 64) Lb[q * stl + row * ldl] = live ? W[j * r + row] * inv : 0.0;
  for (int64_t row = ln; row < c; row += 64) Rb[q * str + row * ldr] = Q[j * c + row];
}
// completion: zero columns have the largest ranks c - nzero .. c - 1
__global__ __launch_bounds__(JT) void jac_complete(const JacMat* __restrict__ mats, const int* __restrict__ nzero,
                                                   double* __restrict__ U, double* __restrict__ Vt) {
  __shared__ double x[JF_MAX];
  __shared__ double red[JT / 64];
  const int nz = nzero[blockIdx.x];
  if (nz == 0) return;
  const JacMat M = mats[blockIdx.x];
  const int64_t r = M.r, c = M.c, k = c;
  const int tid = threadIdx.x, wave = tid / 64, ln = tid % 64;
  const int64_t ldl = M.trans ? k : 1, stl = M.trans ? 1 : M.m;
  double* Lb = M.trans ? Vt + M.v_off : U + M.u_off;
  for (int64_t q = c - nz; q < c; ++q) {
    for (int64_t cand = 0; cand < r; ++cand) {
      for (int64_t row = tid; row < r; row += JT) x[row] = row == cand ? 1.0 : 0.0;
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        for (int64_t o = 0; o < q; ++o) {
          double d = 0.0;
          for (int64_t row = tid; row < r; row += JT) d = fma(Lb[o * stl + row * ldl], x[row], d);
          for (int s2 = 32; s2 > 0; s2 >>= 1) d += __shfl_xor(d, s2, 64);
          if (ln == 0) red[wave] = d;
          __syncthreads();
          const double dot = red[0] + red[1] + red[2] + red[3];
          for (int64_t row = tid; row < r; row += JT) x[row] -= dot * Lb[o * stl + row * ldl];
          __syncthreads();
        }
      }
      double nn = 0.0;
      for (int64_t row = tid; row < r; row += JT) nn = fma(x[row], x[row], nn);
      for (int s2 = 32; s2 > 0; s2 >>= 1) nn += __shfl_xor(nn, s2, 64);
      if (ln == 0) red[wave] = nn;
      __syncthreads();
      const double nrm = sqrt(red[0] + red[1] + red[2] + red[3]);
      __syncthreads();
      if (nrm > 0.5) {
        for (int64_t row = tid; row < r; row += JT) Lb[q * stl + row * ldl] = x[row] / nrm;
        __syncthreads();
        break;
      }
    }
  }
}

// Per device: NLANE side streams, each with its own rocBLAS handle, so the
// size buckets' gesdd calls (latency-bound chains of small kernels) run
// concurrently.  Used under a lock (handles are not thread-safe).
constexpr int NLANE = 4;
struct Lanes {
  hipStream_t st[NLANE];
  rocblas_handle h[NLANE];
  hipEvent_t done[NLANE];
  hipEvent_t fork;
};
std::mutex g_handle_mu;
std::map<int, Lanes> g_lanes;

Lanes& lanes_for(int dev) {
  auto it = g_lanes.find(dev);
  if (it != g_lanes.end()) return it->second;
  Lanes L{};
  for (int i = 0; i < NLANE; ++i) {
    ck(hipStreamCreateWithFlags(&L.st[i], hipStreamNonBlocking), "hipStreamCreate");
    ckb(rocblas_create_handle(&L.h[i]), "rocblas_create_handle");
    ckb(rocblas_set_stream(L.h[i], L.st[i]), "rocblas_set_stream");
    ck(hipEventCreateWithFlags(&L.done[i], hipEventDisableTiming), "hipEventCreate");
  }
  ck(hipEventCreateWithFlags(&L.fork, hipEventDisableTiming), "hipEventCreate");
  return g_lanes.emplace(dev, L).first->second;
}

}  // namespace

struct mpsh_svd_s {
  int64_t nblocks = 0;
  std::vector<int64_t> m, n, k, a_off, s_off, u_off, v_off;
  int64_t s_total = 0, u_total = 0, v_total = 0;
  std::vector<Bucket> buckets;
  std::vector<int64_t> bucket_of, slot_of;
  int64_t ws_doubles = 0, info_ints = 0;
  CopyPlan pad;               // block -> padded slot
  CopyPlan unpad_u, unpad_v, unpad_s;  // padded outputs -> packed U, Vt, S
  // one-sided Jacobi path
  std::vector<JacMat> jm;
  std::vector<JacTask> jsmall, jbig;  // tasks of 2-block matrices / of one round of the others,
  int32_t jsmall_cls[4] = {0, 0, 0, 0}, jbig_cls[4] = {0, 0, 0, 0};  // by width class (16, 8, 4): [cls, cls+1)
  std::vector<int32_t> jbig_idx;      // matrices with more than two column blocks
  int32_t jrounds = 0;                // rounds per sweep (max over matrices)
  int64_t jcmax = 0;                  // largest column count of the Jacobi matrices
  std::vector<GramBucket> gb;         // preconditioning: syevd size buckets (ascending P)
  std::vector<int32_t> gmat;          // their matrices, bucket by bucket
  int64_t jac_off = 0, jac_meta_off = 0;  // workspace (doubles): W | Q, then offmax | done | left, flags
  CopyPlan jinit;                     // blocks -> W (transposed when m < n)
  struct Dev {
    std::vector<std::unique_ptr<DevBuf>> bufs;
    DevCopy pad, u, v, s, jinit;
    const JacMat* jm = nullptr;
    const JacTask* jsmall = nullptr;
    const JacTask* jbig = nullptr;
    const int32_t* jbig_idx = nullptr;
    const int32_t* gmat = nullptr;
  };
  std::mutex mu;
  std::map<int, std::unique_ptr<Dev>> dev;

  template <class T>
  const T* up(std::vector<std::unique_ptr<DevBuf>>& bufs, const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    auto b = std::make_unique<DevBuf>();
    ck(hipMalloc(&b->p, v.size() * sizeof(T)), "hipMalloc(svd tables)");
    ck(upload_tables(b->p, v.data(), v.size() * sizeof(T)), "upload(svd tables)");
    const T* r = (const T*)b->p;
    bufs.push_back(std::move(b));
    return r;
  }
  DevCopy upc(std::vector<std::unique_ptr<DevBuf>>& bufs, const CopyPlan& c) {
    DevCopy d;
    d.descs = up(bufs, c.descs);
    d.tile_map = up(bufs, c.tile_map);
    d.ndesc = (int)c.descs.size();
    d.ntiles = c.ntiles;
    return d;
  }
  const Dev& get() {
    std::lock_guard<std::mutex> g(mu);
    int d = 0;
    ck(hipGetDevice(&d), "hipGetDevice");
    auto it = dev.find(d);
    if (it == dev.end()) {
      auto D = std::make_unique<Dev>();
      D->pad = upc(D->bufs, pad);
      D->u = upc(D->bufs, unpad_u);
      D->v = upc(D->bufs, unpad_v);
      D->s = upc(D->bufs, unpad_s);
      D->jinit = upc(D->bufs, jinit);
      D->jm = up(D->bufs, jm);
      D->jsmall = up(D->bufs, jsmall);
      D->jbig = up(D->bufs, jbig);
      D->jbig_idx = up(D->bufs, jbig_idx);
      D->gmat = up(D->bufs, gmat);
      it = dev.emplace(d, std::move(D)).first;
    }
    return *it->second;
  }
};

namespace {

// shapes of the matrices not marked done, for the error message
std::string jacobi_unconverged(const mpsh_svd_s& P, const int* done, hipStream_t s) {
  std::vector<int> h(P.jm.size());
  if (hipMemcpyAsync(h.data(), done, h.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return "";
  std::string out;
  int n = 0;
  for (size_t i = 0; i < h.size(); ++i)
    if (!h[i] && n++ < 8)
      out += " " + std::to_string(P.jm[i].m) + "x" + std::to_string(P.jm[i].n) + "(w" + std::to_string(P.jm[i].w) + ")";
  return " (" + std::to_string(n) + " blocks:" + out + ")";
}

void jacobi_execute(const mpsh_svd_s& P, const mpsh_svd_s::Dev& cp, const double* A, double* S, double* U,
                    double* Vt, double* ws, hipStream_t s) {
  const int nj = (int)P.jm.size();
  double* jw = ws + P.jac_off;
  // meta: offmax | amax (u64) | done | sexp | bad | nzero (int32, even counts) | left, flags[2], rounds
  unsigned long long* offmax = (unsigned long long*)(ws + P.jac_meta_off);
  unsigned long long* amax = offmax + nj;
  const int nje = nj + (nj & 1);
  int* done = (int*)(amax + nj);
  int* sexp = done + nje;
  int* bad = sexp + nje;
  int* nzero = bad + nje;
  int* left = nzero + nje;
  int* flags = left + 1;
  ck(hipMemsetAsync(offmax, 0, (size_t)nj * 16 + (size_t)nje * 16 + 16, s), "hipMemsetAsync(jacobi)");
  ck(launch_copy(cp.jinit, 8, A, jw, s), "jacobi init W");
  hipLaunchKernelGGL(jac_absmax, dim3(nj, JSPLIT), dim3(JT), 0, s, cp.jm, (const double*)jw, amax, bad);
  ck(hipGetLastError(), "jac_absmax");
  hipLaunchKernelGGL(jac_rescale, dim3(nj, JSPLIT), dim3(JT), 0, s, cp.jm, jw, (const unsigned long long*)amax, sexp);
  ck(hipGetLastError(), "jac_rescale");
  hipLaunchKernelGGL(jac_init_q, dim3(nj), dim3(JT), 0, s, cp.jm, jw);
  ck(hipGetLastError(), "jac_init_q");
  if (!P.gb.empty()) {
    // preconditioning, size buckets spread over the side lanes (largest first):
    // G = W0^T W0 (upper), padding, syevd, Q = eigenvectors, W = W0 Q
    rocblas_int* info = (rocblas_int*)(ws + P.ws_doubles);
    int dev = 0;
    ck(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> g(g_handle_mu);
    Lanes& L = lanes_for(dev);
    ck(hipEventRecord(L.fork, s), "hipEventRecord");
    const int ngb = (int)P.gb.size();
    for (int i = 0; i < ngb && i < NLANE; ++i) ck(hipStreamWaitEvent(L.st[i], L.fork, 0), "hipStreamWaitEvent");
    const double one = 1.0, zero = 0.0;
    for (int i = 0; i < ngb; ++i) {
      const GramBucket& B = P.gb[ngb - 1 - i];
      rocblas_handle h = L.h[i % NLANE];
      hipStream_t ls = L.st[i % NLANE];
      for (int64_t k = 0; k < B.count; ++k) {
        const JacMat& J = P.jm[P.gmat[B.mat_start + k]];
        ckb(rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_transpose, (rocblas_int)J.c, (rocblas_int)J.r,
                          &one, jw + J.w0_off, (rocblas_int)J.r, &zero, jw + J.g_off, (rocblas_int)B.P),
            "rocblas_dsyrk");
      }
      hipLaunchKernelGGL(jac_gram_pad, dim3((unsigned)B.count, JSPLIT), dim3(JT), 0, ls, cp.jm, cp.gmat + B.mat_start,
                         (const int*)bad, jw, B.P);
      ck(hipGetLastError(), "jac_gram_pad");
      ckb(rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)B.P,
                                           jw + B.g_off, (rocblas_int)B.P, B.P * B.P, jw + B.d_off, B.P,
                                           jw + B.e_off, B.P, info + B.info_start, (rocblas_int)B.count),
          "rocsolver_dsyevd_strided_batched");
      hipLaunchKernelGGL(jac_take_v, dim3((unsigned)B.count, JSPLIT), dim3(JT), 0, ls, cp.jm, cp.gmat + B.mat_start,
                         jw, (const rocblas_int*)info, B.P);
      ck(hipGetLastError(), "jac_take_v");
      for (int64_t k = 0; k < B.count; ++k) {
        const JacMat& J = P.jm[P.gmat[B.mat_start + k]];
        ckb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)J.r, (rocblas_int)J.c,
                          (rocblas_int)J.c, &one, jw + J.w0_off, (rocblas_int)J.r, jw + J.q_off, (rocblas_int)J.c,
                          &zero, jw + J.w_off, (rocblas_int)J.r),
            "rocblas_dgemm");
      }
    }
    for (int i = 0; i < ngb && i < NLANE; ++i) {
      ck(hipEventRecord(L.done[i], L.st[i]), "hipEventRecord");
      ck(hipStreamWaitEvent(s, L.done[i], 0), "hipStreamWaitEvent");
    }
  }
  // one launch per width class present
  auto pairs = [&](const JacTask* tasks, const int32_t* cls, int t, int iterate) {
    for (int k = 0; k < 3; ++k) {
      const int n = cls[k + 1] - cls[k];
      if (n == 0) continue;
      const dim3 g((unsigned)n);
      if (k == 0)
        hipLaunchKernelGGL((jac_pairs<16, 64>), g, dim3(1024), 0, s, cp.jm, tasks + cls[k], t, iterate, jw, offmax,
                           done, flags);
      else if (k == 1)
        hipLaunchKernelGGL((jac_pairs<8, 64>), g, dim3(512), 0, s, cp.jm, tasks + cls[k], t, iterate, jw, offmax,
                           done, flags);
      else
        hipLaunchKernelGGL((jac_pairs<4, 64>), g, dim3(256), 0, s, cp.jm, tasks + cls[k], t, iterate, jw, offmax,
                           done, flags);
    }
    ck(hipGetLastError(), "jac_pairs");
  };
  if (!P.jsmall.empty()) pairs(cp.jsmall, P.jsmall_cls, 0, 1);
  if (!P.jbig.empty()) {
    // a sweep runs the rounds of the largest matrix not yet done
    int hl = 1, sweeps = 0, rounds = P.jrounds, total_rounds = 0;
    for (int sweep = 0; sweep < JAC_MAX_SWEEPS && hl != 0; ++sweep, ++sweeps) {
      for (int t = 0; t < rounds; ++t) pairs(cp.jbig, P.jbig_cls, t, 0);
      total_rounds += rounds;
      hipLaunchKernelGGL(jac_check, dim3(1), dim3(64), 0, s, cp.jm, cp.jbig_idx, (int)P.jbig_idx.size(), offmax,
                         done, left);
      ck(hipGetLastError(), "jac_check");
      int h4[4] = {0, 0, 0, 0};
      ck(hipMemcpyAsync(h4, left, 16, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(jacobi)");
      ck(hipStreamSynchronize(s), "hipStreamSynchronize");
      hl = h4[0];
      rounds = h4[3];
    }
    if (hl != 0) throw Error(ERR_NONFINITE, "Jacobi SVD did not converge" + jacobi_unconverged(P, done, s));
    static const bool stats = getenv("MPSH_SVD_STATS") != nullptr;
    if (stats)
      fprintf(stderr, "mpsh_svd: %d Jacobi sweeps, %d rounds (%d per full sweep), %d preconditioned in %d buckets\n",
              sweeps, total_rounds, P.jrounds, (int)P.gmat.size(), (int)P.gb.size());
  }
  const unsigned ycols = (unsigned)((P.jcmax + JF_COLS - 1) / JF_COLS);
  hipLaunchKernelGGL(jac_sigma, dim3(nj, ycols), dim3(JT), 0, s, cp.jm, jw);
  ck(hipGetLastError(), "jac_sigma");
  hipLaunchKernelGGL(jac_rank, dim3(nj, (unsigned)((P.jcmax + JT - 1) / JT)), dim3(JT), 0, s, cp.jm, jw,
                     (const int*)sexp, S, nzero, flags);
  ck(hipGetLastError(), "jac_rank");
  hipLaunchKernelGGL(jac_vectors, dim3(nj, ycols), dim3(JT), 0, s, cp.jm, (const double*)jw, U, Vt);
  ck(hipGetLastError(), "jac_vectors");
  hipLaunchKernelGGL(jac_complete, dim3(nj), dim3(JT), 0, s, cp.jm, (const int*)nzero, U, Vt);
  ck(hipGetLastError(), "jac_complete");
  int hf[2] = {0, 0};
  ck(hipMemcpyAsync(hf, flags, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(jacobi flags)");
  ck(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (hf[0]) throw Error(ERR_NONFINITE, "SVD input is not finite");
  if (hf[1]) throw Error(ERR_NONFINITE, "Jacobi SVD did not converge" + jacobi_unconverged(P, done, s));
}

}  // namespace

extern "C" {

mpsh_status mpsh_svd_plan(int64_t nblocks, const int64_t* m, const int64_t* n, const int64_t* a_off,
                          mpsh_svd* plan) {
  try {
    if (nblocks < 0 || (nblocks > 0 && (!m || !n || !a_off)) || !plan) throw Error(ERR_INVALID, "bad svd arguments");
    auto P = std::make_unique<mpsh_svd_s>();
    P->nblocks = nblocks;
    P->m.assign(m, m + nblocks);
    P->n.assign(n, n + nblocks);
    P->a_off.assign(a_off, a_off + nblocks);
    std::map<std::pair<int64_t, int64_t>, std::vector<int64_t>> by;
    for (int64_t b = 0; b < nblocks; ++b) {
      if (m[b] <= 0 || n[b] <= 0 || a_off[b] < 0) throw Error(ERR_INVALID, "svd block with empty shape");
      if (m[b] > 16384 || n[b] > 16384) throw Error(ERR_UNSUPPORTED, "svd block larger than 16384");
      const int64_t k = std::min(m[b], n[b]);
      P->k.push_back(k);
      P->s_off.push_back(P->s_total);
      P->u_off.push_back(P->u_total);
      P->v_off.push_back(P->v_total);
      P->s_total += k;
      P->u_total += m[b] * k;
      P->v_total += k * n[b];
    }
    const char* algo = getenv("MPSH_SVD_ALGO");
    const bool jacobi = !(algo && std::string(algo) == "gesdd");
    const char* pre_env = getenv("MPSH_SVD_PRECOND");
    const bool precond = !(pre_env && std::string(pre_env) == "0");
    std::vector<int64_t> order;
    std::vector<int64_t> jac_blocks;
    for (int64_t b = 0; b < nblocks; ++b) {
      if (jacobi && std::max(m[b], n[b]) <= JAC_MAX_ROWS)
        jac_blocks.push_back(b);
      else
        order.push_back(b);
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
      return std::max(m[x], n[x]) < std::max(m[y], n[y]);
    });
    const int64_t nroc = (int64_t)order.size();
    std::vector<int64_t> Ps(nroc);
    for (int64_t i = 0; i < nroc; ++i) Ps[i] = pad16(std::max(m[order[i]], n[order[i]]));
    int64_t start = 0;
    for (int64_t end : choose_buckets(Ps)) {
      const int64_t P = Ps[end - 1];
      for (int64_t i = start; i < end; ++i) by[{P, P}].push_back(order[i]);
      start = end;
    }
    P->bucket_of.assign(nblocks, -1);
    P->slot_of.assign(nblocks, -1);
    int64_t w = 0, info = 0;
    for (auto& kv : by) {
      Bucket B;
      B.Mp = kv.first.first;
      B.Np = kv.first.second;
      B.kp = std::min(B.Mp, B.Np);
      B.count = (int64_t)kv.second.size();
      B.a_off = w;
      w += B.count * B.Mp * B.Np;
      B.s_off = w;
      w += B.count * B.kp;
      B.u_off = w;
      w += B.count * B.Mp * B.kp;
      B.v_off = w;
      w += B.count * B.kp * B.Np;
      w = (w + 1) / 2 * 2;
      B.info_off = info;
      info += B.count;
      for (int64_t i = 0; i < B.count; ++i) {
        const int64_t b = kv.second[i];
        P->bucket_of[b] = (int64_t)P->buckets.size();
        P->slot_of[b] = i;
        // input block (m x n, ld m) -> padded slot (ld Mp)
        const int64_t dims[2] = {m[b], n[b]}, sst[2] = {1, m[b]}, dst[2] = {1, B.Mp};
        P->pad.add(2, dims, sst, dst, a_off[b], B.a_off + i * B.Mp * B.Np);
        const int64_t k = P->k[b];
        // U: first m rows, first k columns of the padded Mp x kp slot
        const int64_t ud[2] = {m[b], k}, us[2] = {1, B.Mp}, uo[2] = {1, m[b]};
        P->unpad_u.add(2, ud, us, uo, B.u_off + i * B.Mp * B.kp, P->u_off[b]);
        // Vt: first k rows, first n columns of the padded kp x Np slot
        const int64_t vd[2] = {k, n[b]}, vs[2] = {1, B.kp}, vo[2] = {1, k};
        P->unpad_v.add(2, vd, vs, vo, B.v_off + i * B.kp * B.Np, P->v_off[b]);
        // S: first k values
        const int64_t sd[1] = {k}, ss[1] = {1}, so[1] = {1};
        P->unpad_s.add(1, sd, ss, so, B.s_off + i * B.kp, P->s_off[b]);
      }
      P->buckets.push_back(B);
    }
    // Jacobi matrices: W | Q per block, then the per-matrix flags
    P->jac_off = w;
    int64_t jw = 0;
    std::vector<JacTask> small_by[3], big_by[3];
    for (int64_t b : jac_blocks) {
      JacMat J{};
      J.m = m[b];
      J.n = n[b];
      J.trans = m[b] < n[b];
      J.r = std::max(m[b], n[b]);
      J.c = std::min(m[b], n[b]);
      J.w = jac_width(J.r);
      const int cls = J.w == 16 ? 0 : (J.w == 8 ? 1 : 2);
      J.nb = (int32_t)((J.c + J.w - 1) / J.w);
      J.nb += J.nb & 1;
      if (J.nb < 2) J.nb = 2;
      J.ctol = std::sqrt((double)J.r) * 2.220446049250313e-16;
      if (2 * J.w * (J.r | 1) + 4 * J.w * J.w > JAC_LDS || J.r > JF_MAX)
        throw Error(ERR_INVALID, "svd plan: Jacobi block exceeds the LDS budget");
      J.u_off = P->u_off[b];
      J.s_off = P->s_off[b];
      J.v_off = P->v_off[b];
      J.pre = precond && J.nb > 2 && J.c >= JAC_PRE_MIN;
      J.w_off = jw;
      jw += J.r * J.c;
      J.q_off = jw;
      jw += J.c * J.c;
      J.w0_off = J.w_off;
      if (J.pre) {
        J.w0_off = jw;
        jw += J.r * J.c;
      }
      J.sg_off = jw;
      jw += J.c;
      J.rk_off = jw;
      jw += (J.c + 1) / 2;
      jw = (jw + 1) / 2 * 2;
      P->jcmax = std::max(P->jcmax, J.c);
      const int32_t idx = (int32_t)P->jm.size();
      if (J.nb == 2) {
        small_by[cls].push_back({idx, 0});
      } else {
        P->jbig_idx.push_back(idx);
        for (int32_t k = 0; k < J.nb / 2; ++k) big_by[cls].push_back({idx, k});
        P->jrounds = std::max(P->jrounds, J.nb - 1);
      }
      if (!J.trans) {
        const int64_t dims[2] = {J.m, J.n}, sst[2] = {1, J.m}, dst[2] = {1, J.r};
        P->jinit.add(2, dims, sst, dst, a_off[b], J.w0_off);
      } else {  // W(i, j) = A(j, i)
        const int64_t dims[2] = {J.n, J.m}, sst[2] = {J.m, 1}, dst[2] = {1, J.r};
        P->jinit.add(2, dims, sst, dst, a_off[b], J.w0_off);
      }
      P->jm.push_back(J);
    }
    {
      // preconditioning buckets over the matrices sorted by column count
      std::vector<int32_t> pre;
      for (int32_t i = 0; i < (int32_t)P->jm.size(); ++i)
        if (P->jm[i].pre) pre.push_back(i);
      std::stable_sort(pre.begin(), pre.end(), [&](int32_t x, int32_t y) { return P->jm[x].c < P->jm[y].c; });
      std::vector<int64_t> Pg(pre.size());
      for (size_t i = 0; i < pre.size(); ++i) Pg[i] = pad16(P->jm[pre[i]].c);
      int64_t start = 0;
      for (int64_t end : choose_gram_buckets(Pg)) {
        GramBucket B{};
        B.P = Pg[end - 1];
        B.count = end - start;
        B.g_off = jw;
        jw += B.count * B.P * B.P;
        B.d_off = jw;
        jw += B.count * B.P;
        B.e_off = jw;
        jw += B.count * B.P;
        jw = (jw + 1) / 2 * 2;
        B.mat_start = (int32_t)P->gmat.size();
        B.info_start = (int32_t)info;
        for (int64_t i = start; i < end; ++i) {
          JacMat& J = P->jm[pre[i]];
          J.g_off = B.g_off + (i - start) * B.P * B.P;
          J.gld = B.P;
          J.g_info = (int32_t)(info + (i - start));
          P->gmat.push_back(pre[i]);
        }
        info += B.count;
        P->gb.push_back(B);
        start = end;
      }
    }
    for (int k = 0; k < 3; ++k) {
      P->jsmall.insert(P->jsmall.end(), small_by[k].begin(), small_by[k].end());
      P->jbig.insert(P->jbig.end(), big_by[k].begin(), big_by[k].end());
      P->jsmall_cls[k + 1] = (int32_t)P->jsmall.size();
      P->jbig_cls[k + 1] = (int32_t)P->jbig.size();
    }
    w += jw;
    P->jac_meta_off = w;
    const int64_t nj = (int64_t)P->jm.size();
    w += 2 * nj + 4 * ((nj + 1) / 2) + 2;  // offmax, amax (u64) | done, sexp, bad, nzero (int32) | left, flags, rounds
    w = (w + 1) / 2 * 2;
    P->jinit.finalize();
    P->ws_doubles = w;
    P->info_ints = info;
    P->pad.finalize();
    P->unpad_u.finalize();
    P->unpad_v.finalize();
    P->unpad_s.finalize();
    *plan = P.release();
    return MPSH_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return (mpsh_status)e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MPSH_ERR_INVALID;
  }
}

mpsh_status mpsh_svd_sizes(mpsh_svd plan, int64_t* k, int64_t* s_off, int64_t* u_off, int64_t* v_off,
                           int64_t* totals) {
  if (!plan) return MPSH_ERR_INVALID;
  for (int64_t b = 0; b < plan->nblocks; ++b) {
    if (k) k[b] = plan->k[b];
    if (s_off) s_off[b] = plan->s_off[b];
    if (u_off) u_off[b] = plan->u_off[b];
    if (v_off) v_off[b] = plan->v_off[b];
  }
  if (totals) {
    totals[0] = plan->s_total;
    totals[1] = plan->u_total;
    totals[2] = plan->v_total;
  }
  return MPSH_OK;
}

size_t mpsh_svd_workspace(mpsh_svd plan) {
  if (!plan) return 0;
  return (size_t)plan->ws_doubles * 8 + (size_t)plan->info_ints * 4 + 64;
}

mpsh_status mpsh_svd_execute(mpsh_svd plan, const double* A, double* S, double* U, double* Vt, void* workspace,
                             void* stream) {
  try {
    if (!plan) throw Error(ERR_INVALID, "null plan");
    if (plan->nblocks == 0) return MPSH_OK;
    if (!A || !S || !U || !Vt || !workspace) throw Error(ERR_INVALID, "null argument");
    if (((uintptr_t)workspace) & 15) throw Error(ERR_INVALID, "workspace must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    double* ws = (double*)workspace;
    rocblas_int* info = (rocblas_int*)(ws + plan->ws_doubles);
    const mpsh_svd_s::Dev& cp = plan->get();
    if (!plan->jm.empty()) jacobi_execute(*plan, cp, A, S, U, Vt, ws, s);
    if (plan->buckets.empty()) return MPSH_OK;
    // zero the padded inputs, then place every block in its slot
    for (const Bucket& B : plan->buckets)
      ck(hipMemsetAsync(ws + B.a_off, 0, (size_t)B.count * B.Mp * B.Np * 8, s), "hipMemsetAsync");
    ck(launch_copy(cp.pad, 8, A, ws, s), "svd pad");
    int dev = 0;
    ck(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> g(g_handle_mu);
    Lanes& L = lanes_for(dev);
    // fork: largest bucket first, bucket i on lane i mod NLANE; join back into s
    ck(hipEventRecord(L.fork, s), "hipEventRecord");
    const int nb = (int)plan->buckets.size();
    for (int i = 0; i < nb && i < NLANE; ++i) ck(hipStreamWaitEvent(L.st[i], L.fork, 0), "hipStreamWaitEvent");
    for (int i = 0; i < nb; ++i) {
      const Bucket& B = plan->buckets[nb - 1 - i];
      ckb(rocsolver_dgesdd_strided_batched(L.h[i % NLANE], rocblas_svect_singular, rocblas_svect_singular,
                                           (rocblas_int)B.Mp, (rocblas_int)B.Np, ws + B.a_off, (rocblas_int)B.Mp,
                                           B.Mp * B.Np, ws + B.s_off, B.kp, ws + B.u_off, (rocblas_int)B.Mp,
                                           B.Mp * B.kp, ws + B.v_off, (rocblas_int)B.kp, B.kp * B.Np,
                                           info + B.info_off, (rocblas_int)B.count),
          "rocsolver_dgesdd_strided_batched");
    }
    for (int i = 0; i < nb && i < NLANE; ++i) {
      ck(hipEventRecord(L.done[i], L.st[i]), "hipEventRecord");
      ck(hipStreamWaitEvent(s, L.done[i], 0), "hipStreamWaitEvent");
    }
    rocblas_handle h = L.h[0];
    std::vector<rocblas_int> hinfo(plan->info_ints);
    ck(hipMemcpyAsync(hinfo.data(), info, hinfo.size() * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(info)");
    ck(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (const Bucket& B : plan->buckets) {
      bool bad = false;
      for (int64_t i = 0; i < B.count; ++i) bad |= hinfo[B.info_off + i] != 0;
      if (!bad) continue;
      // gesdd overwrote the padded inputs: rebuild them for this bucket, retry with gesvd
      ck(hipMemsetAsync(ws + B.a_off, 0, (size_t)B.count * B.Mp * B.Np * 8, s), "hipMemsetAsync");
      ck(launch_copy(cp.pad, 8, A, ws, s), "svd pad");
      DevBuf eb;  // gesvd's superdiagonal scratch (this retry path only)
      ck(hipMalloc(&eb.p, (size_t)B.count * B.kp * 8 + 8), "hipMalloc(gesvd E)");
      // the padded inputs are rebuilt on s; lane 0 runs gesvd after them
      ck(hipEventRecord(L.fork, s), "hipEventRecord");
      ck(hipStreamWaitEvent(L.st[0], L.fork, 0), "hipStreamWaitEvent");
      ckb(rocsolver_dgesvd_strided_batched(h, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)B.Mp,
                                           (rocblas_int)B.Np, ws + B.a_off, (rocblas_int)B.Mp, B.Mp * B.Np,
                                           ws + B.s_off, B.kp, ws + B.u_off, (rocblas_int)B.Mp, B.Mp * B.kp,
                                           ws + B.v_off, (rocblas_int)B.kp, B.kp * B.Np, (double*)eb.p, B.kp,
                                           rocblas_outofplace, info + B.info_off, (rocblas_int)B.count),
          "rocsolver_dgesvd_strided_batched");
      ck(hipEventRecord(L.done[0], L.st[0]), "hipEventRecord");
      ck(hipStreamWaitEvent(s, L.done[0], 0), "hipStreamWaitEvent");
      std::vector<rocblas_int> hi(B.count);
      ck(hipMemcpyAsync(hi.data(), info + B.info_off, hi.size() * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      ck(hipStreamSynchronize(s), "hipStreamSynchronize");
      for (auto x : hi)
        if (x != 0) throw Error(ERR_NONFINITE, "SVD did not converge (gesdd and gesvd)");
    }
    ck(launch_copy(cp.u, 8, ws, U, s), "svd unpad U");
    ck(launch_copy(cp.v, 8, ws, Vt, s), "svd unpad Vt");
    ck(launch_copy(cp.s, 8, ws, S, s), "svd unpad S");
    return MPSH_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return (mpsh_status)e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MPSH_ERR_INVALID;
  }
}

void mpsh_svd_destroy(mpsh_svd plan) { delete plan; }

}  // extern "C"
