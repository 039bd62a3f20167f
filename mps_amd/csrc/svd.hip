// Batched per-block FP64 SVD on the device: the per-sector SVD of
// svdtrunc / svd (reference src/mpsutils.jl:116-179, src/symtools.jl:1-35).
// Blocks of different shapes are zero-padded into square shape buckets and
// each bucket is one rocSOLVER gesdd strided-batched
// call -- the divide-and-conquer SVD the reference gets from LAPACK gesdd.
// Zero padding adds only zero singular values; the leading min(m, n) singular
// triplets of a block are read back from its padded slot.  A bucket whose
// gesdd reports non-convergence is redone with gesvd (QR iteration), as the
// reference retries gesdd failures with gesvd (src/symtools.jl:14-19).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
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
  struct Dev {
    std::vector<std::unique_ptr<DevBuf>> bufs;
    DevCopy pad, u, v, s;
  };
  std::mutex mu;
  std::map<int, std::unique_ptr<Dev>> dev;

  template <class T>
  const T* up(std::vector<std::unique_ptr<DevBuf>>& bufs, const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    auto b = std::make_unique<DevBuf>();
    ck(hipMalloc(&b->p, v.size() * sizeof(T)), "hipMalloc(svd tables)");
    ck(hipMemcpy(b->p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(svd tables)");
    const T* r = (const T*)b->p;
    bufs.push_back(std::move(b));
    return r;
  }
  DevCopy upc(std::vector<std::unique_ptr<DevBuf>>& bufs, const CopyPlan& c) {
    DevCopy d;
    d.descs = up(bufs, c.descs);
    d.tile_start = up(bufs, c.tile_start);
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
      it = dev.emplace(d, std::move(D)).first;
    }
    return *it->second;
  }
};

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
    std::vector<int64_t> order(nblocks);
    for (int64_t b = 0; b < nblocks; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
      return std::max(m[x], n[x]) < std::max(m[y], n[y]);
    });
    std::vector<int64_t> Ps(nblocks);
    for (int64_t i = 0; i < nblocks; ++i) Ps[i] = pad16(std::max(m[order[i]], n[order[i]]));
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
