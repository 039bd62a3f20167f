// C ABI (include/mpsh.h): plan objects around the host planner, lazy
// per-device upload of their descriptor tables, and the execute calls that
// enqueue the HIP kernels on the caller's stream.  No call here allocates
// device memory except the one-time table upload of a plan.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <cmath>

#include "../../include/mpsh.h"
#include "kernels.hpp"
#include "planner.hpp"

using namespace mpsh;

static thread_local std::string g_err;

namespace mpsh {
void set_last_error(const std::string& m) { g_err = m; }

// Plan tables are uploaded on a per-device non-blocking stream of the
// library's own and waited for there: a plain hipMemcpy runs on the null
// stream, which first waits for every kernel queued on the caller's streams.
hipError_t upload_tables(void* dst, const void* src, size_t bytes) {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  int d = 0;
  hipError_t e = hipGetDevice(&d);
  if (e != hipSuccess) return e;
  hipStream_t s;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = streams.find(d);
    if (it == streams.end()) {
      e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
      if (e != hipSuccess) return e;
      it = streams.emplace(d, s).first;
    }
    s = it->second;
  }
  e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}
}  // namespace mpsh

static mpsh_status fail(int code, const std::string& m) {
  g_err = m;
  return (mpsh_status)code;
}

#define API_TRY try {
#define API_CATCH                                                      \
  }                                                                    \
  catch (const Error& e) { return fail(e.code, e.what()); }            \
  catch (const std::bad_alloc&) { return fail(ERR_OOM, "host OOM"); } \
  catch (const std::exception& e) { return fail(ERR_INVALID, e.what()); }

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(e == hipErrorOutOfMemory ? ERR_OOM : ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Every table of one plan on one device in a single allocation and a single
// upload: fill(f) runs the builder f twice, first only sizing the tables, then
// staging them at their final offsets, and copies the staged bytes once.
struct DevTables {
  DevBuf buf;
  std::vector<char> host;
  size_t off = 0;
  bool sizing = false;
  template <class T>
  const T* up(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    const size_t bytes = v.size() * sizeof(T);
    off = (off + 255) / 256 * 256;
    const size_t o = off;
    off += bytes;
    if (sizing) return nullptr;
    std::memcpy(host.data() + o, v.data(), bytes);
    return (const T*)((char*)buf.p + o);
  }
  template <class F>
  void fill(F&& f) {
    sizing = true;
    off = 0;
    f(*this);
    const size_t total = std::max<size_t>(off, 256);
    check(hipMalloc(&buf.p, total), "hipMalloc(plan tables)");
    host.assign(total, 0);
    sizing = false;
    off = 0;
    f(*this);
    check(upload_tables(buf.p, host.data(), std::max<size_t>(off, 1)), "upload(plan tables)");
    std::vector<char>().swap(host);
  }
  DevCopy copy(const CopyPlan& c) {
    DevCopy d;
    d.descs = up(c.descs);
    d.tile_map = up(c.tile_map);
    d.ndesc = (int)c.descs.size();
    d.ntiles = c.ntiles;
    return d;
  }
  // mode >= 0: the balanced per-work-group schedule for this kernel variant's
  // resident work groups (mode as in launch_gemm); mode < 0: shared queues
  // A stream-K plan (GemmPlan::use_streamk) always gets its per-work-group piece
  // lists, for min(resident work groups, the plan's bound) work groups per XCD.
  // kmode >= 0: the kernel mode the launch runs in; DMA variants then get the
  // per-tile records of that mode (GemmPlan::tile_records)
  DevGemm gemm(const GemmPlan& g, int mode = -1, int kmode = -1) {
    DevGemm d;
    std::vector<int32_t> qoff, woff;
    if (g.streamk) {
      if (g.KS != 0 || g.sk_mode < 0) throw Error(ERR_INVALID, "stream-K needs a DMA-streamed kernel variant");
      const int cap = sk_max_slots(g.sk_mode, g.BM, g.BN);
      int slots = gemm_slots_per_xcd(g.sk_mode, g.BM, g.BN, g.KS);
      if (slots <= 0 || slots > cap) slots = cap;
      std::vector<TileRed> tr;
      std::vector<GemmTile> t = g.streamk_schedule(slots, chunk_k(g.sk_mode, g.KS), &woff, &tr);
      (void)g.tiles(&qoff);
      d.q.slots = slots;
      d.q.woff = up(woff);
      for (int x = 0; x <= NXCD; ++x) d.q.off[x] = qoff[x];
      d.probs = up(g.probs);
      d.segs = up(g.segs);
      d.atab = up(g.atab);
      d.btab = up(g.btab);
      d.crow = up(g.crow);
      d.cld = up(g.cld);
      d.tiles = up(t);
      d.ntiles = (int)t.size();
      d.treds = up(tr);
      d.ntred = (int)tr.size();
      d.recs = up(g.tile_records(t, chunk_k(g.sk_mode, g.KS)));
      d.rec_mode = g.sk_mode;
      d.BM = g.BM;
      d.BN = g.BN;
      d.KS = g.KS;
      return d;
    }
    const int slots = mode >= 0 ? gemm_slots_per_xcd(mode, g.BM, g.BN, g.KS) : 0;
    std::vector<GemmTile> t = slots > 0 ? g.schedule(slots, chunk_k(mode, g.KS), &woff) : g.tiles(&qoff);
    if (slots > 0) {
      (void)g.tiles(&qoff);
      d.q.slots = slots;
      d.q.woff = up(woff);
    }
    for (int x = 0; x <= NXCD; ++x) d.q.off[x] = qoff[x];
    std::vector<RedDesc> r = g.reductions();
    d.probs = up(g.probs);
    d.segs = up(g.segs);
    d.atab = up(g.atab);
    d.btab = up(g.btab);
    d.crow = up(g.crow);
    d.cld = up(g.cld);
    d.tiles = up(t);
    d.ntiles = (int)t.size();
    if (kmode >= 0 && g.KS == 0) {
      d.recs = up(g.tile_records(t, chunk_k(kmode, g.KS)));
      d.rec_mode = kmode;
    }
    d.nred = (int)r.size() - 1;
    d.red_groups = r.back().start;
    if (d.nred > 0) {
      d.reds = up(r);
      std::vector<int32_t> rmap(2 * d.red_groups);
      for (int i = 0; i < d.nred; ++i)
        for (int64_t gi = r[i].start; gi < r[i + 1].start; ++gi) {
          rmap[2 * gi] = r[i].prob;
          rmap[2 * gi + 1] = (int32_t)(gi - r[i].start);
        }
      d.red_map = up(rmap);
    }
    d.BM = g.BM;
    d.BN = g.BN;
    d.KS = g.KS;
    return d;
  }
};

int current_device() {
  int d = 0;
  check(hipGetDevice(&d), "hipGetDevice");
  return d;
}

int elem_size(mpsh_dtype dt) { return dt == MPSH_F64 ? 8 : 4; }

void need_aligned(const void* p, const char* what) {
  if (((uintptr_t)p) & 15) throw Error(ERR_INVALID, std::string(what) + " must be 16-byte aligned");
}

// Optional override of the H_eff kernel variants, read at plan creation:
// MPSH_HEFF_TILES="s1=64x128x1,s4=64x64x2,s1h=128x64x1,s4h=64x64x2,s1d=64x64x1,s4d=64x64x1,split=1024,splith=2048,splitd=-1"
// (split targets < 0: stream-K)
HeffTiles tiles_from_env() {
  HeffTiles t;
  const char* e = std::getenv("MPSH_HEFF_TILES");
  if (!e) return t;
  std::string str(e);
  size_t pos = 0;
  while (pos < str.size()) {
    size_t comma = str.find(',', pos);
    std::string item = str.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
    pos = comma == std::string::npos ? str.size() : comma + 1;
    size_t eq = item.find('=');
    if (eq == std::string::npos) continue;
    std::string k = item.substr(0, eq), v = item.substr(eq + 1);
    if (k == "split") {
      t.split_target = std::atoll(v.c_str());
      continue;
    }
    if (k == "splitd") {
      t.split_target_d = std::atoll(v.c_str());
      continue;
    }
    if (k == "splith") {
      t.split_target_h = std::atoll(v.c_str());
      continue;
    }
    int bm = 0, bn = 0, ks = 1;
    if (std::sscanf(v.c_str(), "%dx%dx%d", &bm, &bn, &ks) < 2) throw Error(ERR_INVALID, "bad MPSH_HEFF_TILES");
    TileCfg c{bm, bn, ks};
    if (k == "s1") t.s1 = c;
    else if (k == "s4") t.s4 = c;
    else if (k == "s1h") t.s1h = c;
    else if (k == "s4h") t.s4h = c;
    else if (k == "s1d") t.s1d = c;
    else if (k == "s4d") t.s4d = c;
    else throw Error(ERR_INVALID, "bad MPSH_HEFF_TILES key");
  }
  return t;
}
int gemm_mode(mpsh_dtype dt) { return dt == MPSH_F64 ? 0 : (dt == MPSH_F32 ? 1 : 2); }

// GEMM tile schedule, read once from MPSH_GEMM_SCHEDULE: "balanced" gives each
// work group its own cost-balanced tile list (GemmPlan::schedule), anything else
// keeps the shared per-XCD queues (measured equal or faster on the H_eff GEMMs).
// Returns the mode to build balanced tables for, or -1 for the shared queues.
int schedule_mode(int mode) {
  static const int on = [] {
    const char* e = std::getenv("MPSH_GEMM_SCHEDULE");
    return (e && std::string(e) == "balanced") ? 1 : 0;
  }();
  return on ? mode : -1;
}

void check_dtype(int dt) {
  if (dt != MPSH_F64 && dt != MPSH_F32 && dt != MPSH_BF16) throw Error(ERR_INVALID, "unknown dtype");
}

std::vector<Leg> parse_legs(int32_t rank, const uint8_t* isdual, const int32_t* nch, const int64_t* ch,
                            const int64_t* dims) {
  if (rank <= 0 || !isdual || !nch || !ch || !dims) throw Error(ERR_INVALID, "bad leg arrays");
  std::vector<Leg> legs;
  int64_t p = 0;
  for (int i = 0; i < rank; ++i) {
    std::vector<std::pair<int64_t, int64_t>> pairs;
    if (nch[i] <= 0) throw Error(ERR_INVALID, "leg with no charges");
    for (int j = 0; j < nch[i]; ++j, ++p) pairs.push_back({ch[p], dims[p]});
    legs.push_back(make_leg(isdual[i] != 0, pairs));
  }
  return legs;
}

}  // namespace

struct mpsh_struct_s {
  Struct s;
};

static mpsh_struct wrap(const Struct& s) {
  mpsh_struct r = new mpsh_struct_s;
  r->s = s;
  return r;
}

struct mpsh_releg_s {
  RelegPlan p;
  std::mutex mu;
  std::map<int, std::pair<std::unique_ptr<DevTables>, DevCopy>> dev;
  const DevCopy& get() {
    std::lock_guard<std::mutex> g(mu);
    int d = current_device();
    auto it = dev.find(d);
    if (it == dev.end()) {
      auto t = std::make_unique<DevTables>();
      DevCopy c;
      t->fill([&](DevTables& T) { c = T.copy(p.copy); });
      it = dev.emplace(d, std::make_pair(std::move(t), c)).first;
    }
    return it->second.second;
  }
};

struct ContractDev {
  std::unique_ptr<DevTables> t;
  DevCopy ga, gb, sc;
  DevGemm gm;   // register-staged operands (BF16)
  DevGemm gmd;  // the same tables with DMA-streamed operands (FP64, FP32): the
                // fused operands' rows are padded to 4 elements, 16-byte aligned
};

struct mpsh_contract_s {
  ContractPlan p;
  std::mutex mu;
  std::map<int, ContractDev> dev;
  const ContractDev& get() {
    std::lock_guard<std::mutex> g(mu);
    int d = current_device();
    auto it = dev.find(d);
    if (it == dev.end()) {
      ContractDev c;
      c.t = std::make_unique<DevTables>();
      c.t->fill([&](DevTables& T) {
        c.ga = T.copy(p.FA.gather);
        c.gb = T.copy(p.FB.gather);
        c.sc = T.copy(p.scatter);
        c.gm = T.gemm(p.gemm);
      });
      c.gmd = c.gm;
      c.gmd.KS = 0;
      it = dev.emplace(d, std::move(c)).first;
    }
    return it->second;
  }
};

struct HeffDev {
  std::unique_ptr<DevTables> t;
  DevMpo mpo;
  DevCopy pack_v, pack_envL, pack_envR;
};

DevMpo upload_mpo(DevTables& t, const Step13& p) {
  DevMpo m;
  m.sigs = t.up(p.sigs);
  m.terms = t.up(p.terms);
  m.pairs = t.up(p.pairs);
  m.groups = t.up(p.groups);
  m.inrows = t.up(p.inrows);
  m.outrows = t.up(p.outrows);
  for (int i = 0; i < 5; ++i) m.work[i] = t.up(p.work[i]);
  m.nterms = (int)p.terms.size();
  for (int i = 0; i < 5; ++i) m.nwork[i] = (int)p.work[i].size();
  m.vec = p.padr > 1;
  m.cmax = (int)p.maxC;
  m.rmax = (int)p.maxR;
  m.t1_elems = p.t1_elems;
  m.t3_elems = p.t3_elems;
  mpo_options(m);
  return m;
}

// device tables of one environment core
struct EnvCoreDev {
  DevCopy pack_ket, pack_env, pack_bra;
  DevMpo mpo;
  DevGemm g1, g1d, gF, gF64;  // step 1 and the final GEMM, FP32 and FP64 schedules
};
struct EnvDev {
  std::unique_ptr<DevTables> t;
  EnvCoreDev core[2];
  DevCopy perm[2][3];
  const double* wr_f64 = nullptr;  // identity WR of one-site plans
  const float* wr_f32 = nullptr;
};

// a GEMM kernel variant of one H_eff stage: (stage 1|4, mode, BM, BN, KS, split target)
typedef std::tuple<int, int, int, int, int, int64_t> VariantKey;

// Tuned variant choice per plan signature (process-wide): plans with the same
// block-size histogram signature reuse the timing of the first one.
static std::mutex g_tuned_mu;
static std::map<std::string, std::pair<VariantKey, VariantKey>> g_tuned;

struct mpsh_heff_s {
  HeffPlan p;
  std::mutex mu;
  std::map<int, HeffDev> dev;
  // per device: kernel variants uploaded so far; per mode: the active step-1 / step-4 variant
  std::map<std::pair<int, VariantKey>, std::pair<std::unique_ptr<DevTables>, DevGemm>> var;
  VariantKey sel1[3], sel4[3];
  bool tuned[3] = {false, false, false};
  const DevGemm& variant(const VariantKey& k) {
    std::lock_guard<std::mutex> g(mu);
    const int d = current_device();
    auto it = var.find({d, k});
    if (it == var.end()) {
      const int stage = std::get<0>(k), mode = std::get<1>(k);
      const TileCfg c{std::get<2>(k), std::get<3>(k), std::get<4>(k)};
      GemmPlan gp = gemm_variant(stage == 1 ? p.g1 : p.g4, c, std::get<5>(k), mode);
      auto t = std::make_unique<DevTables>();
      DevGemm dg;
      t->fill([&](DevTables& T) { dg = T.gemm(gp, schedule_mode(mode), mode); });
      it = var.emplace(std::make_pair(d, k), std::make_pair(std::move(t), dg)).first;
    }
    return it->second.second;
  }
  // the tables already on the current device, or nullptr (no upload)
  const HeffDev* peek() {
    std::lock_guard<std::mutex> g(mu);
    auto it = dev.find(current_device());
    return it == dev.end() ? nullptr : &it->second;
  }
  const DevGemm* peek_variant(const VariantKey& k) {
    std::lock_guard<std::mutex> g(mu);
    auto it = var.find({current_device(), k});
    return it == var.end() ? nullptr : &it->second.second;
  }
  const HeffDev& get() {
    std::lock_guard<std::mutex> g(mu);
    int d = current_device();
    auto it = dev.find(d);
    if (it == dev.end()) {
      HeffDev h;
      h.t = std::make_unique<DevTables>();
      h.t->fill([&](DevTables& T) {
        h.mpo = upload_mpo(T, p);
        h.pack_v = T.copy(p.pack_v);
        h.pack_envL = T.copy(p.pack_envL);
        h.pack_envR = T.copy(p.pack_envR);
      });
      it = dev.emplace(d, std::move(h)).first;
    }
    return it->second;
  }
};

// Scratch of the allocation-free H_eff application: one copy of this array per
// GPU, held until the process ends.  Host code uses only the device address that
// hipGetSymbolAddress returns for the current device (kept per device below),
// never the array's name.  Every call on one GPU shares it: calls that overlap
// on one GPU race.  Nothing in it needs to start at zero (the kernels write every
// element they later read).
static __device__ __align__(256) unsigned char scratch[MPSH_NOALLOC_SCRATCH_BYTES];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];  // per device: address of `scratch`, null until looked up

// address of the current device's `scratch` (looked up once per device, no allocation)
static void* noalloc_scratch() {
  const int d = current_device();
  if (d < 0 || d >= kMaxDevices) throw Error(ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[d]) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr || ((uintptr_t)ptr & 255)) throw Error(ERR_DEVICE, "scratch symbol address is not 256-byte aligned");
    g_scratch_ptr[d] = ptr;
  }
  return g_scratch_ptr[d];
}

static size_t align16(size_t b) { return (b + 15) / 16 * 16; }

// Byte layout of an H_eff plan's prepared operator [envL_p | envR_p | W] and
// per-matvec workspace [v_p | T1 | T3 | split-K partials] for one dtype.  FP64 /
// FP32 keep everything in that type; BF16 keeps v_p, envL_p, envR_p, T1 and T3
// in BF16 (the packs convert) with the combined MPO W and the split-K partials
// in FP32, and runs the BF16-storage GEMM kernels (modes 3 and 4).
struct HeffBytes {
  size_t es;                          // bytes per packed / T1 / T3 element
  size_t vp, t1, t3, part;            // workspace offsets
  size_t el, er, w;                   // prepared-operator offsets
  size_t prep, ws;                    // total sizes (16-byte multiples)
  int wes;                            // bytes per W element
  int mode1, mode4;                   // kernel modes of the step-1 / step-4 GEMMs
};
static HeffBytes heff_bytes(const HeffPlan& p, mpsh_dtype dt) {
  HeffBytes b;
  const bool h = dt == MPSH_BF16;
  b.es = h ? 2 : (size_t)elem_size(dt);
  const size_t ps = h ? 4 : b.es;  // partials
  b.wes = h ? 4 : (int)b.es;
  b.vp = p.vp_off * b.es;
  b.t1 = p.t1_off * b.es;
  b.t3 = p.t3_off * b.es;
  b.part = p.part_off * b.es;
  b.ws = align16(b.part + (size_t)std::max<int64_t>(p.ws_elems - p.part_off, 1) * ps);
  b.el = p.el_off * b.es;
  b.er = p.er_off * b.es;
  b.w = h ? (p.pw_off * b.es + 255) / 256 * 256 : p.pw_off * b.es;
  b.prep = align16(b.w + (size_t)p.w_elems * b.wes);
  b.mode1 = h ? 3 : gemm_mode(dt);
  b.mode4 = h ? 4 : gemm_mode(dt);
  return b;
}

// The kernel stages of one prepared H_eff application on a workspace of
// mpsh_heff_apply_workspace bytes (mask: 1 pack v, 2 step-1 GEMM, 4 MPO pass,
// 8 step-4 GEMM).
static void heff_stages(const HeffPlan& p, const HeffDev& d, const DevGemm& g1, const DevGemm& g4, mpsh_dtype dt,
                        int32_t stages, const void* prepared, const void* v, void* out, char* ws, hipStream_t s) {
  const HeffBytes b = heff_bytes(p, dt);
  const char* pr = (const char*)prepared;
  void* vp = ws + b.vp;
  void* T1 = ws + b.t1;
  void* T3 = ws + b.t3;
  void* Pt = ws + b.part;
  if (stages & 1)
    check(dt == MPSH_BF16 ? launch_copy_to_bf16(d.pack_v, (const float*)v, vp, s)
                          : launch_copy(d.pack_v, (int)b.es, v, vp, s),
          "pack v");
  if (stages & 2) check(launch_gemm(g1, b.mode1, vp, pr + b.el, T1, Pt, s), "step-1 gemm");
  if (stages & 4) check(launch_mpo_apply(d.mpo, (int)b.es, pr + b.w, T1, T3, s), "mpo");
  if (stages & 8) check(launch_gemm(g4, b.mode4, T3, pr + b.er, out, Pt, s), "step-4 gemm");
}

// environment updates and the H_eff adjoint share one plan object
struct mpsh_env_s {
  EnvPlan p;
  std::mutex mu;
  std::map<int, EnvDev> dev;
  const EnvDev& get() {
    std::lock_guard<std::mutex> g(mu);
    int d = current_device();
    auto it = dev.find(d);
    if (it == dev.end()) {
      EnvDev e;
      e.t = std::make_unique<DevTables>();
      const std::vector<float> wr32(p.wr_identity.begin(), p.wr_identity.end());
      e.t->fill([&](DevTables& T) {
        for (int c = 0; c < p.ncore; ++c) {
          const EnvCore& C = p.core[c];
          EnvCoreDev& D = e.core[c];
          D.pack_ket = T.copy(C.pack_v);
          D.pack_env = T.copy(C.pack_envL);
          D.pack_bra = T.copy(C.pack_bra);
          D.mpo = upload_mpo(T, C);
          D.g1 = T.gemm(C.g1, schedule_mode(1), 1);
          D.g1d = T.gemm(C.g1, schedule_mode(0), 0);
          D.gF = T.gemm(C.gF, schedule_mode(1), 1);
          D.gF64 = T.gemm(p.core_f64[c].gF, schedule_mode(0), 0);
          for (int k = 0; k < 3; ++k)
            if (p.has_perm[c][k]) e.perm[c][k] = T.copy(p.perm[c][k].copy);
        }
        if (p.wr_is_identity) {
          e.wr_f64 = T.up(p.wr_identity);
          e.wr_f32 = T.up(wr32);
        }
      });
      it = dev.emplace(d, std::move(e)).first;
    }
    return it->second;
  }
};
struct mpsh_envupdate_s : mpsh_env_s {};
struct mpsh_heff_adjoint_s : mpsh_env_s {};

// One environment core on device buffers: packs, combined MPO, step-1 GEMM, MPO
// pass, final GEMM against the bra (the ket's data).
static void run_env_core(const mpsh_env_s& h, const EnvDev& d, int c, mpsh_dtype dt, const void* env,
                         const void* ket, const void* WL, const void* WR, void* out, char* ws, hipStream_t s) {
  const EnvCore& C = dt == MPSH_F64 ? h.p.core_f64[c] : h.p.core[c];
  const EnvCoreDev& D = d.core[c];
  const int es = elem_size(dt), mode = gemm_mode(dt);
  check(launch_copy(D.pack_ket, es, ket, ws + C.kp_off * es, s), "env: pack ket");
  check(launch_copy(D.pack_env, es, env, ws + C.ep_off * es, s), "env: pack env");
  check(launch_copy(D.pack_bra, es, ket, ws + C.bp_off * es, s), "env: pack bra");
  check(launch_mpo_combine(D.mpo, es, WL, WR, ws + C.w_off * es, s), "env: mpo combine");
  check(launch_gemm(dt == MPSH_F64 ? D.g1d : D.g1, mode, ws + C.kp_off * es, ws + C.ep_off * es, ws + C.t1_off * es,
                    ws + C.part_off * es, s),
        "env: step-1 gemm");
  check(launch_mpo_apply(D.mpo, es, ws + C.w_off * es, ws + C.t1_off * es, ws + C.t3_off * es, s), "env: mpo");
  check(launch_gemm(dt == MPSH_F64 ? D.gF64 : D.gF, mode, ws + C.t3_off * es, ws + C.bp_off * es, out,
                    ws + C.part_off * es, s),
        "env: final gemm");
}

static void env_check_dtype(mpsh_dtype dt) {
  if (dt != MPSH_F64 && dt != MPSH_F32) throw Error(ERR_UNSUPPORTED, "environment plans run in FP64 or FP32");
}

static size_t env_workspace(const mpsh_env_s* h, mpsh_dtype dt) {
  return h ? (size_t)std::max<int64_t>(h->p.ws_elems, 1) * (size_t)elem_size(dt) + 16 : 0;
}

static void env_stats(const mpsh_env_s* h, int64_t* st) {
  if (!h || !st) throw Error(ERR_INVALID, "null argument");
  st[0] = h->p.ref_flops;
  st[1] = h->p.ref_elems;
  st[2] = 0;
  st[3] = 0;
  for (int c = 0; c < h->p.ncore; ++c) {
    st[2] += h->p.core[c].g1.flops + h->p.core[c].gF.flops;
    st[3] += h->p.core[c].mpo_flops;
  }
}

extern "C" {

const char* mpsh_last_error(void) { return g_err.c_str(); }
const char* mpsh_version(void) { return "mpsh 0.1 gfx950"; }

// ---------------------------------------------------------------------------
// structures
// ---------------------------------------------------------------------------
mpsh_status mpsh_struct_create(int64_t charge, int32_t rank, const uint8_t* isdual, const int32_t* ncharges,
                               const int64_t* charges, const int64_t* dims, mpsh_struct* out) {
  API_TRY
  if (!out) throw Error(ERR_INVALID, "null out");
  *out = wrap(make_struct(charge, parse_legs(rank, isdual, ncharges, charges, dims)));
  return MPSH_OK;
  API_CATCH
}

void mpsh_struct_destroy(mpsh_struct s) { delete s; }
int32_t mpsh_struct_rank(mpsh_struct s) { return s ? s->s.rank() : -1; }
int64_t mpsh_struct_charge(mpsh_struct s) { return s ? s->s.charge : 0; }
int64_t mpsh_struct_nblocks(mpsh_struct s) { return s ? s->s.nblocks : -1; }
int64_t mpsh_struct_numel(mpsh_struct s) { return s ? s->s.numel() : -1; }

int32_t mpsh_struct_leg_ncharges(mpsh_struct s, int32_t leg) {
  if (!s || leg < 0 || leg >= s->s.rank()) return -1;
  return (int32_t)s->s.legs[leg].charges.size();
}

mpsh_status mpsh_struct_leg(mpsh_struct s, int32_t leg, uint8_t* isdual, int64_t* charges, int64_t* dims) {
  if (!s || leg < 0 || leg >= s->s.rank()) return fail(ERR_INVALID, "bad leg");
  const Leg& l = s->s.legs[leg];
  if (isdual) *isdual = l.isdual ? 1 : 0;
  if (charges) std::memcpy(charges, l.charges.data(), l.charges.size() * 8);
  if (dims) std::memcpy(dims, l.dims.data(), l.dims.size() * 8);
  return MPSH_OK;
}

mpsh_status mpsh_struct_blocks(mpsh_struct s, int64_t* sectors, int64_t* sizes, int64_t* offsets) {
  if (!s) return fail(ERR_INVALID, "null structure");
  if (sectors) std::memcpy(sectors, s->s.sectors.data(), s->s.sectors.size() * 8);
  if (sizes) std::memcpy(sizes, s->s.sizes.data(), s->s.sizes.size() * 8);
  if (offsets) std::memcpy(offsets, s->s.offsets.data(), s->s.offsets.size() * 8);
  return MPSH_OK;
}

int64_t mpsh_struct_find(mpsh_struct s, const int64_t* sector) {
  if (!s || !sector) return -1;
  try {
    return s->s.find(std::vector<int64_t>(sector, sector + s->s.rank()));
  } catch (...) {
    return -1;  // host allocation failure: reported as not found, nothing crosses the C boundary
  }
}

// ---------------------------------------------------------------------------
// re-legging
// ---------------------------------------------------------------------------
static mpsh_status finish_releg(RelegPlan&& p, mpsh_releg* plan, mpsh_struct* out) {
  if (!plan || !out) return fail(ERR_INVALID, "null out");
  mpsh_releg r = new mpsh_releg_s;
  r->p = std::move(p);
  *out = wrap(r->p.out);
  *plan = r;
  return MPSH_OK;
}

mpsh_status mpsh_permutelegs_plan(mpsh_struct A, const int32_t* perm, mpsh_releg* plan, mpsh_struct* out) {
  API_TRY
  if (!A || !perm) throw Error(ERR_INVALID, "null argument");
  return finish_releg(permutelegs_plan(A->s, std::vector<int>(perm, perm + A->s.rank())), plan, out);
  API_CATCH
}

mpsh_status mpsh_fuselegs_plan(mpsh_struct A, int32_t l, int32_t n, int32_t dual, mpsh_releg* plan,
                               mpsh_struct* out) {
  API_TRY
  if (!A) throw Error(ERR_INVALID, "null argument");
  return finish_releg(fuselegs_plan(A->s, l, n, dual != 0), plan, out);
  API_CATCH
}

mpsh_status mpsh_splitleg_plan(mpsh_struct A, int32_t l, int32_t nspaces, const uint8_t* isdual,
                               const int32_t* ncharges, const int64_t* charges, const int64_t* dims,
                               mpsh_releg* plan, mpsh_struct* out) {
  API_TRY
  if (!A) throw Error(ERR_INVALID, "null argument");
  return finish_releg(splitleg_plan(A->s, l, parse_legs(nspaces, isdual, ncharges, charges, dims)), plan, out);
  API_CATCH
}

mpsh_status mpsh_symmatrix_plan(mpsh_struct A, int32_t nrows, const int32_t* rows, int32_t ncols,
                                const int32_t* cols, mpsh_releg* plan, mpsh_struct* out) {
  API_TRY
  if (!A || (nrows && !rows) || (ncols && !cols)) throw Error(ERR_INVALID, "null argument");
  return finish_releg(symmatrix_releg(A->s, std::vector<int>(rows, rows + nrows), std::vector<int>(cols, cols + ncols)),
                      plan, out);
  API_CATCH
}

mpsh_status mpsh_releg_execute(mpsh_releg plan, mpsh_dtype dt, const void* src, void* dst, void* stream) {
  API_TRY
  check_dtype(dt);
  if (!plan || (!src && plan->p.in.numel()) || (!dst && plan->p.out.numel())) throw Error(ERR_INVALID, "null argument");
  const DevCopy& c = plan->get();
  check(launch_copy(c, elem_size(dt), src, dst, (hipStream_t)stream), "copy kernel");
  return MPSH_OK;
  API_CATCH
}

void mpsh_releg_destroy(mpsh_releg plan) { delete plan; }

// ---------------------------------------------------------------------------
// contract
// ---------------------------------------------------------------------------
mpsh_status mpsh_contract_plan(mpsh_struct A, const int32_t* idxA, mpsh_struct B, const int32_t* idxB,
                               mpsh_contract* plan, mpsh_struct* C) {
  API_TRY
  if (!A || !B || !idxA || !idxB || !plan || !C) throw Error(ERR_INVALID, "null argument");
  mpsh_contract p = new mpsh_contract_s;
  try {
    p->p = contract_plan(A->s, std::vector<int>(idxA, idxA + A->s.rank()), B->s,
                         std::vector<int>(idxB, idxB + B->s.rank()));
  } catch (...) {
    delete p;
    throw;
  }
  *C = wrap(p->p.C);
  *plan = p;
  return MPSH_OK;
  API_CATCH
}

size_t mpsh_contract_workspace(mpsh_contract plan, mpsh_dtype dt) {
  if (!plan) return 0;
  return (size_t)plan->p.ws_elems * (size_t)elem_size(dt);
}

int64_t mpsh_contract_flops(mpsh_contract plan) { return plan ? plan->p.gemm.flops : 0; }

mpsh_status mpsh_contract_execute(mpsh_contract plan, mpsh_dtype dt, const void* A, const void* B, void* C,
                                  void* workspace, void* stream) {
  API_TRY
  check_dtype(dt);
  if (!plan || (!A && plan->p.A.numel()) || (!B && plan->p.B.numel()) || (!C && plan->p.C.numel()) ||
      (!workspace && plan->p.ws_elems))
    throw Error(ERR_INVALID, "null argument");
  need_aligned(workspace, "workspace");
  const ContractDev& d = plan->get();
  hipStream_t s = (hipStream_t)stream;
  const int es = elem_size(dt);
  char* ws = (char*)workspace;
  check(launch_copy(d.ga, es, A, ws, s), "gather A");
  check(launch_copy(d.gb, es, B, ws, s), "gather B");
  check(launch_gemm(dt == MPSH_BF16 ? d.gm : d.gmd, gemm_mode(dt), ws + plan->p.fa_off * es, ws + plan->p.fb_off * es,
                    ws + plan->p.fc_off * es, nullptr, s),
        "grouped gemm");
  check(launch_copy(d.sc, es, ws, C, s), "scatter C");
  return MPSH_OK;
  API_CATCH
}

void mpsh_contract_destroy(mpsh_contract plan) { delete plan; }

// ---------------------------------------------------------------------------
// H_eff
// ---------------------------------------------------------------------------
mpsh_status mpsh_heff_plan(mpsh_struct v, mpsh_struct envL, mpsh_struct envR, mpsh_struct WL, mpsh_struct WR,
                           mpsh_heff* plan, mpsh_struct* out) {
  API_TRY
  if (!v || !envL || !envR || !WL || !WR || !plan || !out) throw Error(ERR_INVALID, "null argument");
  mpsh_heff h = new mpsh_heff_s;
  try {
    const HeffTiles t = tiles_from_env();
    h->p = heff_plan(v->s, envL->s, envR->s, WL->s, WR->s, t);
    if (h->p.maxR > 64 || h->p.maxC > 64)
      throw Error(ERR_UNSUPPORTED, "combined MPO group wider than 64 channels");
    const TileCfg* c1[3] = {&t.s1d, &t.s1, &t.s1h};
    const TileCfg* c4[3] = {&t.s4d, &t.s4, &t.s4h};
    // index m = gemm_mode(dt); BF16 (m = 2) runs the BF16-storage kernels (modes 3 and 4)
    for (int m = 0; m < 3; ++m) {
      h->sel1[m] = VariantKey{1, m == 2 ? 3 : m, c1[m]->BM, c1[m]->BN, c1[m]->KS, 0};
      h->sel4[m] = VariantKey{4, m == 2 ? 4 : m, c4[m]->BM, c4[m]->BN, c4[m]->KS,
                              m == 0 ? t.split_target_d : (m == 2 ? t.split_target_h : t.split_target)};
      if (!tune_tile_ok(std::get<1>(h->sel1[m]), *c1[m]) || !tune_tile_ok(std::get<1>(h->sel4[m]), *c4[m]))
        throw Error(ERR_INVALID, "MPSH_HEFF_TILES: no such kernel variant for this precision");
    }
  } catch (...) {
    delete h;
    throw;
  }
  *out = wrap(h->p.out);
  *plan = h;
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_heff_engine_stats(mpsh_heff plan, int64_t* stats) {
  if (!plan || !stats) return fail(ERR_INVALID, "null argument");
  const HeffPlan& p = plan->p;
  stats[0] = p.t1_elems;
  stats[1] = p.t3_elems;
  stats[2] = p.w_elems;
  stats[3] = p.mpo_flops;
  stats[4] = p.g1.flops;
  stats[5] = p.g4.flops;
  stats[6] = p.vp_elems;
  stats[7] = p.ws_elems;
  return MPSH_OK;
}

size_t mpsh_heff_prepared_size(mpsh_heff plan, mpsh_dtype dt) {
  if (!plan || (dt != MPSH_F64 && dt != MPSH_F32 && dt != MPSH_BF16)) return 0;
  return heff_bytes(plan->p, dt).prep;
}

size_t mpsh_heff_apply_workspace(mpsh_heff plan, mpsh_dtype dt) {
  if (!plan || (dt != MPSH_F64 && dt != MPSH_F32 && dt != MPSH_BF16)) return 0;
  return heff_bytes(plan->p, dt).ws;
}

size_t mpsh_heff_workspace(mpsh_heff plan, mpsh_dtype dt) {
  return mpsh_heff_apply_workspace(plan, dt) + mpsh_heff_prepared_size(plan, dt);
}

mpsh_status mpsh_heff_stats(mpsh_heff plan, int64_t* stats) {
  if (!plan || !stats) return fail(ERR_INVALID, "null argument");
  const HeffPlan& p = plan->p;
  stats[0] = p.ref_flops;
  stats[1] = p.ref_elems;
  for (int i = 0; i < 4; ++i) {
    stats[2 + i] = p.step_flops[i];
    stats[6 + i] = p.step_elems[i];
  }
  return MPSH_OK;
}

mpsh_status mpsh_heff_prepare(mpsh_heff plan, mpsh_dtype dt, const void* envL, const void* envR, const void* WL,
                              const void* WR, void* prepared, void* stream) {
  API_TRY
  check_dtype(dt);
  if (!plan) throw Error(ERR_INVALID, "null plan");
  const HeffPlan& p = plan->p;
  if ((!envL && p.envL.numel()) || (!envR && p.envR.numel()) || (!WL && p.WL.numel()) || (!WR && p.WR.numel()) ||
      (!prepared && p.prep_elems))
    throw Error(ERR_INVALID, "null argument");
  need_aligned(prepared, "prepared");
  const HeffDev& d = plan->get();
  // the GEMM tables of the selected variants are uploaded here as well, so that
  // applications on this device need no allocation
  plan->variant(plan->sel1[gemm_mode(dt)]);
  plan->variant(plan->sel4[gemm_mode(dt)]);
  hipStream_t s = (hipStream_t)stream;
  const HeffBytes b = heff_bytes(p, dt);
  char* pr = (char*)prepared;
  // padded plans: the K padding inside envR_p's rows is read by the step-4 GEMM
  // (against zeros in T3) and must be finite: the region is zeroed first
  if (p.padr > 1) check(launch_zero(pr + b.er, (size_t)(p.pw_off - p.er_off) * b.es, s), "zero envR");
  if (dt == MPSH_BF16) {  // FP32 operands packed into BF16, the combined MPO in FP32
    check(launch_copy_to_bf16(d.pack_envL, (const float*)envL, pr + b.el, s), "pack envL");
    check(launch_copy_to_bf16(d.pack_envR, (const float*)envR, pr + b.er, s), "pack envR");
  } else {
    check(launch_copy(d.pack_envL, (int)b.es, envL, pr + b.el, s), "pack envL");
    check(launch_copy(d.pack_envR, (int)b.es, envR, pr + b.er, s), "pack envR");
  }
  check(launch_mpo_combine(d.mpo, b.wes, WL, WR, pr + b.w, s), "mpo combine");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_heff_apply_prepared_stages(mpsh_heff plan, mpsh_dtype dt, int32_t stages, const void* prepared,
                                            const void* v, void* out, void* workspace, void* stream) {
  API_TRY
  check_dtype(dt);
  if (!plan) throw Error(ERR_INVALID, "null plan");
  const HeffPlan& p = plan->p;
  if ((!v && p.v.numel()) || (!out && p.out.numel()) || (!prepared && p.prep_elems) || !workspace)
    throw Error(ERR_INVALID, "null argument");
  need_aligned(prepared, "prepared");
  need_aligned(workspace, "workspace");
  const int mode = gemm_mode(dt);
  heff_stages(p, plan->get(), plan->variant(plan->sel1[mode]), plan->variant(plan->sel4[mode]), dt, stages,
              prepared, v, out, (char*)workspace, (hipStream_t)stream);
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_heff_apply_prepared(mpsh_heff plan, mpsh_dtype dt, const void* prepared, const void* v,
                                     void* out, void* workspace, void* stream) {
  return mpsh_heff_apply_prepared_stages(plan, dt, 15, prepared, v, out, workspace, stream);
}

mpsh_status mpsh_heff_apply_prepared_noalloc(mpsh_heff plan, mpsh_dtype dt, const void* prepared, const void* v,
                                             void* out, void* stream) {
  API_TRY
  check_dtype(dt);
  if (!plan) throw Error(ERR_INVALID, "null plan");
  const HeffPlan& p = plan->p;
  if ((!v && p.v.numel()) || (!out && p.out.numel()) || (!prepared && p.prep_elems))
    throw Error(ERR_INVALID, "null argument");
  need_aligned(prepared, "prepared");
  const size_t need = mpsh_heff_apply_workspace(plan, dt);
  if (need > (size_t)MPSH_NOALLOC_SCRATCH_BYTES)
    throw Error(ERR_INVALID, "mpsh_heff_apply_prepared_noalloc: the plan needs " + std::to_string(need) +
                                 " bytes of scratch, more than the fixed " +
                                 std::to_string((size_t)MPSH_NOALLOC_SCRATCH_BYTES) +
                                 " (use mpsh_heff_apply_prepared with a workspace)");
  const int mode = gemm_mode(dt);
  const HeffDev* d = plan->peek();
  const DevGemm* g1 = plan->peek_variant(plan->sel1[mode]);
  const DevGemm* g4 = plan->peek_variant(plan->sel4[mode]);
  if (!d || !g1 || !g4)
    throw Error(ERR_INVALID, "mpsh_heff_apply_prepared_noalloc: the plan's tables are not on this device "
                             "(call mpsh_heff_prepare for this dtype on this device first)");
  heff_stages(p, *d, *g1, *g4, dt, 15, prepared, v, out, (char*)noalloc_scratch(), (hipStream_t)stream);
  return MPSH_OK;
  API_CATCH
}

// Block-size histogram signature of a plan's two GEMMs (totals and averages in
// factor-2 buckets), the key under which tuned variants are shared between plans.
static std::string heff_signature(const HeffPlan& p, int mode) {
  auto q = [](double x) { return x > 0 ? (long)std::lround(std::log2(x)) : -1L; };  // factor-2 buckets
  auto stat = [&](const GemmPlan& g) {
    double M = 0, N = 0, K = 0, n = 0;
    for (const GemmProb& pr : g.probs) {
      double k = 0;
      for (int s2 = 0; s2 < pr.nseg; ++s2) k += (double)g.segs[pr.seg0 + s2].K;
      M += pr.M;
      N += pr.N;
      K += k;
      n += 1;
    }
    n = std::max(n, 1.0);
    return std::to_string(q(g.flops)) + "," + std::to_string(q(M / n)) + "," + std::to_string(q(N / n)) + "," +
           std::to_string(q(K / n)) + "," + std::to_string(q(n));
  };
  return std::to_string(mode) + "|" + stat(p.g1) + "|" + stat(p.g4);
}

mpsh_status mpsh_heff_tune(mpsh_heff plan, mpsh_dtype dt, const void* prepared, const void* v, void* out,
                           void* workspace, int32_t reps, int32_t force, int32_t* choice, void* stream) {
  API_TRY
  check_dtype(dt);
  if (!plan) throw Error(ERR_INVALID, "null plan");
  const HeffPlan& p = plan->p;
  if ((!v && p.v.numel()) || (!out && p.out.numel()) || (!prepared && p.prep_elems) || !workspace)
    throw Error(ERR_INVALID, "null argument");
  need_aligned(prepared, "prepared");
  need_aligned(workspace, "workspace");
  const int mode = gemm_mode(dt);
  const std::string sig = heff_signature(p, mode);
  bool cached = false;
  if (!force) {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    auto it = g_tuned.find(sig);
    if (it != g_tuned.end()) {
      plan->sel1[mode] = it->second.first;
      plan->sel4[mode] = it->second.second;
      cached = true;
    }
  }
  if (!cached && p.g1.probs.size() + p.g4.probs.size() > 0) {
    hipStream_t s = (hipStream_t)stream;
    const HeffDev& d = plan->get();
    const HeffBytes b = heff_bytes(p, dt);
    const char* pr = (const char*)prepared;
    char* ws = (char*)workspace;
    void* vp = ws + b.vp;
    void* T1 = ws + b.t1;
    void* T3 = ws + b.t3;
    void* Pt = ws + b.part;
    reps = std::max(1, std::min(reps, 50));
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    auto timeit = [&](const DevGemm& g, int kmode, const void* A, const void* B, void* C) {
      check(launch_gemm(g, kmode, A, B, C, Pt, s), "tune gemm");  // warm
      check(hipEventRecord(e0, s), "hipEventRecord");
      for (int r = 0; r < reps; ++r) check(launch_gemm(g, kmode, A, B, C, Pt, s), "tune gemm");
      check(hipEventRecord(e1, s), "hipEventRecord");
      check(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0;
      check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      return ms;
    };
    heff_stages(p, d, plan->variant(plan->sel1[mode]), plan->variant(plan->sel4[mode]), dt, 1, prepared, v, out, ws,
                s);
    float best = 1e30f;
    for (const TileCfg& c : kTuneTiles) {
      if (!tune_tile_ok(b.mode1, c)) continue;
      const VariantKey k{1, b.mode1, c.BM, c.BN, c.KS, 0};
      const float t = timeit(plan->variant(k), b.mode1, vp, pr + b.el, T1);
      if (t < best) {
        best = t;
        plan->sel1[mode] = k;
      }
    }
    check(launch_mpo_apply(d.mpo, (int)b.es, pr + b.w, T1, T3, s), "mpo");
    best = 1e30f;
    for (const TileCfg& c : kTuneTiles) {
      if (!tune_tile_ok(b.mode4, c)) continue;
      for (int64_t sp : kTuneSplits) {
        const VariantKey k{4, b.mode4, c.BM, c.BN, c.KS, sp};
        const float t = timeit(plan->variant(k), b.mode4, T3, pr + b.er, out);
        if (t < best) {
          best = t;
          plan->sel4[mode] = k;
        }
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> g(g_tuned_mu);
    g_tuned[sig] = {plan->sel1[mode], plan->sel4[mode]};
  }
  plan->tuned[mode] = true;
  if (choice) {
    const VariantKey* ks[2] = {&plan->sel1[mode], &plan->sel4[mode]};
    for (int i = 0; i < 2; ++i) {
      choice[4 * i + 0] = std::get<2>(*ks[i]);
      choice[4 * i + 1] = std::get<3>(*ks[i]);
      choice[4 * i + 2] = std::get<4>(*ks[i]);
      choice[4 * i + 3] = (int32_t)std::get<5>(*ks[i]);
    }
    choice[8] = cached ? 1 : 0;
  }
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_heff_apply(mpsh_heff plan, mpsh_dtype dt, const void* v, const void* envL, const void* envR,
                            const void* WL, const void* WR, void* out, void* workspace, void* stream) {
  if (!plan || !workspace) return fail(ERR_INVALID, "null argument");
  char* ws = (char*)workspace;
  void* prepared = ws + mpsh_heff_apply_workspace(plan, dt);
  mpsh_status st = mpsh_heff_prepare(plan, dt, envL, envR, WL, WR, prepared, stream);
  if (st != MPSH_OK) return st;
  return mpsh_heff_apply_prepared(plan, dt, prepared, v, out, workspace, stream);
}

void mpsh_heff_destroy(mpsh_heff plan) { delete plan; }

// ---------------------------------------------------------------------------
// environment updates
// ---------------------------------------------------------------------------
static mpsh_status envupdate_plan_c(int side, mpsh_struct env, mpsh_struct A, mpsh_struct W, mpsh_envupdate* plan,
                                    mpsh_struct* out) {
  API_TRY
  if (!env || !A || !W || !plan || !out) throw Error(ERR_INVALID, "null argument");
  auto h = std::make_unique<mpsh_envupdate_s>();
  h->p = envupdate_plan(side, env->s, A->s, W->s);
  *out = wrap(h->p.outs[0]);
  *plan = h.release();
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_envupdate_left_plan(mpsh_struct env, mpsh_struct A, mpsh_struct W, mpsh_envupdate* plan,
                                     mpsh_struct* out) {
  return envupdate_plan_c(0, env, A, W, plan, out);
}

mpsh_status mpsh_envupdate_right_plan(mpsh_struct env, mpsh_struct A, mpsh_struct W, mpsh_envupdate* plan,
                                      mpsh_struct* out) {
  return envupdate_plan_c(1, env, A, W, plan, out);
}

size_t mpsh_envupdate_workspace(mpsh_envupdate plan, mpsh_dtype dt) { return env_workspace(plan, dt); }

mpsh_status mpsh_envupdate_stats(mpsh_envupdate plan, int64_t* stats) {
  API_TRY
  env_stats(plan, stats);
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_envupdate_execute(mpsh_envupdate plan, mpsh_dtype dt, const void* env, const void* A,
                                   const void* W, void* out, void* workspace, void* stream) {
  API_TRY
  env_check_dtype(dt);
  if (!plan) throw Error(ERR_INVALID, "null plan");
  const EnvPlan& p = plan->p;
  if ((!env && p.ins[0].numel()) || (!A && p.ins[1].numel()) || (!W && p.ins[2].numel()) ||
      (!out && p.outs[0].numel()) || !workspace)
    throw Error(ERR_INVALID, "null argument");
  need_aligned(workspace, "workspace");
  const EnvDev& d = plan->get();
  hipStream_t s = (hipStream_t)stream;
  const int es = elem_size(dt);
  char* ws = (char*)workspace;
  const void* ket = A;
  const void* WL = W;
  if (p.has_perm[0][0]) {
    check(launch_copy(d.perm[0][0], es, A, ws + p.perm_off[0][0] * es, s), "env: mirror A");
    ket = ws + p.perm_off[0][0] * es;
  }
  if (p.has_perm[0][1]) {
    check(launch_copy(d.perm[0][1], es, W, ws + p.perm_off[0][1] * es, s), "env: mirror W");
    WL = ws + p.perm_off[0][1] * es;
  }
  const void* WR = dt == MPSH_F64 ? (const void*)d.wr_f64 : (const void*)d.wr_f32;
  run_env_core(*plan, d, 0, dt, env, ket, WL, WR, out, ws, s);
  return MPSH_OK;
  API_CATCH
}

void mpsh_envupdate_destroy(mpsh_envupdate plan) { delete plan; }

// ---------------------------------------------------------------------------
// H_eff adjoint
// ---------------------------------------------------------------------------
mpsh_status mpsh_heff_adjoint_plan(mpsh_struct v, mpsh_struct envL, mpsh_struct envR, mpsh_struct WL,
                                   mpsh_struct WR, mpsh_heff_adjoint* plan, mpsh_struct* gL, mpsh_struct* gR) {
  API_TRY
  if (!v || !envL || !envR || !WL || !WR || !plan || !gL || !gR) throw Error(ERR_INVALID, "null argument");
  auto h = std::make_unique<mpsh_heff_adjoint_s>();
  h->p = heff_adjoint_plan(v->s, envL->s, envR->s, WL->s, WR->s);
  *gL = wrap(h->p.outs[0]);
  *gR = wrap(h->p.outs[1]);
  *plan = h.release();
  return MPSH_OK;
  API_CATCH
}

size_t mpsh_heff_adjoint_workspace(mpsh_heff_adjoint plan, mpsh_dtype dt) { return env_workspace(plan, dt); }

mpsh_status mpsh_heff_adjoint_stats(mpsh_heff_adjoint plan, int64_t* stats) {
  API_TRY
  env_stats(plan, stats);
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_heff_adjoint_execute(mpsh_heff_adjoint plan, mpsh_dtype dt, const void* v, const void* envL,
                                      const void* envR, const void* WL, const void* WR, void* gL, void* gR,
                                      void* workspace, void* stream) {
  API_TRY
  env_check_dtype(dt);
  if (!plan) throw Error(ERR_INVALID, "null plan");
  const EnvPlan& p = plan->p;
  if ((!v && p.ins[0].numel()) || (!envL && p.ins[1].numel()) || (!envR && p.ins[2].numel()) ||
      (!WL && p.ins[3].numel()) || (!WR && p.ins[4].numel()) || (!gL && p.outs[0].numel()) ||
      (!gR && p.outs[1].numel()) || !workspace)
    throw Error(ERR_INVALID, "null argument");
  need_aligned(workspace, "workspace");
  const EnvDev& d = plan->get();
  hipStream_t s = (hipStream_t)stream;
  const int es = elem_size(dt);
  char* ws = (char*)workspace;
  // dE/denvR: the two-site core on the operands as given
  run_env_core(*plan, d, 0, dt, envL, v, WL, WR, gR, ws, s);
  // dE/denvL: the two-site core of the mirror image
  const void* src[3] = {v, WR, WL};
  void* mir[3];
  for (int k = 0; k < 3; ++k) {
    mir[k] = ws + p.perm_off[1][k] * es;
    check(launch_copy(d.perm[1][k], es, src[k], mir[k], s), "adjoint: mirror operand");
  }
  run_env_core(*plan, d, 1, dt, envR, mir[0], mir[1], mir[2], gL, ws, s);
  return MPSH_OK;
  API_CATCH
}

void mpsh_heff_adjoint_destroy(mpsh_heff_adjoint plan) { delete plan; }

// ---------------------------------------------------------------------------
// vector ops
// ---------------------------------------------------------------------------
mpsh_status mpsh_dot(mpsh_dtype dt, int64_t n, const void* x, const void* y, double* result, double* partials,
                     void* stream) {
  API_TRY
  check_dtype(dt);
  if (n < 0 || !result || !partials || (n && (!x || !y))) throw Error(ERR_INVALID, "bad dot arguments");
  check(launch_dot(elem_size(dt), n, x, y, result, partials, MPSH_DOT_PARTIALS, (hipStream_t)stream), "dot");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_axpby(mpsh_dtype dt, int64_t n, double a, const void* x, double b, void* y, void* stream) {
  API_TRY
  check_dtype(dt);
  if (n < 0 || (n && (!x || !y))) throw Error(ERR_INVALID, "bad axpby arguments");
  check(launch_axpby(elem_size(dt), n, a, x, b, y, (hipStream_t)stream), "axpby");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_scal(mpsh_dtype dt, int64_t n, double a, void* x, void* stream) {
  API_TRY
  check_dtype(dt);
  if (n < 0 || (n && !x)) throw Error(ERR_INVALID, "bad scal arguments");
  check(launch_scal(elem_size(dt), n, a, x, (hipStream_t)stream), "scal");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_lanczos_dots(int32_t k, int64_t n, const double* V, int64_t ldv, const double* w, double* c,
                              double* partials, void* stream) {
  API_TRY
  if (k < 0 || n < 0 || k > 4096) throw Error(ERR_INVALID, "bad Lanczos dimensions");
  if (k > 0 && n > 0 && (!V || !w || !c || !partials)) throw Error(ERR_INVALID, "null argument");
  if (k > 1 && ldv < n) throw Error(ERR_INVALID, "ldv < n");
  check(launch_lz_dots(k, n, V, ldv, w, c, partials, (hipStream_t)stream), "lanczos dots");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_lanczos_update(int32_t k, int64_t n, const double* V, int64_t ldv, const double* c, double sign,
                                double* w, void* stream) {
  API_TRY
  if (k < 0 || n < 0) throw Error(ERR_INVALID, "bad Lanczos dimensions");
  if (k > 0 && n > 0 && (!V || !w || !c)) throw Error(ERR_INVALID, "null argument");
  if (k > 1 && ldv < n) throw Error(ERR_INVALID, "ldv < n");
  check(launch_lz_update(k, n, V, ldv, c, sign, w, (hipStream_t)stream), "lanczos update");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_lanczos_normalize(int64_t n, const double* w, double* out, double* beta, double* partials,
                                   void* stream) {
  API_TRY
  if (n < 0) throw Error(ERR_INVALID, "bad Lanczos dimensions");
  if (!w || !out || !beta || !partials) throw Error(ERR_INVALID, "null argument");
  check(launch_lz_normalize(n, w, out, beta, partials, (hipStream_t)stream), "lanczos normalize");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_convert(mpsh_dtype src_dt, const void* src, mpsh_dtype dst_dt, void* dst, int64_t n,
                         void* stream) {
  API_TRY
  check_dtype(src_dt);
  check_dtype(dst_dt);
  if (n < 0 || (n && (!src || !dst))) throw Error(ERR_INVALID, "bad convert arguments");
  check(launch_convert(elem_size(src_dt), src, elem_size(dst_dt), dst, n, (hipStream_t)stream), "convert");
  return MPSH_OK;
  API_CATCH
}

mpsh_status mpsh_gather_blocks(int32_t nblk, int64_t ntiles, const int64_t* table, const double* src,
                               const double* scale, double* dst, void* stream) {
  API_TRY
  if (nblk < 0 || ntiles < 0 || ntiles > INT32_MAX || (ntiles && (!table || !src || !dst)))
    throw Error(ERR_INVALID, "bad gather arguments");
  if (((uintptr_t)table) & 7) throw Error(ERR_INVALID, "table must be 8-byte aligned");
  check(launch_gather_blocks(nblk, ntiles, table, src, scale, dst, (hipStream_t)stream), "gather blocks");
  return MPSH_OK;
  API_CATCH
}

}  // extern "C"
