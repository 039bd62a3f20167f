// CPU check of the fused H_eff plan's T1 / T3 layout (heff_plan, step1_layout,
// mpo_layout in mps_amd/csrc/planner.cpp), built and run by
// tests/test_heff_layout_cpu.py.  For seeded random two-site bonds (uneven
// sector dims, several MPO bond sizes), padded (the default) and unpadded
// (MPSH_HEFF_PADR=0) plans:
//   * every T1 element the MPO pass reads (channel rows over db x dRp pixels) is
//     an element of some step-1 GEMM output (problem rows M include the padding);
//   * every T3 element the step-4 GEMM reads (A operand rows over each K
//     segment) is written by some MPO-pass output row (pixels 0 .. dRp), so the
//     padding the GEMM reads is always the MPO pass's zeros;
//   * padded plans: dRp = roundup(dR, 4), every channel row and T3 row has unit
//     pixel stride and base and row stride on multiples of 4 elements, so the
//     pass's vector runs are aligned; unpadded plans: dRp == dR;
//   * every MPO work list covers each group's runs exactly once.
// Prints one line per case and exits non-zero on the first violation.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../mps_amd/csrc/planner.hpp"

using namespace mpsh;

static int fail(const char* what, int c) {
  std::printf("FAIL case %d: %s\n", c, what);
  return 1;
}

// the step-1 output elements (T1): problem rows m < M, columns n < N, affine rows
static std::vector<char> t1_written(const HeffPlan& H, int c, bool& ok) {
  std::vector<char> w((size_t)H.t1_elems, 0);
  ok = true;
  for (const GemmProb& p : H.g1.probs) {
    for (int64_t m = 0; m < p.M; ++m)
      for (int64_t n = 0; n < p.N; ++n) {
        const int64_t e = p.crow0 >= 0 ? H.g1.crow[p.crow0 + m] + n * H.g1.cld[p.crow0 + m] : p.c_off + m + n * p.ldc;
        if (e < 0 || e >= H.t1_elems) {
          ok = false;
          fail("step-1 output outside T1", c);
          return w;
        }
        w[(size_t)e] = 1;
      }
  }
  return w;
}

static int check(const HeffPlan& H, bool padded, int c) {
  if (H.padr != (padded ? kMpoPadR : 1)) return fail("row alignment", c);
  bool ok = true;
  const std::vector<char> w1 = t1_written(H, c, ok);
  if (!ok) return 1;
  std::vector<char> w3((size_t)H.t3_elems, 0);
  auto aligned = [&](int64_t x) { return !padded || x % kMpoPadR == 0; };
  std::vector<std::vector<int64_t>> runs_exp(5, std::vector<int64_t>(H.groups.size(), 0));
  for (size_t gi = 0; gi < H.groups.size(); ++gi) {
    const MpoGroup& g = H.groups[gi];
    const MpoSig& sg = H.sigs[g.sig];
    if (padded ? (g.dRp % kMpoPadR != 0 || g.dRp < g.dR || g.dRp - g.dR >= kMpoPadR) : g.dRp != g.dR)
      return fail("group padded range", c);
    for (int ch = 0; ch < sg.C; ++ch) {
      const MpoInRow& q = H.inrows[g.in0 + ch];
      if (q.nstride != 1 || !aligned(q.base) || !aligned(q.mstride)) return fail("channel row alignment", c);
      for (int64_t m = 0; m < g.db; ++m)
        for (int64_t n = 0; n < g.dRp; ++n) {
          const int64_t e = q.base + m * q.mstride + n;
          if (e < 0 || e >= H.t1_elems || !w1[(size_t)e]) return fail("MPO pass reads T1 outside step 1's output", c);
        }
    }
    for (int r = 0; r < sg.R; ++r) {
      const MpoOutRow& o = H.outrows[g.out0 + r];
      if (o.nstride != 1 || !aligned(o.base) || !aligned(o.mstride)) return fail("T3 row alignment", c);
      for (int64_t m = 0; m < g.db; ++m)
        for (int64_t n = 0; n < g.dRp; ++n) {
          const int64_t e = o.base + m * o.mstride + n;
          if (e < 0 || e >= H.t3_elems) return fail("MPO pass writes outside T3", c);
          w3[(size_t)e] = 1;
        }
    }
    for (int i = 0; i < 5; ++i) {
      const int P = i < 3 ? kMpoRuns[i] : (padded ? mpo_vec_run(sg.C, i == 3 ? 4 : 8) : mpo_mixed_run(sg.C));
      runs_exp[i][gi] = i == 4 && !padded ? 0 : (int64_t)g.db * ((g.dRp + P - 1) / P);
    }
  }
  // each work list: a group's items start 256 runs apart, in order, and cover its runs
  for (int i = 0; i < 5; ++i) {
    std::vector<int64_t> runs(H.groups.size(), 0);
    for (const MpoWork& w : H.work[i]) {
      if (w.group < 0 || w.group >= (int)H.groups.size() || w.run0 != runs[w.group]) return fail("work list order", c);
      runs[w.group] += 256;
    }
    for (size_t gi = 0; gi < H.groups.size(); ++gi)
      if (runs_exp[i][gi] == 0 ? runs[gi] != 0 : (runs[gi] < runs_exp[i][gi] || runs[gi] - 256 >= runs_exp[i][gi]))
        return fail("work list coverage", c);
  }
  // every T3 element the step-4 GEMM reads is written by the MPO pass
  int64_t read = 0;
  for (const GemmProb& p : H.g4.probs)
    for (int s = 0; s < p.nseg; ++s) {
      const GemmSeg& sg = H.g4.segs[p.seg0 + s];
      for (int64_t m = 0; m < p.M; ++m) {
        const int64_t row = sg.atab0 >= 0 ? H.g4.atab[sg.atab0 + m] : sg.a_off + m * sg.lda;
        for (int64_t k = 0; k < sg.K; ++k) {
          const int64_t e = row + k;
          if (e < 0 || e >= H.t3_elems || !w3[(size_t)e]) return fail("step 4 reads a T3 element the MPO pass does not write", c);
          ++read;
        }
      }
    }
  std::printf("case %d: %s, %zu groups, T1 %lld, T3 %lld (step 4 reads %lld)\n", c, padded ? "padded" : "unpadded",
              H.groups.size(), (long long)H.t1_elems, (long long)H.t3_elems, (long long)read);
  return 0;
}

int main() {
  std::mt19937_64 rng(2024);
  int c = 0;
  const std::vector<std::vector<std::pair<int64_t, int64_t>>> mpos = {
      {{-1, 1}, {0, 3}, {1, 1}}, {{-1, 2}, {0, 2}, {1, 2}}, {{-1, 6}, {0, 8}, {1, 6}}};
  for (int rep = 0; rep < 12; ++rep) {
    const int nsec = 3 + (int)(rng() % 9);
    std::vector<std::pair<int64_t, int64_t>> left, right;
    for (int q = 0; q < nsec; ++q) {
      left.push_back({q, 1 + (int64_t)(rng() % 37)});
      right.push_back({q + 1, 1 + (int64_t)(rng() % 37)});
    }
    const Leg Vl = make_leg(false, left), Vr = make_leg(false, right), Vw = make_leg(false, mpos[rep % 3]);
    const Leg Vd = make_leg(false, {{0, 1}, {1, 1}});
    const Struct v = make_struct(0, {Vl, Vd, Vd, Vr.dual()});
    const Struct envL = make_struct(0, {Vl.dual(), Vw.dual(), Vl});
    const Struct envR = make_struct(0, {Vr, Vw, Vr.dual()});
    const Struct W = make_struct(0, {Vw, Vd, Vw.dual(), Vd.dual()});
    for (int pad = 1; pad >= 0; --pad) {
      if (pad)
        unsetenv("MPSH_HEFF_PADR");
      else
        setenv("MPSH_HEFF_PADR", "0", 1);
      const HeffPlan H = heff_plan(v, envL, envR, W, W);
      if (check(H, pad == 1, ++c)) return 1;
    }
  }
  unsetenv("MPSH_HEFF_PADR");
  std::printf("OK %d cases\n", c);
  return 0;
}
