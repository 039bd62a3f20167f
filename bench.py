#!/usr/bin/env python3
"""Benchmark: two-site H_eff matvecs/s at chi=2048 (6x24 Heisenberg cylinder bond, D=20).

One step = one H_eff application v -> H_eff v with all inputs resident in
HBM (fused path: step-1 grouped GEMM, combined-MPO pass, step-4 grouped
GEMM).  Prints ONE JSON line.  ``--gpus N`` with N > 1 is not supported:
this engine runs on one GPU.

Roofline: the dominant kernel is the grouped MFMA GEMM (two launches per
matvec: step 1 and step 4); its algorithmic FLOPs per launch are the
reference-order 2*M*N*K of those two contracts, its duration comes from HIP
events recorded on the launch stream around each stage.
"""
from __future__ import annotations

import argparse
import contextlib
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK = {"f32": 157.3e12, "f64": 78.6e12, "bf16": 2.5e15}
HBM_PEAK = 8.0e12
METRIC = "H_eff matvecs/sec + DMRG sweep time at χ=2048; |ΔE| vs FP64 CPU oracle"
# HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a plain
# copy's readings: x2.000 / x1.000) for the autotuned variants of this workload;
# step 4 = its GEMM plus the split-K reduction
TRAFFIC = {
    ("cyl6x24_chi2048", "f32", "step1"): 211.8e6,   # gemm_dma<1, 64, 128, 4>
    ("cyl6x24_chi2048", "f32", "step4"): 187.7e6,   # gemm_dma<1, 64, 64, 4> 156.6 + splitk_reduce 31.1
    ("cyl6x24_chi2048", "f32", "mpo"): 201.5e6,     # mpo_apply<float, 32, 0, float, float, unsigned int, true>
    ("cyl6x24_chi2048", "bf16", "step1"): 103.5e6,  # gemm_dma<3, 64, 128, 4>
    ("cyl6x24_chi2048", "bf16", "step4"): 97.7e6,   # gemm_dma<4, 64, 128, 4> 80.5 + splitk_reduce 17.2
    ("cyl6x24_chi2048", "bf16", "mpo"): 110.6e6,    # mpo_apply<float, 32, 0, __bf16, __bf16, unsigned int, true>
    ("cyl6x24_chi2048", "f64", "step1"): 434.9e6,   # gemm_dma<0, 64, 128, 3>
    ("cyl6x24_chi2048", "f64", "step4"): 390.2e6,   # gemm_dma<0, 64, 64, 3> 325.4 + splitk_reduce 64.8
    ("cyl6x24_chi2048", "f64", "mpo"): 387.0e6,     # mpo_apply<double, 32, 0, double, double, unsigned int, true>
}



def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cyl6x24_chi2048")
    ap.add_argument("--prec", default="f32", choices=["f32", "f64", "bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra-modes", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work for the CPU baseline")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the GEMM autotuner (variants from MPSH_HEFF_TILES or the defaults); for profiling")
    ap.add_argument("--lanczos-graphs", type=int, default=0, choices=[0, 1],
                    help="capture each Lanczos iteration of the DMRG window in a hipGraph")
    ap.add_argument("--dmrg-bonds", type=int, default=0,
                    help="timed two-site updates at chi=2048 in the bulk of the 6x24 cylinder (0: skip)")
    ap.add_argument("--dmrg-sweep", type=int, default=1, choices=[0, 1],
                    help="time one full two-site DMRG sweep at chi=2048 on the 6x24 cylinder")
    ap.add_argument("--precision", default="none", choices=["none", "chain", "tv", "all"],
                    help="precision studies: XXZ L=100 chi=512 four sweeps (chain), t-V L=128 chi=4096 "
                         "middle-bond eigensolve (tv)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (FP32) for HBM-counter calibration")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE per dispatch
    of this run give the counter scale (expected: mib * 1024 KiB each way)."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "bytes_each_way": n * 4}), flush=True)


@contextlib.contextmanager
def gc_paused():
    """Python's cyclic collector collected just before a timed region and paused
    inside it (as timeit does): a full collection in a process holding torch takes
    tens of ms of host time and would otherwise land in whichever short timed loop
    it happens to hit."""
    gc.collect()
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def time_fused(plan, prec, ts, steps, warmup, stage_events=True, tune=True):
    """Prepared-operator matvecs (GEMM variants autotuned first): returns
    (ms_per_step, per-stage ms, out, prepare_ms, tuned variants).  ms_per_step
    times ``steps`` plain matvecs (one mpsh_heff_apply_prepared call each,
    nothing else on the stream); with ``stage_events`` a separate run of as
    many matvecs, launched stage by stage with HIP events between the stages,
    gives each stage's average duration (the roofline's per-launch times)."""
    import torch
    v, L, R, WL, WR = ts
    out = torch.empty(plan.out.numel, dtype=v.dtype, device=v.device)
    ws = plan.workspace(prec, v.device, "apply")
    prepared = plan.prepare(prec, L.data, R.data, WL.data, WR.data)
    tuned = plan.tune(prec, prepared, v.data, out, ws, reps=5) if tune else None
    torch.cuda.synchronize()
    p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    p0.record()
    plan.prepare(prec, L.data, R.data, WL.data, WR.data, prepared)
    p1.record()
    for _ in range(warmup):
        plan.apply_prepared(prec, prepared, v.data, out, ws)
    torch.cuda.synchronize()
    prep_ms = p0.elapsed_time(p1)
    stages = [0.0, 0.0, 0.0, 0.0]
    if stage_events:
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(steps)]
        with gc_paused():
            torch.cuda.synchronize()
            for e in evs:
                e[0].record()
                plan.apply_stages(prec, 1, prepared, v.data, out, ws)
                e[1].record()
                plan.apply_stages(prec, 2, prepared, v.data, out, ws)
                e[2].record()
                plan.apply_stages(prec, 4, prepared, v.data, out, ws)
                e[3].record()
                plan.apply_stages(prec, 8, prepared, v.data, out, ws)
                e[4].record()
            torch.cuda.synchronize()
        for e in evs:
            for k in range(4):
                stages[k] += e[k].elapsed_time(e[k + 1])
        stages = [s / steps for s in stages]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    with gc_paused():
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        t_start.record()
        for _ in range(steps):
            plan.apply_prepared(prec, prepared, v.data, out, ws)
        t_end.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_time(t_end)
    return max(total_ms, wall * 1e3) / steps, stages, out, prep_ms, tuned


def gemm_kernel(prec, tuned, step):
    """Name of the step-1 / step-4 GEMM kernel variant: the autotuned one, else the
    one MPSH_HEFF_TILES or the defaults select (DMA-streamed 64x128 / 64x64 in every
    precision; BF16 runs the BF16-storage kernels, mode 3 writing BF16 T1, mode 4
    the FP32 output); DMA kernels carry their LDS ring depth (FP64 3, else 4)."""
    key = {"f32": "s1", "bf16": "s1h", "f64": "s1d"}[prec]
    cfg = "64x128x0"
    if step == "step4":
        key = key.replace("s1", "s4")
        cfg = "64x64x0"
    if tuned:
        cfg = tuned[step]
    else:
        for item in os.environ.get("MPSH_HEFF_TILES", "").split(","):
            if item.startswith(key + "="):
                cfg = item.split("=", 1)[1]
    parts = (cfg.split("x") + ["1"])[:3]
    bm, bn, ks = parts
    mode = {"f64": 0, "f32": 1, "bf16": 3 if step == "step1" else 4}[prec]
    return (f"gemm_dma<{mode}, {bm}, {bn}, {3 if mode == 0 else 4}>" if ks == "0"
            else f"gemm_grouped<{mode}, {bm}, {bn}, {ks}>")


def env_update_rates(ts64, steps):
    """_mpsupdateleft / _mpsupdateright (src/apply.jl:49-58, :17-25) on the workload's
    bond: site tensor A = (Vl, Vd, Vr†) with seeded Gaussian blocks, the workload's
    envL / envR and MPO tensors; FP64 and FP32, ms per update (inputs in HBM)."""
    import torch
    import mps_amd as M
    from mps_amd.tensornetalgs import envupdate_plan
    v, L, R, WL, WR = ts64
    A = M.randn((v.space[0], v.space[1], v.space[3]), 0, seed=1911, stream=21, device=v.device)
    out = {"path": "fused plan (mpsh_envupdate_*)", "note": "GFLOP/s at the reference chain's FLOPs"}
    for dt, name in ((torch.float64, "f64"), (torch.float32, "f32")):
        a, l, r, wl, wr = (t.to(dt) for t in (A, L, R, WL, WR))
        for side, fn, args in (("left", M._mpsupdateleft, (l, a, wl)), ("right", M._mpsupdateright, (r, a, wr))):
            for _ in range(2):
                fn(*args)
            with gc_paused():
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    fn(*args)
                torch.cuda.synchronize()
                ms = 1e3 * (time.perf_counter() - t0) / steps
            fl = envupdate_plan(side, *args).ref_flops
            out[f"{side}_{name}_ms"] = round(ms, 4)
            out[f"{side}_{name}_gflops"] = round(fl / (ms * 1e-3) / 1e9, 1)
            out["ref_gflop_per_update"] = round(fl / 1e9, 4)
            if name == "f64":
                for _ in range(1):
                    fn(*args, path="plain")
                with gc_paused():
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    for _ in range(max(2, steps // 4)):
                        fn(*args, path="plain")
                    torch.cuda.synchronize()
                    out[f"{side}_plain_f64_ms"] = round(1e3 * (time.perf_counter() - t0) / max(2, steps // 4), 4)
    return out


def refbench(dev, steps):
    """The reference's own benchmark shape (benchmarkings.ipynb:48-73: chi=1000, 8
    sectors per bond, D=15): fused FP64 / FP32 and plain FP64 H_eff on this GPU,
    beside the one published number for it (3.152 s per call, FP64, CPU not stated)."""
    import torch
    from mps_amd.tensornetalgs import heff_plan, heff_workload
    ts64 = heff_workload("refbench_chi1000", dtype=torch.float64, device=dev)
    out = {"published_reference_s_per_call": 3.152}
    for prec in ("f64", "f32"):
        ts = ts64 if prec == "f64" else [t.to(torch.float32) for t in ts64]
        pl = heff_plan(*ts)
        ms, _, _, _, _ = time_fused(pl, prec, ts, steps, 5)
        out[f"fused_{prec}_ms"] = round(ms, 5)
        out["gflops_per_call"] = round(pl.ref_flops / 1e9, 4)
    out["plain_f64_ms"] = round(time_plain(ts64, max(3, steps // 4), 2), 4)
    out["fused_f64_speedup_vs_published"] = round(3152.0 / out["fused_f64_ms"], 1)
    return out


def largest_chi(dev, steps):
    """The largest-chi workload: one two-site bond of the spinless t-V chain L=128
    at chi=4096 (~70 sectors of dims <= 400, Jordan-Wigner MPO D=6; seeded
    Gaussian blocks): fused H_eff rates with BF16 storage (FP32 accumulation),
    FP32 and FP64, the plain FP64 path, and each reduced precision's Rayleigh
    quotient <v|Hv>/<v|v> against the FP64 one."""
    import torch
    from mps_amd.tensornetalgs import heff_plan, heff_workload
    ts64 = heff_workload("tv_L128_chi4096", dtype=torch.float64, device=dev)
    v64 = ts64[0].data
    vv = float(torch.dot(v64, v64))
    out = {"workload": "tv_L128_chi4096", "chi": 4096, "mpo_D": 6}
    e = {}
    for prec in ("bf16", "f32", "f64"):
        ts = ts64 if prec == "f64" else [t.to(torch.float32) for t in ts64]
        pl = heff_plan(*ts)
        ms, _, o, _, _ = time_fused(pl, prec, ts, steps, 5, stage_events=False)
        out["gflop_per_matvec"] = round(pl.ref_flops / 1e9, 4)
        out[f"fused_{prec}_matvecs_per_s"] = round(1e3 / ms, 3)
        e[prec] = float(torch.dot(v64, o.double())) / vv
    for prec in ("bf16", "f32"):
        out[f"{prec}_rayleigh_rel_err_vs_f64"] = abs(e[prec] - e["f64"]) / abs(e["f64"])
    ms_plain = time_plain(ts64, max(3, steps // 4), 2)
    out["plain_f64_matvecs_per_s"] = round(1e3 / ms_plain, 3)
    out["bf16_speedup_vs_plain_f64"] = round(out["fused_bf16_matvecs_per_s"] * ms_plain / 1e3, 3)
    out["f32_speedup_vs_plain_f64"] = round(out["fused_f32_matvecs_per_s"] * ms_plain / 1e3, 3)
    return out


def unpadded_t1_t3(ts):
    """Elements of T1 plus T3 of the H_eff on these operands without the plan's
    ket_r row padding (a host-side plan made with MPSH_HEFF_PADR=0)."""
    from mps_amd.tensornetalgs import HeffPlan
    old = os.environ.get("MPSH_HEFF_PADR")
    os.environ["MPSH_HEFF_PADR"] = "0"
    try:
        p = HeffPlan(*[t.struct for t in ts])
    finally:
        if old is None:
            os.environ.pop("MPSH_HEFF_PADR", None)
        else:
            os.environ["MPSH_HEFF_PADR"] = old
    return p.t1_elems + p.t3_elems


def time_plain(ts, steps, warmup):
    import torch
    import mps_amd as M
    for _ in range(warmup):
        M._applymps2site(*ts, path="plain")
    with gc_paused():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            M._applymps2site(*ts, path="plain")
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps


DMRG_CFG = dict(W=6, Lx=24, chi=2048, nsec=60, dmax=200, seed=1911, maxdim=2048, tol=1e-10, lanczostol=1e-6,
                krylovdim=8, krylovmaxiter=2, prec="f32", refine_tol=1e-9)


def dmrg_window(nbonds, dev):
    """Two-site DMRG updates at chi=2048 on the 6x24 Heisenberg cylinder (snake order, D=20):
    a random U(1) MPS (N = L/2) canonical at the middle site, initial environments, one
    warm-up update, then ``nbonds`` timed left-to-right updates (eigsolve with FP32
    matvecs + FP64 refinement, batched FP64 SVD truncation, energy, env update)."""
    import torch
    from mps_amd import mpstools as mt
    from mps_amd import tensornetalgs as ta
    from mps_amd.symtensors import contract
    c = DMRG_CFG
    L, N = c["W"] * c["Lx"], c["W"] * c["Lx"] // 2
    t0 = time.perf_counter()
    mpo = mt.heisenberg_cylinder_mpo(c["W"], c["Lx"], device=dev)
    mps = mt.random_mps(L, N, c["chi"], seed=c["seed"], nsec=c["nsec"], dmax=c["dmax"], device=dev)
    print(f"dmrg: random chi={c['chi']} MPS ready ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    mt.center_at(mps, L // 2)
    env = ta.initialenv(mps, mpo, at=L // 2)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    print(f"dmrg: center {L // 2}, environments ready ({setup_s:.1f} s)", file=sys.stderr, flush=True)
    kw = dict(maxdim=c["maxdim"], tol=c["tol"], lanczostol=c["lanczostol"], krylovdim=c["krylovdim"],
              krylovmaxiter=c["krylovmaxiter"], prec=c["prec"], refine_tol=c["refine_tol"])
    ta.dmrg2site_window(mps, mpo, env, 1, **kw)
    timing = []
    torch.cuda.synchronize()
    tc0 = dict(ta.TUNE_COUNTS)
    t1 = time.perf_counter()
    energies = ta.dmrg2site_window(mps, mpo, env, nbonds, timing=timing, **kw)
    torch.cuda.synchronize()
    print(f"dmrg: {nbonds} timed bond updates done", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t1
    ms_bond = 1e3 * wall / nbonds
    stages = {k: round(sum(t[k] for t in timing) / nbonds, 3)
              for k in ("ms_theta", "ms_eigsolve", "ms_svdtrunc", "ms_energy", "ms_envupdate")}
    cen = mps.center
    AA = contract(mps.As[cen - 1], (1, 2, -1), mps.As[cen], (-1, 3, 4))
    state = (AA, env[cen - 1], env[cen + 2], mpo.Ws[cen - 1], mpo.Ws[cen])
    out = {"model": "Heisenberg 6x24 cylinder (snake), L=144, N=72, MPO D=20", "chi": c["chi"],
           "bonds_timed": nbonds, "first_bond": L // 2 + 1, "ms_per_bond_update": round(ms_bond, 3),
           "stage_ms_per_bond": stages,
           "matvecs_per_bond": round(sum(t["matvecs"] for t in timing) / nbonds, 2),
           "chi_bulk": [timing[0]["chi_left"], timing[0]["chi_right"]],
           "sweep_s_extrapolated": round(2 * (L - 1) * ms_bond / 1e3, 2),
           "sweep_note": "2(L-1) bulk two-site updates x measured ms per bulk update (edge bonds are cheaper)",
           "setup_s": round(setup_s, 2), "energies": [round(e, 10) for e in energies],
           "autotune_in_window": {k: ta.TUNE_COUNTS[k] - tc0[k] for k in tc0},
           "settings": {k: c[k] for k in ("maxdim", "tol", "lanczostol", "krylovdim", "krylovmaxiter", "prec",
                                           "refine_tol")}}
    return out, state


PREC_CHAIN = dict(L=100, N=50, chi=512, nsec=30, dmax=60, delta=1.0, sweeps=4, seed=1911, maxdim=512, tol=1e-10,
                  lanczostol=1e-8, krylovdim=8, krylovmaxiter=2, refine_tol=1e-10)


def precision_chain(dev):
    """XXZ chain L=100, Delta=1, U(1) Sz, D=5, chi=512 (~30 sectors, dims <= 60): four
    two-site DMRG sweeps from one seeded random MPS with FP64 matvecs, FP32-MFMA
    matvecs, and FP32 matvecs + FP64 refinement (Lanczos orthogonalisation,
    tridiagonal and SVD are FP64 in all runs).  Per sweep: the final energy
    (FP64 <psi|H|psi> of the truncated state) and the truncation error (sum and
    max of the discarded weight 1 - sum kept s^2 over the sweep's bonds)."""
    import torch
    from mps_amd import mpstools as mt
    from mps_amd import tensornetalgs as ta
    c = PREC_CHAIN
    kw = {k: c[k] for k in ("maxdim", "tol", "lanczostol", "krylovdim", "krylovmaxiter")}
    runs = {}
    for name, prec, refine in (("f64", "f64", None), ("f32", "f32", None), ("f32_refined", "f32", c["refine_tol"])):
        t0 = time.perf_counter()
        mpo = mt.xxz_u1mpo(c["L"], c["delta"], device=dev)
        mps = mt.random_mps(c["L"], c["N"], c["chi"], seed=c["seed"], nsec=c["nsec"], dmax=c["dmax"], device=dev)
        env = ta.initialenv(mps, mpo)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        sweeps = []
        for k in range(c["sweeps"]):
            log = []
            t1 = time.perf_counter()
            es = ta.dmrg2sitesweep(mps, mpo, env, prec=prec, refine_tol=refine, log=log, **kw)
            torch.cuda.synchronize()
            dw = [x["discarded_weight"] for x in log if "discarded_weight" in x]
            sweeps.append({"energy": es[-1], "sweep_s": round(time.perf_counter() - t1, 3),
                           "trunc_err_sum": sum(dw), "trunc_err_max": max(dw),
                           "matvecs": sum(x["matvecs"] for x in log if "matvecs" in x)})
            print(f"precision chain {name}: sweep {k + 1} E={es[-1]:.12f} ({sweeps[-1]['sweep_s']} s)",
                  file=sys.stderr, flush=True)
        runs[name] = {"setup_s": round(setup, 2), "sweeps": sweeps, "max_bond": max(mps.bonddims())}
    e64 = runs["f64"]["sweeps"][-1]["energy"]
    for name, r in runs.items():
        e = r["sweeps"][-1]["energy"]
        r["energy_rel_diff_vs_f64"] = abs(e - e64) / abs(e64)
        r["trunc_err_sum_diff_vs_f64"] = r["sweeps"][-1]["trunc_err_sum"] - runs["f64"]["sweeps"][-1]["trunc_err_sum"]
    return {"model": f"XXZ chain L={c['L']}, Delta={c['delta']}, N={c['N']}, MPO D=5", "chi": c["chi"],
            "settings": {k: c[k] for k in ("sweeps", "maxdim", "tol", "lanczostol", "krylovdim", "krylovmaxiter",
                                           "refine_tol", "nsec", "dmax", "seed")},
            "runs": runs}


PREC_TV = dict(L=128, N=64, chi=4096, nsec=70, dmax=400, t=1.0, V=2.0, seed=1911, lanczostol=1e-12, krylovdim=24,
               krylovmaxiter=12, refine_tol=1e-11, maxdim=4096, tol=1e-10)


def precision_tv(dev):
    """Spinless-fermion t-V chain L=128, t=1, V=2, half filling, Jordan-Wigner MPO
    D=6, chi=4096 (~70 sectors, dims <= 400): the two-site eigenproblem at the
    middle bond of a seeded random MPS (canonical there) solved with FP64, FP32
    and BF16-operand (FP32 accumulate) matvecs, each with and without FP64
    refinement cycles; the eigenvalue, the FP64 Rayleigh quotient of the
    returned vector, and the truncation error of its SVD to chi are compared
    with the FP64 solve."""
    import torch
    from mps_amd import mpstools as mt
    from mps_amd import tensornetalgs as ta
    from mps_amd.symtensors import contract, dot, symmatrix
    c = PREC_TV
    t0 = time.perf_counter()
    mpo = mt.tv_u1mpo(c["L"], c["t"], c["V"], device=dev)
    mps = mt.random_mps(c["L"], c["N"], c["chi"], seed=c["seed"], nsec=c["nsec"], dmax=c["dmax"], device=dev)
    print(f"precision t-V: random chi={c['chi']} MPS ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    cen = c["L"] // 2
    mt.center_at(mps, cen)
    env = ta.initialenv(mps, mpo, at=cen)
    AA = contract(mps.As[cen - 1], (1, 2, -1), mps.As[cen], (-1, 3, 4))
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    print(f"precision t-V: center {cen}, environments ready ({setup:.1f} s)", file=sys.stderr, flush=True)
    args = (env[cen - 1], env[cen + 2], mpo.Ws[cen - 1], mpo.Ws[cen])
    H64 = ta.HeffOperator(*args, "f64")
    runs = {}
    for name, prec, refine in (("f64", "f64", None), ("f32", "f32", None), ("f32_refined", "f32", c["refine_tol"]),
                               ("bf16", "bf16", None), ("bf16_refined", "bf16", c["refine_tol"])):
        H = H64 if prec == "f64" else ta.HeffOperator(*args, prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        theta, v, info = ta.eigsolve(H, AA, c["lanczostol"], c["krylovdim"], c["krylovmaxiter"], refine)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t1
        rq = dot(v, H64(v)) / dot(v, v)
        r = mt.svdtrunc(symmatrix(v, [1, 2], [3, 4]), maxdim=c["maxdim"], tol=c["tol"])
        s = r.singular_values()
        runs[name] = {"eigenvalue": theta, "rayleigh_f64": rq, "normres": info.normres, "matvecs": info.numops,
                      "solve_s": round(secs, 3), "trunc_err": max(0.0, 1.0 - float((s ** 2).sum()) / float(dot(v, v))),
                      "kept": int(r.Vl.dim())}
        print(f"precision t-V {name}: E={theta:.14f} rq={rq:.14f} ({secs:.2f} s, {info.numops} matvecs)",
              file=sys.stderr, flush=True)
    e64 = runs["f64"]["rayleigh_f64"]
    for r in runs.values():
        r["eigenvalue_rel_diff_vs_f64"] = abs(r["eigenvalue"] - runs["f64"]["eigenvalue"]) / abs(runs["f64"]["eigenvalue"])
        r["rayleigh_rel_diff_vs_f64"] = abs(r["rayleigh_f64"] - e64) / abs(e64)
        r["trunc_err_diff_vs_f64"] = r["trunc_err"] - runs["f64"]["trunc_err"]
    return {"model": f"spinless t-V chain L={c['L']}, t={c['t']}, V={c['V']}, N={c['N']}, Jordan-Wigner MPO D=6",
            "chi": c["chi"], "bond": [cen, cen + 1], "setup_s": round(setup, 2),
            "dim_two_site": AA.struct.numel, "settings": {k: c[k] for k in ("lanczostol", "krylovdim", "krylovmaxiter",
                                                                    "refine_tol", "maxdim", "tol", "nsec", "dmax")},
            "runs": runs}


def dmrg_sweep(dev, prec=None, refine_tol="default", **overrides):
    """One full two-site DMRG sweep (L->R->L, dmrg2sitesweep!) at chi=2048 on the
    6x24 Heisenberg cylinder from a seeded random U(1) MPS canonical at site 1:
    wall time of the sweep with its stage totals; the final energy is checked
    against an independent device <psi|H|psi> (measure) of the swept MPS.
    Returns (record, swept MPS, MPO)."""
    import torch
    from mps_amd import mpstools as mt
    from mps_amd import tensornetalgs as ta
    c = dict(DMRG_CFG)
    if prec is not None:
        c["prec"] = prec
    if refine_tol != "default":
        c["refine_tol"] = refine_tol
    c.update(overrides)
    L, N = c["W"] * c["Lx"], c["W"] * c["Lx"] // 2
    t0 = time.perf_counter()
    mpo = mt.heisenberg_cylinder_mpo(c["W"], c["Lx"], device=dev)
    mps = mt.random_mps(L, N, c["chi"], seed=c["seed"], nsec=c["nsec"], dmax=c["dmax"], device=dev)
    env = ta.initialenv(mps, mpo)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    env_bytes = sum(t.data.numel() * t.data.element_size() for t in env)
    mps_bytes = sum(t.data.numel() * t.data.element_size() for t in mps.As)
    torch.cuda.reset_peak_memory_stats()
    print(f"dmrg sweep ({c['prec']} matvecs): random chi={c['chi']} MPS and environments ready ({setup_s:.1f} s)",
          file=sys.stderr, flush=True)
    kw = dict(maxdim=c["maxdim"], tol=c["tol"], lanczostol=c["lanczostol"], krylovdim=c["krylovdim"],
              krylovmaxiter=c["krylovmaxiter"], prec=c["prec"], refine_tol=c["refine_tol"])
    stages, log = {}, []
    tc0 = dict(ta.TUNE_COUNTS)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    energies = ta.dmrg2sitesweep(mps, mpo, env, log=log, stage_ms=stages, **kw)
    torch.cuda.synchronize()
    sweep_s = time.perf_counter() - t1
    print(f"dmrg sweep ({c['prec']} matvecs): done ({sweep_s:.1f} s)", file=sys.stderr, flush=True)
    e_measure = mt.measure(mps, mpo)
    updates = sum(1 for x in log if "matvecs" in x)
    out = {"model": "Heisenberg 6x24 cylinder (snake), L=144, N=72, MPO D=20", "chi": c["chi"],
           "sweep_s": round(sweep_s, 3), "setup_s": round(setup_s, 2), "bond_updates": updates,
           "ms_per_update": round(1e3 * sweep_s / max(1, updates), 3),
           "stage_s": {k: round(v / 1e3, 3) for k, v in stages.items()},
           "matvecs": sum(x["matvecs"] for x in log if "matvecs" in x),
           "max_bond": max(mps.bonddims()), "energy_final": energies[-1], "energy_first": energies[0],
           "energy_measure": e_measure, "energy_rel_diff_vs_measure": abs(energies[-1] - e_measure) / abs(e_measure),
           "max_discarded_weight": max(x["discarded_weight"] for x in log if "discarded_weight" in x),
           "memory_gb": {"environments_initial": round(env_bytes / 1e9, 3), "mps_initial": round(mps_bytes / 1e9, 3),
                         "device_peak_allocated_during_sweep": round(torch.cuda.max_memory_allocated() / 1e9, 3)},
           "autotune_in_sweep": {k: ta.TUNE_COUNTS[k] - tc0[k] for k in tc0},
           "settings": {k: c[k] for k in ("maxdim", "tol", "lanczostol", "krylovdim", "krylovmaxiter", "prec",
                                           "refine_tol")}}
    return out, mps, mpo


def oracle_measure(mps, mpo):
    """<psi|H|psi> of a device MPS computed on the host by the oracle (numpy FP64
    restatement of measure, src/symmps.jl:354-371): every site tensor and MPO
    tensor copied to the host, one left sweep of oracle environment updates."""
    import numpy as np
    from oracle import symtensors as ost
    from oracle import tensornetalgs as ota

    def host(t):
        sp = tuple(ost.Space(list(V.items()), V.isdual) for V in t.space)
        return ost.SymTensor(t.charge, sp, {k: np.asarray(b) for k, b in t.blocks().items()})
    t0 = time.perf_counter()
    omps = ota.MPState([host(A) for A in mps.As], mps.center)
    ompo = ota.MPOperator([host(W) for W in mpo.Ws])
    e = ota.measure(omps, ompo)
    return e, time.perf_counter() - t0


def cpu_baseline(name, budget):
    """Oracle (numpy restatement of the reference algorithm) on the host cores."""
    import numpy as np
    from oracle import symtensors as ost
    from oracle import tensornetalgs as ota
    from mps_amd.workloads import gaussian, heff_spaces
    left, right, mpo = heff_spaces(name)
    Vl, Vr, Vw = ost.Space(left), ost.Space(right), ost.Space(mpo)
    Vd = ost.U1Space((0, 1), (1, 1))
    spaces = [(Vl, Vd, Vd, Vr.dual()), (Vl.dual(), Vw.dual(), Vl), (Vr, Vw, Vr.dual()),
              (Vw, Vd, Vw.dual(), Vd.dual()), (Vw, Vd, Vw.dual(), Vd.dual())]
    ts = []
    for stream, sp in enumerate(spaces, start=1):
        sects, sizes = ost.allsectorsandsizes(0, sp)
        n = sum(int(np.prod(z)) for z in sizes)
        flat = gaussian(1911, stream, n)
        blocks, p = {}, 0
        for s, z in zip(sects, sizes):
            k = int(np.prod(z))
            blocks[s] = flat[p:p + k].reshape(z, order="F")
            p += k
        ts.append(ost.SymTensor(0, sp, blocks))
    calls, t0 = 0, time.perf_counter()
    out = None
    while True:
        out = ota.applymps2site(*ts)
        calls += 1
        if time.perf_counter() - t0 > budget or calls >= 1000:
            break
    dt = time.perf_counter() - t0
    threads = 1
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"] or [1])
    except Exception:
        pass
    flat_out = np.concatenate([b.reshape(-1, order="F") for b in out.blocks.values()])
    flat_v = np.concatenate([b.reshape(-1, order="F") for b in ts[0].blocks.values()])
    return calls / dt, calls, dt, threads, flat_out, flat_v


def bond_heff_energy_check(state):
    """<AA|H_eff AA>/<AA|AA> of one two-site state under its bond's H_eff (the
    environments given): the device FP64 fused path vs the oracle's four-contract
    chain (numpy FP64) on the host.  A check of one H_eff, not of a sweep energy."""
    import numpy as np
    import torch
    import mps_amd as M
    from oracle import symtensors as ost
    from oracle import tensornetalgs as ota
    AA = state[0]
    Hd = M._applymps2site(*state, prec="f64")
    nn = M.dot(AA, AA)
    e_dev = M.dot(AA, Hd) / nn

    def host(t):
        sp = tuple(ost.Space(list(V.items()), V.isdual) for V in t.space)
        return ost.SymTensor(t.charge, sp, {k: np.asarray(b) for k, b in t.blocks().items()})
    ho = [host(t) for t in state]
    t0 = time.perf_counter()
    Ho = ota.applymps2site(*ho)
    e_or = ost.dot(ho[0], Ho) / ost.dot(ho[0], ho[0])
    return {"energy_device_f64": e_dev, "energy_cpu_oracle": e_or, "abs_diff": abs(e_dev - e_or),
            "rel_diff": abs(e_dev - e_or) / abs(e_or), "oracle_s": round(time.perf_counter() - t0, 2)}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding a bond over GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    import numpy as np
    import torch
    import mps_amd as M
    from mps_amd.tensornetalgs import heff_plan, heff_workload

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    prec = args.prec
    dt = torch.float64 if prec == "f64" else torch.float32
    ts64 = heff_workload(args.workload, dtype=torch.float64, device=dev)
    ts = [t.to(dt) for t in ts64] if dt != torch.float64 else ts64
    plan = heff_plan(*ts)
    ms, stages, out, prep_ms, tuned = time_fused(plan, prec, ts, args.steps, args.warmup, tune=not args.no_tune)
    print(f"bench: {args.workload} {prec} fused H_eff {1e3 / ms:.1f} matvecs/s", file=sys.stderr, flush=True)
    value = 1e3 / ms
    F = plan.ref_flops
    ebytes = {"f64": 8, "f32": 4, "bf16": 2}[prec]  # storage bytes (BF16 mode: the H_eff output is FP32)
    # the two grouped-GEMM launches of a matvec (steps 1 and 4 of the reference
    # order, src/apply.jl:99,102); algorithmic FLOPs = 2*M*N*K of those contracts,
    # durations from HIP events on the launch stream (step 4 includes its split-K
    # or stream-K reduction launch)
    launches = {}
    skey = {"f32": "split", "bf16": "splith", "f64": "splitd"}[prec]
    split = tuned.get("split", 2048) if tuned else 2048
    for item in ("" if tuned else os.environ.get("MPSH_HEFF_TILES", "")).split(","):
        if item.startswith(skey + "="):
            split = int(item.split("=", 1)[1])
    red = "tile_reduce" if split < 0 else "splitk_reduce"
    for key, si, stage, what in (("step1", 0, 1, "v^T envL"), ("step4", 3, 3, f"T3^T envR, with {red}")):
        fl = plan.step_flops[si]
        t = stages[stage]
        launches[key] = {"kernel": f"{gemm_kernel(prec, tuned, key)} ({what})", "flops_per_launch": fl,
                         "avg_launch_ms": round(t, 5),
                         "achieved": round(fl / (t * 1e-3) / 1e12, 3) if t > 0 else 0.0,
                         "frac": round(fl / (t * 1e-3) / PEAK[prec], 4) if t > 0 else 0.0,
                         "algorithmic_bytes_per_launch": plan.step_elems[si] * ebytes,
                         "traffic": TRAFFIC.get((args.workload, prec, key))}
    dom = max(launches, key=lambda k: launches[k]["avg_launch_ms"])
    mpo_bytes = unpadded_t1_t3(ts) * ebytes  # algorithmic: the logical T1 and T3, no row padding
    D = launches[dom]
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic: seeded Gaussian blocks (SplitMix64/Box-Muller), random env and MPO tensors",
        "config": {"workload": args.workload, "chi": sum(V for V in ts[0].space[0].dims),
                   "sectors_per_bond": len(ts[0].space[0].charges), "mpo_D": ts[1].space[1].dim(),
                   "path": "fused", "precision": {"f32": "FP32 MFMA", "f64": "FP64 MFMA",
                                                  "bf16": "BF16 storage and MFMA operands, FP32 accumulate"}[prec]},
        "gflops_per_matvec": round(F / 1e9, 4),
        "achieved_tflops": round(F / (ms * 1e-3) / 1e12, 3),
        "matvec_frac_of_peak": round(F / (ms * 1e-3) / PEAK[prec], 4),
        "per_kernel_ms": {"pack_v": round(stages[0], 5), "step1_gemm": round(stages[1], 5),
                          "mpo_pass": round(stages[2], 5), "step4_gemm": round(stages[3], 5)},
        "prepare_ms_per_bond": round(prep_ms, 5),
        "autotuned_gemm_variants": tuned,
        "roofline": {"bound": "mfma", "kernel": f"{D['kernel']} {dom} launch", "achieved": D["achieved"],
                     "peak": PEAK[prec] / 1e12, "unit": "TFLOP/s", "frac": D["frac"], "traffic": D["traffic"],
                     "flops_per_launch": D["flops_per_launch"], "avg_launch_ms": D["avg_launch_ms"],
                     "algorithmic_bytes_per_launch": D["algorithmic_bytes_per_launch"],
                     "dominant_of": "the longer of the two GEMM launches per matvec", "launches": launches},
        "mpo_pass_roofline": {"bound": "hbm", "achieved": round(mpo_bytes / (stages[2] * 1e-3) / 1e9, 1)
                              if stages[2] > 0 else 0.0, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                              "algorithmic_bytes": mpo_bytes, "avg_launch_ms": round(stages[2], 5),
                              "traffic": TRAFFIC.get((args.workload, prec, "mpo")),
                              "note": "T1 read once + T3 written once (the combined-MPO matrices stay in LDS)"},
    }
    # E = <v|H v>/<v|v> of this matvec in the benchmark precision, FP64 device path as reference
    v64 = ts64[0]
    Hv64 = M._applymps2site(*ts64, prec="f64")
    torch.cuda.synchronize()
    vv = float(torch.dot(v64.data, v64.data))
    e64 = float(torch.dot(v64.data, Hv64.data)) / vv
    e_prec = float(torch.dot(ts[0].data.double(), out.double())) / vv
    result["rayleigh_rel_err_vs_f64"] = abs(e_prec - e64) / abs(e64)
    if not args.no_extra_modes:
        extra = {}
        for p2 in ("f64", "f32", "bf16"):
            if p2 == prec:
                continue
            t2 = ts64 if p2 == "f64" else [t.to(torch.float32) for t in ts64]
            pl2 = heff_plan(*t2)
            ms2, st2, _, _, _ = time_fused(pl2, p2, t2, max(10, args.steps // 4), 5)
            extra[f"fused_{p2}_matvecs_per_s"] = round(1e3 / ms2, 3)
            extra[f"fused_{p2}_per_kernel_ms"] = {"pack_v": round(st2[0], 5), "step1_gemm": round(st2[1], 5),
                                                  "mpo_pass": round(st2[2], 5), "step4_gemm": round(st2[3], 5)}
        ms_plain = time_plain(ts64, max(3, args.steps // 20), 2)
        extra["plain_f64_matvecs_per_s"] = round(1e3 / ms_plain, 3)
        extra["speedup_vs_plain_f64"] = round(ms_plain / ms, 3)
        result["other_paths"] = extra
        result["env_update"] = env_update_rates(ts64, max(5, args.steps // 10))
        result["refbench_chi1000"] = refbench(dev, max(10, args.steps // 4))
        result["largest_chi"] = largest_chi(dev, max(10, args.steps // 4))
        print("bench: other paths done", file=sys.stderr, flush=True)
    dmrg_state = None
    if args.dmrg_bonds > 0:
        from mps_amd import tensornetalgs as ta
        ta.LANCZOS_GRAPHS = bool(args.lanczos_graphs)
        result["dmrg"], dmrg_state = dmrg_window(args.dmrg_bonds, dev)
        result["dmrg"]["lanczos_hipgraph"] = ta.LANCZOS_GRAPHS
    swept = None
    if args.dmrg_sweep:
        from mps_amd import tensornetalgs as ta
        ta.LANCZOS_GRAPHS = bool(args.lanczos_graphs)
        result["dmrg_sweep"], mps, mpo = dmrg_sweep(dev)
        swept = (mps, mpo)
        # the same sweep from the same start with FP64 matvecs throughout
        r64, mps64, _ = dmrg_sweep(dev, prec="f64", refine_tol=None, lanczostol=DMRG_CFG["refine_tol"],
                                   krylovmaxiter=DMRG_CFG["krylovmaxiter"] + 1)
        result["dmrg_sweep_f64"] = {k: r64[k] for k in ("sweep_s", "stage_s", "matvecs", "energy_final",
                                                        "energy_measure", "max_discarded_weight", "settings")}
        e32, e64s = result["dmrg_sweep"]["energy_final"], r64["energy_final"]
        result["dmrg_sweep_f64"]["energy_rel_diff_vs_refined_f32_sweep"] = abs(e32 - e64s) / abs(e64s)
        result["dmrg_sweep_f64"]["note"] = ("truncation to chi=2048 cuts through near-degenerate singular "
                                            "values, so a rounding-level change of the matvec can flip which "
                                            "state one sweep from a random start keeps")
        del mps64
    if args.precision in ("chain", "all"):
        result["precision_chain"] = precision_chain(dev)
    if args.precision in ("tv", "all"):
        result["precision_tv"] = precision_tv(dev)
    if not args.no_cpu_baseline:
        rate, calls, secs, threads, oflat, vflat = cpu_baseline(args.workload, args.cpu_budget)
        e_or = float(np.dot(vflat, oflat) / np.dot(vflat, vflat))
        result["cpu_baseline"] = {"value": round(rate, 5), "unit": "matvecs/s", "cores": threads,
                                  "kind": "port",
                                  "sample": f"{calls} H_eff call(s) of the full {args.workload} workload in "
                                            f"{secs:.1f} s (oracle: numpy restatement of the reference "
                                            f"fuse-GEMM-split-permute chain, FP64, OpenBLAS threads={threads}, "
                                            f"host nproc={os.cpu_count()})"}
        # one unrefined matvec on the random workload, Rayleigh quotient vs the oracle's
        result["matvec_rayleigh_rel_err_vs_cpu_oracle"] = abs(e_prec - e_or) / abs(e_or)
        result["f64_matvec_rayleigh_rel_err_vs_cpu_oracle"] = abs(e64 - e_or) / abs(e_or)
        if dmrg_state is not None:
            result["dmrg"]["edge_or_window_bond_heff_energy_vs_cpu_oracle"] = bond_heff_energy_check(dmrg_state)
        if swept is not None:
            # the swept chi=2048 MPS copied to the host; its energy computed there by the oracle
            e_or_sweep, secs_or = oracle_measure(*swept)
            e_sw = result["dmrg_sweep"]["energy_final"]
            result["dmrg_sweep"]["energy_cpu_oracle_measure"] = e_or_sweep
            result["dmrg_sweep"]["oracle_measure_s"] = round(secs_or, 2)
            result["dmrg_energy_rel_err_vs_cpu_oracle"] = abs(e_sw - e_or_sweep) / abs(e_or_sweep)
            print(f"bench: oracle measure of the swept MPS done ({secs_or:.1f} s)", file=sys.stderr, flush=True)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
