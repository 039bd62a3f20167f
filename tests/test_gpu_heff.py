"""Device parity of the two-site H_eff (fused and plain paths) and the environment updates.

Small and mid-size workloads are compared with the oracle's reference-order
chain (FP64: 1e-12 relative, norm-wise per block).  The full-size workloads
(chi=1000, 2048) are checked through properties: fused == plain (two
different device algorithms), linearity, and FP32/BF16 vs FP64 agreement.
"""
import numpy as np
import pytest
import torch

from oracle import tensornetalgs as ota
from tests.helpers import global_rel_err, o_from_flat, o_from_product, o_space, rel_block_err

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import mps_amd as M  # noqa: E402
from mps_amd import tensornetalgs as ta  # noqa: E402
from mps_amd.tensornetalgs import heff_workload  # noqa: E402
from mps_amd.workloads import gaussian  # noqa: E402

DEV = torch.device("cuda", 0)


def oracle_of(P):
    return o_from_product(P)


def _workload(name, dtype=torch.float64, seed=1911):
    return heff_workload(name, seed=seed, dtype=dtype, device=DEV)


def _check_structure(P, O):
    assert P.charge == O.charge
    assert tuple(o_space(V) for V in P.space) == O.space
    assert P.sectors() == list(O.blocks.keys())


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L64_chi256", "xxz_L100_chi512"])
def test_heff_fused_and_plain_f64_vs_oracle(name):
    ts = _workload(name)
    O = [oracle_of(t) for t in ts]
    want = ota.applymps2site(*O)
    for path in ("fused", "plain"):
        got = M._applymps2site(*ts, path=path)
        _check_structure(got, want)
        err = rel_block_err(oracle_of(got), want)
        print(f"{name} {path}: max block rel err {err:.2e}", flush=True)
        assert err <= 1e-12


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512"])
def test_heff_fused_f32_bf16_vs_oracle(name):
    ts64 = _workload(name)
    want = ota.applymps2site(*[oracle_of(t) for t in ts64])
    ts32 = [t.to(torch.float32) for t in ts64]
    got32 = M._applymps2site(*ts32, prec="f32")
    _check_structure(got32, want)
    e32 = global_rel_err(oracle_of(got32), want)
    got16 = M._applymps2site(*ts32, prec="bf16")
    e16 = global_rel_err(oracle_of(got16), want)
    print(f"{name}: f32 rel err {e32:.2e}, bf16 rel err {e16:.2e}", flush=True)
    assert e32 < 1e-5
    assert e16 < 3e-2


@pytest.mark.parametrize("name", ["refbench_chi1000", "cyl6x24_chi2048"])
def test_heff_full_size_properties(name):
    ts = _workload(name)
    fused = M._applymps2site(*ts, path="fused")
    plain = M._applymps2site(*ts, path="plain")
    assert fused.sectors() == plain.sectors()
    d = (fused.data - plain.data).norm().item() / plain.data.norm().item()
    print(f"{name}: fused vs plain f64 rel diff {d:.2e}", flush=True)
    assert d < 1e-12
    # linearity: H(a x + b y) = a Hx + b Hy
    v, L, R, WL, WR = ts
    y = M.randn(v.space, v.charge, seed=7, stream=11, device=DEV)
    z = v.copy()
    M.axpby_(0.75, y, -1.25, z)  # z = 0.75 y - 1.25 v
    Hz = M._applymps2site(z, L, R, WL, WR)
    Hy = M._applymps2site(y, L, R, WL, WR)
    comb = Hy.copy()
    M.axpby_(-1.25, fused, 0.75, comb)
    lin = (Hz.data - comb.data).norm().item() / Hz.data.norm().item()
    assert lin < 1e-12
    # reduced precision against FP64
    t32 = [t.to(torch.float32) for t in ts]
    for prec, tol in (("f32", 1e-5), ("bf16", 3e-2)):
        got = M._applymps2site(*t32, prec=prec)
        e = (got.data.double() - fused.data).norm().item() / fused.data.norm().item()
        print(f"{name}: {prec} vs f64 rel err {e:.2e}", flush=True)
        assert e < tol


def test_heff_repeated_calls_are_deterministic():
    ts = _workload("xxz_L100_chi512")
    a = M._applymps2site(*ts)
    b = M._applymps2site(*ts)
    assert torch.equal(a.data, b.data)


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512"])
def test_env_updates_vs_oracle(name):
    v, L, R, WL, WR = _workload(name)
    # site tensors A(l, p, r̄) living on the left and on the right bond of v
    Vl, Vd, Vr = v.space[0], v.space[1], v.space[3].dual()
    Al = M.randn((Vl, Vd, Vl.dual()), 0, seed=3, stream=22, device=DEV)
    Ar = M.randn((Vr, Vd, Vr.dual()), 0, seed=3, stream=23, device=DEV)
    gotL = M._mpsupdateleft(L, Al, WL)
    wantL = ota.mpsupdateleft(oracle_of(L), oracle_of(Al), oracle_of(WL))
    _check_structure(gotL, wantL)
    assert rel_block_err(oracle_of(gotL), wantL) <= 1e-12
    gotR = M._mpsupdateright(R, Ar, WR)
    wantR = ota.mpsupdateright(oracle_of(R), oracle_of(Ar), oracle_of(WR))
    _check_structure(gotR, wantR)
    assert rel_block_err(oracle_of(gotR), wantR) <= 1e-12


@pytest.mark.parametrize("prec,bar", [("f64", 1e-12), ("f32", 1e-5), ("bf16", 3e-2)])
def test_autotuned_variants_match_default(prec, bar):
    """Every variant the autotuner may pick computes the same H_eff: the tuned plan's
    output equals the default variants' output (FP64 1e-12; reduced precisions
    against FP64), and a second plan with the same signature reuses the choice."""
    from mps_amd.tensornetalgs import HeffPlan
    ts64 = _workload("xxz_L100_chi512")
    ts = ts64 if prec == "f64" else [t.to(torch.float32) for t in ts64]
    plan = HeffPlan(*[t.struct for t in ts])
    prepared = plan.prepare(prec, *(t.data for t in ts[1:]))
    dt = ts[0].dtype
    ref = torch.empty(plan.out.numel, dtype=dt, device=DEV)
    plan.apply_prepared(prec, prepared, ts[0].data, ref)
    want = M._applymps2site(*ts64, prec="f64").data
    scratch = torch.empty_like(ref)
    choice = plan.tune(prec, prepared, ts[0].data, scratch, reps=2, force=True)
    assert not choice["from_cache"]
    got = torch.empty_like(ref)
    plan.apply_prepared(prec, prepared, ts[0].data, got)
    for x in (ref, got):
        assert float(torch.linalg.norm(x.double() - want)) <= bar * float(torch.linalg.norm(want))
    if prec == "f64":
        assert float(torch.linalg.norm(got - ref)) <= 1e-13 * float(torch.linalg.norm(ref))
    plan2 = HeffPlan(*[t.struct for t in ts])
    prep2 = plan2.prepare(prec, *(t.data for t in ts[1:]))
    assert plan2.tune(prec, prep2, ts[0].data, scratch)["from_cache"]


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512", "cyl6x24_chi2048"])
@pytest.mark.parametrize("tiles", [(64, 64), (128, 64), (64, 128)])
def test_dma_gemm_variants_bitwise(name, tiles):
    """The FP32 GEMM with DMA-streamed operands (tile variant KS = 0) sums the same
    products in the same order as the register-staged kernel with K chunk 16
    (KS = 1): step 1 and step 4 (K segments, K tails, split-K) are bitwise equal."""
    import os
    from mps_amd.tensornetalgs import HeffPlan
    ts = [t.to(torch.float32) for t in _workload(name)]
    bm, bn = tiles
    outs = []
    old = os.environ.get("MPSH_HEFF_TILES")
    try:
        for ks in (1, 0):
            os.environ["MPSH_HEFF_TILES"] = f"s1={bm}x{bn}x{ks},s4={bm}x{bn}x{ks},split=1024"
            plan = HeffPlan(*[t.struct for t in ts])
            prepared = plan.prepare("f32", *(t.data for t in ts[1:]))
            out = torch.full((plan.out.numel,), float("nan"), dtype=torch.float32, device=DEV)
            plan.apply_prepared("f32", prepared, ts[0].data, out)
            outs.append(out)
    finally:
        if old is None:
            os.environ.pop("MPSH_HEFF_TILES", None)
        else:
            os.environ["MPSH_HEFF_TILES"] = old
    assert torch.isfinite(outs[1]).all()
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512", "cyl6x24_chi2048"])
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
def test_mpo_pass_variants_bitwise(name, prec):
    """The MPO pass sums each output's channels in the same order whatever the
    run length (unpadded plans, MPSH_HEFF_PADR=0: the default mixed list, 4
    pixels per thread in groups of at most 8 channels, 1 otherwise; or 1, 2 or 4
    everywhere, FP64 at most 2) and offset width (32-bit byte offsets or 64-bit
    element offsets), so every variant's H_eff output is bitwise equal to the
    default's; padded plans (the default layout, vector-access runs) likewise
    over the offset width."""
    import os
    from mps_amd.tensornetalgs import HeffPlan
    dt = torch.float64 if prec == "f64" else torch.float32
    ts = [t.to(dt) for t in _workload(name)]
    keys = ("MPSH_MPO_RUN", "MPSH_MPO_OFF64", "MPSH_HEFF_PADR")
    old = {k: os.environ.get(k) for k in keys}
    outs = {}
    try:
        for padr, runs in (("0", (0, 1, 2, 4)), ("1", (0,))):
            for run in runs:
                for off64 in (0, 1):
                    os.environ["MPSH_HEFF_PADR"] = padr
                    os.environ["MPSH_MPO_RUN"] = str(run)
                    os.environ["MPSH_MPO_OFF64"] = str(off64)
                    plan = HeffPlan(*[t.struct for t in ts])  # options are read when the plan is made / uploaded
                    prepared = plan.prepare(prec, *(t.data for t in ts[1:]))
                    out = torch.full((plan.out.numel,), float("nan"), dtype=dt, device=DEV)
                    plan.apply_prepared(prec, prepared, ts[0].data, out)
                    torch.cuda.synchronize()
                    outs[(padr, run, off64)] = out
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for padr in ("0", "1"):
        ref = outs[(padr, 0, 0)]
        assert torch.isfinite(ref).all()
        for key, out in outs.items():
            if key[0] == padr:
                assert torch.equal(out, ref), key


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512", "cyl6x24_chi2048"])
@pytest.mark.parametrize("tiles", [(64, 64), (128, 64), (64, 128)])
def test_dma_gemm_f64_variants(name, tiles):
    """The FP64 GEMM with DMA-streamed operands (K chunk 8 doubles) against the
    register-staged FP64 kernel (K chunk 16): the same products summed in another
    order, so step 1 and step 4 (K segments, K tails, split-K) agree to FP64
    rounding (1e-13 relative)."""
    import os
    from mps_amd.tensornetalgs import HeffPlan
    ts = _workload(name)
    bm, bn = tiles
    outs = []
    old = os.environ.get("MPSH_HEFF_TILES")
    try:
        for ks in (1, 0):
            os.environ["MPSH_HEFF_TILES"] = f"s1d={bm}x{bn}x{ks},s4d={bm}x{bn}x{ks},split=1024"
            plan = HeffPlan(*[t.struct for t in ts])
            prepared = plan.prepare("f64", *(t.data for t in ts[1:]))
            out = torch.full((plan.out.numel,), float("nan"), dtype=torch.float64, device=DEV)
            plan.apply_prepared("f64", prepared, ts[0].data, out)
            outs.append(out)
    finally:
        if old is None:
            os.environ.pop("MPSH_HEFF_TILES", None)
        else:
            os.environ["MPSH_HEFF_TILES"] = old
    assert torch.isfinite(outs[1]).all()
    err = float(torch.linalg.vector_norm(outs[1] - outs[0]) / torch.linalg.vector_norm(outs[0]))
    assert err < 1e-13, err


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512", "cyl6x24_chi2048"])
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("tiles", [(64, 64), (64, 128)])
def test_streamk_step4_vs_unsplit(name, prec, tiles):
    """Stream-K step 4 (tiles cut between work groups into K pieces whose partial
    tiles tile_reduce sums in piece order) against the unsplit GEMM and split-K:
    the same products summed in another order, so the outputs agree to rounding
    (FP64 1e-13, FP32 and BF16 storage with FP32 accumulation 1e-5 relative)."""
    import os
    from mps_amd.tensornetalgs import HeffPlan
    dt = torch.float64 if prec == "f64" else torch.float32
    ts = [t.to(dt) for t in _workload(name)]
    key, skey = {"f64": ("s4d", "splitd"), "f32": ("s4", "split"), "bf16": ("s4h", "splith")}[prec]
    bm, bn = tiles
    outs = {}
    old = os.environ.get("MPSH_HEFF_TILES")
    try:
        for split in (0, 1024, -1):
            os.environ["MPSH_HEFF_TILES"] = f"{key}={bm}x{bn}x0,{skey}={split}"
            plan = HeffPlan(*[t.struct for t in ts])
            prepared = plan.prepare(prec, *(t.data for t in ts[1:]))
            out = torch.full((plan.out.numel,), float("nan"), dtype=dt, device=DEV)
            plan.apply_prepared(prec, prepared, ts[0].data, out)
            plan.apply_prepared(prec, prepared, ts[0].data, out)  # a second call: partial tiles rewritten
            torch.cuda.synchronize()
            outs[split] = out
    finally:
        if old is None:
            os.environ.pop("MPSH_HEFF_TILES", None)
        else:
            os.environ["MPSH_HEFF_TILES"] = old
    tol = 1e-13 if prec == "f64" else 1e-5
    ref = outs[0]
    assert torch.isfinite(ref).all() and torch.isfinite(outs[-1]).all()
    for split in (1024, -1):
        err = float(torch.linalg.vector_norm(outs[split] - ref) / torch.linalg.vector_norm(ref))
        assert err < tol, (split, err)


def _site_tensor(v, seed=3, stream=22):
    """A = (Vl, Vd, Vr†) with the workload's bonds: left update input of envL, right of envR."""
    return M.randn((v.space[0], v.space[1], v.space[3]), 0, seed=seed, stream=stream, device=DEV)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", ["xxz_L100_chi512", "refbench_chi1000", "cyl6x24_chi2048"])
def test_env_updates_fused_vs_oracle_and_plain(name):
    """Fused environment updates (one plan: packs, step-1 GEMM, MPO pass, GEMM with the
    bra) against the oracle's three contracts (src/apply.jl:17-58) block by block
    (FP64 1e-12) and against the device's three-contract path; FP32 to 1e-5."""
    v, L, R, WL, WR = _workload(name)
    A = _site_tensor(v)
    for side, fn, ofn, args in (("left", M._mpsupdateleft, ota.mpsupdateleft, (L, A, WL)),
                                ("right", M._mpsupdateright, ota.mpsupdateright, (R, A, WR))):
        want = ofn(*[oracle_of(t) for t in args])
        got = fn(*args)
        _check_structure(got, want)
        err = rel_block_err(oracle_of(got), want)
        plain = fn(*args, path="plain")
        d = (got.data - plain.data).norm().item() / plain.data.norm().item()
        got32 = fn(*[t.to(torch.float32) for t in args])
        e32 = global_rel_err(oracle_of(got32), want)
        print(f"{name} {side}: {len(want.blocks)} blocks, f64 max block rel err {err:.2e}, fused vs plain {d:.2e}, "
              f"f32 rel err {e32:.2e}", flush=True)
        assert err <= 1e-12
        assert d <= 1e-12
        assert e32 <= 1e-5


def test_env_updates_edge_bonds_and_charges():
    """Edge environments (dimension-1 bond legs, as at the chain ends) and a site
    tensor of nonzero charge: fused equals the oracle."""
    Vd = M.U1Space((0, 1), (1, 1))
    Vw = M.U1Space((-1, 1), (0, 3), (1, 1))
    V1 = M.U1Space((0, 1))
    V2 = M.U1Space((0, 1), (1, 1))
    V3 = M.U1Space((0, 2), (1, 3), (2, 1))
    for charge in (0, 1):
        A = M.randn((V2, Vd, V3.dual()), charge, seed=5, stream=30 + charge, device=DEV)
        W = M.randn((Vw, Vd, Vw.dual(), Vd.dual()), 0, seed=5, stream=40, device=DEV)
        Le = M.fill(1.0, (V2.dual(), Vw.dual(), V2), device=DEV)
        Re = M.randn((V3, Vw, V3.dual()), 0, seed=5, stream=41, device=DEV)
        for fn, ofn, args in ((M._mpsupdateleft, ota.mpsupdateleft, (Le, A, W)),
                              (M._mpsupdateright, ota.mpsupdateright, (Re, A, W))):
            want = ofn(*[oracle_of(t) for t in args])
            got = fn(*args)
            _check_structure(got, want)
            assert rel_block_err(oracle_of(got), want) <= 1e-12
    # one-dimensional boundary legs on both sides
    A = M.randn((V1, Vd, V2.dual()), 0, seed=6, stream=1, device=DEV)
    Le = M.fill(1.0, (V1.dual(), M.U1Space((0, 1)).dual(), V1), device=DEV)
    W1 = M.randn((M.U1Space((0, 1)), Vd, Vw.dual(), Vd.dual()), 0, seed=6, stream=2, device=DEV)
    want = ota.mpsupdateleft(*[oracle_of(t) for t in (Le, A, W1)])
    got = M._mpsupdateleft(Le, A, W1)
    _check_structure(got, want)
    assert rel_block_err(oracle_of(got), want) <= 1e-12


def _oracle_adjoint(v, L, R, WL, WR):
    """d<v|H v>/denvL and d<v|H v>/denvR from the oracle's contract (the reference's
    contraction on the host), the same label chains as the device's plain path."""
    from oracle import symtensors as ost
    bra = ost.dual(v)
    t = ost.contract(v, (1, 2, 3, -1), R, (-1, 4, 5))
    t = ost.contract(t, (1, 2, -1, -2, 3), WR, (4, 5, -2, -1))
    t = ost.contract(t, (1, -1, 2, -2, 3), WL, (4, 5, -2, -1))
    X = ost.contract(t, (1, -1, -2, 2, -3), bra, (3, -3, -2, -1))
    t = ost.contract(L, (-1, 2, 1), v, (-1, 3, 4, 5))
    t = ost.contract(t, (1, -1, -2, 4, 5), WL, (-1, 2, 3, -2))
    t = ost.contract(t, (1, 2, -1, -2, 5), WR, (-1, 3, 4, -2))
    Y = ost.contract(t, (-1, -2, -3, 2, 1), bra, (-1, -2, -3, 3))
    return X, Y


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512", "cyl6x24_chi2048"])
def test_heff_adjoint_kernels_vs_oracle(name):
    """Adjoint kernels of <v|H_eff v> (mpsh_heff_adjoint_*) against the oracle's
    contraction chain (FP64 1e-12 per block) and the device's plain contract chain;
    FP32 against FP64 to 1e-5."""
    from mps_amd.tensornetalgs import heff_adjoint_plan
    ts = _workload(name)
    P = heff_adjoint_plan(*ts)
    gL, gR = P(*ts)
    X, Y = _oracle_adjoint(*[oracle_of(t) for t in ts])
    v = ts[0]
    for g, st, want in ((gL, P.gL, X), (gR, P.gR, Y)):
        assert st.sectors() == list(want.blocks.keys())
        assert tuple(o_space(V) for V in st.spaces) == want.space
        got = o_from_flat(st.charge, tuple(o_space(V) for V in st.spaces), g.cpu().numpy())
        err = rel_block_err(got, want)
        print(f"{name} adjoint: {len(want.blocks)} blocks, max block rel err {err:.2e}", flush=True)
        assert err <= 1e-12
    # the structures are those of envL / envR (what the gradient is returned as)
    assert P.gL.sectors() == ts[1].struct.sectors() and P.gR.sectors() == ts[2].struct.sectors()
    E, gv, dL, dR = ta.rayleigh_gradient(*ts, path="plain")
    E2, gv2, dL2, dR2 = ta.rayleigh_gradient(*ts)
    for a, b in ((dL2, dL), (dR2, dR)):
        assert (a.data - b.data).norm().item() <= 1e-12 * b.data.norm().item()
    gL32, gR32 = P(*[t.to(torch.float32) for t in ts])
    for a, b in ((gL32, gL), (gR32, gR)):
        assert (a.double() - b).norm().item() <= 1e-5 * b.norm().item()


@pytest.mark.parametrize("name", ["xxz_L32_chi64", "xxz_L100_chi512", "cyl6x24_chi2048"])
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
def test_padded_and_unpadded_layouts_agree(name, prec):
    """Padded H_eff plans (the default: every ket_r row segment of T1, T3 and the
    prepared envR rounded up to 4 elements, the MPO pass moving its pixel runs by
    aligned vector access and writing T3's padding as zeros, the prepare
    zero-filling envR's) against unpadded plans (MPSH_HEFF_PADR=0): the prepared
    operator and the workspace start filled with NaN bytes, so padding that
    reached the output through the step-4 GEMM would show; both outputs are
    finite and agree to rounding (FP64 1e-13; FP32 / BF16, whose step-4 K chunks
    fall differently, 1e-5 relative)."""
    import os
    from mps_amd.tensornetalgs import HeffPlan
    dt = torch.float64 if prec == "f64" else torch.float32
    ts = [t.to(dt) for t in _workload(name)]
    old = os.environ.get("MPSH_HEFF_PADR")
    outs = {}
    try:
        for padr in ("0", "1"):
            os.environ["MPSH_HEFF_PADR"] = padr
            plan = HeffPlan(*[t.struct for t in ts])
            prepared = torch.full((max(16, plan.prepared_bytes(prec)),), 255, dtype=torch.uint8, device=DEV)
            plan.prepare(prec, *(t.data for t in ts[1:]), prepared=prepared)
            ws = torch.full((max(16, plan.apply_workspace_bytes(prec)),), 255, dtype=torch.uint8, device=DEV)
            out = torch.full((plan.out.numel,), float("nan"), dtype=dt, device=DEV)
            plan.apply_prepared(prec, prepared, ts[0].data, out, ws=ws)
            torch.cuda.synchronize()
            outs[padr] = out.double()
    finally:
        if old is None:
            os.environ.pop("MPSH_HEFF_PADR", None)
        else:
            os.environ["MPSH_HEFF_PADR"] = old
    assert torch.isfinite(outs["1"]).all() and torch.isfinite(outs["0"]).all()
    d = float((outs["1"] - outs["0"]).norm() / outs["0"].norm())
    print(f"{name} {prec}: padded vs unpadded rel diff {d:.2e}", flush=True)
    assert d <= (1e-13 if prec == "f64" else 1e-5)
