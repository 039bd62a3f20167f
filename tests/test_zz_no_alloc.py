"""Allocation-free H_eff application (mpsh_heff_apply_prepared_noalloc) through the C ABI.

Its per-matvec scratch is a fixed device array shared by every call on a GPU,
so two calls in a row (the second on a smaller bond) must each give bitwise the
output of mpsh_heff_apply_prepared with a caller-allocated workspace; the FP64
outputs also match the oracle's four-contract chain to 1e-12.  Plans that need
more scratch than the array holds, or whose tables are not on the device yet,
are rejected before any kernel runs.
"""
import pytest
import torch

from oracle import tensornetalgs as ota
from tests.helpers import o_from_product, rel_block_err

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

from mps_amd import _lib as L  # noqa: E402
from mps_amd import tensornetalgs as ta  # noqa: E402
from mps_amd.symtensors import _stream  # noqa: E402

DEV = torch.device("cuda", 0)
MODE = {"f64": L.MPSH_F64, "f32": L.MPSH_F32, "bf16": L.MPSH_BF16}


def _noalloc(plan, prec, prepared, v, out):
    return L.lib().mpsh_heff_apply_prepared_noalloc(plan.plan, MODE[prec], prepared.data_ptr(), v.data_ptr(),
                                                    out.data_ptr(), _stream())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_equals_workspace_path_on_consecutive_calls(prec):
    dt = torch.float64 if prec == "f64" else torch.float32
    for name in ("xxz_L100_chi512", "xxz_L32_chi64"):  # the second call on a smaller bond
        ts = [t.to(dt) for t in ta.heff_workload(name, dtype=torch.float64, device=DEV)]
        plan = ta.heff_plan(*ts)
        assert plan.apply_workspace_bytes(prec) <= L.NOALLOC_SCRATCH_BYTES
        prepared = plan.prepare(prec, *(t.data for t in ts[1:]))
        want = torch.empty(plan.out.numel, dtype=dt, device=DEV)
        plan.apply_prepared(prec, prepared, ts[0].data, want)
        got = torch.full_like(want, float("nan"))
        L.check(_noalloc(plan, prec, prepared, ts[0].data, got))
        torch.cuda.synchronize()
        print(f"{name} {prec}: noalloc == workspace path: {torch.equal(got, want)}", flush=True)
        assert torch.equal(got, want)
        if prec == "f64":
            oracle = ota.applymps2site(*[o_from_product(t) for t in ts])
            res = ta.SymTensor(plan.out.charge, plan.out.spaces, got, plan.out)
            assert rel_block_err(o_from_product(res), oracle) <= 1e-12


def test_noalloc_rejects_plans_beyond_the_scratch():
    ts = [t.to(torch.float32) for t in ta.heff_workload("cyl6x24_chi2048", dtype=torch.float64, device=DEV)]
    plan = ta.heff_plan(*ts)
    assert plan.apply_workspace_bytes("f32") > L.NOALLOC_SCRATCH_BYTES
    dummy = torch.zeros(16, dtype=torch.float32, device=DEV)
    st = _noalloc(plan, "f32", dummy, dummy, dummy)
    assert st == 1  # MPSH_ERR_INVALID
    assert b"scratch" in L.lib().mpsh_last_error()


def test_noalloc_needs_the_plans_tables_on_the_device():
    ts = ta.heff_workload("xxz_L64_chi256", dtype=torch.float64, device=DEV)
    plan = ta.HeffPlan(*(t.struct for t in ts))  # a fresh plan: nothing uploaded yet
    out = torch.empty(plan.out.numel, dtype=torch.float64, device=DEV)
    prepared = torch.zeros(max(16, plan.prepared_bytes("f64")), dtype=torch.uint8, device=DEV)
    st = _noalloc(plan, "f64", prepared, ts[0].data, out)
    assert st == 1
    assert b"mpsh_heff_prepare" in L.lib().mpsh_last_error()
    plan.prepare("f64", *(t.data for t in ts[1:]), prepared=prepared)
    L.check(_noalloc(plan, "f64", prepared, ts[0].data, out))
    want = torch.empty_like(out)
    plan.apply_prepared("f64", prepared, ts[0].data, want)
    torch.cuda.synchronize()
    assert torch.equal(out, want)
