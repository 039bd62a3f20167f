"""Two-site DMRG hot path on the GPU: H_eff application and environment updates.

Mirrors ``src/apply.jl`` of amiragha/mps.  ``_applymps2site`` has two device
paths with identical output structure:

* ``path="fused"`` (default): one plan per bond; step-1 grouped MFMA GEMM,
  one combined-MPO pass, step-4 grouped MFMA GEMM (libmpsh ``mpsh_heff_*``).
* ``path="plain"``: the reference's own four ``contract`` calls
  (apply.jl:99-103), each a gather-fuse / grouped GEMM / scatter-split on
  the device.  This is the per-sector FP64 baseline path.
"""
from __future__ import annotations

from typing import Dict, Optional

import torch

from . import _lib as L
from ._lib import check, lib
from .symtensors import (SymTensor, Struct, _DT, _ptr, _stream, _struct_from_handle, contract, dual)

_PREC = {"f64": L.MPSH_F64, "f32": L.MPSH_F32, "bf16": L.MPSH_BF16}
_TORCH = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.float32}


class HeffPlan:
    """Cached fused H_eff plan for one bond structure (v, envL, envR, WL, WR)."""

    def __init__(self, sv: Struct, sL: Struct, sR: Struct, sWL: Struct, sWR: Struct):
        plan, out = L._p(), L._p()
        check(lib().mpsh_heff_plan(sv.handle, sL.handle, sR.handle, sWL.handle, sWR.handle, plan, out))
        self.plan = plan
        self.out = _struct_from_handle(out)
        st = (L._i64 * 10)()
        check(lib().mpsh_heff_stats(plan, st))
        self.ref_flops = int(st[0])
        self.ref_elems = int(st[1])
        self.step_flops = [int(st[2 + i]) for i in range(4)]
        self.step_elems = [int(st[6 + i]) for i in range(4)]
        self._ws: Dict = {}

    def workspace_bytes(self, prec: str) -> int:
        return int(lib().mpsh_heff_workspace(self.plan, _PREC[prec]))

    def workspace(self, prec: str, device) -> torch.Tensor:
        key = (prec, device)
        ws = self._ws.get(key)
        if ws is None:
            ws = torch.empty(max(1, self.workspace_bytes(prec)), dtype=torch.uint8, device=device)
            self._ws[key] = ws
        return ws

    def apply(self, prec: str, v: torch.Tensor, envL: torch.Tensor, envR: torch.Tensor, WL: torch.Tensor,
              WR: torch.Tensor, out: torch.Tensor, ws: torch.Tensor = None, stream=None):
        """Raw apply on packed device buffers (all of the precision's storage dtype)."""
        ws = ws if ws is not None else self.workspace(prec, v.device)
        check(lib().mpsh_heff_apply(self.plan, _PREC[prec], _ptr(v), _ptr(envL), _ptr(envR), _ptr(WL), _ptr(WR),
                                    _ptr(out), _ptr(ws), stream if stream is not None else _stream()))
        return out

    def apply_stages(self, prec: str, stages: int, v, envL, envR, WL, WR, out, ws, stream=None):
        """Run a subset of the three kernel stages (1 step-1 GEMM, 2 MPO pass, 4 step-4 GEMM)."""
        check(lib().mpsh_heff_apply_stages(self.plan, _PREC[prec], int(stages), _ptr(v), _ptr(envL), _ptr(envR),
                                           _ptr(WL), _ptr(WR), _ptr(out), _ptr(ws),
                                           stream if stream is not None else _stream()))
        return out

    def __del__(self):
        try:
            lib().mpsh_heff_destroy(self.plan)
        except Exception:
            pass


_HEFF: Dict = {}


def heff_plan(v: SymTensor, envL: SymTensor, envR: SymTensor, WL: SymTensor, WR: SymTensor) -> HeffPlan:
    key = (v.struct.key, envL.struct.key, envR.struct.key, WL.struct.key, WR.struct.key)
    p = _HEFF.get(key)
    if p is None:
        p = HeffPlan(v.struct, envL.struct, envR.struct, WL.struct, WR.struct)
        _HEFF[key] = p
    return p


def _applymps2site(v: SymTensor, envL: SymTensor, envR: SymTensor, hmpoL: SymTensor, hmpoR: SymTensor,
                   path: str = "fused", prec: Optional[str] = None) -> SymTensor:
    """``_applymps2site(v, envL, envR, hmpoL, hmpoR)`` (apply.jl:93-105).

    ``prec`` selects the device arithmetic of the fused path: "f64", "f32"
    or "bf16" (bf16 MFMA operands, f32 accumulation); default follows the
    tensors' dtype.  Inputs must share one dtype (float64 or float32); for
    "bf16" they are float32.
    """
    if path == "plain":
        T1 = contract(envL, (-1, 2, 1), v, (-1, 3, 4, 5))
        T2 = contract(T1, (1, -1, -2, 4, 5), hmpoL, (-1, 2, 3, -2))
        T3 = contract(T2, (1, 2, -1, -2, 5), hmpoR, (-1, 3, 4, -2))
        return contract(T3, (1, 2, 3, -1, -2), envR, (-2, -1, 4))
    if path != "fused":
        raise ValueError(f"unknown path {path!r}")
    if prec is None:
        prec = "f64" if v.dtype == torch.float64 else "f32"
    need = _TORCH[prec]
    for t in (v, envL, envR, hmpoL, hmpoR):
        if t.dtype != need:
            raise TypeError(f"prec={prec} needs {need} tensors, got {t.dtype}")
    p = heff_plan(v, envL, envR, hmpoL, hmpoR)
    out = torch.empty(p.out.numel, dtype=need, device=v.device)
    p.apply(prec, v.data, envL.data, envR.data, hmpoL.data, hmpoR.data, out)
    return SymTensor(p.out.charge, p.out.spaces, out, p.out)


def _mpsupdateleft(lenv: SymTensor, mat: SymTensor, hmpo: SymTensor) -> SymTensor:
    """``_mpsupdateleft(lenv, mat, hmpo)`` (apply.jl:49-58): three device contracts."""
    return contract(contract(contract(lenv, (-1, 3, 4), mat, (-1, 2, 1)),
                             (1, -1, -2, 4), hmpo, (-2, 2, 3, -1)),
                    (1, -1, 2, -2), dual(mat), (-2, -1, 3))


def _mpsupdateright(renv: SymTensor, mat: SymTensor, hmpo: SymTensor) -> SymTensor:
    """``_mpsupdateright(renv, mat, hmpo)`` (apply.jl:17-25): three device contracts."""
    return contract(contract(contract(renv, (-1, 3, 4), mat, (1, 2, -1)),
                             (1, -1, -2, 4), hmpo, (3, 2, -2, -1)),
                    (1, -1, 2, -2), dual(mat), (3, -1, -2))


def heff_workload(name: str, seed: int = 1911, dtype=torch.float64, device=None):
    """Random (v, envL, envR, WL, WR) with the leg layout of a two-site DMRG bond.

    Blocks are seeded Gaussians (like the reference's own benchmark,
    benchmarkings.ipynb:48-73, which uses random env/MPO tensors).
    """
    from .symtensors import U1Space, randn
    from .workloads import heff_spaces
    left, right, mpo = heff_spaces(name)
    Vl, Vr, Vw = U1Space(*left), U1Space(*right), U1Space(*mpo)
    Vd = U1Space((0, 1), (1, 1))
    v = randn((Vl, Vd, Vd, Vr.dual()), 0, seed, 1, dtype, device)
    envL = randn((Vl.dual(), Vw.dual(), Vl), 0, seed, 2, dtype, device)
    envR = randn((Vr, Vw, Vr.dual()), 0, seed, 3, dtype, device)
    WL = randn((Vw, Vd, Vw.dual(), Vd.dual()), 0, seed, 4, dtype, device)
    WR = randn((Vw, Vd, Vw.dual(), Vd.dual()), 0, seed, 5, dtype, device)
    return v, envL, envR, WL, WR
