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

from collections import OrderedDict

import numpy as np
import torch

from . import _lib as L
from ._lib import check, lib
from .symtensors import (PlanHandle, PlanLRU, SymTensor, Struct, _DT, _ptr, _stream, _struct_from_handle, contract,
                         dual)

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
        es = (L._i64 * 8)()
        check(lib().mpsh_heff_engine_stats(plan, es))
        self.t1_elems, self.t3_elems, self.w_elems, self.mpo_flops = (int(x) for x in es[:4])
        self.vp_elems = int(es[6])

    def workspace_bytes(self, prec: str) -> int:
        """Workspace of the one-shot apply (prepared operator included)."""
        return int(lib().mpsh_heff_workspace(self.plan, _PREC[prec]))

    def apply_workspace_bytes(self, prec: str) -> int:
        return int(lib().mpsh_heff_apply_workspace(self.plan, _PREC[prec]))

    def prepared_bytes(self, prec: str) -> int:
        return int(lib().mpsh_heff_prepared_size(self.plan, _PREC[prec]))

    def workspace(self, prec: str, device, kind: str = "oneshot") -> torch.Tensor:
        """Scratch of the one-shot apply ("oneshot") or of the prepared apply ("apply"):
        a view of the device's shared H_eff scratch (applies run in stream order)."""
        n = self.workspace_bytes(prec) if kind == "oneshot" else self.apply_workspace_bytes(prec)
        return _scratch(device, max(16, n))

    def apply(self, prec: str, v: torch.Tensor, envL: torch.Tensor, envR: torch.Tensor, WL: torch.Tensor,
              WR: torch.Tensor, out: torch.Tensor, ws: torch.Tensor = None, stream=None):
        """One-shot apply on packed device buffers (all of the precision's storage dtype)."""
        ws = ws if ws is not None else self.workspace(prec, v.device)
        check(lib().mpsh_heff_apply(self.plan, _PREC[prec], _ptr(v), _ptr(envL), _ptr(envR), _ptr(WL), _ptr(WR),
                                    _ptr(out), _ptr(ws), stream if stream is not None else _stream()))
        return out

    def prepare(self, prec: str, envL: torch.Tensor, envR: torch.Tensor, WL: torch.Tensor, WR: torch.Tensor,
                prepared: torch.Tensor = None, stream=None) -> torch.Tensor:
        """Repack the bond's environments and form the combined MPO (once per bond)."""
        if prepared is None:
            prepared = torch.empty(max(16, self.prepared_bytes(prec)), dtype=torch.uint8, device=envL.device)
        check(lib().mpsh_heff_prepare(self.plan, _PREC[prec], _ptr(envL), _ptr(envR), _ptr(WL), _ptr(WR),
                                      _ptr(prepared), stream if stream is not None else _stream()))
        return prepared

    def tune(self, prec: str, prepared: torch.Tensor, v: torch.Tensor, out: torch.Tensor, ws: torch.Tensor = None,
             reps: int = 3, force: bool = False, stream=None) -> dict:
        """Autotune the two GEMMs' tile / split-K variants for ``prec`` on this plan's
        real problems (``mpsh_heff_tune``); plans with the same block-size histogram
        signature reuse the result.  Overwrites ``out`` and the workspace."""
        ws = ws if ws is not None else self.workspace(prec, v.device, "apply")
        ch = (L._i32 * 9)()
        check(lib().mpsh_heff_tune(self.plan, _PREC[prec], _ptr(prepared), _ptr(v), _ptr(out), _ptr(ws), reps,
                                   int(force), ch, stream if stream is not None else _stream()))
        TUNE_COUNTS["cached" if ch[8] else "timed"] += 1
        return {"step1": f"{ch[0]}x{ch[1]}x{ch[2]}", "step4": f"{ch[4]}x{ch[5]}x{ch[6]}", "split": int(ch[7]),
                "from_cache": bool(ch[8])}

    def apply_prepared(self, prec: str, prepared: torch.Tensor, v: torch.Tensor, out: torch.Tensor,
                       ws: torch.Tensor = None, stream=None):
        ws = ws if ws is not None else self.workspace(prec, v.device, "apply")
        check(lib().mpsh_heff_apply_prepared(self.plan, _PREC[prec], _ptr(prepared), _ptr(v), _ptr(out), _ptr(ws),
                                             stream if stream is not None else _stream()))
        return out

    def apply_stages(self, prec: str, stages: int, prepared, v, out, ws, stream=None):
        """Run a subset of the kernel stages (1 pack v, 2 step-1 GEMM, 4 MPO pass, 8 step-4 GEMM)."""
        check(lib().mpsh_heff_apply_prepared_stages(self.plan, _PREC[prec], int(stages), _ptr(prepared), _ptr(v),
                                                    _ptr(out), _ptr(ws),
                                                    stream if stream is not None else _stream()))
        return out

    def __del__(self):
        try:
            lib().mpsh_heff_destroy(self.plan)
        except Exception:
            pass


# One H_eff scratch buffer per device, shared by every plan: a sweep meets a new
# bond structure at every bond, and per-plan buffers (~1 GB each at chi=2048)
# would pile up.  The buffer only grows; a replaced buffer is kept alive so a
# pointer captured earlier (hipGraph replay) never points at freed memory.
_SCRATCH: Dict = {}
_SCRATCH_OLD: list = []


def _scratch(device, nbytes: int) -> torch.Tensor:
    device = torch.device(device)
    buf = _SCRATCH.get(device)
    if buf is None or buf.numel() < nbytes:
        if buf is not None:
            _SCRATCH_OLD.append(buf)
        buf = torch.empty(int(nbytes * 1.25) + 256, dtype=torch.uint8, device=device)
        _SCRATCH[device] = buf
    return buf[:nbytes]


# plans (host and device tables, a few hundred KB each at chi=2048), least
# recently used dropped beyond HEFF_PLAN_CACHE
HEFF_PLAN_CACHE = 512
_HEFF: "OrderedDict" = PlanLRU(HEFF_PLAN_CACHE)


def heff_plan(v: SymTensor, envL: SymTensor, envR: SymTensor, WL: SymTensor, WR: SymTensor) -> HeffPlan:
    key = (v.struct.key, envL.struct.key, envR.struct.key, WL.struct.key, WR.struct.key)
    p = _HEFF.get(key)
    if p is None:
        p = _HEFF[key] = HeffPlan(v.struct, envL.struct, envR.struct, WL.struct, WR.struct)
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


class EnvUpdatePlan:
    """Fused environment update of one structure (``mpsh_envupdate_*``): operand
    packs, the H_eff's step-1 grouped GEMM and combined-MPO pass over one site,
    and one grouped GEMM against the bra dual(A)."""

    def __init__(self, side: str, s_env: Struct, sA: Struct, sW: Struct):
        plan, out = L._p(), L._p()
        make = lib().mpsh_envupdate_left_plan if side == "left" else lib().mpsh_envupdate_right_plan
        check(make(s_env.handle, sA.handle, sW.handle, plan, out))
        self.owner = PlanHandle(plan, "mpsh_envupdate_destroy")
        self.out = _struct_from_handle(out)
        st = (L._i64 * 4)()
        check(lib().mpsh_envupdate_stats(plan, st))
        self.ref_flops, self.ref_elems, self.gemm_flops, self.mpo_flops = (int(x) for x in st)

    def __call__(self, env: SymTensor, A: SymTensor, W: SymTensor) -> SymTensor:
        dt = env.dtype
        if A.dtype != dt or W.dtype != dt:
            raise TypeError("environment update of mixed dtypes")
        h = self.owner.h
        ws = _scratch(env.device, max(16, int(lib().mpsh_envupdate_workspace(h, _DT[dt]))))
        out = torch.empty(self.out.numel, dtype=dt, device=env.device)
        check(lib().mpsh_envupdate_execute(h, _DT[dt], _ptr(env.data), _ptr(A.data), _ptr(W.data), _ptr(out),
                                           _ptr(ws), _stream()))
        return SymTensor(self.out.charge, self.out.spaces, out, self.out)


ENV_PLAN_CACHE = 1024
_ENV: "OrderedDict" = PlanLRU(ENV_PLAN_CACHE)


def envupdate_plan(side: str, env: SymTensor, A: SymTensor, W: SymTensor) -> EnvUpdatePlan:
    key = (side, env.struct.key, A.struct.key, W.struct.key)
    p = _ENV.get(key)
    if p is None:
        p = _ENV[key] = EnvUpdatePlan(side, env.struct, A.struct, W.struct)
    return p


def _mpsupdateleft(lenv: SymTensor, mat: SymTensor, hmpo: SymTensor, path: str = "fused") -> SymTensor:
    """``_mpsupdateleft(lenv, mat, hmpo)`` (apply.jl:49-58).  ``path="fused"``: one
    fused device plan; ``path="plain"``: the reference's three device contracts."""
    if path == "plain":
        return contract(contract(contract(lenv, (-1, 3, 4), mat, (-1, 2, 1)),
                                 (1, -1, -2, 4), hmpo, (-2, 2, 3, -1)),
                        (1, -1, 2, -2), dual(mat), (-2, -1, 3))
    if path != "fused":
        raise ValueError(f"unknown path {path!r}")
    return envupdate_plan("left", lenv, mat, hmpo)(lenv, mat, hmpo)


def _mpsupdateright(renv: SymTensor, mat: SymTensor, hmpo: SymTensor, path: str = "fused") -> SymTensor:
    """``_mpsupdateright(renv, mat, hmpo)`` (apply.jl:17-25).  ``path="fused"``: one
    fused device plan (the left update of the mirror image); ``path="plain"``: the
    reference's three device contracts."""
    if path == "plain":
        return contract(contract(contract(renv, (-1, 3, 4), mat, (1, 2, -1)),
                                 (1, -1, -2, 4), hmpo, (3, 2, -2, -1)),
                        (1, -1, 2, -2), dual(mat), (3, -1, -2))
    if path != "fused":
        raise ValueError(f"unknown path {path!r}")
    return envupdate_plan("right", renv, mat, hmpo)(renv, mat, hmpo)


class HeffAdjointPlan:
    """Adjoint kernels of <v|H_eff v> w.r.t. both environments (``mpsh_heff_adjoint_*``)."""

    def __init__(self, sv: Struct, sL: Struct, sR: Struct, sWL: Struct, sWR: Struct):
        plan, gL, gR = L._p(), L._p(), L._p()
        check(lib().mpsh_heff_adjoint_plan(sv.handle, sL.handle, sR.handle, sWL.handle, sWR.handle, plan, gL, gR))
        self.owner = PlanHandle(plan, "mpsh_heff_adjoint_destroy")
        self.gL, self.gR = _struct_from_handle(gL), _struct_from_handle(gR)
        st = (L._i64 * 4)()
        check(lib().mpsh_heff_adjoint_stats(plan, st))
        self.ref_flops, self.ref_elems, self.gemm_flops, self.mpo_flops = (int(x) for x in st)

    def __call__(self, v, envL, envR, WL, WR):
        """(d<v|H v>/d envL, d<v|H v>/d envR) as packed buffers of the dtype of the operands."""
        dt = v.dtype
        for t in (envL, envR, WL, WR):
            if t.dtype != dt:
                raise TypeError("adjoint of mixed dtypes")
        h = self.owner.h
        ws = _scratch(v.device, max(16, int(lib().mpsh_heff_adjoint_workspace(h, _DT[dt]))))
        gL = torch.empty(self.gL.numel, dtype=dt, device=v.device)
        gR = torch.empty(self.gR.numel, dtype=dt, device=v.device)
        check(lib().mpsh_heff_adjoint_execute(h, _DT[dt], _ptr(v.data), _ptr(envL.data), _ptr(envR.data),
                                              _ptr(WL.data), _ptr(WR.data), _ptr(gL), _ptr(gR), _ptr(ws),
                                              _stream()))
        return gL, gR


_ADJ: "OrderedDict" = PlanLRU(256)


def heff_adjoint_plan(v, envL, envR, WL, WR) -> HeffAdjointPlan:
    key = (v.struct.key, envL.struct.key, envR.struct.key, WL.struct.key, WR.struct.key)
    p = _ADJ.get(key)
    if p is None:
        p = _ADJ[key] = HeffAdjointPlan(v.struct, envL.struct, envR.struct, WL.struct, WR.struct)
    return p


def rayleigh_gradient(v: SymTensor, envL: SymTensor, envR: SymTensor, WL: SymTensor, WR: SymTensor,
                      prec: str = None, path: str = "fused"):
    """Gradient of E = <v|H_eff v>/<v|v> (H_eff of ``_applymps2site``, apply.jl:93-105)
    with respect to the two-site tensor and both environments.

    Returns (E, dE/dv, dE/denvL, dE/denvR), each gradient with the structure of
    its variable.  E is linear in each environment, so dE/denvL is the adjoint
    contraction of (bra v, WL, WR, envR, ket v) with the envL legs left open,
    divided by <v|v> (likewise for envR): with ``path="fused"`` both come from
    the adjoint kernels (``mpsh_heff_adjoint_*``, two environment cores: the
    operands and their mirror image), with ``path="plain"`` from chains of
    block-sparse device contracts (the adjoints of the reference's four
    contracts).  dE/dv = 2 (H v - E v)/<v|v> holds for the Hermitian H_eff of
    DMRG (environments of one MPS and a Hermitian MPO).  ``prec`` selects the
    H_eff arithmetic of the v-gradient; the environment adjoints run in the
    tensors' dtype.  Nothing is stored for a backward pass: the environments are
    inputs, recomputed by the caller (``initialenv``) when they are needed.
    """
    from .symtensors import axpby_, dot, rmul_
    nn = dot(v, v)
    Hv = _applymps2site(v, envL, envR, WL, WR, prec=prec) if prec else _applymps2site(v, envL, envR, WL, WR)
    E = dot(v, Hv) / nn
    gv = Hv.copy()
    axpby_(-2.0 * E / nn, v, 2.0 / nn, gv)

    def as_grad(G, Gs, like):
        # the adjoint's legs are the duals of the variable's: same sectors, same block order
        if Gs.sectors() != like.struct.sectors() or Gs.numel != like.struct.numel:
            raise ValueError("adjoint structure does not match the environment")
        return rmul_(SymTensor(like.charge, like.space, G.to(like.dtype), like.struct), 1.0 / nn)

    if path == "fused":
        P = heff_adjoint_plan(v, envL, envR, WL, WR)
        gL, gR = P(v, envL, envR, WL, WR)
        return E, gv, as_grad(gL, P.gL, envL), as_grad(gR, P.gR, envR)
    if path != "plain":
        raise ValueError(f"unknown path {path!r}")
    bra = dual(v)
    # dE/denvL[a,m,b] = sum v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] envR[c,k,e] v[b,q,r,e] / <v|v>
    t = contract(v, (1, 2, 3, -1), envR, (-1, 4, 5))
    t = contract(t, (1, 2, -1, -2, 3), WR, (4, 5, -2, -1))
    t = contract(t, (1, -1, 2, -2, 3), WL, (4, 5, -2, -1))
    X = contract(t, (1, -1, -2, 2, -3), bra, (3, -3, -2, -1))
    # dE/denvR[c,k,e] = sum envL[a,m,b] v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] v[b,q,r,e] / <v|v>
    t = contract(envL, (-1, 2, 1), v, (-1, 3, 4, 5))
    t = contract(t, (1, -1, -2, 4, 5), WL, (-1, 2, 3, -2))
    t = contract(t, (1, 2, -1, -2, 5), WR, (-1, 3, 4, -2))
    Y = contract(t, (-1, -2, -3, 2, 1), bra, (-1, -2, -3, 3))
    return E, gv, as_grad(X.data.clone(), X.struct, envL), as_grad(Y.data.clone(), Y.struct, envR)


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


# ---------------------------------------------------------------------------
# Lanczos on device vectors (stands in for KrylovKit.eigsolve, dmrg.jl:77-82)
# ---------------------------------------------------------------------------
class EigInfo:
    def __init__(self, normres, numiter, numops, converged):
        self.normres, self.numiter, self.numops, self.converged = normres, numiter, numops, converged


LANCZOS_GRAPHS = False  # capture each Lanczos iteration in a hipGraph (eigsolve default)
# HeffOperator tunes the FP32/BF16 GEMM variants of each new plan (signature-cached)
# when set.  Off by default: in a chi=2048 sweep each of ~30 new signatures costs
# ~34 ms of timing runs while the default variants are close to the tuned ones
# (measured sweep 17.8 s tuned, 16.8 s untuned); runs of many sweeps, which reuse
# the cached signatures, may turn it on.
AUTOTUNE = False
AUTOTUNE_MIN_FLOPS = 2e8  # ... when one matvec costs at least this many reference FLOPs
STAGE_RTOL = {"f32": 1e-6}  # relative residual at which a refined eigsolve leaves its reduced-precision stage
TUNE_COUNTS = {"timed": 0, "cached": 0}  # autotuner calls that timed variants / reused a signature


class _LanczosBasis:
    """Device-resident Lanczos basis: V[0..kdim] (FP64 rows), the device scalars of
    each iteration (block j: c1 | c2 | beta), their pinned host copies and the
    partial-sum scratch of the ``mpsh_lanczos_*`` kernels."""

    def __init__(self, x: SymTensor, kdim: int):
        self.like = x
        self.n = x.struct.numel
        self.kdim = kdim
        dev = x.device
        self.ld = max(1, self.n)
        kd = kdim + 1
        self.blk = 2 * kd + 2  # doubles per iteration block: c1 (kd) | c2 (kd) | beta | pad
        self.V = torch.empty((kdim + 1, self.ld), dtype=torch.float64, device=dev)
        # blocks 0..kdim-1: iterations; block kdim: the start / Ritz normalisations
        self.sc = torch.zeros((kdim + 1, self.blk), dtype=torch.float64, device=dev)
        self.host = torch.zeros((kdim + 1, self.blk), dtype=torch.float64, pin_memory=dev.type == "cuda")
        self.events = [torch.cuda.Event() if dev.type == "cuda" else None for _ in range(kdim)]
        self.part = torch.empty((kdim + 1) * L.LZ_PARTS, dtype=torch.float64, device=dev)
        self.graphs = {}
        self.outs = {}

    def vec(self, j) -> SymTensor:
        return SymTensor(self.like.charge, self.like.space, self.V[j, :self.n], self.like.struct)

    def _blk(self, j) -> int:
        return self.sc[j].data_ptr()

    def iterate(self, apply, j):
        """w = apply(V_j); two classical Gram-Schmidt passes against V_0..V_j (scalars
        c1, c2 of block j on the device); V_{j+1} = w / ||w||.  No host synchronisation."""
        w = apply(self.vec(j))
        lb, st, k, kd = lib(), _stream(), j + 1, self.kdim + 1
        V, wp, b = _ptr(self.V), _ptr(w.data), self._blk(j)
        check(lb.mpsh_lanczos_dots(k, self.n, V, self.ld, wp, b, _ptr(self.part), st))
        check(lb.mpsh_lanczos_update(k, self.n, V, self.ld, b, -1.0, wp, st))
        check(lb.mpsh_lanczos_dots(k, self.n, V, self.ld, wp, b + 8 * kd, _ptr(self.part), st))
        check(lb.mpsh_lanczos_update(k, self.n, V, self.ld, b + 8 * kd, -1.0, wp, st))
        check(lb.mpsh_lanczos_normalize(self.n, wp, self.V[j + 1].data_ptr(), b + 16 * kd, _ptr(self.part), st))
        return w

    def step(self, apply, j, graph: bool):
        """Enqueue iteration j and the copy of its scalars to the host (no wait)."""
        if not graph:
            self.outs[j] = self.iterate(apply, j)
        else:
            g = self.graphs.get(j)
            if g is None:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self.outs[j] = self.iterate(apply, j)
                self.graphs[j] = g
            g.replay()
        if self.events[j] is not None:
            self.host[j].copy_(self.sc[j], non_blocking=True)
            self.events[j].record()

    def scalars(self, j):
        """(alpha_j, beta_j) of iteration j: waits for that iteration only."""
        kd = self.kdim + 1
        if self.events[j] is not None:
            self.events[j].synchronize()
            s = self.host[j].numpy()
        else:
            s = self.sc[j].cpu().numpy()
        return float(s[j] + s[kd + j]), float(s[2 * kd])

    def normalize_into_v0(self, x_ptr: int) -> None:
        check(lib().mpsh_lanczos_normalize(self.n, x_ptr, self.V[0].data_ptr(), self._blk(self.kdim) + 16 * (self.kdim + 1),
                                           _ptr(self.part), _stream()))

    def combine(self, y) -> None:
        """V_0 <- normalised sum_i y_i V_i (the Ritz vector)."""
        k = len(y)
        c = torch.tensor(np.asarray(y, dtype=np.float64), device=self.V.device)
        x = torch.zeros(self.ld, dtype=torch.float64, device=self.V.device)
        check(lib().mpsh_lanczos_update(k, self.n, _ptr(self.V), self.ld, _ptr(c), 1.0, _ptr(x), _stream()))
        self.normalize_into_v0(_ptr(x))


def lanczos_lowest(apply, x0: SymTensor, tol: float = 1e-7, krylovdim: int = 5, maxiter: int = 8,
                   graphs: bool = False, rtol: float = 0.0):
    """Restarted Lanczos for the lowest eigenpair of a Hermitian map on FP64 device vectors.

    The basis lives on the device; every iteration applies the map, runs two
    classical Gram-Schmidt passes against the whole basis and normalises, with
    all scalars in device memory (``mpsh_lanczos_*``); the host reads (alpha,
    beta) of each iteration from a pinned copy to solve the tridiagonal
    eigenproblem in FP64 and test |beta_k y_k| < max(tol, rtol |theta|).
    Iteration j+1 is enqueued before the host waits for iteration j's scalars,
    so the device never idles on the host's test; it is skipped when the
    residuals' geometric trend predicts that iteration j converges (an
    iteration enqueued past convergence is discarded: it writes only V_{j+2}
    and its own scalar block).  With ``graphs`` each iteration index is
    captured once in a hipGraph and replayed (the first application runs
    eagerly, so the operator prepares itself outside the capture).
    """
    kd = max(1, int(krylovdim))
    B = _LanczosBasis(x0, kd)
    x = x0 if x0.dtype == torch.float64 else x0.to(torch.float64)
    B.normalize_into_v0(_ptr(x.data))
    numops, theta, res, it = 0, None, float("inf"), 0
    warm = not graphs
    stop = tol
    for it in range(1, maxiter + 1):
        alphas, betas, hist = [], [], []
        y = np.array([1.0])
        launched = 0
        for j in range(kd):
            if launched <= j:
                B.step(apply, j, graph=graphs and warm)
                warm = True
                launched = j + 1
            predicted = hist[-1] * hist[-1] / hist[-2] if len(hist) >= 2 and hist[-2] > 0 else float("inf")
            if j + 1 < kd and predicted >= stop:
                B.step(apply, j + 1, graph=graphs and warm)
                launched = j + 2
            a, b = B.scalars(j)
            numops += 1
            alphas.append(a)
            T = np.diag(alphas) + np.diag(betas, 1) + np.diag(betas, -1)
            evals, evecs = np.linalg.eigh(T)
            theta, y = float(evals[0]), evecs[:, 0]
            res = abs(b * y[-1])
            stop = max(tol, rtol * abs(theta))
            hist.append(res)
            if res < stop or j == kd - 1 or b < 1e-14:
                break
            betas.append(b)
        numops += launched - len(alphas)  # applications enqueued past the exit
        B.combine(y)
        if res < stop:
            break
    out = SymTensor(x0.charge, x0.space, B.V[0, :B.n].clone(), x0.struct)
    return theta, out, EigInfo(res, it, numops, res < stop)


class HeffOperator:
    """H_eff of one bond as a map on FP64 device vectors; matvec in ``prec`` (f64/f32/bf16).

    The environments and MPO are prepared once (repacked rows, combined MPO);
    every matvec then runs the three kernel stages on the prepared operator.
    """

    def __init__(self, envL: SymTensor, envR: SymTensor, WL: SymTensor, WR: SymTensor, prec: str = "f64"):
        self.prec = prec
        self.ops64 = (envL, envR, WL, WR)
        self.ops = self.ops64 if prec == "f64" else tuple(t.to(torch.float32) for t in self.ops64)
        self._prep = {}
        self.matvecs = 0

    def _apply(self, x: SymTensor, prec: str, ops) -> SymTensor:
        plan = heff_plan(x, *ops)
        key = (prec, x.struct.key)
        out = torch.empty(plan.out.numel, dtype=x.dtype, device=x.device)
        if key not in self._prep:
            self._prep[key] = plan.prepare(prec, *(t.data for t in ops))
            # the bulk (reduced-precision) matvecs are tuned; FP64 refinement runs few
            # matvecs per bond, so it keeps the default variants
            if AUTOTUNE and prec != "f64" and plan.ref_flops >= AUTOTUNE_MIN_FLOPS:
                plan.tune(prec, self._prep[key], x.data, out)
        plan.apply_prepared(prec, self._prep[key], x.data, out)
        return SymTensor(plan.out.charge, plan.out.spaces, out, plan.out)

    def __call__(self, x: SymTensor) -> SymTensor:
        self.matvecs += 1
        if self.prec == "f64":
            return self._apply(x, "f64", self.ops)
        return self._apply(x.to(torch.float32), self.prec, self.ops).to(torch.float64)

    def exact(self, x: SymTensor) -> SymTensor:
        self.matvecs += 1
        return self._apply(x, "f64", self.ops64)


def eigsolve(H: HeffOperator, x0: SymTensor, tol=1e-7, krylovdim=5, maxiter=8, refine_tol=None, graphs=None):
    """Lowest eigenpair; reduced-precision cycles then (optionally) FP64 refinement cycles.
    ``graphs`` (default ``LANCZOS_GRAPHS``) captures each Lanczos iteration in a hipGraph."""
    graphs = LANCZOS_GRAPHS if graphs is None else graphs
    # with refinement to follow, the FP32 stage stops at its own noise floor
    # (residuals below ~1e-6 |theta| are not resolved by FP32 matvecs)
    rtol = STAGE_RTOL.get(H.prec, 0.0) if refine_tol is not None else 0.0
    theta, x, info = lanczos_lowest(H, x0, tol, krylovdim, maxiter, graphs=graphs, rtol=rtol)
    if H.prec != "f64" and refine_tol is not None:
        theta, x, info2 = lanczos_lowest(H.exact, x, refine_tol, krylovdim, max(1, maxiter // 2), graphs=graphs)
        info = EigInfo(info2.normres, info.numiter + info2.numiter, info.numops + info2.numops, info2.converged)
    return theta, x, info


# ---------------------------------------------------------------------------
# environments and the sweep   (src/dmrg.jl)
# ---------------------------------------------------------------------------
class EnvStore:
    """The environments of a sweep (env[0..L+1], as ``initialenv`` returns them)
    under a device-memory budget.

    The reference keeps environments in files and prefetches them
    (``_initialenv_tempfiles`` / ``dmrg2sitesweep_envio!``, src/dmrg_envio.jl:5-39,
    151-261).  Here they stay in HBM up to ``budget`` bytes; past it, the resident
    environment farthest from the last one accessed is dropped (never the two
    boundary environments, nor the ``protect`` most recently accessed), and a
    dropped environment is recomputed when it is read, from its neighbour towards
    its chain end with the fused update: a left environment env[i] =
    _mpsupdateleft(env[i-1], A_i, W_i), a right one env[i] =
    _mpsupdateright(env[i+1], A_i, W_i), recursively.  An environment update at
    chi=2048 costs well under a millisecond on the GPU, less than moving the
    ~50 MB environment over PCIe and back, so dropping beats offloading.

    The two boundary environments and the ``protect`` most recently accessed
    ones always stay resident, so a budget below that many environments is
    exceeded by them.  Recomputation reads the MPS tensors as they are now: it is exact for the
    sweep's access pattern, in which an environment is read only while the site
    tensors it was built from are unchanged (a left environment env[i] from
    A_1..A_i, a right one from A_i..A_L).  ``set(i, T, side)`` records which
    side an environment belongs to (``initialenv`` and the sweeps call it);
    ``env[i] = T`` works it out from the legs: a left environment env[i]
    carries A_i's right bond leg first, a right one A_i's left bond leg.
    Supports ``len``, ``env[i]`` and assignment, so ``dmrg2sitesweep`` and
    ``dmrg2site_window`` take it in place of the list."""

    def __init__(self, n: int, mps, mpo, budget: int, protect: int = 4, update_left=None, update_right=None):
        self.t = [None] * n
        self.kind = [None] * n
        self.mps, self.mpo = mps, mpo
        self.budget = int(budget)
        self.protect = int(protect)
        self.recent = []
        self.last = 0
        self.recomputed = 0
        self.dropped = 0
        self.peak = 0
        self._left = update_left or _mpsupdateleft
        self._right = update_right or _mpsupdateright

    def __len__(self):
        return len(self.t)

    @staticmethod
    def _bytes(T) -> int:
        return T.data.numel() * T.data.element_size()

    def resident_bytes(self) -> int:
        return sum(self._bytes(T) for T in self.t if T is not None)

    def _touch(self, i):
        self.last = i
        if i in self.recent:
            self.recent.remove(i)
        self.recent.append(i)
        del self.recent[:-self.protect]

    def _enforce(self):
        n = len(self.t)
        used = self.resident_bytes()
        while used > self.budget:
            cands = [i for i in range(1, n - 1) if self.t[i] is not None and i not in self.recent]
            if not cands:
                break
            v = max(cands, key=lambda i: (abs(i - self.last), i))
            used -= self._bytes(self.t[v])
            self.t[v] = None
            self.dropped += 1
        self.peak = max(self.peak, used)

    def set(self, i, T, side: str):
        if side not in ("L", "R"):
            raise ValueError("side must be 'L' or 'R'")
        self.t[i] = T
        self.kind[i] = side
        self._touch(i)
        self._enforce()

    def __setitem__(self, i, T):
        n = len(self.t)
        if i == 0 or i == n - 1:
            side = "L" if i == 0 else "R"
        else:
            A = self.mps.As[i - 1]
            is_left, is_right = T.space[0] == A.space[2], T.space[0] == A.space[0]
            if is_left == is_right:
                raise ValueError(f"environment {i}: side not determined by its legs; use set(i, T, side)")
            side = "L" if is_left else "R"
        self.set(i, T, side)

    def __getitem__(self, i):
        T = self.t[i]
        if T is None:
            if self.kind[i] is None:
                raise KeyError(f"environment {i} was never set")
            if self.kind[i] == "L":
                T = self._left(self[i - 1], self.mps.As[i - 1], self.mpo.Ws[i - 1])
            else:
                T = self._right(self[i + 1], self.mps.As[i - 1], self.mpo.Ws[i - 1])
            self.recomputed += 1
            self.t[i] = T
            self._touch(i)
            self._enforce()
        else:
            self._touch(i)
        return T


def _set_env(env, i, T, side):
    """env[i] = T for a list; for an :class:`EnvStore` also records the side."""
    if isinstance(env, EnvStore):
        env.set(i, T, side)
    else:
        env[i] = T


def initialenv(mps, mpo, at: int = 1, budget: int = None):
    """``initialenv(mps, mpo; at)`` (dmrg.jl:9-36); returns env[1..L+2] as a 0-based
    list, or with ``budget`` (bytes) as an :class:`EnvStore` that keeps at most
    about that much of them in device memory and recomputes the rest on demand."""
    from .symtensors import fill
    if mps.center != at:
        raise ValueError(f"MPS center has to be {at}")
    lx = len(mps)
    env = [None] * (lx + 2) if budget is None else EnvStore(lx + 2, mps, mpo, budget)
    lVa, rVa = mps.leftspace(), mps.rightspace()
    lVw, rVw = mpo.leftspace(), mpo.rightspace()
    dt, dev = mps.As[0].dtype, mps.As[0].device
    _set_env(env, 0, fill(1.0, (lVa.dual(), lVw.dual(), lVa), dtype=dt, device=dev), "L")
    _set_env(env, lx + 1, fill(1.0, (rVa, rVw, rVa.dual()), dtype=dt, device=dev), "R")
    for l in range(lx, at - 1, -1):
        _set_env(env, l, _mpsupdateright(env[l + 1], mps.As[l - 1], mpo.Ws[l - 1]), "R")
    for l in range(1, at):
        _set_env(env, l, _mpsupdateleft(env[l - 1], mps.As[l - 1], mpo.Ws[l - 1]), "L")
    return env


def dmrg2site_window(mps, mpo, env, nbonds: int, maxdim: int = 200, tol: float = 1e-9, lanczostol: float = 1e-7,
                     krylovdim: int = 5, krylovmaxiter: int = 8, prec: str = "f64", refine_tol: float = None,
                     timing: list = None):
    """``nbonds`` consecutive left-to-right two-site updates starting at the MPS
    center c, exactly the L->R loop body of ``dmrg2sitesweep!`` (dmrg.jl:70-106):
    eigsolve on bonds (c, c+1) ... (c+nbonds-1, c+nbonds), truncate, update the
    left environments; the center ends at c + nbonds.  ``env`` must hold the
    left environments up to c-1 and the right ones from c+2 (``initialenv(at=c)``).
    With ``timing`` a list, one dict of per-stage wall milliseconds per bond is
    appended (the device is synchronised at stage boundaries)."""
    import time
    from .mpstools import svdtrunc
    from .symtensors import dot, splitleg, symmatrix
    c = mps.center
    if c < 1 or c + nbonds > len(mps):
        raise ValueError("window leaves the chain")
    Ws = mpo.Ws
    energies = []
    sync = torch.cuda.synchronize if mps.As[0].data.is_cuda else (lambda: None)
    A = mps.As[c - 1]
    for l in range(c, c + nbonds):
        sync()
        t0 = time.perf_counter()
        AA = contract(A, (1, 2, -1), mps.As[l], (-1, 3, 4))
        sync()
        t1 = time.perf_counter()
        H = HeffOperator(env[l - 1], env[l + 2], Ws[l - 1], Ws[l], prec)
        theta, v, info = eigsolve(H, AA, lanczostol, krylovdim, krylovmaxiter, refine_tol)
        sync()
        t2 = time.perf_counter()
        r = svdtrunc(symmatrix(v, [1, 2], [3, 4]), maxdim=maxdim, tol=tol)
        r.normalize_s()
        AA2 = splitleg(splitleg(r.usv(), 2, AA.space[2:4]), 1, AA.space[0:2])
        sync()
        t3 = time.perf_counter()
        energies.append(dot(AA2, _applymps2site(AA2, env[l - 1], env[l + 2], Ws[l - 1], Ws[l])))
        sync()
        t4 = time.perf_counter()
        mps.As[l - 1] = splitleg(r.u(), 1, A.space[0:2])
        _set_env(env, l, _mpsupdateleft(env[l - 1], mps.As[l - 1], Ws[l - 1]), "L")
        A = splitleg(r.sv(), 2, AA2.space[2:4])
        sync()
        t5 = time.perf_counter()
        if timing is not None:
            timing.append({"bond": l, "chi_left": AA.space[0].dim(), "chi_right": AA.space[3].dim(),
                           "kept": r.Vl.dim(), "lanczos_energy": theta, "matvecs": info.numops,
                           "ms_theta": 1e3 * (t1 - t0), "ms_eigsolve": 1e3 * (t2 - t1), "ms_svdtrunc": 1e3 * (t3 - t2),
                           "ms_energy": 1e3 * (t4 - t3), "ms_envupdate": 1e3 * (t5 - t4), "ms_total": 1e3 * (t5 - t0)})
    mps.As[c + nbonds - 1] = A
    mps.center = c + nbonds
    return energies


def dmrg2sitesweep(mps, mpo, env, maxdim: int = 200, tol: float = 1e-9, lanczostol: float = 1e-7,
                   krylovdim: int = 5, krylovmaxiter: int = 8, verbose: bool = False, prec: str = "f64",
                   refine_tol: float = None, log: list = None, stage_ms: dict = None):
    """``dmrg2sitesweep!(mps, mpo, env; ...)`` (dmrg.jl:58-179) with every contraction on the device.

    Keeps the reference's energy list including quirk Q6 (each R->L energy
    pushed twice, dmrg.jl:132,156).  ``prec`` is the H_eff arithmetic of the
    Lanczos matvecs; with ``refine_tol`` the eigenpair of a reduced-precision
    solve is refined by FP64 Lanczos cycles.  Energies after truncation are
    always FP64 matvecs.  With ``stage_ms`` a dict, the wall milliseconds of
    the eigensolves, truncations (SVD), energies and the rest (two-site tensors
    and environment updates) are accumulated into it (the device is
    synchronised at stage boundaries).
    """
    import time
    from .mpstools import svdtrunc
    from .symtensors import dot, splitleg, symmatrix
    lx = len(mps)
    if mps.center != 1:
        raise ValueError("The center of MPS has to be 1 for dmrg L->R->L")
    Ws = mpo.Ws
    energies = []
    timed = stage_ms is not None
    sync = torch.cuda.synchronize if (timed and mps.As[0].data.is_cuda) else (lambda: None)

    def stage(name, fn, *a):
        if not timed:
            return fn(*a)
        sync()
        t0 = time.perf_counter()
        out = fn(*a)
        sync()
        stage_ms[name] = stage_ms.get(name, 0.0) + 1e3 * (time.perf_counter() - t0)
        return out

    def _solve(AA, l):
        H = HeffOperator(env[l - 1], env[l + 2], Ws[l - 1], Ws[l], prec)
        theta, v, info = eigsolve(H, AA, lanczostol, krylovdim, krylovmaxiter, refine_tol)
        if log is not None:
            log.append({"bond": l, "lanczos": theta, "normres": info.normres, "matvecs": info.numops})
        if verbose:
            print(f"bond {l},{l + 1}: energy {theta} normres {info.normres} #applications {info.numops}")
        return v

    def _trunc(v, AA):
        r = svdtrunc(symmatrix(v, [1, 2], [3, 4]), maxdim=maxdim, tol=tol)
        if log is not None:
            s = r.singular_values()
            log.append({"trunc_1_minus_norm_s": 1 - r.norm_s(), "discarded_weight": max(0.0, 1 - float((s ** 2).sum()))})
        r.normalize_s()
        AA2 = splitleg(splitleg(r.usv(), 2, AA.space[2:4]), 1, AA.space[0:2])
        return r, AA2

    def _energy(AA2, l):
        return dot(AA2, _applymps2site(AA2, env[l - 1], env[l + 2], Ws[l - 1], Ws[l]))

    def solve(AA, l):
        return stage("eigsolve", _solve, AA, l)

    def trunc(v, AA):
        return stage("svdtrunc", _trunc, v, AA)

    def energy(AA2, l):
        return stage("energy", _energy, AA2, l)

    A = mps.As[0]
    for l in range(1, lx - 1):
        AA = contract(A, (1, 2, -1), mps.As[l], (-1, 3, 4))
        v = solve(AA, l)
        r, AA2 = trunc(v, AA)
        energies.append(energy(AA2, l))
        mps.As[l - 1] = splitleg(r.u(), 1, A.space[0:2])
        _set_env(env, l, _mpsupdateleft(env[l - 1], mps.As[l - 1], Ws[l - 1]), "L")
        A = splitleg(r.sv(), 2, AA2.space[2:4])
    l = lx - 1
    AA = contract(A, (1, 2, -1), mps.As[l], (-1, 3, 4))
    v = solve(AA, l)
    e = None
    for l in range(lx - 1, 1, -1):
        r, AA2 = trunc(v, AA)
        e = energy(AA2, l)
        energies.append(e)
        mps.As[l] = splitleg(r.v(), 2, AA2.space[2:4])
        _set_env(env, l + 1, _mpsupdateright(env[l + 2], mps.As[l], Ws[l]), "R")
        A = splitleg(r.us(), 1, AA2.space[0:2])
        AA = contract(mps.As[l - 2], (1, 2, -1), A, (-1, 3, 4))
        v = solve(AA, l - 1)
        energies.append(e)
    l = 1
    r, AA2 = trunc(v, AA)
    energies.append(energy(AA2, l))
    mps.As[l] = splitleg(r.v(), 2, AA2.space[2:4])
    _set_env(env, l + 1, _mpsupdateright(env[l + 2], mps.As[l], Ws[l]), "R")
    mps.As[l - 1] = splitleg(r.us(), 1, AA2.space[0:2])
    return energies
