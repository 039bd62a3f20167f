"""MPS / MPO containers, model MPOs and the SVD truncation step on the device.

Mirrors ``src/symmps.jl``, ``src/symmpo.jl`` and ``src/mpsutils.jl`` of
amiragha/mps for the parts the two-site sweep needs:

* ``MPState`` / ``MPOperator`` containers (symmps.jl:1-4, symmpo.jl:1-3);
* ``fsm_u1mpo``: U(1) MPO of a spin-1/2 Hamiltonian sum_{i<j} Jz SzSz +
  Jxy/2 (S+S- + S-S+) as a finite-state machine; for nearest-neighbour
  chains it reproduces ``xxz_u1mpo`` (symmpo.jl:45-97) block for block;
  ``heisenberg_cylinder_mpo`` builds the width-W cylinder in snake order;
* ``svd_`` / ``truncatedsizes`` / ``svdtrunc`` (symtools.jl:1-35,
  mpsutils.jl:89-179) with all FP64 per-sector SVDs of a matrix in one
  batched device call (``mpsh_svd_*``: rocSOLVER gesdd over padded shape
  buckets) and the reference's greedy global truncation;
* ``random_mps``, ``center_at``, ``normalize``, ``measure``.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from ._lib import check, lib
from .symtensors import (PlanLRU, SymTensor, U1Space, _ptr, _stream, contract, default_device, dual, fuselegs,
                         get_struct, splitleg)
from .workloads import binom_capped_profile, gaussian


class MPState:
    """``MPState`` (symmps.jl:1-4): site tensors (l, p, r̄) and the orthogonality center."""

    def __init__(self, As: List[SymTensor], center: int = 0):
        self.As = list(As)
        self.center = center

    def __len__(self):
        return len(self.As)

    def leftspace(self):
        return self.As[0].space[0]

    def rightspace(self):
        return self.As[-1].space[2].dual()

    def bonddims(self):
        return [A.space[2].dim() for A in self.As[:-1]]


class MPOperator:
    """``MPOperator`` (symmpo.jl:1-3): site tensors (wl, p_out, w̄r, p̄_in)."""

    def __init__(self, Ws: List[SymTensor]):
        self.Ws = list(Ws)

    def __len__(self):
        return len(self.Ws)

    def leftspace(self):
        return self.Ws[0].space[0]

    def rightspace(self):
        return self.Ws[-1].space[2].dual()


# ---------------------------------------------------------------------------
# finite-state-machine U(1) MPO for spin-1/2 two-site terms
# ---------------------------------------------------------------------------
_SZ = {(0, 0): -0.5, (1, 1): 0.5}
_SP = {(1, 0): 1.0}
_SM = {(0, 1): 1.0}
_ID = {(0, 0): 1.0, (1, 1): 1.0}


def fsm_u1mpo(L: int, terms: Sequence[Tuple[int, int, float, float]], dtype=torch.float64, device=None,
              fields: Dict[int, float] = None, const: float = 0.0) -> MPOperator:
    """MPO of H = sum over terms (i, j, Jz, Jxy), 1 <= i < j <= L, of Jz Sz_i Sz_j + Jxy/2 (S+_i S-_j + h.c.),
    plus on-site fields sum_i h_i Sz_i and a constant (both on the before -> after channel).

    Bond states per charge: 0 -> [after, Sz channels by source site, before],
    +1 -> S+ channels, -1 -> S- channels; the two boundary bonds keep only
    "before" (left) and "after" (right).  Physical charge 1 is spin up.
    """
    partners: Dict[int, Dict[int, Tuple[float, float]]] = {}
    for i, j, jz, jxy in terms:
        if not (1 <= i < j <= L):
            raise ValueError(f"bad term sites ({i}, {j})")
        a, b = partners.setdefault(i, {}).get(j, (0.0, 0.0))
        partners[i][j] = (a + jz, b + jxy)
    maxp = {i: max(ps) for i, ps in partners.items()}
    has_z = {i: any(v[0] != 0.0 for v in ps.values()) for i, ps in partners.items()}
    has_xy = {i: any(v[1] != 0.0 for v in ps.values()) for i, ps in partners.items()}

    def bond_states(k):
        """ordered list of (charge, label) for bond k (between sites k and k+1)."""
        if k == 0:
            return [(0, "before")]
        if k == L:
            return [(0, "after")]
        srcs = sorted(i for i in partners if i <= k < maxp[i])
        st = [(0, "after")] + [(0, ("z", i)) for i in srcs if has_z[i]] + [(0, "before")]
        st += [(1, ("p", i)) for i in srcs if has_xy[i]]
        st += [(-1, ("m", i)) for i in srcs if has_xy[i]]
        return st

    bonds = [bond_states(k) for k in range(L + 1)]

    def space_of(states):
        cnt: Dict[int, int] = {}
        for c, _ in states:
            cnt[c] = cnt.get(c, 0) + 1
        return U1Space(*cnt.items())

    def index_of(states):
        pos, cnt = {}, {}
        for c, lab in states:
            pos[lab] = (c, cnt.get(c, 0))
            cnt[c] = cnt.get(c, 0) + 1
        return pos

    Vd = U1Space((0, 1), (1, 1))
    Ws = []
    for s in range(1, L + 1):
        left, right = bonds[s - 1], bonds[s]
        Vl, Vr = space_of(left), space_of(right)
        li, ri = index_of(left), index_of(right)
        # the last tensor keeps a non-dual charge-0 boundary leg, like Wend (symmpo.jl:55)
        space = (Vl, Vd, Vr if s == L else Vr.dual(), Vd.dual())
        st = get_struct(0, space)
        sec, siz, off = st.tables()
        flat = np.zeros(st.numel)
        bidx = {tuple(int(x) for x in r): b for b, r in enumerate(sec)}

        def put(lab_l, lab_r, op, coef):
            if lab_l not in li or lab_r not in ri or coef == 0.0:
                return
            (cl, il), (cr, ir) = li[lab_l], ri[lab_r]
            for (so, si), val in op.items():
                b = bidx[(cl, so, cr, si)]
                dl, dr = int(siz[b][0]), int(siz[b][2])
                flat[off[b] + il + dl * (0 + 1 * (ir + dr * 0))] += coef * val

        put("before", "before", _ID, 1.0)
        put("after", "after", _ID, 1.0)
        if fields and fields.get(s, 0.0) != 0.0:
            put("before", "after", _SZ, fields[s])
        if const != 0.0 and s == 1:
            put("before", "after", _ID, const)
        if s in partners:
            if has_xy[s]:
                put("before", ("p", s), _SP, 1.0)
                put("before", ("m", s), _SM, 1.0)
            if has_z[s]:
                put("before", ("z", s), _SZ, 1.0)
        for i in partners:
            if i < s < maxp[i]:
                for t in ("p", "m", "z"):
                    put((t, i), (t, i), _ID, 1.0)
            if i < s and s in partners[i]:
                jz, jxy = partners[i][s]
                put(("p", i), "after", _SM, 0.5 * jxy)
                put(("m", i), "after", _SP, 0.5 * jxy)
                put(("z", i), "after", _SZ, jz)
        Ws.append(SymTensor.from_flat(0, space, flat, dtype, device))
    return MPOperator(Ws)


def tv_u1mpo(L: int, t: float = 1.0, V: float = 2.0, dtype=torch.float64, device=None) -> MPOperator:
    """Spinless-fermion t-V chain H = -t sum (c+_i c_{i+1} + h.c.) + V sum n_i n_{i+1}, open,
    U(1) particle number.  Jordan-Wigner (occupied = physical charge 1; nearest-neighbour
    hopping carries no string): c+_i c_{i+1} + h.c. = S+_i S-_{i+1} + h.c. and
    n_i n_j = Sz_i Sz_j + (Sz_i + Sz_j)/2 + 1/4, i.e. Jxy = -2t, Jz = V, fields V/2 per bond
    end, constant V/4 per bond.  (No reference counterpart: the reference has no t-V MPO.)"""
    fields: Dict[int, float] = {}
    for i in range(1, L):
        fields[i] = fields.get(i, 0.0) + V / 2
        fields[i + 1] = fields.get(i + 1, 0.0) + V / 2
    return fsm_u1mpo(L, [(i, i + 1, V, -2.0 * t) for i in range(1, L)], dtype, device, fields=fields,
                     const=V / 4 * (L - 1))


def xxz_u1mpo(lx: int, delta: float = 1.0, dtype=torch.float64, device=None) -> MPOperator:
    """``xxz_u1mpo(T, lx, d, delta)`` (symmpo.jl:45-97): open XXZ chain, D = 5."""
    return fsm_u1mpo(lx, [(i, i + 1, delta, 1.0) for i in range(1, lx)], dtype, device)


def cylinder_sites(W: int, Lx: int, ordering: str = "snake"):
    """Site numbering (1-based) of a W x Lx cylinder: x major; 'snake' reverses odd columns."""
    idx = {}
    for x in range(Lx):
        for y in range(W):
            yy = y if (ordering != "snake" or x % 2 == 0) else W - 1 - y
            idx[(x, y)] = x * W + yy + 1
    return idx


def heisenberg_cylinder_bonds(W: int, Lx: int, ordering: str = "snake"):
    """Nearest-neighbour bonds (i < j) of a cylinder periodic in y (width W), open in x."""
    idx = cylinder_sites(W, Lx, ordering)
    bonds = set()
    for x in range(Lx):
        for y in range(W):
            a = idx[(x, y)]
            if W > 2 or y + 1 < W:
                b = idx[(x, (y + 1) % W)]
                if a != b:
                    bonds.add((min(a, b), max(a, b)))
            if x + 1 < Lx:
                b = idx[(x + 1, y)]
                bonds.add((min(a, b), max(a, b)))
    return sorted(bonds)


def heisenberg_cylinder_mpo(W: int, Lx: int, J: float = 1.0, ordering: str = "snake", dtype=torch.float64,
                            device=None) -> MPOperator:
    """Heisenberg J1 model on a W x Lx cylinder as a U(1) MPO (channels per open source site)."""
    return fsm_u1mpo(W * Lx, [(i, j, J, J) for i, j in heisenberg_cylinder_bonds(W, Lx, ordering)], dtype, device)


# ---------------------------------------------------------------------------
# SVD step   (symtools.jl:1-35, mpsutils.jl:89-179)
# ---------------------------------------------------------------------------
class _SvdPlan:
    """Cached batched-SVD plan of one rank-2 structure (``mpsh_svd_*``)."""

    def __init__(self, st):
        sec, siz, off = st.tables()
        nb = st.nblocks
        self.keys = [(int(sec[b][0]), int(sec[b][1])) for b in range(nb)]
        self.m = [int(siz[b][0]) for b in range(nb)]
        self.n = [int(siz[b][1]) for b in range(nb)]
        self.handle = L._p()
        check(lib().mpsh_svd_plan(nb, L.i64arr(self.m), L.i64arr(self.n), L.i64arr([int(x) for x in off[:nb]]),
                                  self.handle))
        k, so, uo, vo, tot = (L._i64 * max(1, nb))(), (L._i64 * max(1, nb))(), (L._i64 * max(1, nb))(), \
            (L._i64 * max(1, nb))(), (L._i64 * 3)()
        check(lib().mpsh_svd_sizes(self.handle, k, so, uo, vo, tot))
        self.k = [int(k[b]) for b in range(nb)]
        self.s_off, self.u_off, self.v_off = [list(map(int, x[:nb])) for x in (so, uo, vo)]
        self.totals = [int(t) for t in tot]
        self.ws_bytes = int(lib().mpsh_svd_workspace(self.handle))

    def __del__(self):
        try:
            lib().mpsh_svd_destroy(self.handle)
        except Exception:
            pass


SVD_PLAN_CACHE = 512
_SVD_PLANS: Dict[object, _SvdPlan] = PlanLRU(SVD_PLAN_CACHE)


def _svd_blocks(M: SymTensor):
    """Every per-sector FP64 SVD of a rank-2 SymTensor in one batched device call
    (symtools.jl:1-35): [(sector, U (m x k), S (k), Vh (k x n))], views of packed buffers."""
    st = M.struct
    P = _SVD_PLANS.get(st)
    if P is None:
        P = _SVD_PLANS[st] = _SvdPlan(st)
    x = M.data if M.dtype == torch.float64 else M.data.double()
    dev = x.device
    S = torch.empty(max(1, P.totals[0]), dtype=torch.float64, device=dev)
    U = torch.empty(max(1, P.totals[1]), dtype=torch.float64, device=dev)
    V = torch.empty(max(1, P.totals[2]), dtype=torch.float64, device=dev)
    ws = torch.empty(max(16, P.ws_bytes), dtype=torch.uint8, device=dev)
    check(lib().mpsh_svd_execute(P.handle, _ptr(x), _ptr(S), _ptr(U), _ptr(V), _ptr(ws), _stream()))
    out = []
    for b, key in enumerate(P.keys):
        m, n, k = P.m[b], P.n[b], P.k[b]
        out.append((key, U[P.u_off[b]:P.u_off[b] + m * k].view(k, m).t(), S[P.s_off[b]:P.s_off[b] + k],
                    V[P.v_off[b]:P.v_off[b] + k * n].view(n, k).t()))
    return out


def truncatedsizes(S: List[np.ndarray], maxdim: int = 200, tol: float = 1e-14) -> List[int]:
    """``truncatedsizes`` (mpsutils.jl:89-112): greedy largest-first, absolute ``tol``.

    The greedy loop takes, one at a time, the largest remaining head of the
    per-sector descending lists (first sector on ties) until ``maxdim`` values
    are taken or the head drops below ``tol``.  That is the prefix of all
    values sorted by (value descending, sector, position), which is how it is
    computed here for ``tol > 0``; ``tol <= 0`` keeps the literal loop (an
    exhausted sector then keeps being chosen at value 0, as in the reference).
    """
    if not S:
        return []
    total = sum(len(s) for s in S)
    if tol <= 0:
        pointers = [1] * len(S)
        c = [float(s[0]) for s in S]
        for n in range(1, min(maxdim, total) + 1):
            index = int(np.argmax(c))
            if c[index] < tol:
                if n == 1:
                    raise ValueError(f"The largest singular value is smaller than tol! {c[index]} < {tol}")
                break
            pointers[index] += 1
            c[index] = float(S[index][pointers[index] - 1]) if pointers[index] - 1 < len(S[index]) else 0.0
        return [p - 1 for p in pointers]
    vals = np.concatenate([np.asarray(s, dtype=np.float64) for s in S])
    blk = np.concatenate([np.full(len(s), i) for i, s in enumerate(S)])
    pos = np.concatenate([np.arange(len(s)) for s in S])
    order = np.lexsort((pos, blk, -vals))[:min(maxdim, total)]
    below = np.nonzero(vals[order] < tol)[0]
    if len(below) and below[0] == 0:
        raise ValueError(f"The largest singular value is smaller than tol! {vals[order[0]]} < {tol}")
    take = order[:below[0]] if len(below) else order
    return [int(x) for x in np.bincount(blk[take], minlength=len(S))]


def _root(t: torch.Tensor) -> torch.Tensor:
    return t if t._base is None else t._base


def _gather_table(blocks, src: torch.Tensor, sc: torch.Tensor, smode: int) -> Tuple[np.ndarray, int]:
    """``mpsh_gather_blocks`` table for blocks [(matrix view of src, scale view of sc or None, dst_off)]."""
    nb = len(blocks)
    rec = np.zeros((nb, L.GATHER_FIELDS), dtype=np.int64)
    tiles = []
    for b, (x, sv, doff) in enumerate(blocks):
        rows, cols = x.shape
        rec[b] = ((x.data_ptr() - src.data_ptr()) // 8, x.stride(0), x.stride(1), rows, cols, doff,
                  0 if sv is None else (sv.data_ptr() - sc.data_ptr()) // 8, 0 if sv is None else smode)
        nt = (rows * cols + L.GATHER_TILE - 1) // L.GATHER_TILE
        tiles.append((b << 32) + np.arange(nt, dtype=np.int64))
    t = np.concatenate(tiles) if tiles else np.zeros(0, dtype=np.int64)
    return np.concatenate([rec.reshape(-1), t]), len(t)


def _assemble(charge, space, parts: Dict[Tuple[int, int], torch.Tensor], dtype, device,
              scales: Dict[Tuple[int, int], torch.Tensor] = None, smode: int = 0) -> SymTensor:
    """Pack column-major blocks (dict sector -> matrix) into a SymTensor of (charge, space);
    with ``scales``, each block's rows (smode 1) or columns (smode 2) are multiplied by
    the sector's vector first.  FP64 views of one device buffer (the SVD's packed
    factors) are gathered in one ``mpsh_gather_blocks`` launch."""
    st = get_struct(charge, space)
    sec, siz, off = st.tables()
    keys = [(int(sec[b][0]), int(sec[b][1])) for b in range(st.nblocks)]
    mats = [parts[k] for k in keys]
    scs = [None] * len(keys) if scales is None else [scales[k] for k in keys]
    src = _root(mats[0]) if mats else None
    sc = _root(scs[0]) if mats and scales is not None else None
    fast = (bool(mats) and dtype == torch.float64 and src.is_cuda
            and all(x.dtype == torch.float64 and _root(x) is src for x in mats)
            and (scales is None or all(v.dtype == torch.float64 and _root(v) is sc for v in scs)))
    if fast:
        for b, x in enumerate(mats):
            if x.shape[0] * x.shape[1] != off[b + 1] - off[b]:
                raise ValueError(f"block {keys[b]} has shape {tuple(x.shape)}, the structure expects "
                                 f"{off[b + 1] - off[b]} elements")
        table, ntiles = _gather_table(list(zip(mats, scs, off[:-1])), src, sc, smode)
        dev_table = torch.from_numpy(table).pin_memory().to(src.device, non_blocking=True)
        data = torch.empty(st.numel, dtype=dtype, device=src.device)
        check(lib().mpsh_gather_blocks(len(mats), ntiles, _ptr(dev_table), _ptr(src),
                                       None if sc is None else _ptr(sc), _ptr(data), _stream()))
        return SymTensor(charge, space, data, st)
    data = torch.empty(st.numel, dtype=dtype, device=device)
    for b in range(st.nblocks):
        blk = mats[b]
        if scales is not None:
            blk = scs[b][:, None] * blk if smode == 1 else blk * scs[b]
        data[off[b]:off[b + 1]] = blk.t().reshape(-1)
    return SymTensor(charge, space, data, st)


class SVDResult:
    """u, s, v of ``svdtrunc`` kept as device blocks; products are formed on request.

    The kept singular values are views of one device vector (``S_pack``) and
    are also held on the host (``S_host``, copied once for the truncation
    decision), so norms and normalisation need no device round trips."""

    def __init__(self, M: SymTensor, keys, U, S, Vh, Vl: U1Space, Vr: U1Space, S_pack=None, S_host=None):
        self.M, self.keys, self.U, self.S, self.Vh, self.Vl, self.Vr = M, keys, U, S, Vh, Vl, Vr
        self.S_pack, self.S_host = S_pack, S_host
        self._cache = {}

    def singular_values(self) -> np.ndarray:
        if not self.S:
            return np.zeros(0)
        if self.S_host is not None:
            return np.sort(np.concatenate(self.S_host))[::-1]
        return np.sort(torch.cat(self.S).double().cpu().numpy())[::-1]

    def norm_s(self) -> float:
        if self.S_host is not None:
            return math.sqrt(sum(float(np.dot(s, s)) for s in self.S_host))
        return math.sqrt(sum(float((s.double() ** 2).sum()) for s in self.S))

    def normalize_s(self):
        """``normalize!(s)`` (symtensor.jl:325)."""
        nrm = self.norm_s()
        if self.S_pack is not None:
            self.S_pack.div_(nrm)  # the per-block views follow
            self.S_host = [s / nrm for s in self.S_host]
        else:
            self.S = [s / nrm for s in self.S]
        self._cache.clear()

    def u(self) -> SymTensor:
        if "u" not in self._cache:
            self._cache["u"] = _assemble(0, (self.M.space[0], self.Vl.dual()),
                                         {(k[0], k[0]): U for k, U in zip(self.keys, self.U)}, self.M.dtype,
                                         self.M.device)
        return self._cache["u"]

    def v(self) -> SymTensor:
        return _assemble(0, (self.Vr.dual(), self.M.space[1]),
                         {(k[1], k[1]): V for k, V in zip(self.keys, self.Vh)}, self.M.dtype, self.M.device)

    def us(self) -> SymTensor:
        """u*s (charge of s, legs (space1, Vr))."""
        return _assemble(self.M.charge, (self.M.space[0], self.Vr),
                         {(k[0], k[1]): U for k, U in zip(self.keys, self.U)}, self.M.dtype, self.M.device,
                         scales={(k[0], k[1]): s for k, s in zip(self.keys, self.S)}, smode=2)

    def sv(self) -> SymTensor:
        """s*v (legs (Vl, space2))."""
        if "sv" not in self._cache:
            self._cache["sv"] = _assemble(self.M.charge, (self.Vl, self.M.space[1]),
                                          {(k[0], k[1]): V for k, V in zip(self.keys, self.Vh)}, self.M.dtype,
                                          self.M.device,
                                          scales={(k[0], k[1]): s for k, s in zip(self.keys, self.S)}, smode=1)
        return self._cache["sv"]

    def usv(self) -> SymTensor:
        """u*s*v: same structure as the factorised matrix (all its sectors; dropped
        ones are zero), one device contraction of u and s*v."""
        if not self.keys:
            data = torch.zeros(self.M.struct.numel, dtype=self.M.dtype, device=self.M.device)
            return SymTensor(self.M.charge, self.M.space, data, self.M.struct)
        return contract(self.u(), (1, -1), self.sv(), (-1, 2))


def _pack_s(facts, dt):
    """The blocks' singular values as one device vector of dtype dt (S_pack),
    its block offsets, and the FP64 values on the host (one copy)."""
    if not facts:
        return None, [0], []
    S_cat = torch.cat([f[2] for f in facts])
    offs = [0] + np.cumsum([len(f[2]) for f in facts]).tolist()
    svals = np.split(S_cat.cpu().numpy(), offs[1:-1])
    return (S_cat if S_cat.dtype == dt else S_cat.to(dt)), offs, svals


def svdtrunc(M: SymTensor, maxdim: int = None, tol: float = 1e-12) -> SVDResult:
    """``svdtrunc(A; maxdim, tol)`` (mpsutils.jl:116-179) on a rank-2 SymTensor."""
    if maxdim is None:
        maxdim = max(V.dim() for V in M.space)
    facts = _svd_blocks(M)
    dt = M.dtype
    S_pack, offs, svals = _pack_s(facts, dt)
    ns = truncatedsizes(svals, maxdim=maxdim, tol=tol) if facts else []
    keys, Us, Ss, Vs, S_host, midl, midr = [], [], [], [], [], [], []
    for b, ((key, U, S, Vh), k) in enumerate(zip(facts, ns)):
        if k <= 0:
            continue
        keys.append(key)
        Us.append(U[:, :k].to(dt))
        Ss.append(S_pack[offs[b]:offs[b] + k])
        S_host.append(svals[b][:k].copy())
        Vs.append(Vh[:k, :].to(dt))
        midl.append((key[0], k))
        midr.append((key[1], k))
    Vl = U1Space(*midl, isdual=M.space[0].isdual)
    Vr = U1Space(*midr, isdual=M.space[1].isdual)
    return SVDResult(M, keys, Us, Ss, Vs, Vl, Vr, S_pack=S_pack, S_host=S_host)


def svd_(M: SymTensor) -> SVDResult:
    """``_svd_(A)`` (symtools.jl:1-35): untruncated per-sector SVD."""
    keys, Us, Ss, Vs, S_host, midl, midr = [], [], [], [], [], [], []
    dt = M.dtype
    facts = _svd_blocks(M)
    S_pack, offs, svals = _pack_s(facts, dt)
    for b, (key, U, S, Vh) in enumerate(facts):
        keys.append(key)
        Us.append(U.to(dt))
        Ss.append(S_pack[offs[b]:offs[b + 1]])
        S_host.append(svals[b].copy())
        Vs.append(Vh.to(dt))
        midl.append((key[0], S.numel()))
        midr.append((key[1], S.numel()))
    Vl = U1Space(*midl, isdual=M.space[0].isdual)
    Vr = U1Space(*midr, isdual=M.space[1].isdual)
    return SVDResult(M, keys, Us, Ss, Vs, Vl, Vr, S_pack=S_pack, S_host=S_host if facts else None)


# ---------------------------------------------------------------------------
# MPS construction and gauge   (symmps.jl:44-70, 150-200, 354-371)
# ---------------------------------------------------------------------------
def bond_spaces(L: int, N: int, chi: int, nsec: int = None, dmax: int = None) -> List[U1Space]:
    """U(1) bond legs 0..L at total up-count N with Gaussian sector profiles capped at chi."""
    out = [U1Space((0, 1))]
    for l in range(1, L):
        out.append(U1Space(*binom_capped_profile(l, L, N, chi, dmax=dmax, nsec=nsec)))
    out.append(U1Space((N, 1)))
    return out


def state_footprint(L: int, N: int, chi: int, mpo_leg: U1Space, nsec: int = None, dmax: int = None,
                    itemsize: int = 8) -> Dict[str, int]:
    """Bytes of an MPS on bond_spaces(L, N, chi, nsec, dmax) and of its full set of
    environments (legs (bond dual, MPO bond dual, bond), src/dmrg.jl:9-36) with
    the bulk MPO bond leg ``mpo_leg`` (the boundary environments have a trivial
    MPO leg), from the structures alone (host-only)."""
    bs = bond_spaces(L, N, chi, nsec, dmax)
    Vd = U1Space((0, 1), (1, 1))
    triv = U1Space((0, 1))
    env = mps = largest = 0
    for l in range(L + 1):
        Vw = triv if l in (0, L) else mpo_leg
        n = get_struct(0, (bs[l].dual(), Vw.dual(), bs[l])).numel
        env += n
        largest = max(largest, n)
    for l in range(1, L + 1):
        mps += get_struct(0, (bs[l - 1], Vd, bs[l].dual())).numel
    return {"mps": mps * itemsize, "environments": env * itemsize, "largest_environment": largest * itemsize}


def random_mps(L: int, N: int, chi: int, seed: int = 1911, nsec: int = None, dmax: int = None,
               dtype=torch.float64, device=None) -> MPState:
    """Seeded Gaussian U(1) MPS (legs (l, p, r̄), charge 0 per site), right-canonical at 1, normalised."""
    device = device or default_device()
    Vd = U1Space((0, 1), (1, 1))
    bs = bond_spaces(L, N, chi, nsec, dmax)
    As = []
    for l in range(1, L + 1):
        space = (bs[l - 1], Vd, bs[l].dual())
        st = get_struct(0, space)
        if st.numel == 0:
            raise ValueError(f"empty site tensor at {l}: widen the sector window")
        As.append(SymTensor.from_flat(0, space, gaussian(seed, 1000 + l, st.numel), dtype, device))
    # right-canonicalise (center_at(mps, 1) of symmps.jl:185-200), rescaling the
    # moving center to unit norm after every step: the product of L random
    # Gaussian tensors otherwise overflows FP64 at large chi.  Only the overall
    # scale changes, and normalize() fixes it below.
    from .symtensors import norm, rmul_
    mps = MPState(As, 0)
    for p in range(L, 1, -1):
        _pushleft(mps, p)
        rmul_(mps.As[p - 2], 1.0 / norm(mps.As[p - 2]))
    mps.center = 1
    normalize(mps)
    return mps


def _pushleft(mps: MPState, l: int):
    """``_pushleft!`` (symmps.jl:167-174)."""
    A = mps.As[l - 1]
    r = svd_(fuselegs(A, 2, 2))
    mps.As[l - 1] = splitleg(r.v(), 2, A.space[1:3])
    mps.As[l - 2] = contract(mps.As[l - 2], (1, 2, -1), r.us(), (-1, 3))


def _pushright(mps: MPState, l: int):
    """``_pushright!`` (symmps.jl:176-183)."""
    A = mps.As[l - 1]
    r = svd_(fuselegs(A, 1, 2))
    mps.As[l - 1] = splitleg(r.u(), 1, A.space[0:2])
    mps.As[l] = contract(r.sv(), (1, -1), mps.As[l], (-1, 2, 3))


def center_at(mps: MPState, center: int) -> MPState:
    """``center_at!`` (symmps.jl:185-200)."""
    x1, xL = 1, len(mps)
    if mps.center > 0:
        x1 = xL = mps.center
    for p in range(x1, center):
        _pushright(mps, p)
    for p in range(xL, center, -1):
        _pushleft(mps, p)
    mps.center = center
    return mps


def normalize(mps: MPState) -> MPState:
    """``normalize!(mps)`` (symmps.jl:150-165)."""
    if mps.center < 1:
        center_at(mps, 1)
    A = mps.As[mps.center - 1]
    if mps.center > len(mps) // 2:
        r = svd_(fuselegs(A, 1, 2, False))
        r.normalize_s()
        mps.As[mps.center - 1] = splitleg(r.usv(), 1, A.space[0:2])
    else:
        r = svd_(fuselegs(A, 2, 2, True))
        r.normalize_s()
        mps.As[mps.center - 1] = splitleg(r.usv(), 2, A.space[1:3])
    return mps


def measure(mps: MPState, mpo: MPOperator) -> float:
    """``measure(mps, mpo)`` (symmps.jl:354-371): <psi|H|psi> by a left sweep of device contracts."""
    from .symtensors import fill
    dummy = U1Space((0, 1))
    Lt = fill(1.0, (dummy, dummy, dummy), dtype=mps.As[0].dtype, device=mps.As[0].device)
    for A, W in zip(mps.As, mpo.Ws):
        Lt = contract(contract(contract(Lt, (-1, 3, 4), A, (-1, 2, 1)), (1, -1, -2, 4), W, (-2, 3, 2, -1)),
                      (1, 2, -1, -2), dual(A), (-2, -1, 3))
    return float(contract(Lt, (-1, -2, -3), fill(1.0, (dummy, dummy, dummy), dtype=Lt.dtype, device=Lt.device),
                          (-1, -2, -3)))
