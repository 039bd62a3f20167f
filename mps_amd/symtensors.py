"""U(1) block-sparse tensors on the GPU: the SymTensors API of the reference.

Names, 1-based leg numbers, label conventions and error behaviour follow
amiragha/mps ``src/SymTensors.jl``.  A :class:`SymTensor` holds its charge,
its legs and ONE packed device buffer (all blocks, column-major sector
order, each block column-major); sector bookkeeping lives in libmpsh
(integer, exact) and every data operation is a HIP kernel behind the C ABI.
"""
from __future__ import annotations

import math
from collections import OrderedDict
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from ._lib import check, i32arr, i64arr, u8arr, lib


# ---------------------------------------------------------------------------
# legs   (src/vspace.jl)
# ---------------------------------------------------------------------------
class U1Space:
    """``VectorSpace{U1Charge}``: sorted charge->dim map and ``isdual`` (vspace.jl:5-18)."""

    __slots__ = ("charges", "dims", "isdual", "_key")

    def __init__(self, *pairs, isdual: bool = False):
        acc = {}
        for c, d in pairs:
            c, d = int(c), int(d)
            if c in acc:
                raise ValueError(f"duplicate charges! {c}")
            if d <= 0:
                raise ValueError("space dimension can only be positive!")
            acc[c] = d
        cs = sorted(acc)
        self.charges = tuple(cs)
        self.dims = tuple(acc[c] for c in cs)
        self.isdual = bool(isdual)
        self._key = (self.charges, self.dims, self.isdual)

    @classmethod
    def from_pairs(cls, pairs, isdual=False):
        return cls(*pairs, isdual=isdual)

    def items(self):
        return zip(self.charges, self.dims)

    def dim(self, c=None) -> int:
        if c is None:
            return sum(self.dims)
        try:
            return self.dims[self.charges.index(int(c))]
        except ValueError:
            return 0

    def dual(self) -> "U1Space":
        return U1Space(*self.items(), isdual=not self.isdual)

    def key(self):
        return self._key

    def __eq__(self, o):
        return isinstance(o, U1Space) and self._key == o._key

    def __hash__(self):
        return hash(self._key)

    def __repr__(self):
        body = ", ".join(f"{c}=>{d}" for c, d in self.items())
        return f"U1Space{'†' if self.isdual else ''}({body})"


def dual_space(V: U1Space) -> U1Space:
    return V.dual()


def _leg_arrays(spaces: Sequence[U1Space]):
    isd = [1 if V.isdual else 0 for V in spaces]
    nch = [len(V.charges) for V in spaces]
    ch = [c for V in spaces for c in V.charges]
    dm = [d for V in spaces for d in V.dims]
    return u8arr(isd), i32arr(nch), i64arr(ch), i64arr(dm)


# ---------------------------------------------------------------------------
# bounded plan caches
# ---------------------------------------------------------------------------
class PlanLRU(OrderedDict):
    """A dict that keeps at most ``maxsize`` entries, dropping the least recently
    used (``get`` and assignment count as use).  Plan entries own their library
    handles (:class:`PlanHandle`), so a dropped plan releases its tables; a sweep
    meets new bond structures at every truncation, so unbounded caches would
    grow with every sweep."""

    def __init__(self, maxsize: int):
        super().__init__()
        self.maxsize = int(maxsize)

    def get(self, key, default=None):
        if key in self:
            self.move_to_end(key)
            return self[key]
        return default

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        self.move_to_end(key)
        while len(self) > self.maxsize:
            self.popitem(last=False)


class PlanHandle:
    """Owner of one plan handle of the library; ``destroy`` names its
    ``mpsh_*_destroy`` entry point, called when the owner is collected."""

    __slots__ = ("h", "destroy")

    def __init__(self, h, destroy: str):
        self.h = h
        self.destroy = destroy

    def __del__(self):
        try:
            if self.h:
                getattr(lib(), self.destroy)(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# structures (interned by value)
# ---------------------------------------------------------------------------
class Struct:
    """Owns an ``mpsh_struct``: the sorted sector table of (charge, spaces)."""

    def __init__(self, handle, charge: int, spaces: Tuple[U1Space, ...]):
        self.handle = handle
        self.charge = int(charge)
        self.spaces = tuple(spaces)
        self.rank = len(spaces)
        self.nblocks = int(lib().mpsh_struct_nblocks(handle))
        self.numel = int(lib().mpsh_struct_numel(handle))
        self._tables = None

    @property
    def key(self):
        return (self.charge, tuple(V.key() for V in self.spaces))

    def tables(self):
        """(sectors [nb, rank], sizes [nb, rank], offsets [nb+1]) as int64 numpy arrays."""
        if self._tables is None:
            nb, r = self.nblocks, self.rank
            s = np.zeros(max(1, nb * r), dtype=np.int64)
            z = np.zeros(max(1, nb * r), dtype=np.int64)
            o = np.zeros(nb + 1, dtype=np.int64)
            P = lambda a: a.ctypes.data_as(L._pi64)
            check(lib().mpsh_struct_blocks(self.handle, P(s), P(z), P(o)))
            self._tables = (s[: nb * r].reshape(nb, r), z[: nb * r].reshape(nb, r), o)
        return self._tables

    def sectors(self):
        return [tuple(int(x) for x in row) for row in self.tables()[0]]

    def __del__(self):
        try:
            if self.handle:
                lib().mpsh_struct_destroy(self.handle)
        except Exception:
            pass


# interned structures (host tables only; a dropped entry lives on in the tensors using it)
_STRUCTS: Dict = PlanLRU(8192)


def _struct_from_handle(handle) -> Struct:
    """Wrap a structure returned by a plan; reuse the interned one when equal."""
    r = int(lib().mpsh_struct_rank(handle))
    spaces = []
    for leg in range(r):
        n = int(lib().mpsh_struct_leg_ncharges(handle, leg))
        isd = (L._u8 * 1)()
        ch = np.zeros(n, dtype=np.int64)
        dm = np.zeros(n, dtype=np.int64)
        check(lib().mpsh_struct_leg(handle, leg, isd, ch.ctypes.data_as(L._pi64), dm.ctypes.data_as(L._pi64)))
        spaces.append(U1Space(*zip(ch.tolist(), dm.tolist()), isdual=bool(isd[0])))
    charge = int(lib().mpsh_struct_charge(handle))
    key = (charge, tuple(V.key() for V in spaces))
    if key in _STRUCTS:
        lib().mpsh_struct_destroy(handle)
        return _STRUCTS[key]
    s = Struct(handle, charge, tuple(spaces))
    _STRUCTS[key] = s
    return s


def get_struct(charge: int, spaces: Sequence[U1Space]) -> Struct:
    spaces = tuple(spaces)
    key = (int(charge), tuple(V.key() for V in spaces))
    s = _STRUCTS.get(key)
    if s is None:
        h = L._p()
        isd, nch, ch, dm = _leg_arrays(spaces)
        check(lib().mpsh_struct_create(int(charge), len(spaces), isd, nch, ch, dm, h))
        s = Struct(h, charge, spaces)
        _STRUCTS[key] = s
    return s


def allsectorsandsizes(charge: int, spaces: Sequence[U1Space]):
    """``_allsectorsandsizes`` (src/sector.jl:69-90), computed by libmpsh."""
    s = get_struct(charge, spaces)
    sec, siz, _ = s.tables()
    return [tuple(int(x) for x in r) for r in sec], [tuple(int(x) for x in r) for r in siz]


# ---------------------------------------------------------------------------
# device helpers
# ---------------------------------------------------------------------------
_DT = {torch.float64: L.MPSH_F64, torch.float32: L.MPSH_F32}


def _stream():
    """The raw handle of torch's current stream on the current device (the
    low-level accessors: this runs several times per matvec)."""
    return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())


def _ptr(t: torch.Tensor):
    return t.data_ptr()


def default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("mps_amd needs a HIP device (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


# ---------------------------------------------------------------------------
# SymTensor
# ---------------------------------------------------------------------------
class SymTensor:
    """``SymTensor{U1Charge,T,N}`` (src/symtensor.jl:12-30) with one packed device buffer."""

    __slots__ = ("charge", "space", "data", "struct")

    def __init__(self, charge: int, space: Sequence[U1Space], data: torch.Tensor, struct: Struct = None):
        self.charge = int(charge)
        self.space = tuple(space)
        self.struct = struct or get_struct(self.charge, self.space)
        if data.dim() != 1 or data.numel() != self.struct.numel:
            raise ValueError(f"data must be 1-D with {self.struct.numel} elements, got {tuple(data.shape)}")
        if data.dtype not in _DT:
            raise TypeError("SymTensor data must be float64 or float32")
        self.data = data

    # -- basic properties ---------------------------------------------------
    @property
    def rank(self):
        return len(self.space)

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def device(self):
        return self.data.device

    def size(self):
        return tuple(V.dim() for V in self.space)

    def sectors(self):
        return self.struct.sectors()

    def blocks(self) -> Dict[Tuple[int, ...], np.ndarray]:
        """Host copy of every block (column-major reshape), in sector order."""
        sec, siz, off = self.struct.tables()
        flat = self.data.detach().cpu().numpy()
        out = {}
        for b in range(self.struct.nblocks):
            out[tuple(int(x) for x in sec[b])] = flat[off[b]:off[b + 1]].reshape(tuple(int(x) for x in siz[b]),
                                                                               order="F")
        return out

    def __repr__(self):
        return f"SymTensor(charge={self.charge}, space={self.space}, nblocks={self.struct.nblocks}, dtype={self.dtype})"

    # -- constructors ---------------------------------------------------------
    @classmethod
    def from_flat(cls, charge, space, flat, dtype=torch.float64, device=None):
        st = get_struct(charge, space)
        t = torch.as_tensor(np.asarray(flat), dtype=dtype).reshape(-1)
        if t.numel() != st.numel:
            raise ValueError(f"flat buffer has {t.numel()} elements, structure needs {st.numel}")
        return cls(charge, space, t.to(device or default_device()), st)

    @classmethod
    def from_blocks(cls, charge, space, blocks: Dict, dtype=torch.float64, device=None):
        """Checked constructor (symtensor.jl:17-28): keys must be exactly the allowed sectors."""
        st = get_struct(charge, space)
        sec, siz, off = st.tables()
        keys = [tuple(int(x) for x in r) for r in sec]
        if sorted(blocks, key=lambda s: s[::-1]) != keys:
            raise ValueError("SectorMismatch(): blocks do not match the allowed sectors")
        flat = np.empty(st.numel, dtype=np.float64)
        for b, k in enumerate(keys):
            a = np.asarray(blocks[k], dtype=np.float64)
            if a.shape != tuple(int(x) for x in siz[b]):
                raise ValueError("SizeMismatch")
            flat[off[b]:off[b + 1]] = a.reshape(-1, order="F")
        return cls.from_flat(charge, space, flat, dtype, device)

    def similar(self, dtype=None) -> "SymTensor":
        """``similar(A)`` (symtensor.jl:252-255): same structure, uninitialised data."""
        return SymTensor(self.charge, self.space, torch.empty_like(self.data, dtype=dtype or self.dtype), self.struct)

    def copy(self) -> "SymTensor":
        return SymTensor(self.charge, self.space, self.data.clone(), self.struct)

    def to(self, dtype) -> "SymTensor":
        if dtype == self.dtype:
            return self
        out = torch.empty(self.data.numel(), dtype=dtype, device=self.device)
        check(lib().mpsh_convert(_DT[self.dtype], _ptr(self.data), _DT[dtype], _ptr(out), self.data.numel(),
                                 _stream()))
        return SymTensor(self.charge, self.space, out, self.struct)


def zeros(space, charge=0, dtype=torch.float64, device=None) -> SymTensor:
    st = get_struct(charge, space)
    return SymTensor(charge, space, torch.zeros(st.numel, dtype=dtype, device=device or default_device()), st)


def fill(x, space, charge=0, dtype=torch.float64, device=None) -> SymTensor:
    """``fill(x, charge, space)`` (symtensor.jl:157-163)."""
    st = get_struct(charge, space)
    return SymTensor(charge, space, torch.full((st.numel,), float(x), dtype=dtype, device=device or default_device()),
                     st)


def fill_linearindex(space, charge=0, dtype=torch.float64, device=None) -> SymTensor:
    """``fill_linearindex`` (symtensor.jl:173-185): 1, 2, 3, ... over the packed blocks."""
    st = get_struct(charge, space)
    return SymTensor(charge, space, torch.arange(1, st.numel + 1, dtype=dtype, device=device or default_device()), st)


def randn(space, charge=0, seed=1911, stream=0, dtype=torch.float64, device=None) -> SymTensor:
    """Seeded Gaussian blocks from the counter-based generator of ``workloads``."""
    from .workloads import gaussian
    st = get_struct(charge, space)
    return SymTensor.from_flat(charge, space, gaussian(seed, stream, st.numel), dtype, device)


def dual(A: SymTensor) -> SymTensor:
    """``dual(A)`` (symtensor.jl:201-212): negated charge, flipped legs, SAME data (no conj)."""
    return SymTensor(-A.charge, tuple(V.dual() for V in A.space), A.data)


def issimilar(A: SymTensor, B: SymTensor) -> bool:
    """``issimilar`` (symtensor.jl:93-94)."""
    return A.charge == B.charge and A.space == B.space


def array(A: SymTensor) -> np.ndarray:
    """Dense host embedding ``array(A)`` (symtensor.jl:236-242)."""
    lay = []
    for V in A.space:
        d, p = {}, 0
        for c, n in V.items():
            d[c] = (p, p + n)
            p += n
        lay.append(d)
    out = np.zeros(A.size(), dtype=np.float64)
    for s, b in A.blocks().items():
        out[tuple(slice(*lay[i][c]) for i, c in enumerate(s))] = b
    return out


# ---------------------------------------------------------------------------
# vector-space ops (symtensor.jl:252-323)
# ---------------------------------------------------------------------------
class _Scratch:
    def __init__(self):
        self.bufs = {}

    def get(self, device):
        b = self.bufs.get(device)
        if b is None:
            b = torch.empty(L.DOT_PARTIALS + 1, dtype=torch.float64, device=device)
            self.bufs[device] = b
        return b


_scratch = _Scratch()


def dot_device(A: SymTensor, B: SymTensor, out: torch.Tensor = None) -> torch.Tensor:
    """``dot(A, B)`` left on the device as a 1-element float64 tensor."""
    if not issimilar(A, B):
        raise ValueError("The two SymTensors have to be similar to dot!")
    if A.dtype != B.dtype:
        raise TypeError("dot of mixed dtypes")
    sc = _scratch.get(A.device)
    res = out if out is not None else torch.empty(1, dtype=torch.float64, device=A.device)
    check(lib().mpsh_dot(_DT[A.dtype], A.data.numel(), _ptr(A.data), _ptr(B.data), _ptr(res), _ptr(sc), _stream()))
    return res


def dot(A: SymTensor, B: SymTensor) -> float:
    """``dot(A, B)`` (symtensor.jl:314-318); real data, so no conjugation is needed."""
    return float(dot_device(A, B).item())


def norm(A: SymTensor) -> float:
    """``norm(A) = sqrt(dot(A, A))`` (symtensor.jl:321-323)."""
    return math.sqrt(dot(A, A))


def axpby_(alpha: float, A: SymTensor, beta: float, B: SymTensor) -> SymTensor:
    """``axpby!(α, A, β, B)``: B = α A + β B in place (symtensor.jl:299-308)."""
    if not issimilar(A, B):
        raise ValueError("axpy! The two matrices are not simliar!")
    check(lib().mpsh_axpby(_DT[B.dtype], B.data.numel(), float(alpha), _ptr(A.data), float(beta), _ptr(B.data),
                           _stream()))
    return B


def axpy_(alpha: float, A: SymTensor, B: SymTensor) -> SymTensor:
    """``axpy!(α, A, B)``: B = α A + B (symtensor.jl:288-296)."""
    return axpby_(alpha, A, 1.0, B)


def rmul_(A: SymTensor, alpha: float) -> SymTensor:
    """``rmul!(A, α)`` (symtensor.jl:280-285)."""
    check(lib().mpsh_scal(_DT[A.dtype], A.data.numel(), float(alpha), _ptr(A.data), _stream()))
    return A


def mul_(B: SymTensor, A: SymTensor, alpha: float) -> SymTensor:
    """``mul!(B, A, α)``: B = α A (symtensor.jl:268-276)."""
    if not issimilar(A, B):
        raise ValueError("oops")
    return axpby_(alpha, A, 0.0, B)


def scale(A: SymTensor, alpha: float) -> SymTensor:
    """``A * α`` out of place (symtensor.jl:244-246)."""
    return rmul_(A.copy(), alpha)


# ---------------------------------------------------------------------------
# re-legging   (src/symreleg.jl)
# ---------------------------------------------------------------------------
RELEG_PLAN_CACHE = 2048
_RELEG: Dict = PlanLRU(RELEG_PLAN_CACHE)


def _releg_run(key, make, A: SymTensor) -> SymTensor:
    ent = _RELEG.get(key)
    if ent is None:
        plan, out = L._p(), L._p()
        check(make(plan, out))
        ent = (PlanHandle(plan, "mpsh_releg_destroy"), _struct_from_handle(out))
        _RELEG[key] = ent
    plan, st = ent
    data = torch.empty(st.numel, dtype=A.dtype, device=A.device)
    check(lib().mpsh_releg_execute(plan.h, _DT[A.dtype], _ptr(A.data), _ptr(data), _stream()))
    return SymTensor(st.charge, st.spaces, data, st)


def permutelegs(A: SymTensor, perm: Sequence[int]) -> SymTensor:
    """``permutelegs(A, perm)`` (symreleg.jl:34-51); identity returns A itself."""
    perm = [int(p) for p in perm]
    if sorted(perm) != list(range(1, A.rank + 1)):
        raise ValueError("Not a valid permutation!")
    if perm == list(range(1, A.rank + 1)):
        return A
    return _releg_run(("perm", A.struct.key, tuple(perm)),
                      lambda p, o: lib().mpsh_permutelegs_plan(A.struct.handle, i32arr(perm), p, o), A)


def fuselegs(A: SymTensor, l: int, n: int = 2, _dual: bool = False) -> SymTensor:
    """``fuselegs(A, l, n, _dual)`` (symreleg.jl:79-146)."""
    if n < 2 and bool(_dual) == A.space[l - 1].isdual:
        return A
    return _releg_run(("fuse", A.struct.key, l, n, bool(_dual)),
                      lambda p, o: lib().mpsh_fuselegs_plan(A.struct.handle, l, n, 1 if _dual else 0, p, o), A)


def splitleg(A: SymTensor, l: int, spaces: Sequence[U1Space]) -> SymTensor:
    """``splitleg(A, l, spaces)`` (symreleg.jl:161-214)."""
    spaces = tuple(spaces)
    if len(spaces) == 1:
        if A.space[l - 1] != spaces[0]:
            raise ValueError(f"SpaceMismatch {A.space[l - 1]}, {spaces[0]}")
        return A
    isd, nch, ch, dm = _leg_arrays(spaces)
    return _releg_run(("split", A.struct.key, l, tuple(V.key() for V in spaces)),
                      lambda p, o: lib().mpsh_splitleg_plan(A.struct.handle, l, len(spaces), isd, nch, ch, dm, p, o),
                      A)


def symmatrix(A: SymTensor, rowidxs: Sequence[int], colidxs: Sequence[int]) -> SymTensor:
    """``SymMatrix(A, rowidxs, colidxs)`` (contract.jl:112-171): permute + fuse to rank 2."""
    r, c = [int(x) for x in rowidxs], [int(x) for x in colidxs]
    return _releg_run(("symmat", A.struct.key, tuple(r), tuple(c)),
                      lambda p, o: lib().mpsh_symmatrix_plan(A.struct.handle, len(r), i32arr(r), len(c), i32arr(c),
                                                             p, o), A)


# ---------------------------------------------------------------------------
# contract   (src/contract.jl:11-47)
# ---------------------------------------------------------------------------
CONTRACT_PLAN_CACHE = 1024
_CONTRACT: Dict = PlanLRU(CONTRACT_PLAN_CACHE)


def contract_plan(sA: Struct, idxA, sB: Struct, idxB):
    """(plan owner, output structure) of contract(A, idxA, B, idxB), LRU-cached by structure."""
    key = (sA.key, tuple(idxA), sB.key, tuple(idxB))
    ent = _CONTRACT.get(key)
    if ent is None:
        plan, out = L._p(), L._p()
        check(lib().mpsh_contract_plan(sA.handle, i32arr(idxA), sB.handle, i32arr(idxB), plan, out))
        ent = (PlanHandle(plan, "mpsh_contract_destroy"), _struct_from_handle(out))
        _CONTRACT[key] = ent
    return ent


def contract(A: SymTensor, idxA: Sequence[int], B: SymTensor, idxB: Sequence[int]):
    """``contract(A, idxA, B, idxB)`` (contract.jl:11-47) on the device.

    Negative labels are contracted (ordered by |label|), positive labels are
    output slots.  A full contraction returns a float (the SymVector branch,
    contract.jl:24-28); rank-1 outputs raise like the reference (Q3).
    """
    idxA, idxB = [int(x) for x in idxA], [int(x) for x in idxB]
    if len(idxA) != A.rank or len(idxB) != B.rank:
        raise ValueError("index set length differs from tensor rank")
    if all(x < 0 for x in idxA) and all(x < 0 for x in idxB):
        # SymVector(A, consA, false) * SymVector(B, consB, true) (contract.jl:24-28, :101-110, :226-232)
        consA = [i + 1 for i in sorted(range(len(idxA)), key=lambda i: abs(idxA[i]))]
        consB = [i + 1 for i in sorted(range(len(idxB)), key=lambda i: abs(idxB[i]))]
        Av = fuselegs(permutelegs(A, consA), 1, A.rank, False)
        Bv = fuselegs(permutelegs(B, consB), 1, B.rank, True)
        Va, Vb = Av.space[0], Bv.space[0]
        if not (Va.charges == Vb.charges and Va.dims == Vb.dims and Va.isdual != Vb.isdual):
            raise ValueError(f"not contractible! {Va} and {Vb}")
        if Av.data.numel() == 0:
            return 0.0
        if A.dtype != B.dtype:
            Bv = Bv.to(A.dtype)
        sc = _scratch.get(A.device)
        res = torch.empty(1, dtype=torch.float64, device=A.device)
        check(lib().mpsh_dot(_DT[A.dtype], Av.data.numel(), _ptr(Av.data), _ptr(Bv.data), _ptr(res), _ptr(sc),
                             _stream()))
        return float(res.item())
    if A.dtype != B.dtype:
        raise TypeError("contract of mixed dtypes is not supported on the device path")
    owner, st = contract_plan(A.struct, idxA, B.struct, idxB)
    out = torch.empty(st.numel, dtype=A.dtype, device=A.device)
    ws = torch.empty(max(1, int(lib().mpsh_contract_workspace(owner.h, _DT[A.dtype]))), dtype=torch.uint8,
                     device=A.device)
    check(lib().mpsh_contract_execute(owner.h, _DT[A.dtype], _ptr(A.data), _ptr(B.data), _ptr(out), _ptr(ws),
                                      _stream()))
    return SymTensor(st.charge, st.spaces, out, st)


def matmul(A: SymTensor, B: SymTensor) -> SymTensor:
    """Rank-2 product ``A * B`` via contract(A, (1,-1), B, (-1,2))."""
    return contract(A, (1, -1), B, (-1, 2))
