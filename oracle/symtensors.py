"""Numpy restatement of SymTensors (reference ``src/SymTensors.jl`` and includes).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Leg indices, leg
labels and permutations are 1-based exactly as in the reference API.
Blocks are numpy arrays indexed by the same multi-index as the Julia
column-major blocks; every reshape uses ``order='F'`` so that element
placement matches Julia's ``reshape``.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np


def _prod(t) -> int:
    p = 1
    for x in t:
        p *= int(x)
    return p


# ---------------------------------------------------------------------------
# U(1) charges and vector spaces   (src/group.jl:12-34, src/vspace.jl:5-138)
# ---------------------------------------------------------------------------
class Space:
    """VectorSpace{U1Charge}: sorted charge->dim map plus ``isdual``.

    Follows ``src/vspace.jl:5-18``: charges unique, dims positive, kept
    sorted (SortedDict semantics, ``src/sorteddict.jl:7-12``).
    """

    __slots__ = ("charges", "dims", "isdual")

    def __init__(self, pairs, isdual: bool = False):
        acc: Dict[int, int] = {}
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

    def items(self):
        return zip(self.charges, self.dims)

    def dim(self, c=None) -> int:
        """``dim(V, c)`` (0 if absent) / ``dim(V)`` total (vspace.jl:43-45)."""
        if c is None:
            return sum(self.dims)
        try:
            return self.dims[self.charges.index(int(c))]
        except ValueError:
            return 0

    def dual(self) -> "Space":
        """``dual(V)``: same map, flipped flag (vspace.jl:53)."""
        return Space(self.items(), not self.isdual)

    def layout(self) -> Dict[int, Tuple[int, int]]:
        """``layout(V)``: charge -> [start, stop) in sorted order (vspace.jl:56-73)."""
        out, p = {}, 0
        for c, d in self.items():
            out[c] = (p, p + d)
            p += d
        return out

    def key(self):
        return (self.charges, self.dims, self.isdual)

    def __eq__(self, other):
        return isinstance(other, Space) and self.key() == other.key()

    def __hash__(self):
        return hash(self.key())

    def __repr__(self):
        body = ", ".join(f"{c}{chr(39) if self.isdual else ''}=>{d}" for c, d in self.items())
        return f"U1Space{'†' if self.isdual else ''}({body})"


def U1Space(*pairs) -> Space:
    """``U1Space(c=>d, ...)`` (vspace.jl:75-76)."""
    return Space(pairs, False)


def isdual_pair(V1: Space, V2: Space) -> bool:
    """``isdual(V1, V2)``: same sector map, opposite flags (vspace.jl:24-25)."""
    return V1.charges == V2.charges and V1.dims == V2.dims and V1.isdual != V2.isdual


def flow(charges: Sequence[int], dualinfo: Sequence[bool]) -> int:
    """``sum(s::Sector, dualinfo)`` (sector.jl:14-20)."""
    r = 0
    for c, d in zip(charges, dualinfo):
        r += -c if d else c
    return r


def fuse(_dual: bool, Vs: Sequence[Space]) -> Space:
    """``fuse(_dual, Vs)`` (vspace.jl:112-138): right-to-left fusion."""
    Vs = tuple(Vs)
    N = len(Vs)
    if N < 2:
        V = Vs[0]
        if _dual == V.isdual:
            return V
        # dual(mapcharges(inv, V))
        return Space([(-c, d) for c, d in V.items()], not V.isdual)
    if N > 2:
        return fuse(_dual, Vs[: N - 2] + (fuse(_dual, Vs[N - 2:]),))
    V1, V2 = Vs
    acc: Dict[int, int] = {}
    for c2, d2 in V2.items():
        c2 = -c2 if V2.isdual else c2
        for c1, d1 in V1.items():
            c1 = -c1 if V1.isdual else c1
            c = -(c1 + c2) if _dual else c1 + c2
            acc[c] = acc.get(c, 0) + d1 * d2
    return Space(acc.items(), _dual)


# ---------------------------------------------------------------------------
# Sectors   (src/sector.jl:30-36, 69-90)
# ---------------------------------------------------------------------------
def sector_key(s: Tuple[int, ...]):
    """Column-major sector order: ``reverse(s1) < reverse(s2)`` (sector.jl:35-36)."""
    return s[::-1]


def allsectorsandsizes(charge: int, Vs: Sequence[Space]):
    """``_allsectorsandsizes(charge, Vs)`` (sector.jl:69-90), recursive, sorted."""
    Vs = tuple(Vs)
    N = len(Vs)
    if N < 2:
        V = Vs[0]
        c = -charge if V.isdual else charge
        d = V.dim(c)
        if d > 0:
            return [(c,)], [(d,)]
        return [], []
    sects, sizes = [], []
    V = Vs[-1]
    for c, d in V.items():
        cc = -c if V.isdual else c
        hs, hz = allsectorsandsizes(charge - cc, Vs[:-1])
        for s, z in zip(hs, hz):
            sects.append(s + (c,))
            sizes.append(z + (d,))
    return sects, sizes


# ---------------------------------------------------------------------------
# SymTensor   (src/symtensor.jl:12-30)
# ---------------------------------------------------------------------------
class SymTensor:
    """Block-sparse tensor; ``blocks`` is kept in sorted (column-major) order."""

    __slots__ = ("charge", "space", "blocks")

    def __init__(self, charge: int, space: Sequence[Space], blocks: Dict, check: bool = False):
        self.charge = int(charge)
        self.space = tuple(space)
        keys = sorted(blocks, key=sector_key)
        self.blocks = {k: blocks[k] for k in keys}
        if check:
            # checked constructor (symtensor.jl:17-28)
            sects, sizes = allsectorsandsizes(self.charge, self.space)
            if len(sects) != len(keys):
                raise ValueError("SectorMismatch")
            for s, z, k in zip(sects, sizes, keys):
                if s != k:
                    raise ValueError(f"SectorMismatch(), {k} vs {s}")
                if tuple(self.blocks[k].shape) != tuple(z):
                    raise ValueError("SizeMismatch")

    @property
    def rank(self):
        return len(self.space)

    def sectors(self):
        return list(self.blocks.keys())

    def dtype(self):
        for b in self.blocks.values():
            return b.dtype
        return np.float64

    def __repr__(self):
        return f"SymTensor(charge={self.charge}, space={self.space}, nblocks={len(self.blocks)})"


def issimilar(A: SymTensor, B: SymTensor) -> bool:
    """``issimilar`` (symtensor.jl:93-94)."""
    return A.charge == B.charge and A.space == B.space


def tensors_equal(A: SymTensor, B: SymTensor) -> bool:
    """``==`` on SymTensors: similar, same keys, identical blocks (symtensor.jl:96-102)."""
    if not issimilar(A, B) or list(A.blocks) != list(B.blocks):
        return False
    return all(A.blocks[k].shape == B.blocks[k].shape and np.array_equal(A.blocks[k], B.blocks[k])
               for k in A.blocks)


def from_function(f, charge: int, space: Sequence[Space], dtype=np.float64) -> SymTensor:
    """``SymTensor(f, T, charge, space)`` (symtensor.jl:135-140); f(dtype, sizes)."""
    sects, sizes = allsectorsandsizes(charge, space)
    return SymTensor(charge, space, {s: f(dtype, z) for s, z in zip(sects, sizes)}, check=True)


def fill(x, space: Sequence[Space], charge: int = 0, dtype=None) -> SymTensor:
    """``fill(x, charge, space)`` (symtensor.jl:157-163)."""
    dt = dtype or np.asarray(x).dtype
    return from_function(lambda T, z: np.full(z, x, dtype=T), charge, space, dt)


def fill_linearindex(space: Sequence[Space], charge: int = 0) -> SymTensor:
    """``fill_linearindex`` (symtensor.jl:173-185): consecutive integers over blocks."""
    sects, sizes = allsectorsandsizes(charge, space)
    blocks, p = {}, 0
    for s, z in zip(sects, sizes):
        n = _prod(z)
        blocks[s] = np.arange(p + 1, p + n + 1, dtype=np.int64).reshape(z, order="F")
        p += n
    return SymTensor(charge, space, blocks, check=True)


def eye(V: Space, dtype=np.float64) -> SymTensor:
    """``eye(T, V)`` (symtensor.jl:188-191)."""
    return SymTensor(0, (V, V.dual()), {(c, c): np.eye(d, dtype=dtype) for c, d in V.items()})


def dual(A: SymTensor) -> SymTensor:
    """``dual(A)`` (symtensor.jl:201-212): negate charge, flip legs, SAME blocks, no conj."""
    return SymTensor(-A.charge, tuple(V.dual() for V in A.space), A.blocks)


def array(A: SymTensor) -> np.ndarray:
    """Dense embedding ``array(A)`` (symtensor.jl:236-242) via ``layout``."""
    lay = [V.layout() for V in A.space]
    out = np.zeros(tuple(V.dim() for V in A.space), dtype=A.dtype())
    for s, b in A.blocks.items():
        idx = tuple(slice(*lay[i][c]) for i, c in enumerate(s))
        out[idx] = b
    return out


# ---------------------------------------------------------------------------
# Lanczos vector interface   (symtensor.jl:252-323)
# ---------------------------------------------------------------------------
def dot(A: SymTensor, B: SymTensor):
    """``dot(A, B)``: sum of per-block dots, conjugating the first (symtensor.jl:314-318)."""
    if not issimilar(A, B):
        raise ValueError("The two SymTensors have to be similar to dot!")
    return sum(np.vdot(A.blocks[k], B.blocks[k]) for k in A.blocks)


def norm(A: SymTensor) -> float:
    """``norm(A) = sqrt(dot(A, A))`` (symtensor.jl:321-323)."""
    return math.sqrt(float(np.real(dot(A, A))))


def axpby(alpha, A: SymTensor, beta, B: SymTensor) -> SymTensor:
    """``axpby!(α, A, β, B)`` -> new B (symtensor.jl:299-308)."""
    if not issimilar(A, B):
        raise ValueError("axpy! The two matrices are not simliar!")
    return SymTensor(B.charge, B.space, {k: alpha * A.blocks[k] + beta * B.blocks[k] for k in B.blocks})


def scale(A: SymTensor, a) -> SymTensor:
    """``A * a`` (symtensor.jl:244-246)."""
    return SymTensor(A.charge, A.space, {k: a * b for k, b in A.blocks.items()})


# ---------------------------------------------------------------------------
# Re-legging   (src/symreleg.jl:34-214)
# ---------------------------------------------------------------------------
def invperm(p: Sequence[int]) -> List[int]:
    out = [0] * len(p)
    for i, x in enumerate(p, 1):
        out[x - 1] = i
    return out


def permutelegs(A: SymTensor, perm: Sequence[int]) -> SymTensor:
    """``permutelegs(A, perm)`` (symreleg.jl:34-51): permutedims per block, re-sorted keys."""
    N = A.rank
    perm = [int(p) for p in perm]
    if sorted(perm) != list(range(1, N + 1)):
        raise ValueError("Not a valid permutation!")
    if perm == list(range(1, N + 1)):
        return A
    p0 = [p - 1 for p in perm]
    blocks = {tuple(s[i] for i in p0): np.transpose(b, p0) for s, b in A.blocks.items()}
    return SymTensor(A.charge, tuple(A.space[i] for i in p0), blocks)


def fuselegs(A: SymTensor, l: int, n: int = 2, _dual: bool = False) -> SymTensor:
    """``fuselegs(A, l, n, _dual)`` (symreleg.jl:79-146).

    The ``n < 2`` branch of the reference evaluates a SymTensor without
    returning it (symreleg.jl:88-97) and then runs the general code with
    ``n = 1``; the general code below gives that same result.
    """
    N = A.rank
    if not (0 < l < N + 2 - n):
        raise ValueError(f"fuselegs {l} to {l + n - 1} for size {N}")
    if n < 2 and _dual == A.space[l - 1].isdual:
        return A
    lo, hi = l - 1, l - 1 + n
    dualinfo = [V.isdual for V in A.space[lo:hi]]
    keys = list(A.blocks)
    fsects = []
    for s in keys:
        fc = flow(s[lo:hi], dualinfo)
        if _dual:
            fc = -fc
        fsects.append(s[:lo] + (fc,) + s[hi:])
    sperm = sorted(range(len(keys)), key=lambda i: sector_key(fsects[i]))
    semits = [keys[i] for i in sperm]
    Vf = fuse(_dual, A.space[lo:hi])
    space = A.space[:lo] + (Vf,) + A.space[hi:]
    sects, sizes = allsectorsandsizes(A.charge, space)
    dt = A.dtype()
    blocks, pointer = {}, 0
    for sect, sz in zip(sects, sizes):
        blk = np.empty(sz, dtype=dt)
        p = 0
        while p < sz[lo]:
            src = A.blocks[semits[pointer]]
            s = src.shape
            fd = _prod(s[lo:hi])
            idx = [slice(None)] * len(sz)
            idx[lo] = slice(p, p + fd)
            blk[tuple(idx)] = src.reshape(s[:lo] + (fd,) + s[hi:], order="F")
            p += fd
            pointer += 1
        blocks[sect] = blk
    return SymTensor(A.charge, space, blocks)


def splitleg(A: SymTensor, l: int, spaces: Sequence[Space]) -> SymTensor:
    """``splitleg(A, l, space)`` (symreleg.jl:161-214)."""
    spaces = tuple(spaces)
    N = A.rank
    if not (0 < l <= N):
        raise ValueError(f"integer l not in range {l}, {N}")
    if len(spaces) == 1:
        if A.space[l - 1] != spaces[0]:
            raise ValueError(f"SpaceMismatch {A.space[l - 1]}, {spaces[0]}")
        return A
    keys = list(A.blocks)
    order = sorted(range(len(keys)), key=lambda i: keys[i][l - 1])
    blocks, cache = {}, {}
    for i in order:
        s = keys[i]
        charge = -s[l - 1] if A.space[l - 1].isdual else s[l - 1]
        if charge not in cache:
            cache[charge] = allsectorsandsizes(charge, spaces)
        pats, sizes = cache[charge]
        blk = A.blocks[s]
        sh = blk.shape
        pivot = 0
        for pat, sz in zip(pats, sizes):
            sl = _prod(sz)
            idx = [slice(None)] * N
            idx[l - 1] = slice(pivot, pivot + sl)
            blocks[s[: l - 1] + pat + s[l:]] = blk[tuple(idx)].reshape(sh[: l - 1] + sz + sh[l:], order="F")
            pivot += sl
    new_space = A.space[: l - 1] + spaces + A.space[l:]
    return SymTensor(A.charge, new_space, blocks)


# ---------------------------------------------------------------------------
# Contraction   (src/contract.jl:11-298)
# ---------------------------------------------------------------------------
def contract_index_perm(indexset: Sequence[int]):
    """``_contract_index_perm`` (contract.jl:82-99): (rems, cons, tofinals), 1-based."""
    rems, cons, cons_order, tofinals = [], [], [], []
    for i, idx in enumerate(indexset, 1):
        if idx < 0:
            cons.append(i)
            cons_order.append(idx)
        else:
            rems.append(i)
            tofinals.append(idx)
    perm = sorted(range(len(tofinals)), key=lambda j: tofinals[j])
    cperm = sorted(range(len(cons_order)), key=lambda j: abs(cons_order[j]))
    return [rems[j] for j in perm], [cons[j] for j in cperm], [tofinals[j] for j in perm]


def symmatrix(A: SymTensor, rowidxs: Sequence[int], colidxs: Sequence[int]) -> SymTensor:
    """``SymMatrix(A, rowidxs, colidxs)`` (contract.jl:112-171): permute + fuse."""
    N = A.rank
    rowidxs, colidxs = list(rowidxs), list(colidxs)
    idxperm = rowidxs + colidxs
    if sorted(idxperm) != list(range(1, N + 1)):
        raise ValueError("Incorrect index set for conversion to SymMatrix!")
    if not rowidxs:
        raise ValueError("empty row for matrix not allowed!")
    if not colidxs:
        raise ValueError("empty col for matrix not allowed!")
    n = len(rowidxs)
    p0 = [i - 1 for i in idxperm]
    rowdual = [A.space[i - 1].isdual for i in rowidxs]
    coldual = [A.space[i - 1].isdual for i in colidxs]
    keys = list(A.blocks)
    csects = [tuple(s[i] for i in p0) for s in keys]
    sperm1 = sorted(range(len(csects)), key=lambda i: sector_key(csects[i]))
    fsects = []
    for i in range(len(csects)):
        sector = csects[sperm1[i]]
        fsects.append((flow(sector[:n], rowdual), -flow(sector[n:], coldual)))
    sperm2 = sorted(range(len(fsects)), key=lambda i: sector_key(fsects[i]))
    semits = [keys[sperm1[j]] for j in sperm2]
    space = (fuse(False, [A.space[i - 1] for i in rowidxs]), fuse(True, [A.space[i - 1] for i in colidxs]))
    sects, sizes = allsectorsandsizes(A.charge, space)
    dt = A.dtype()
    blocks, pointer = {}, 0
    for sect, (rlim, clim) in zip(sects, sizes):
        blk = np.empty((rlim, clim), dtype=dt)
        pc = 0
        while pc < clim:
            pr, fdc = 0, 0
            while pr < rlim:
                Ablk = np.transpose(A.blocks[semits[pointer]], p0)
                s = Ablk.shape
                fdr, fdc = _prod(s[:n]), _prod(s[n:])
                blk[pr:pr + fdr, pc:pc + fdc] = Ablk.reshape((fdr, fdc), order="F")
                pr += fdr
                pointer += 1
            pc += fdc
        blocks[sect] = blk
    return SymTensor(A.charge, space, blocks)


def matmul(A: SymTensor, B: SymTensor) -> SymTensor:
    """``*(A::AbstractSymMatrix, B::AbstractSymMatrix)`` (contract.jl:234-298).

    Unmatched output sectors get explicit zero blocks (contract.jl:292-294).
    """
    if not isdual_pair(A.space[1], B.space[0]):
        raise ValueError(f"not contractible! {A.space[1]} and {B.space[0]}")
    charge = A.charge + B.charge
    space = (A.space[0], B.space[1])
    dt = np.result_type(A.dtype(), B.dtype())
    sects, sizes = allsectorsandsizes(charge, space)
    blocks = {}
    for (c1, c2), sz in zip(sects, sizes):
        _c1 = -c1 if A.space[0].isdual else c1
        _c2 = -c2 if B.space[1].isdual else c2
        if A.space[1].isdual:
            sA, sB = (c1, -A.charge + _c1), (B.charge - _c2, c2)
        else:
            sA, sB = (c1, A.charge - _c1), (-B.charge + _c2, c2)
        if sA in A.blocks and sB in B.blocks:
            blocks[(c1, c2)] = A.blocks[sA] @ B.blocks[sB]
        else:
            blocks[(c1, c2)] = np.zeros(sz, dtype=dt)
    return SymTensor(charge, space, blocks)


def split_matrix(C: SymTensor, rlegs: Sequence[Space], clegs: Sequence[Space]) -> SymTensor:
    """``SymTensor(A::SymMatrix, rlegs, clegs)`` (contract.jl:187-224): split fused blocks."""
    rlegs, clegs = tuple(rlegs), tuple(clegs)
    blocks = {}
    for (c1, c2), blk in C.blocks.items():
        c1 = -c1 if C.space[0].isdual else c1
        c2 = -c2 if C.space[1].isdual else c2
        rpats, rsizes = allsectorsandsizes(c1, rlegs)
        cpats, csizes = allsectorsandsizes(c2, clegs)
        pc = 0
        for cp, cz in zip(cpats, csizes):
            csl = _prod(cz)
            pr = 0
            for rp, rz in zip(rpats, rsizes):
                rsl = _prod(rz)
                blocks[rp + cp] = blk[pr:pr + rsl, pc:pc + csl].reshape(rz + cz, order="F")
                pr += rsl
            pc += csl
    return SymTensor(C.charge, rlegs + clegs, blocks)


def symvector(A: SymTensor, perm: Sequence[int], _dual: bool) -> SymTensor:
    """``SymVector(A, perm, _dual)`` (contract.jl:101-110)."""
    N = A.rank
    if sorted(perm) != list(range(1, N + 1)):
        raise ValueError("Incorrect index set for conversion to SymVector!")
    return fuselegs(permutelegs(A, perm), 1, N, _dual)


def vecvec(A: SymTensor, B: SymTensor):
    """``*(A::SymVector, B::SymVector)`` (contract.jl:226-232): first-block product sum."""
    if not isdual_pair(A.space[0], B.space[0]):
        raise ValueError(f"not contractible! {A.space[0]} and {B.space[0]}")
    k = next(iter(A.blocks))
    kb = next(iter(B.blocks))
    return np.sum(A.blocks[k] * B.blocks[kb])


def contract(A: SymTensor, idxA: Sequence[int], B: SymTensor, idxB: Sequence[int]):
    """``contract(A, idxA, B, idxB)`` (contract.jl:11-47)."""
    remsA, consA, tofA = contract_index_perm(idxA)
    remsB, consB, tofB = contract_index_perm(idxB)
    if not remsA:
        _A = symvector(A, consA, False)
        if not remsB:
            return vecvec(_A, symvector(B, consB, True))
        # contract.jl:30-32 reads a removed field (`.legs`) and errors there.
        raise NotImplementedError("rank-1 output contract branch (contract.jl:30-32) is broken in the reference")
    _A = symmatrix(A, remsA, consA)
    if not remsB:
        raise NotImplementedError("rank-1 output contract branch (contract.jl:38-41) is broken in the reference")
    _B = symmatrix(B, consB, remsB)
    C = split_matrix(matmul(_A, _B), [A.space[i - 1] for i in remsA], [B.space[i - 1] for i in remsB])
    return permutelegs(C, invperm(tofA + tofB))


def dense_contract(A: np.ndarray, idxA: Sequence[int], B: np.ndarray, idxB: Sequence[int]) -> np.ndarray:
    """Dense ``contract`` for arrays (contract.jl:49-79), via einsum on label letters."""
    letters = {}

    def lab(x):
        if x not in letters:
            letters[x] = chr(ord("a") + len(letters))
        return letters[x]

    sa = "".join(lab(x) for x in idxA)
    sb = "".join(lab(x) for x in idxB)
    outs = sorted(x for x in list(idxA) + list(idxB) if x > 0)
    so = "".join(lab(x) for x in outs)
    return np.einsum(f"{sa},{sb}->{so}", A, B)


# ---------------------------------------------------------------------------
# Factorisation   (src/symtools.jl:1-35, src/mpsutils.jl:89-179)
# ---------------------------------------------------------------------------
def _svd_block(blk: np.ndarray):
    """``svd(blk, full=false)`` with the QRIteration fallback (symtools.jl:14-19)."""
    try:
        return np.linalg.svd(blk, full_matrices=False)
    except np.linalg.LinAlgError:
        import scipy.linalg
        return scipy.linalg.svd(blk, full_matrices=False, lapack_driver="gesvd")


def svd_(A: SymTensor):
    """``_svd_(A::AbstractSymMatrix)`` (symtools.jl:1-35): per-sector untruncated SVD."""
    ub, sb, vb, midl, midr = {}, {}, {}, [], []
    for (c1, c2), blk in A.blocks.items():
        U, S, Vt = _svd_block(blk)
        ub[(c1, c1)] = U
        sb[(c1, c2)] = np.diag(S)
        vb[(c2, c2)] = Vt
        midl.append((c1, len(S)))
        midr.append((c2, len(S)))
    Vl = Space(midl, A.space[0].isdual)
    Vr = Space(midr, A.space[1].isdual)
    u = SymTensor(0, (A.space[0], Vl.dual()), ub)
    s = SymTensor(A.charge, (Vl, Vr), sb)
    v = SymTensor(0, (Vr.dual(), A.space[1]), vb)
    return u, s, v


def truncatedsizes(S: List[np.ndarray], maxdim: int = 200, tol: float = 1e-14) -> List[int]:
    """``truncatedsizes`` (mpsutils.jl:89-112): greedy global largest-first, ABSOLUTE tol."""
    pointers = [1] * len(S)
    c = [float(S[i][0]) for i in range(len(S))]
    total = sum(len(s) for s in S)
    for n in range(1, min(maxdim, total) + 1):
        index = int(np.argmax(c))  # first max, like findmax
        m = c[index]
        if m < tol:
            if n == 1:
                raise ValueError(f"The largest singular value is smaller than tol! {m} < {tol}")
            break
        if pointers[index] >= len(S[index]):
            pointers[index] += 1
            c[index] = 0.0
        else:
            pointers[index] += 1
            c[index] = float(S[index][pointers[index] - 1])
    return [p - 1 for p in pointers]


def svdtrunc(A: SymTensor, maxdim: int = None, tol: float = 1e-12):
    """``svdtrunc(A::AbstractSymTensor; maxdim, tol)`` (mpsutils.jl:116-179)."""
    if maxdim is None:
        maxdim = max(V.dim() for V in A.space)
    items = list(A.blocks.items())
    facts = [_svd_block(blk) for _, blk in items]
    ns = truncatedsizes([f[1] for f in facts], maxdim=maxdim, tol=tol)
    ub, sb, vb, midl, midr = {}, {}, {}, [], []
    for ((c1, c2), _), (U, S, Vt), k in zip(items, facts, ns):
        if k <= 0:
            continue
        ub[(c1, c1)] = U[:, :k]
        sb[(c1, c2)] = np.diag(S[:k])
        vb[(c2, c2)] = Vt[:k, :]
        midl.append((c1, k))
        midr.append((c2, k))
    Vl = Space(midl, A.space[0].isdual)
    Vr = Space(midr, A.space[1].isdual)
    u = SymTensor(0, (A.space[0], Vl.dual()), ub)
    s = SymTensor(A.charge, (Vl, Vr), sb)
    v = SymTensor(0, (Vr.dual(), A.space[1]), vb)
    return u, s, v


def diag_values(s: SymTensor) -> np.ndarray:
    """``diag(A::SymDiagonal)``: all singular values, descending (symtensor.jl:71)."""
    vals = np.concatenate([np.diag(b) for b in s.blocks.values()]) if s.blocks else np.zeros(0)
    return np.sort(vals)[::-1]


def normalize_diag(s: SymTensor) -> SymTensor:
    """``normalize!(S::SymDiagonal) = rmul!(S, 1/norm(S))`` (symtensor.jl:325)."""
    return scale(s, 1.0 / norm(s))
