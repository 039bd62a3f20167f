"""Numpy restatement of the TensorNetAlgs / MatrixProductStateTools hot path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Covers ``_applymps2site`` / ``_mpsupdateleft`` / ``_mpsupdateright``
(``src/apply.jl``), ``initialenv`` and ``dmrg2sitesweep!`` (``src/dmrg.jl``),
``xxz_u1mpo`` (``src/symmpo.jl:45-97``), ``center_at!`` / ``_pushleft!`` /
``normalize!`` / ``measure`` (``src/symmps.jl``), the dense XXZ Hamiltonian
(``src/xxz.jl:13-38``) and a restarted Lanczos standing in for KrylovKit's
``eigsolve`` (absent from the reference snapshot; trajectory parity unpinned).
"""
from __future__ import annotations

from typing import Callable, List

import numpy as np

from . import symtensors as st
from .symtensors import Space, SymTensor, U1Space, contract, dual


# ---------------------------------------------------------------------------
# H_eff and environment updates   (src/apply.jl)
# ---------------------------------------------------------------------------
def applymps2site(v: SymTensor, envL: SymTensor, envR: SymTensor, WL: SymTensor, WR: SymTensor) -> SymTensor:
    """``_applymps2site`` for SymTensors (apply.jl:93-105): four chained contracts."""
    T1 = contract(envL, (-1, 2, 1), v, (-1, 3, 4, 5))
    T2 = contract(T1, (1, -1, -2, 4, 5), WL, (-1, 2, 3, -2))
    T3 = contract(T2, (1, 2, -1, -2, 5), WR, (-1, 3, 4, -2))
    return contract(T3, (1, 2, 3, -1, -2), envR, (-2, -1, 4))


def applymps2site_steps(v, envL, envR, WL, WR):
    """The same chain, returning every intermediate (T1, T2, T3, out)."""
    T1 = contract(envL, (-1, 2, 1), v, (-1, 3, 4, 5))
    T2 = contract(T1, (1, -1, -2, 4, 5), WL, (-1, 2, 3, -2))
    T3 = contract(T2, (1, 2, -1, -2, 5), WR, (-1, 3, 4, -2))
    return T1, T2, T3, contract(T3, (1, 2, 3, -1, -2), envR, (-2, -1, 4))


def mpsupdateleft(lenv: SymTensor, A: SymTensor, W: SymTensor) -> SymTensor:
    """``_mpsupdateleft`` (apply.jl:49-58)."""
    return contract(contract(contract(lenv, (-1, 3, 4), A, (-1, 2, 1)),
                             (1, -1, -2, 4), W, (-2, 2, 3, -1)),
                    (1, -1, 2, -2), dual(A), (-2, -1, 3))


def mpsupdateright(renv: SymTensor, A: SymTensor, W: SymTensor) -> SymTensor:
    """``_mpsupdateright`` (apply.jl:17-25)."""
    return contract(contract(contract(renv, (-1, 3, 4), A, (1, 2, -1)),
                             (1, -1, -2, 4), W, (3, 2, -2, -1)),
                    (1, -1, 2, -2), dual(A), (3, -1, -2))


def dense_applymps2site(v, envL, envR, WL, WR):
    """Dense ``_applymps2site`` (apply.jl:80-91) as one einsum."""
    # envL[a,m,b] v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] envR[c,k,e] -> out[b,q,r,e],
    # pairwise in the @tensor order (((envL v) WL) WR) envR
    t = np.einsum("amb,apsc->mbpsc", envL, v)
    t = np.einsum("mbpsc,mqnp->bscqn", t, WL)
    t = np.einsum("bscqn,nrks->bcqrk", t, WR)
    return np.einsum("bcqrk,cke->bqre", t, envR)


def dense_mpsupdateleft(lenv, A, W):
    """Dense ``_mpsupdateleft`` (apply.jl:41-47); real data, so conj is a no-op."""
    t = np.einsum("pmq,pou->mqou", lenv, A)
    t = np.einsum("mqou,mzno->quzn", t, W)
    return np.einsum("quzn,qzd->und", t, np.conj(A))


def dense_mpsupdateright(renv, A, W):
    """Dense ``_mpsupdateright`` (apply.jl:9-15)."""
    t = np.einsum("pmq,uop->mquo", renv, A)
    t = np.einsum("mquo,nzmo->quzn", t, W)
    return np.einsum("quzn,dzq->und", t, np.conj(A))


# ---------------------------------------------------------------------------
# MPS / MPO containers   (src/symmps.jl, src/symmpo.jl)
# ---------------------------------------------------------------------------
class MPState:
    """``MPState`` (symmps.jl:1-4): site tensors ``As`` (legs l, p, r̄) and ``center``."""

    def __init__(self, As: List[SymTensor], center: int = 0):
        self.As = list(As)
        self.center = center

    def __len__(self):
        return len(self.As)

    def leftspace(self):
        return self.As[0].space[0]

    def rightspace(self):
        return self.As[-1].space[2].dual()


class MPOperator:
    """``MPOperator`` (symmpo.jl:1-3): site tensors ``Ws`` (legs wl, p_out, w̄r, p̄_in)."""

    def __init__(self, Ws: List[SymTensor]):
        self.Ws = list(Ws)

    def __len__(self):
        return len(self.Ws)

    def leftspace(self):
        return self.Ws[0].space[0]

    def rightspace(self):
        return self.Ws[-1].space[2].dual()


def _setblock(W: SymTensor, sector, blk):
    if sector not in W.blocks:
        raise KeyError("Block doesn't exist")
    W.blocks[sector] = np.asarray(blk, dtype=np.float64)


def xxz_u1mpo(lx: int, delta: float = 1.0) -> MPOperator:
    """``xxz_u1mpo(T, lx, d, delta)`` (symmpo.jl:45-97): D=5 U(1) XXZ MPO."""
    V0 = U1Space((0, 1))
    V = U1Space((-1, 1), (0, 3), (1, 1))
    d = U1Space((0, 1), (1, 1))
    space = (V, d, V.dual(), d.dual())
    Wbeg = st.fill(0.0, (V0,) + space[1:])
    W = st.fill(0.0, space)
    Wend = st.fill(0.0, space[:2] + (V0, space[3]))
    mat0 = np.array([[1.0, 0, 0], [delta * -0.5, 0, 0], [0, -0.5, 1]])
    mat1 = np.array([[1.0, 0, 0], [delta * 0.5, 0, 0], [0, 0.5, 1]])
    F = lambda a, shp: np.asarray(a, dtype=np.float64).reshape(shp, order="F")
    _setblock(W, (0, 0, 0, 0), F(mat0, (3, 1, 3, 1)))
    _setblock(W, (0, 1, 0, 1), F(mat1, (3, 1, 3, 1)))
    _setblock(W, (0, 1, 1, 0), F([0.0, 0.0, 1.0], (3, 1, 1, 1)))
    _setblock(W, (0, 0, -1, 1), F([0.0, 0.0, 1.0], (3, 1, 1, 1)))
    _setblock(W, (1, 0, 0, 1), 0.5 * F([1.0, 0.0, 0.0], (1, 1, 3, 1)))
    _setblock(W, (-1, 1, 0, 0), 0.5 * F([1.0, 0.0, 0.0], (1, 1, 3, 1)))
    _setblock(Wbeg, (0, 0, 0, 0), F(mat0[2, :], (1, 1, 3, 1)))
    _setblock(Wbeg, (0, 1, 0, 1), F(mat1[2, :], (1, 1, 3, 1)))
    _setblock(Wbeg, (0, 1, 1, 0), np.ones((1, 1, 1, 1)))
    _setblock(Wbeg, (0, 0, -1, 1), np.ones((1, 1, 1, 1)))
    _setblock(Wend, (0, 0, 0, 0), F(mat0[:, 0], (3, 1, 1, 1)))
    _setblock(Wend, (0, 1, 0, 1), F(mat1[:, 0], (3, 1, 1, 1)))
    _setblock(Wend, (-1, 1, 0, 0), 0.5 * np.ones((1, 1, 1, 1)))
    _setblock(Wend, (1, 0, 0, 1), 0.5 * np.ones((1, 1, 1, 1)))
    return MPOperator([Wbeg] + [W] * (lx - 2) + [Wend])


def dense_spin_half():
    """``spinoperators(1/2)`` dense (spin_definitions.jl:7-17): basis (up, down)."""
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T.copy()
    return sz, sp, sm


def xxz_dense_hamiltonian(lx: int, delta: float = 1.0) -> np.ndarray:
    """``xxz_explicit(lx, delta, :OBC)`` (xxz.jl:13-38), site 1 most significant."""
    sz, sp, sm = dense_spin_half()
    heis = 0.5 * (np.kron(sp, sm) + np.kron(sm, sp)) + delta * np.kron(sz, sz)
    H = heis
    for i in range(3, lx + 1):
        H = np.kron(H, np.eye(2)) + np.kron(np.eye(2 ** (i - 2)), heis)
    return H


def mpo_dense_chain(mpo: MPOperator) -> np.ndarray:
    """Dense Hamiltonian from a U(1) MPO by chaining ``array()`` embeddings.

    Physical index of each site: row = p_out, col = p_in; site 1 slowest.
    """
    Ws = [st.array(W) for W in mpo.Ws]
    H = Ws[0][0]  # (p_out, wr, p_in)
    H = np.transpose(H, (0, 2, 1))  # (po, pi, w)
    for W in Ws[1:]:
        # H[(P_out), (P_in), w] x W[w, po, w', pi]
        H = np.einsum("abw,wcvd->acbdv", H, W)
        sh = H.shape
        H = H.reshape(sh[0] * sh[1], sh[2] * sh[3], sh[4])
    return H[:, :, 0]


# ---------------------------------------------------------------------------
# MPS tools   (src/symmps.jl:150-200, 354-371)
# ---------------------------------------------------------------------------
def _pushleft(mps: MPState, l: int):
    """``_pushleft!`` (symmps.jl:167-174); l is 1-based."""
    A = mps.As[l - 1]
    u, s, v = st.svd_(st.fuselegs(A, 2, 2))
    mps.As[l - 1] = st.splitleg(v, 2, A.space[1:3])
    mps.As[l - 2] = contract(mps.As[l - 2], (1, 2, -1), st.matmul(u, s), (-1, 3))


def _pushright(mps: MPState, l: int):
    """``_pushright!`` (symmps.jl:176-183)."""
    A = mps.As[l - 1]
    u, s, v = st.svd_(st.fuselegs(A, 1, 2))
    mps.As[l - 1] = st.splitleg(u, 1, A.space[0:2])
    mps.As[l] = contract(st.matmul(s, v), (1, -1), mps.As[l], (-1, 2, 3))


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


def normalize_mps(mps: MPState) -> MPState:
    """``normalize!(mps)`` (symmps.jl:150-165)."""
    if mps.center < 1:
        center_at(mps, 1)
    A = mps.As[mps.center - 1]
    if mps.center > len(mps) // 2:
        u, s, v = st.svd_(st.fuselegs(A, 1, 2, False))
        s = st.normalize_diag(s)
        mps.As[mps.center - 1] = st.splitleg(st.matmul(st.matmul(u, s), v), 1, A.space[0:2])
    else:
        u, s, v = st.svd_(st.fuselegs(A, 2, 2, True))
        s = st.normalize_diag(s)
        mps.As[mps.center - 1] = st.splitleg(st.matmul(st.matmul(u, s), v), 2, A.space[1:3])
    return mps


def measure(mps: MPState, mpo: MPOperator) -> float:
    """``measure(mps, mpo)`` (symmps.jl:354-371): <psi|H|psi> by a left sweep."""
    dummy = U1Space((0, 1))
    L = st.fill(1.0, (dummy, dummy, dummy))
    for A, W in zip(mps.As, mpo.Ws):
        L = contract(contract(contract(L, (-1, 3, 4), A, (-1, 2, 1)),
                              (1, -1, -2, 4), W, (-2, 3, 2, -1)),
                     (1, 2, -1, -2), dual(A), (-2, -1, 3))
    return float(contract(L, (-1, -2, -3), st.fill(1.0, (dummy, dummy, dummy)), (-1, -2, -3)))


def initialenv(mps: MPState, mpo: MPOperator, at: int = 1) -> List[SymTensor]:
    """``initialenv(mps, mpo; at)`` (dmrg.jl:9-36); returns env[1..L+2] as a 0-based list."""
    if mps.center != at:
        raise ValueError(f"MPS center has to be {at}")
    lx = len(mps)
    env: List[SymTensor] = [None] * (lx + 2)
    lVa, rVa = mps.leftspace(), mps.rightspace()
    lVw, rVw = mpo.leftspace(), mpo.rightspace()
    env[0] = st.fill(1.0, (lVa.dual(), lVw.dual(), lVa))
    env[lx + 1] = st.fill(1.0, (rVa, rVw, rVa.dual()))
    for l in range(lx, at - 1, -1):
        env[l] = mpsupdateright(env[l + 1], mps.As[l - 1], mpo.Ws[l - 1])
    for l in range(1, at):
        env[l] = mpsupdateleft(env[l - 1], mps.As[l - 1], mpo.Ws[l - 1])
    return env


# ---------------------------------------------------------------------------
# Lanczos (stands in for KrylovKit.eigsolve; dmrg.jl:77-82 call semantics)
# ---------------------------------------------------------------------------
class EigInfo:
    def __init__(self, normres, numiter, numops, converged):
        self.normres, self.numiter, self.numops, self.converged = normres, numiter, numops, converged


def lanczos_lowest(apply: Callable, x0, dotf, axpby, scal, norm, tol=1e-7, krylovdim=5, maxiter=8):
    """Restarted Lanczos for the lowest eigenpair of a Hermitian map.

    Parity unpinned against KrylovKit (absent from the reference snapshot):
    the trajectory is this build's own.  Each cycle builds up to
    ``krylovdim`` fully re-orthogonalised vectors from the current Ritz
    vector; convergence when the residual norm |beta_k y_k| < tol.  Returns
    (theta, x, EigInfo).  The vector algebra comes in as callables so the
    same routine drives numpy SymTensors here and device vectors elsewhere.
    """
    x = scal(x0, 1.0 / norm(x0))
    numops, theta, res = 0, None, np.inf
    it = 0
    for it in range(1, maxiter + 1):
        V = [x]
        alphas, betas = [], []
        w = apply(x)
        numops += 1
        y = np.array([1.0])
        for j in range(krylovdim):
            a = float(dotf(V[j], w))
            alphas.append(a)
            w = axpby(-a, V[j], 1.0, w)
            if j > 0:
                w = axpby(-betas[j - 1], V[j - 1], 1.0, w)
            for i in range(j + 1):  # full re-orthogonalisation
                c = float(dotf(V[i], w))
                w = axpby(-c, V[i], 1.0, w)
            b = float(norm(w))
            T = np.diag(alphas) + np.diag(betas, 1) + np.diag(betas, -1)
            evals, evecs = np.linalg.eigh(T)
            theta, y = float(evals[0]), evecs[:, 0]
            res = abs(b * y[-1])
            if res < tol or j == krylovdim - 1 or b < 1e-14:
                break
            betas.append(b)
            V.append(scal(w, 1.0 / b))
            w = apply(V[-1])
            numops += 1
        x = scal(V[0], y[0])
        for i in range(1, len(y)):
            x = axpby(y[i], V[i], 1.0, x)
        x = scal(x, 1.0 / norm(x))
        if res < tol:
            break
    return theta, x, EigInfo(res, it, numops, res < tol)


def eigsolve_symtensor(f: Callable, x0: SymTensor, tol=1e-7, krylovdim=5, maxiter=8):
    return lanczos_lowest(f, x0, st.dot, st.axpby, st.scale, st.norm, tol, krylovdim, maxiter)


# ---------------------------------------------------------------------------
# Two-site DMRG sweep   (src/dmrg.jl:58-179)
# ---------------------------------------------------------------------------
def dmrg2sitesweep(mps: MPState, mpo: MPOperator, env: List[SymTensor], maxdim=200, tol=1e-9,
                   lanczostol=1e-7, krylovdim=5, krylovmaxiter=8, verbose=False,
                   heff: Callable = None, log: list = None) -> List[float]:
    """``dmrg2sitesweep!`` (dmrg.jl:58-179), L->R then R->L, mutating mps/env.

    Reproduces quirk Q6: in the R->L loop each energy is pushed twice
    (dmrg.jl:132 and :156, the second one stale).  ``heff`` may replace the
    H_eff map (default: the oracle contraction chain).
    """
    H = heff or applymps2site
    lx = len(mps)
    if mps.center != 1:
        raise ValueError("The center of MPS has to be 1 for dmrg L->R->L")
    Ws = mpo.Ws
    energies: List[float] = []

    def solve(AA, l):  # l is 1-based bond (l, l+1)
        f = lambda x: H(x, env[l - 1], env[l + 2], Ws[l - 1], Ws[l])
        theta, v, info = eigsolve_symtensor(f, AA, lanczostol, krylovdim, krylovmaxiter)
        if log is not None:
            log.append(("eig", l, theta, info.normres, info.numops))
        return v

    def trunc(v, AA):
        u, s, w = st.svdtrunc(st.symmatrix(v, [1, 2], [3, 4]), maxdim=maxdim, tol=tol)
        if log is not None:
            log.append(("trunc", 1 - st.norm(s)))
        s = st.normalize_diag(s)
        AAn = st.splitleg(st.splitleg(st.matmul(st.matmul(u, s), w), 2, AA.space[2:4]), 1, AA.space[0:2])
        return u, s, w, AAn

    A = mps.As[0]
    for l in range(1, lx - 1):
        AA = contract(A, (1, 2, -1), mps.As[l], (-1, 3, 4))
        v = solve(AA, l)
        u, s, w, AA2 = trunc(v, AA)
        e = float(st.dot(AA2, H(AA2, env[l - 1], env[l + 2], Ws[l - 1], Ws[l])))
        energies.append(e)
        mps.As[l - 1] = st.splitleg(u, 1, A.space[0:2])
        env[l] = mpsupdateleft(env[l - 1], mps.As[l - 1], Ws[l - 1])
        A = st.splitleg(st.matmul(s, w), 2, AA2.space[2:4])
    l = lx - 1
    AA = contract(A, (1, 2, -1), mps.As[l], (-1, 3, 4))
    v = solve(AA, l)
    e = None
    for l in range(lx - 1, 1, -1):
        u, s, w, AA2 = trunc(v, AA)
        e = float(st.dot(AA2, H(AA2, env[l - 1], env[l + 2], Ws[l - 1], Ws[l])))
        energies.append(e)
        mps.As[l] = st.splitleg(w, 2, AA2.space[2:4])
        env[l + 1] = mpsupdateright(env[l + 2], mps.As[l], Ws[l])
        A = st.splitleg(st.matmul(u, s), 1, AA2.space[0:2])
        AA = contract(mps.As[l - 2], (1, 2, -1), A, (-1, 3, 4))
        v = solve(AA, l - 1)
        energies.append(e)
    l = 1
    u, s, w, AA2 = trunc(v, AA)
    e = float(st.dot(AA2, H(AA2, env[l - 1], env[l + 2], Ws[l - 1], Ws[l])))
    energies.append(e)
    mps.As[l] = st.splitleg(w, 2, AA2.space[2:4])
    env[l + 1] = mpsupdateright(env[l + 2], mps.As[l], Ws[l])
    mps.As[l - 1] = st.splitleg(st.matmul(u, s), 1, AA2.space[0:2])
    return energies


# ---------------------------------------------------------------------------
# Exact diagonalisation in a fixed-magnetisation block   (src/xxz.jl:56-113)
# ---------------------------------------------------------------------------
def u1_block_hamiltonian(L: int, bonds, M: int, jz=1.0, jxy=1.0):
    """Sparse H on the states with M up spins (bit i = site i+1 up), for two-site terms on ``bonds``.

    Restates ``xxz_u1sym`` (xxz.jl:56-113) for an arbitrary bond list: each
    parallel pair adds +jz/4, each antiparallel pair adds -jz/4 and a
    jxy/2 spin-flip matrix element.  ``bonds`` are 1-based (i, j) pairs.
    """
    import itertools
    import scipy.sparse as sp
    states = [sum(1 << b for b in c) for c in itertools.combinations(range(L), M)]
    states.sort()
    index = {s: k for k, s in enumerate(states)}
    I, J, V = [], [], []
    for k, s in enumerate(states):
        for (i, j) in bonds:
            a, b = i - 1, j - 1
            if ((s >> a) & 1) == ((s >> b) & 1):
                I.append(k); J.append(k); V.append(0.25 * jz)
            else:
                I.append(k); J.append(k); V.append(-0.25 * jz)
                t = s ^ ((1 << a) | (1 << b))
                I.append(k); J.append(index[t]); V.append(0.5 * jxy)
    n = len(states)
    return sp.csr_matrix((V, (I, J)), shape=(n, n))


def u1_ground_energy(L: int, bonds, M: int, jz=1.0, jxy=1.0) -> float:
    H = u1_block_hamiltonian(L, bonds, M, jz, jxy)
    if H.shape[0] <= 2000:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    import scipy.sparse.linalg as sla
    return float(sla.eigsh(H, k=1, which="SA", tol=1e-13)[0][0])


def random_u1mps_oracle(L, N, bond_pairs, flats):
    """Oracle MPS from given bond charge->dim lists and flat Gaussian buffers (for parity tests)."""
    Vd = st.U1Space((0, 1), (1, 1))
    As = []
    for l in range(1, L + 1):
        Vl, Vr = st.Space(bond_pairs[l - 1]), st.Space(bond_pairs[l])
        space = (Vl, Vd, Vr.dual())
        sects, sizes = st.allsectorsandsizes(0, space)
        blocks, p = {}, 0
        for s, z in zip(sects, sizes):
            n = int(np.prod(z))
            blocks[s] = np.asarray(flats[l - 1][p:p + n]).reshape(z, order="F")
            p += n
        As.append(SymTensor(0, space, blocks, check=True))
    return MPState(As, 0)
