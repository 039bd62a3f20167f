"""Device batched per-sector SVD (``mpsh_svd_*``) and svdtrunc vs the oracle.

Pinned by mathematics, not by the oracle's formula: U diag(S) Vh reconstructs
each block, U and Vh have orthonormal columns / rows, S is descending and
equals LAPACK's singular values; svdtrunc keeps exactly the oracle's counts
(mpsutils.jl:89-112 greedy, absolute tol) and its u*s*v equals the oracle's
truncated product (gauge-invariant).  FP64 bar: 1e-12 relative.  The
one-sided Jacobi path (default) and the rocSOLVER gesdd path
(``MPSH_SVD_ALGO=gesdd``) are checked against each other and against LAPACK.
"""
import os

import numpy as np
import pytest
import torch

from oracle import symtensors as ost
from tests.helpers import o_from_product

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import mps_amd as M  # noqa: E402
from mps_amd import mpstools as mt  # noqa: E402

DEV = torch.device("cuda", 0)


def _matrix(rows, cols, charge=0, seed=5, scale=None):
    """Rank-2 SymTensor with legs (rows, cols dual): block (c, c - charge)."""
    V1 = M.U1Space(*rows)
    V2 = M.U1Space(*cols, isdual=True)
    A = M.randn((V1, V2), charge, seed=seed, stream=3, device=DEV)
    if scale is not None:  # graded spectrum so truncation has something to cut
        blocks = A.blocks()
        for k, b in blocks.items():
            m, n = b.shape
            u, _, vt = np.linalg.svd(b, full_matrices=False)
            s = scale ** np.arange(min(m, n)) * (1.0 + 0.1 * abs(k[0]))
            blocks[k] = (u * s) @ vt
        A = M.SymTensor.from_blocks(A.charge, A.space, blocks, dtype=torch.float64, device=DEV)
    return A


SHAPES = [
    ([(0, 1), (1, 3), (2, 2)], [(0, 2), (1, 1), (2, 5)]),                 # tiny, m < n and m > n
    ([(q, 17 + 3 * q) for q in range(6)], [(q, 40 - 2 * q) for q in range(6)]),
    ([(0, 130), (1, 250), (2, 300)], [(0, 129), (1, 520), (2, 301)]),    # crosses padding granularities
    ([(q, 64) for q in range(9)], [(q, 64) for q in range(9)]),          # one bucket, nine blocks
]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_svd_blocks_reconstruct_and_orthonormal(shape):
    rows, cols = SHAPES[shape]
    A = _matrix(rows, cols, seed=11 + shape)
    blocks = A.blocks()
    facts = mt._svd_blocks(A)
    assert [f[0] for f in facts] == list(blocks.keys())
    for key, U, S, Vh in facts:
        b = blocks[key]
        m, n = b.shape
        k = min(m, n)
        U, S, Vh = U.cpu().numpy(), S.cpu().numpy(), Vh.cpu().numpy()
        assert U.shape == (m, k) and S.shape == (k,) and Vh.shape == (k, n)
        assert np.all(np.diff(S) <= 1e-12 * S[0])
        np.testing.assert_allclose(S, np.linalg.svd(b, compute_uv=False), rtol=1e-12, atol=1e-12 * S[0])
        assert np.linalg.norm((U * S) @ Vh - b) <= 1e-12 * np.linalg.norm(b)
        np.testing.assert_allclose(U.T @ U, np.eye(k), atol=1e-12)
        np.testing.assert_allclose(Vh @ Vh.T, np.eye(k), atol=1e-12)


def test_svd_of_charged_matrix():
    A = _matrix([(q, 5 + q) for q in range(5)], [(q, 7) for q in range(5)], charge=1, seed=3)
    r = mt.svd_(A)
    assert np.linalg.norm(M.array(r.usv()) - M.array(A)) <= 1e-12 * np.linalg.norm(M.array(A))


@pytest.mark.parametrize("maxdim,tol", [(40, 1e-12), (12, 1e-12), (1000, 1e-3)])
def test_svdtrunc_matches_oracle(maxdim, tol):
    A = _matrix([(q, 9 + 2 * q) for q in range(5)], [(q, 14 - q) for q in range(5)], seed=21, scale=0.5)
    O = o_from_product(A)
    r = mt.svdtrunc(A, maxdim=maxdim, tol=tol)
    u, s, v = ost.svdtrunc(O, maxdim=maxdim, tol=tol)
    kept_dev = {c: d for c, d in r.Vl.items()}
    kept_or = dict(s.space[0].items())
    assert kept_dev == kept_or
    sv_or = np.sort(np.concatenate([np.diag(b) for b in s.blocks.values()]))[::-1]
    np.testing.assert_allclose(r.singular_values(), sv_or, rtol=1e-12, atol=1e-14)
    usv_or = ost.contract(ost.contract(u, (1, -1), s, (-1, 2)), (1, -1), v, (-1, 2))
    got = M.array(r.usv())
    want = ost.array(usv_or)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("charge,maxdim,normalize", [(0, 12, False), (1, 30, True), (1, 10_000, False)])
def test_truncated_factors_gathered_in_one_launch(charge, maxdim, normalize):
    """u, v, u*s, s*v come from one mpsh_gather_blocks launch each; every block
    equals the truncated factor built block by block with torch (bit for bit:
    one IEEE product per element either way)."""
    A = _matrix([(q, 9 + 2 * q) for q in range(5)], [(q, 14 - q) for q in range(5)], charge=charge, seed=22,
                scale=0.6)
    r = mt.svdtrunc(A, maxdim=maxdim, tol=1e-13)
    if normalize:
        r.normalize_s()
    assert r.keys and any(U.shape[1] < min(U.shape[0], Vh.shape[1]) for U, Vh in zip(r.U, r.Vh)) == (maxdim < 100)
    want = {"u": {}, "v": {}, "us": {}, "sv": {}}
    for k, U, s, Vh in zip(r.keys, r.U, r.S, r.Vh):
        want["u"][(k[0], k[0])] = U.cpu().numpy()
        want["v"][(k[1], k[1])] = Vh.cpu().numpy()
        want["us"][(k[0], k[1])] = (U * s).cpu().numpy()
        want["sv"][(k[0], k[1])] = (s[:, None] * Vh).cpu().numpy()
    for name in want:
        got = getattr(r, name)().blocks()
        assert set(got) == set(want[name]), name
        for key, blk in got.items():
            assert np.array_equal(blk, want[name][key]), (name, key)


def _blocks_with(A, algo, precond=True):
    """_svd_blocks through a fresh plan built with MPSH_SVD_ALGO=algo and
    MPSH_SVD_PRECOND (plans read both when built)."""
    env = {"MPSH_SVD_ALGO": algo, "MPSH_SVD_PRECOND": "1" if precond else "0"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        P = mt._SvdPlan(A.struct)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    saved = mt._SVD_PLANS.get(A.struct)
    mt._SVD_PLANS[A.struct] = P
    try:
        return mt._svd_blocks(A)
    finally:
        if saved is None:
            del mt._SVD_PLANS[A.struct]
        else:
            mt._SVD_PLANS[A.struct] = saved


def _check_facts(blocks, facts, rtol=1e-12):
    for key, U, S, Vh in facts:
        b = blocks[key]
        m, n = b.shape
        k = min(m, n)
        U, S, Vh = U.cpu().numpy(), S.cpu().numpy(), Vh.cpu().numpy()
        assert U.shape == (m, k) and S.shape == (k,) and Vh.shape == (k, n)
        assert np.all(np.diff(S) <= 0)
        s_ref = np.linalg.svd(b, compute_uv=False)
        scale = max(s_ref[0], 1e-300)
        np.testing.assert_allclose(S, s_ref, rtol=rtol, atol=rtol * scale)
        assert np.linalg.norm((U * S) @ Vh - b) <= rtol * max(np.linalg.norm(b), 1e-300)
        np.testing.assert_allclose(U.T @ U, np.eye(k), atol=rtol)
        np.testing.assert_allclose(Vh @ Vh.T, np.eye(k), atol=rtol)


@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_jacobi_and_gesdd_paths_agree(shape):
    rows, cols = SHAPES[shape]
    A = _matrix(rows, cols, seed=31 + shape)
    blocks = A.blocks()
    fj = _blocks_with(A, "jacobi")
    fg = _blocks_with(A, "gesdd")
    _check_facts(blocks, fj)
    _check_facts(blocks, fg)
    for (kj, _, Sj, _), (kg, _, Sg, _) in zip(fj, fg):
        assert kj == kg
        np.testing.assert_allclose(Sj.cpu().numpy(), Sg.cpu().numpy(), rtol=1e-12, atol=1e-12 * float(Sg[0]))


def test_jacobi_graded_spectrum_relative_accuracy():
    """Singular values spanning 1e-13..1: the Jacobi path keeps them to 1e-12 of the largest."""
    A = _matrix([(q, 60 + 30 * q) for q in range(4)], [(q, 70 + 25 * q) for q in range(4)], seed=41, scale=0.8)
    blocks = A.blocks()
    _check_facts(blocks, _blocks_with(A, "jacobi"))


@pytest.mark.parametrize("scale", [1, 5])
def test_jacobi_rank_deficient_and_zero_blocks(scale):
    """Exactly zero singular values: the completed singular vectors stay orthonormal.
    scale 5: the zero and rank-deficient blocks have 200+ columns and take the
    Gram-preconditioned path."""
    A = _matrix([(0, 40 * scale), (1, 12), (2, 70 * scale)],
                [(0, 25 if scale == 1 else 40 * scale + 10), (1, 12), (2, 90 * scale)], seed=43)
    blocks = A.blocks()
    keys = list(blocks.keys())
    blocks[keys[0]] = np.zeros_like(blocks[keys[0]])                  # all-zero block
    b = blocks[keys[2]]
    rk = 45 * scale
    b[:, rk:] = b[:, :rk] @ np.random.default_rng(3).standard_normal((rk, b.shape[1] - rk))  # rank 45 scale
    blocks[keys[1]][:, 5] = 0.0                                         # one zero column
    Z = M.SymTensor.from_blocks(A.charge, A.space, blocks, dtype=torch.float64, device=DEV)
    facts = _blocks_with(Z, "jacobi")
    for key, U, S, Vh in facts:
        bb = blocks[key]
        k = min(bb.shape)
        U, S, Vh = U.cpu().numpy(), S.cpu().numpy(), Vh.cpu().numpy()
        np.testing.assert_allclose(U.T @ U, np.eye(k), atol=1e-12)
        np.testing.assert_allclose(Vh @ Vh.T, np.eye(k), atol=1e-12)
        assert np.linalg.norm((U * S) @ Vh - bb) <= 1e-12 * max(1.0, np.linalg.norm(bb))
        s_ref = np.linalg.svd(bb, compute_uv=False)
        np.testing.assert_allclose(S, s_ref, atol=1e-12 * max(1.0, s_ref[0]))
    assert np.all(facts[0][2].cpu().numpy() == 0.0)


def test_jacobi_wide_blocks_and_rocsolver_fallback():
    """Blocks with more than 2048 rows go to gesdd inside the same plan; w = 8 and 4 classes."""
    A = _matrix([(0, 600), (1, 1500), (2, 2060)], [(0, 700), (1, 40), (2, 24)], seed=47)
    blocks = A.blocks()
    _check_facts(blocks, _blocks_with(A, "jacobi"))


@pytest.mark.parametrize("scale", [1, 3])
def test_svd_nonfinite_input_raises(scale):
    A = _matrix([(0, 50 * scale), (1, 80 * scale)], [(0, 60 * scale), (1, 90 * scale)], seed=53)
    blocks = A.blocks()
    k1 = list(blocks.keys())[1]
    blocks[k1][3, 7] = np.nan
    Z = M.SymTensor.from_blocks(A.charge, A.space, blocks, dtype=torch.float64, device=DEV)
    with pytest.raises(Exception, match="not finite"):
        _blocks_with(Z, "jacobi")


@pytest.mark.parametrize("big", [40, 220])
def test_jacobi_extreme_block_scales(big):
    """Blocks scaled to 1e-120 .. 1e150 (DMRG two-site blocks far from the mean
    charge are ~1e-120): each block's factors match LAPACK relative to its own scale."""
    A = _matrix([(q, [2, 2, 7, big, 2][q]) for q in range(5)], [(q, [2, 3, 7, big - 5, 2][q]) for q in range(5)],
                seed=59)
    blocks = A.blocks()
    for key, f in zip(list(blocks.keys()), (1e-120, 1e-60, 1e-140, 1e150, 3e-200)):
        blocks[key] = blocks[key] * f
    Z = M.SymTensor.from_blocks(A.charge, A.space, blocks, dtype=torch.float64, device=DEV)
    _check_facts(blocks, _blocks_with(Z, "jacobi"))


def test_preconditioned_and_plain_jacobi_agree():
    """Graded blocks of 100-400 columns (a DMRG-like spectrum spanning 1 .. 1e-9):
    the Gram-preconditioned sweeps (default) and the plain sweeps
    (MPSH_SVD_PRECOND=0) both match LAPACK to 1e-12 of the largest singular value
    and reconstruct the block; their singular values agree to the same bar."""
    rows = [(0, 100), (1, 260), (2, 410)]
    cols = [(0, 120), (1, 240), (2, 400)]
    A = _matrix(rows, cols, seed=61)
    blocks = A.blocks()
    rng = np.random.default_rng(7)
    for k, b in blocks.items():
        m, n = b.shape
        u, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
        v, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
        s = np.logspace(0, -9, min(m, n))
        blocks[k] = (u * s) @ v.T
    Z = M.SymTensor.from_blocks(A.charge, A.space, blocks, dtype=torch.float64, device=DEV)
    fp = _blocks_with(Z, "jacobi", precond=True)
    f0 = _blocks_with(Z, "jacobi", precond=False)
    _check_facts(blocks, fp)
    _check_facts(blocks, f0)
    for (_, _, Sp, _), (_, _, S0, _) in zip(fp, f0):
        np.testing.assert_allclose(Sp.cpu().numpy(), S0.cpu().numpy(), rtol=1e-12, atol=1e-12 * float(S0[0]))


@pytest.mark.parametrize("rows,cols", [((924, 300), (792, 280)), ((1500, 120), (300, 110))])
def test_jacobi_dmrg_sized_graded_blocks(rows, cols):
    """Block sizes of a chi=2048 DMRG sweep (924 x 792: w = 8 columns per block,
    rows 1500: w = 4) with spectra graded from 1 to 1e-9 like the sweep's
    two-site matrices: the Gram-preconditioned Jacobi path matches LAPACK to
    1e-12 of the largest singular value, reconstructs each block and returns
    orthonormal factors."""
    A = _matrix([(0, rows[0]), (1, rows[1])], [(0, cols[0]), (1, cols[1])], seed=67)
    blocks = A.blocks()
    rng = np.random.default_rng(11)
    for k, b in blocks.items():
        m, n = b.shape
        u, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
        v, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
        blocks[k] = (u * np.logspace(0, -9, min(m, n))) @ v.T
    Z = M.SymTensor.from_blocks(A.charge, A.space, blocks, dtype=torch.float64, device=DEV)
    _check_facts(blocks, _blocks_with(Z, "jacobi"))
