"""Device parity: SymTensors kernels (gather/transpose, grouped GEMM, BLAS-1) vs the oracle.

Bar: structure bit-exact; integer-valued KATs exact; FP64 values within
1e-12 relative (norm-wise per block); FP32 within 1e-5.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import symtensors as ost
from tests.helpers import (global_rel_err, o_from_flat, o_from_product, o_space, p_space, random_space,
                           rel_block_err)

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import mps_amd as M  # noqa: E402
from mps_amd import symtensors as pst  # noqa: E402

DEV = torch.device("cuda", 0)


def to_dev(O, dtype=torch.float64):
    flat = np.concatenate([np.asarray(b, np.float64).reshape(-1, order="F") for b in O.blocks.values()]) \
        if O.blocks else np.zeros(0)
    return pst.SymTensor.from_flat(O.charge, tuple(p_space(V) for V in O.space), flat, dtype, DEV)


def assert_same(P, O, tol=0.0):
    assert P.charge == O.charge
    assert tuple(o_space(V) for V in P.space) == O.space
    assert P.sectors() == list(O.blocks.keys())
    Q = o_from_product(P)
    if tol == 0.0:
        for k in O.blocks:
            assert np.array_equal(Q.blocks[k], np.asarray(O.blocks[k], np.float64)), k
    else:
        assert rel_block_err(Q, O) <= tol, rel_block_err(Q, O)


def test_contract_matrices_exact_kat():
    """test/symtensortests.jl:94-110 on the device (exact)."""
    vlist = (ost.U1Space((0, 2), (1, 3), (2, 4)), ost.U1Space((0, 4), (1, 2), (2, 2), (3, 3)).dual(),
             ost.U1Space((0, 4), (1, 2), (2, 2), (3, 3)), ost.U1Space((0, 3), (1, 5), (2, 2)).dual())
    A = ost.fill_linearindex(vlist[0:2])
    B = ost.fill(1.0, vlist[2:4])
    C = M.contract(to_dev(A), (1, -1), to_dev(B), (-1, 2))
    assert np.array_equal(M.array(C), ost.array(A) @ ost.array(B))


def test_u1_unitary_kat_device():
    """test/symtensortests.jl:41-53 on the device (exact values after splitleg o splitleg)."""
    A = ost.eye(ost.U1Space((0, 1), (1, 2), (2, 1)))
    A.blocks[(1, 1)] = np.array([[2.0, 3.0], [4.0, 5.0]])
    d = M.U1Space((0, 1), (1, 1))
    B = M.splitleg(M.splitleg(to_dev(A), 2, (d.dual(), d.dual())), 1, (d, d))
    want = sorted([(1, 1, 1, 1), (1, 0, 0, 1), (0, 1, 0, 1), (1, 0, 1, 0), (0, 1, 1, 0), (0, 0, 0, 0)],
                  key=lambda s: s[::-1])
    assert B.sectors() == want
    assert [float(b.reshape(-1)[0]) for b in B.blocks().values()] == [1.0, 2.0, 4.0, 3.0, 5.0, 1.0]


@pytest.mark.parametrize("perm", list(itertools.permutations([1, 2, 3, 4])))
def test_symmatrix_kat_device(perm):
    """test/symtensortests.jl:76-92 on the device, every permutation (exact)."""
    vlist = (ost.U1Space((0, 2), (1, 3), (2, 4)), ost.U1Space((-2, 4), (-1, 3), (0, 2), (1, 1)),
             ost.U1Space((-1, 1), (0, 2), (1, 3), (2, 4)), ost.U1Space((-2, 3), (-1, 2), (0, 1)))
    O = ost.fill_linearindex(vlist)
    A = to_dev(O)
    p = list(perm)
    want = ost.fuselegs(ost.fuselegs(ost.permutelegs(O, p), 3, 2, True), 1, 2)
    assert_same(M.symmatrix(A, p[:2], p[2:]), want)
    assert_same(M.fuselegs(M.fuselegs(M.permutelegs(A, p), 3, 2, True), 1, 2), want)


def test_fuse_split_roundtrip_device():
    """test/symtensortests.jl:3-29 and :126-128 on the device (exact round trips)."""
    vlist = (ost.U1Space((0, 2), (1, 2), (2, 2)), ost.U1Space((0, 2), (-1, 2), (-2, 2), (-3, 2)),
             ost.U1Space((0, 2), (1, 2), (2, 2)), ost.U1Space((0, 2), (1, 2)).dual())
    rng = np.random.default_rng(5)
    O = ost.from_function(lambda T, z: rng.standard_normal(z), 0, vlist)
    A = to_dev(O)
    for l, n in ((1, 3), (2, 2), (2, 3), (1, 2), (3, 2)):
        for d in (False, True):
            F = M.fuselegs(A, l, n, d)
            assert_same(F, ost.fuselegs(O, l, n, d))
            R = M.splitleg(F, l, A.space[l - 1:l - 1 + n])
            assert_same(R, O)


def _rand_case(rng):
    from tests.test_struct_cpu import _random_contract_case
    while True:
        case = _random_contract_case(rng)
        if case is not None:
            A, idxA, B, idxB = case
            A = ost.from_function(lambda T, z: rng.standard_normal(z), A.charge, A.space)
            B = ost.from_function(lambda T, z: rng.standard_normal(z), B.charge, B.space)
            return A, idxA, B, idxB


@pytest.mark.parametrize("seed", range(40))
def test_contract_random_f64(seed):
    rng = np.random.default_rng(1000 + seed)
    A, idxA, B, idxB = _rand_case(rng)
    want = ost.contract(A, tuple(idxA), B, tuple(idxB))
    got = M.contract(to_dev(A), idxA, to_dev(B), idxB)
    assert_same(got, want, tol=1e-12)


@pytest.mark.parametrize("seed", range(10))
def test_contract_random_f32(seed):
    rng = np.random.default_rng(2000 + seed)
    A, idxA, B, idxB = _rand_case(rng)
    want = ost.contract(A, tuple(idxA), B, tuple(idxB))
    got = M.contract(to_dev(A, torch.float32), idxA, to_dev(B, torch.float32), idxB)
    assert got.sectors() == list(want.blocks.keys())
    assert global_rel_err(o_from_product(got), want) < 1e-5


def test_contract_multileg_matches_dense_einsum():
    """test/symtensortests.jl:112-137 on the device: block-sparse == dense einsum."""
    vlist = (ost.U1Space((0, 2), (1, 2), (2, 2)), ost.U1Space((0, 2), (-1, 2), (-2, 2), (-3, 2)),
             ost.U1Space((0, 2), (1, 2), (2, 2)), ost.U1Space((0, 2), (1, 2)).dual(), ost.U1Space((0, 2), (1, 2)),
             ost.U1Space((0, 2), (1, 2), (2, 2), (3, 2), (4, 2)), ost.U1Space((0, 2), (1, 2)),
             ost.U1Space((0, 2), (-1, 2), (-2, 2)))
    rng = np.random.default_rng(1911)
    A = ost.from_function(lambda T, z: rng.random(z), 0, vlist[0:4])
    B = ost.from_function(lambda T, z: rng.random(z), 0, vlist[4:8])
    C = M.contract(to_dev(A), (1, 2, 3, -1), to_dev(B), (-1, 4, 5, 6))
    Cd = np.einsum("abcd,defg->abcefg", ost.array(A), ost.array(B))
    np.testing.assert_allclose(M.array(C), Cd, rtol=1e-12, atol=1e-12)


def test_empty_operand_gives_zero_blocks():
    """An operand with no allowed sector contracts to explicit zero blocks (contract.jl:292-294)."""
    V = ost.U1Space((0, 2), (1, 3))
    W = ost.U1Space((2, 2))
    U = ost.U1Space((-2, 2), (0, 1))
    A = ost.fill(1.0, (V, V.dual(), W), 0)  # c1 - c2 + 2 = 0 has no solution: no sectors
    assert not A.blocks
    B = ost.fill(1.0, (V, U), 0)
    want = ost.contract(A, (1, -1, 2), B, (-1, 3))
    assert want.blocks and all(not np.any(b) for b in want.blocks.values())
    got = M.contract(to_dev(A), (1, -1, 2), to_dev(B), (-1, 3))
    assert got.sectors() == list(want.blocks.keys())
    assert torch.count_nonzero(got.data).item() == 0


def test_xxz_mpo_kat_device():
    """test/mpotests.jl:3-59 through the device contract + fuselegs (exact)."""
    from oracle import tensornetalgs as ota
    Ws = [to_dev(W) for W in ota.xxz_u1mpo(4, 1.0).Ws]
    h = M.contract(M.contract(M.contract(Ws[0], (1, 2, -1, 5), Ws[1], (-1, 3, 4, 6)),
                              (1, 2, 3, -1, 6, 7), Ws[2], (-1, 4, 5, 8)),
                   (1, 2, 3, 4, -1, 7, 8, 9), Ws[3], (-1, 5, 6, 10))
    h = M.fuselegs(M.fuselegs(h, 1, 5), 2, 5, True)
    H4 = ota.xxz_dense_hamiltonian(4, 1.0)
    ones = np.array([bin(x).count("1") for x in range(16)])
    blocks = h.blocks()
    for k in range(5):
        idx = ones == k
        assert np.array_equal(H4[np.ix_(idx, idx)], blocks[(k, k)])


def test_vector_ops():
    rng = np.random.default_rng(9)
    V = ost.U1Space((0, 30), (1, 40), (2, 25))
    OA = ost.from_function(lambda T, z: rng.standard_normal(z), 0, (V, V.dual()))
    OB = ost.from_function(lambda T, z: rng.standard_normal(z), 0, (V, V.dual()))
    A, B = to_dev(OA), to_dev(OB)
    assert abs(M.dot(A, B) - ost.dot(OA, OB)) < 1e-12 * max(1.0, abs(ost.dot(OA, OB)))
    assert abs(M.norm(A) - ost.norm(OA)) < 1e-12 * ost.norm(OA)
    C = B.copy()
    M.axpby_(0.3, A, -1.7, C)
    assert_same(C, ost.axpby(0.3, OA, -1.7, OB), tol=1e-15)
    D = A.copy()
    M.rmul_(D, 2.5)
    assert_same(D, ost.scale(OA, 2.5), tol=1e-15)
    A32 = A.to(torch.float32)
    assert A32.dtype == torch.float32
    back = A32.to(torch.float64)
    assert global_rel_err(o_from_product(back), OA) < 1e-7
    with pytest.raises(ValueError):
        M.axpby_(1.0, A, 1.0, M.zeros((V, V)))


def test_dual_aliases_data_without_conjugation():
    V = M.U1Space((0, 2), (1, 1))
    A = M.fill(3.0, (V, V), 1, device=DEV)
    B = M.dual(A)
    assert B.charge == -1 and all(s.isdual for s in B.space)
    assert B.data.data_ptr() == A.data.data_ptr()
