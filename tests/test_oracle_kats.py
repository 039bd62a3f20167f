"""The oracle against the reference's own known-answer tests.

Each test ports one reference test set; expected values are the ones the
reference asserts (exact ``==`` where the reference uses ``==``).
"""
import itertools

import numpy as np
import pytest

from oracle import symtensors as st
from oracle import tensornetalgs as ta
from oracle.symtensors import U1Space, dual


def test_fuse_split_roundtrip():
    """test/symtensortests.jl:3-29 (exact)."""
    space = (U1Space((0, 2), (1, 3)), U1Space((0, 1), (1, 1)), U1Space((0, 3), (1, 4)).dual())
    sects = [(0, 0, 0), (0, 1, 1), (1, 0, 1)]
    F = lambda a, z: np.array(a, dtype=np.int64).reshape(z, order="F")
    nz = [F(range(1, 7), (2, 1, 3)), F(range(7, 15), (2, 1, 4)), F(range(15, 27), (3, 1, 4))]
    A0 = st.SymTensor(0, space, dict(zip(sects, nz)), check=True)
    A1 = st.fuselegs(A0, 1, 2)
    assert st.tensors_equal(A0, st.splitleg(A1, 1, space[0:2]))
    A1 = st.fuselegs(A0, 2, 2, True)
    assert st.tensors_equal(A0, st.splitleg(A1, 2, space[1:3]))


def test_u1_unitary_kat():
    """test/symtensortests.jl:41-53 (exact sector list and block values)."""
    A = st.eye(U1Space((0, 1), (1, 2), (2, 1)))
    A.blocks[(1, 1)] = np.array([[2.0, 3.0], [4.0, 5.0]])
    d = U1Space((0, 1), (1, 1))
    cod, dom = (d, d), (d.dual(), d.dual())
    B = st.splitleg(st.splitleg(A, 2, dom), 1, cod)
    expected = sorted([(1, 1, 1, 1), (1, 0, 0, 1), (0, 1, 0, 1), (1, 0, 1, 0), (0, 1, 1, 0), (0, 0, 0, 0)],
                      key=st.sector_key)
    assert B.sectors() == expected
    i = np.ones((1, 1, 1, 1))
    vals = list(B.blocks.values())
    for got, want in zip(vals, [i, 2 * i, 4 * i, 3 * i, 5 * i, 1 * i]):
        assert np.array_equal(got, want)


@pytest.mark.parametrize("perm", list(itertools.permutations([1, 2, 3, 4])))
def test_symmatrix_equals_fuselegs(perm):
    """test/symtensortests.jl:76-92 (exact), for every permutation instead of one random one."""
    vlist = (U1Space((0, 2), (1, 3), (2, 4)), U1Space((-2, 4), (-1, 3), (0, 2), (1, 1)),
             U1Space((-1, 1), (0, 2), (1, 3), (2, 4)), U1Space((-2, 3), (-1, 2), (0, 1)))
    A = st.fill_linearindex(vlist)
    if perm == (1, 2, 3, 4):
        M = st.symmatrix(A, [1, 2], [3, 4])
        assert st.tensors_equal(M, st.fuselegs(st.fuselegs(A, 3, 2, True), 1, 2))
        assert st.tensors_equal(M, st.fuselegs(st.fuselegs(A, 1, 2), 2, 2, True))
    p = list(perm)
    assert st.tensors_equal(st.symmatrix(A, p[:2], p[2:]),
                            st.fuselegs(st.fuselegs(st.permutelegs(A, p), 3, 2, True), 1, 2))


def test_contract_matrices_exact():
    """test/symtensortests.jl:94-110 (exact integer-valued product)."""
    vlist = (U1Space((0, 2), (1, 3), (2, 4)), U1Space((0, 4), (1, 2), (2, 2), (3, 3)).dual(),
             U1Space((0, 4), (1, 2), (2, 2), (3, 3)), U1Space((0, 3), (1, 5), (2, 2)).dual())
    A = st.fill_linearindex(vlist[0:2])
    B = st.fill(1.0, vlist[2:4])
    C = st.contract(A, (1, -1), B, (-1, 2))
    assert np.array_equal(st.array(A) @ st.array(B), st.array(C))


def _rand_tensor(rng, space, charge=0):
    return st.from_function(lambda T, z: rng.random(z), charge, space)


def test_contract_multileg_dense():
    """test/symtensortests.jl:112-137: block-sparse contract == dense einsum; split o fuse == id."""
    rng = np.random.default_rng(1911)
    vlist = (U1Space((0, 2), (1, 2), (2, 2)), U1Space((0, 2), (-1, 2), (-2, 2), (-3, 2)),
             U1Space((0, 2), (1, 2), (2, 2)), U1Space((0, 2), (1, 2)).dual(), U1Space((0, 2), (1, 2)),
             U1Space((0, 2), (1, 2), (2, 2), (3, 2), (4, 2)), U1Space((0, 2), (1, 2)),
             U1Space((0, 2), (-1, 2), (-2, 2)))
    A = _rand_tensor(rng, vlist[0:4])
    B = _rand_tensor(rng, vlist[4:8])
    for l, n in ((1, 3), (2, 2), (2, 3)):
        R = st.splitleg(st.fuselegs(A, l, n), l, vlist[l - 1:l - 1 + n])
        assert st.tensors_equal(R, A)
    C = st.contract(A, (1, 2, 3, -1), B, (-1, 4, 5, 6))
    Cd = np.einsum("abcd,defg->abcefg", st.array(A), st.array(B))
    np.testing.assert_allclose(st.array(C), Cd, rtol=1e-13, atol=1e-13)


def test_dot_matches_dense():
    """test/symtensortests.jl:140-156 (complex dot)."""
    rng = np.random.default_rng(7)
    vlist = (U1Space((0, 2), (1, 3), (2, 4)), U1Space((0, 4), (-1, 2), (-2, 2), (-3, 3)),
             U1Space((0, 3), (1, 5), (2, 2)), U1Space((0, 6), (-1, 7)))
    A = st.from_function(lambda T, z: rng.random(z) + 1j * rng.random(z), 0, vlist, np.complex128)
    Ad = st.array(A)
    assert np.isclose(np.vdot(Ad, Ad), st.dot(A, A))


@pytest.mark.parametrize("charge", [0, 1, -1])
def test_svd_reconstructs(charge):
    """test/symtensortests.jl:158-190: u*s*v == A; _svd_ == svdtrunc (no truncation)."""
    rng = np.random.default_rng(3 + charge)
    V1 = U1Space((0, 2), (1, 3), (2, 4))
    V2 = U1Space((0, 4), (-1, 2), (-2, 2), (-3, 3))
    V3 = U1Space((0, 4), (1, 2), (2, 2), (3, 3))
    for sp in ((V1, V2), (V1, V3.dual())):
        A = _rand_tensor(rng, sp, charge)
        u, s, v = st.svd_(A)
        R = st.matmul(st.matmul(u, s), v)
        assert st.issimilar(R, A)
        for k in A.blocks:
            np.testing.assert_allclose(R.blocks[k], A.blocks[k], rtol=1e-12, atol=1e-12)
        u2, s2, v2 = st.svdtrunc(A)
        for k in s.blocks:
            np.testing.assert_allclose(s.blocks[k], s2.blocks[k], rtol=1e-12, atol=1e-14)


def _dense_xxz_mpo(lx, delta=1.0):
    """Dense ``xxz_mpo`` (src/mpo.jl:27-51) with Julia index order (wl, po, wr, pi)."""
    sz, sp, sm = ta.dense_spin_half()
    W = np.zeros((5, 2, 5, 2))
    W[0, :, 0, :] = np.eye(2)
    W[1, :, 0, :] = 0.5 * sp
    W[2, :, 0, :] = 0.5 * sm
    W[3, :, 0, :] = delta * sz
    W[4, :, 1, :] = sm
    W[4, :, 2, :] = sp
    W[4, :, 3, :] = sz
    W[4, :, 4, :] = np.eye(2)
    return [W[4:5]] + [W] * (lx - 2) + [W[:, :, 0:1, :]]


def test_xxz_mpo_kat():
    """test/mpotests.jl:3-59: dense MPO == ED H; U(1) MPO blocks == H restricted by count_ones, exactly."""
    smpo = ta.xxz_u1mpo(4, 1.0)
    Ws = smpo.Ws
    h4sym = st.contract(st.contract(st.contract(Ws[0], (1, 2, -1, 5), Ws[1], (-1, 3, 4, 6)),
                                    (1, 2, 3, -1, 6, 7), Ws[2], (-1, 4, 5, 8)),
                        (1, 2, 3, 4, -1, 7, 8, 9), Ws[3], (-1, 5, 6, 10))
    h4sym = st.fuselegs(st.fuselegs(h4sym, 1, 5), 2, 5, True)
    D = _dense_xxz_mpo(4, 1.0)
    h4 = np.einsum("aAxi,xBy j,yCzk,zDEl->aABCDEijkl".replace(" ", ""), D[0], D[1], D[2], D[3])
    h4 = h4.reshape((16, 16), order="F")  # Julia reshape: site 1 is the fastest bit
    H4 = ta.xxz_dense_hamiltonian(4, 1.0)
    assert np.allclose(h4, H4)
    nums = np.arange(16)
    ones = np.array([bin(x).count("1") for x in nums])
    for k in range(5):
        idx = ones == k
        assert np.array_equal(h4[np.ix_(idx, idx)], h4sym.blocks[(k, k)])
    # also: the U(1) MPO chained densely equals the ED Hamiltonian (flip-symmetric basis)
    assert np.allclose(ta.mpo_dense_chain(smpo), ta.xxz_dense_hamiltonian(4, 1.0))


def test_dual_shares_blocks_no_conjugation():
    """dual(A) flips charge/duals, keeps keys and blocks (src/symtensor.jl:201-212)."""
    V = U1Space((0, 2), (1, 1))
    A = st.from_function(lambda T, z: np.full(z, 1 + 2j), 1, (V, V), np.complex128)
    B = dual(A)
    assert B.charge == -1 and all(x.isdual for x in B.space)
    assert list(B.blocks) == list(A.blocks)
    assert all(B.blocks[k] is A.blocks[k] for k in A.blocks)
