"""Host-side model builders: U(1) MPOs (no kernel runs; data stays on the CPU)."""
import numpy as np
import pytest
import torch

from oracle import symtensors as ost
from oracle import tensornetalgs as ota
from tests.helpers import o_from_flat, o_space

from mps_amd import mpstools as mt  # noqa: E402

CPU = torch.device("cpu")


def _to_oracle(P):
    return o_from_flat(P.charge, tuple(o_space(V) for V in P.space), P.data.numpy())


@pytest.mark.parametrize("L", [3, 4, 6])
@pytest.mark.parametrize("delta", [1.0, 0.5])
def test_fsm_mpo_equals_reference_xxz_u1mpo(L, delta):
    """The FSM builder reproduces xxz_u1mpo (symmpo.jl:45-97) block for block, exactly."""
    got = mt.xxz_u1mpo(L, delta, device=CPU)
    want = ota.xxz_u1mpo(L, delta)
    for Wg, Ww in zip(got.Ws, want.Ws):
        O = _to_oracle(Wg)
        assert O.space == Ww.space and list(O.blocks) == list(Ww.blocks)
        for k in Ww.blocks:
            assert np.array_equal(O.blocks[k], Ww.blocks[k]), k


def _dense_from_bonds(L, bonds, J=1.0):
    sz, sp, sm = ota.dense_spin_half()
    H = np.zeros((2 ** L, 2 ** L))

    def op(o, i):
        return np.kron(np.kron(np.eye(2 ** (i - 1)), o), np.eye(2 ** (L - i)))
    for i, j in bonds:
        H += J * (op(sz, i) @ op(sz, j) + 0.5 * (op(sp, i) @ op(sm, j) + op(sm, i) @ op(sp, j)))
    return H


@pytest.mark.parametrize("W,Lx", [(3, 2), (2, 3), (4, 2)])
def test_cylinder_mpo_equals_dense_hamiltonian(W, Lx):
    bonds = mt.heisenberg_cylinder_bonds(W, Lx)
    mpo = mt.heisenberg_cylinder_mpo(W, Lx, device=CPU)
    Hm = ota.mpo_dense_chain(ota.MPOperator([_to_oracle(Wt) for Wt in mpo.Ws]))
    Hd = _dense_from_bonds(W * Lx, bonds)
    # MPO basis: charge 0 (down) first; dense basis: up first -> flip every site
    n = W * Lx
    perm = np.array([(~i) & ((1 << n) - 1) for i in range(2 ** n)])
    np.testing.assert_allclose(Hm, Hd[np.ix_(perm, perm)], atol=1e-13)


def test_cylinder_6x24_mpo_dimension():
    """Snake-ordered 6-leg cylinder: interior MPO bonds carry six open sources (D = 20)."""
    mpo = mt.heisenberg_cylinder_mpo(6, 24, device=CPU)
    dims = [W.space[2].dim() for W in mpo.Ws[:-1]]
    assert max(dims) in (20, 23)
    assert len(mt.heisenberg_cylinder_bonds(6, 24)) == 6 * 24 + 6 * 23


def test_truncatedsizes_matches_oracle_greedy():
    """The sorted-prefix form of truncatedsizes equals the oracle's literal greedy loop
    (mpsutils.jl:89-112), ties between sectors included."""
    import numpy as np
    from mps_amd.mpstools import truncatedsizes
    from oracle.symtensors import truncatedsizes as o_trunc
    rng = np.random.default_rng(4)
    for trial in range(200):
        nb = int(rng.integers(1, 8))
        S = []
        for _ in range(nb):
            k = int(rng.integers(1, 12))
            v = np.sort(rng.choice([1.0, 0.5, 0.25, 0.125, 1e-3, 1e-9], size=k) if trial % 2 else rng.random(k))[::-1]
            S.append(v)
        maxdim = int(rng.integers(1, 40))
        tol = [1e-12, 1e-2, 0.3][trial % 3]
        try:
            want = o_trunc([s.copy() for s in S], maxdim=maxdim, tol=tol)
        except ValueError:
            import pytest
            with pytest.raises(ValueError):
                truncatedsizes(S, maxdim=maxdim, tol=tol)
            continue
        assert truncatedsizes(S, maxdim=maxdim, tol=tol) == list(want)


def _dense_tv(L, t, V):
    """Spinless fermions with explicit Jordan-Wigner strings, basis (empty, occupied) per
    site, site 1 slowest: H = -t sum (c+_i c_{i+1} + h.c.) + V sum n_i n_{i+1}."""
    a = np.array([[0.0, 1.0], [0.0, 0.0]])
    n1 = np.diag([0.0, 1.0])
    P = np.diag([1.0, -1.0])

    def chain(ops):
        out = np.eye(1)
        for o in ops:
            out = np.kron(out, o)
        return out
    c = [chain([P] * i + [a] + [np.eye(2)] * (L - i - 1)) for i in range(L)]
    n = [chain([np.eye(2)] * i + [n1] + [np.eye(2)] * (L - i - 1)) for i in range(L)]
    H = np.zeros((2 ** L, 2 ** L))
    for i in range(L - 1):
        H += -t * (c[i].T @ c[i + 1] + c[i + 1].T @ c[i]) + V * n[i] @ n[i + 1]
    return H


@pytest.mark.parametrize("L,t,V", [(4, 1.0, 2.0), (6, 1.0, 2.0), (5, 0.7, -1.3)])
def test_tv_mpo_equals_fermion_hamiltonian(L, t, V):
    """The t-V MPO (Jordan-Wigner, fields and constant on the FSM) equals the fermionic
    Hamiltonian built with explicit JW strings, matrix element by matrix element."""
    mpo = mt.tv_u1mpo(L, t, V, device=CPU)
    Hm = ota.mpo_dense_chain(ota.MPOperator([_to_oracle(Wt) for Wt in mpo.Ws]))
    np.testing.assert_allclose(Hm, _dense_tv(L, t, V), atol=1e-13)


def test_tv_free_fermion_closed_form():
    """V = 0: the N-particle ground energy of the open chain is the sum of the N lowest
    single-particle levels -2t cos(k pi / (L + 1))."""
    L, N, t = 8, 3, 1.0
    H = _dense_tv(L, t, 0.0)
    occ = np.array([bin(i).count("1") for i in range(2 ** L)])
    idx = np.nonzero(occ == N)[0]
    e = np.linalg.eigvalsh(H[np.ix_(idx, idx)])[0]
    levels = np.sort(-2 * t * np.cos(np.arange(1, L + 1) * np.pi / (L + 1)))
    assert abs(e - levels[:N].sum()) < 1e-12


def test_pack_s_offsets_and_host_copy():
    """svdtrunc / svd_ keep the blocks' singular values as views of one packed
    device vector (normalised in place) plus one host copy: offsets, views and
    the host values agree, and an empty factorisation packs to nothing."""
    import torch
    from mps_amd import mpstools as mt
    rng = np.random.default_rng(3)
    sv = [np.sort(rng.random(k))[::-1].copy() for k in (3, 1, 5)]
    facts = [((q, q), None, torch.tensor(s), None) for q, s in enumerate(sv)]
    for dt in (torch.float64, torch.float32):
        S_pack, offs, svals = mt._pack_s(facts, dt)
        assert S_pack.dtype == dt and offs == [0, 3, 4, 9]
        for b, s in enumerate(sv):
            np.testing.assert_array_equal(svals[b], s)
            view = S_pack[offs[b]:offs[b + 1]]
            np.testing.assert_allclose(view.double().numpy(), s, rtol=1e-7)
        S_pack.div_(2.0)  # the per-block views follow an in-place normalisation
        np.testing.assert_allclose(S_pack[offs[2]:offs[3]].double().numpy(), sv[2] / 2, rtol=1e-7)
    assert mt._pack_s([], torch.float64) == (None, [0], [])


def _svd_ws(m, n, env):
    """Workspace bytes of a host-side mpsh_svd_plan for blocks m[i] x n[i] under env."""
    import os
    from mps_amd import _lib as L
    from mps_amd._lib import check, lib
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        offs = [0]
        for a, b in zip(m, n):
            offs.append(offs[-1] + a * b)
        h = L._p()
        check(lib().mpsh_svd_plan(len(m), L.i64arr(m), L.i64arr(n), L.i64arr(offs[:-1]), h))
        ws = int(lib().mpsh_svd_workspace(h))
        lib().mpsh_svd_destroy(h)
        return ws
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def test_svd_plan_reserves_gram_preconditioner_workspace():
    """The Jacobi plan (host side, no device) reserves for every block of at
    least 96 columns a copy of the scaled block and a padded Gram / eigenvector
    slot with its eigenvalue and syevd scratch vectors; the 900 x 800 and
    500 x 400 blocks share one 800 x 800 size bucket (the bucket cost model
    prefers one syevd call), the 60 x 50 block gets nothing, and
    MPSH_SVD_PRECOND=0 drops the whole reservation."""
    m, n = [900, 500, 60], [800, 400, 50]
    pre = _svd_ws(m, n, {"MPSH_SVD_ALGO": "jacobi", "MPSH_SVD_PRECOND": "1"})
    plain = _svd_ws(m, n, {"MPSH_SVD_ALGO": "jacobi", "MPSH_SVD_PRECOND": "0"})
    want = 8 * (900 * 800 + 500 * 400 + 2 * 800 * 800 + 2 * 800 + 2 * 800) + 4 * 2
    assert want <= pre - plain <= want + 128
    small = _svd_ws([60], [50], {"MPSH_SVD_ALGO": "jacobi", "MPSH_SVD_PRECOND": "1"})
    assert small == _svd_ws([60], [50], {"MPSH_SVD_ALGO": "jacobi", "MPSH_SVD_PRECOND": "0"})


def test_gather_table_replayed_on_host_builds_the_truncated_factors():
    """mpsh_gather_blocks' table (records + tile list), replayed here with the
    kernel's formula, rebuilds u (column prefix), v (strided row prefix) and the
    s-scaled products block by block; the tiles cover every element once."""
    from mps_amd import _lib as L
    g = torch.Generator().manual_seed(7)
    buf = torch.randn(40_000, generator=g, dtype=torch.float64)
    sbuf = torch.rand(200, generator=g, dtype=torch.float64)
    # (m x k) column-major U blocks cut to their first kk columns; (k x n) Vh blocks cut to kk rows
    shapes = [(70, 60, 45, 90), (5, 5, 5, 3), (130, 40, 1, 77)]
    off = so = 0
    U, Vh, S = [], [], []
    for m, k, kk, n in shapes:
        U.append(buf[off:off + m * k].view(k, m).t()[:, :kk])
        off += m * k
        Vh.append(buf[off:off + k * n].view(n, k).t()[:kk, :])
        off += k * n
        S.append(sbuf[so:so + kk])
        so += kk
    for mats, scs, smode in ((U, None, 0), (Vh, None, 0), (U, S, 2), (Vh, S, 1)):
        doffs = np.cumsum([0] + [x.numel() for x in mats])
        table, ntiles = mt._gather_table(list(zip(mats, scs or [None] * 3, doffs[:-1])), buf,
                                         None if scs is None else sbuf, smode)
        nb = len(mats)
        rec = table[:nb * L.GATHER_FIELDS].reshape(nb, L.GATHER_FIELDS)
        tiles = table[nb * L.GATHER_FIELDS:]
        assert len(tiles) == ntiles == sum(-(-x.numel() // L.GATHER_TILE) for x in mats)
        src, sc = buf.numpy(), sbuf.numpy()
        dst = np.full(int(doffs[-1]), np.nan)
        hits = np.zeros(int(doffs[-1]), dtype=int)
        for t in tiles:
            s_off, rs, cs, rows, cols, d_off, sc_off, sm = rec[int(t) >> 32]
            e = (int(t) & 0xFFFFFFFF) * L.GATHER_TILE + np.arange(L.GATHER_TILE)
            e = e[e < rows * cols]
            j, i = e // rows, e % rows
            f = 1.0 if sm == 0 else sc[sc_off + (i if sm == 1 else j)]
            dst[d_off + e] = src[s_off + i * rs + j * cs] * f
            hits[d_off + e] += 1
        assert (hits == 1).all()
        for b, x in enumerate(mats):
            want = x if scs is None else (x * scs[b] if smode == 2 else scs[b][:, None] * x)
            got = dst[doffs[b]:doffs[b + 1]].reshape(x.shape, order="F")
            assert np.array_equal(got, want.numpy()), (smode, b)


def test_state_footprint_counts_every_block():
    """state_footprint sums the structures' block sizes: for a small chain it
    equals the element count of explicitly enumerated MPS and environment
    structures, and at chi=2048 on the 6x24 cylinder (the sweep the bench
    runs) the environments take 5.63 GB and the MPS 0.57 GB in FP64."""
    from mps_amd import mpstools as mt
    from mps_amd.symtensors import U1Space, get_struct
    leg = U1Space((-1, 1), (0, 3), (1, 1))
    f = mt.state_footprint(8, 4, 16, leg)
    bs = mt.bond_spaces(8, 4, 16)
    Vd = U1Space((0, 1), (1, 1))
    mps = sum(int(np.prod(sz)) for l in range(1, 9)
              for sz in get_struct(0, (bs[l - 1], Vd, bs[l].dual())).tables()[1])
    triv = U1Space((0, 1))
    env = sum(int(np.prod(sz)) for l in range(9)
              for sz in get_struct(0, (bs[l].dual(), (triv if l in (0, 8) else leg).dual(), bs[l])).tables()[1])
    assert f["environments"] == 8 * env
    assert f["mps"] == 8 * mps
    cyl = mt.state_footprint(144, 72, 2048, U1Space((-1, 6), (0, 8), (1, 6)), nsec=60, dmax=200)
    assert abs(cyl["environments"] / 1e9 - 5.634) < 0.01 and abs(cyl["mps"] / 1e9 - 0.565) < 0.01
