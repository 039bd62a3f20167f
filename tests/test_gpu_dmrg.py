"""Device DMRG path: canonical MPS, measure, H_eff hermiticity, Lanczos, sweeps vs ED and vs the oracle sweep."""
import numpy as np
import pytest
import torch

from oracle import symtensors as ost
from oracle import tensornetalgs as ota
from tests.helpers import o_from_product

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import mps_amd as M  # noqa: E402
from mps_amd import mpstools as mt  # noqa: E402
from mps_amd import tensornetalgs as ta  # noqa: E402

DEV = torch.device("cuda", 0)


def _chain_bonds(L):
    return [(i, i + 1) for i in range(1, L)]


def test_random_mps_is_right_canonical_and_normalised():
    mps = mt.random_mps(10, 5, 12, seed=3, device=DEV)
    for A in mps.As[1:]:
        # sum_{p,r} A[l,p,r] A[l',p,r] = delta_{l,l'}
        I = M.contract(A, (1, -1, -2), M.dual(A), (2, -1, -2))
        for (c1, c2), b in I.blocks().items():
            assert np.allclose(b, np.eye(b.shape[0]), atol=1e-12)
    assert abs(M.norm(mps.As[0]) - 1.0) < 1e-12


def test_measure_matches_oracle_and_heff_expectation():
    L = 10
    mps = mt.random_mps(L, 5, 12, seed=4, device=DEV)
    mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
    e = mt.measure(mps, mpo)
    e_or = ota.measure(ota.MPState([o_from_product(A) for A in mps.As], 1),
                       ota.MPOperator([o_from_product(W) for W in mpo.Ws]))
    assert abs(e - e_or) < 1e-12 * max(1.0, abs(e_or))
    env = ta.initialenv(mps, mpo)
    AA = M.contract(mps.As[0], (1, 2, -1), mps.As[1], (-1, 3, 4))
    e_heff = M.dot(AA, M._applymps2site(AA, env[0], env[3], mpo.Ws[0], mpo.Ws[1]))
    assert abs(e_heff - e) < 1e-11 * max(1.0, abs(e))


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_heff_is_hermitian_on_real_environments(path):
    L = 12
    mps = mt.random_mps(L, 6, 16, seed=5, device=DEV)
    mpo = mt.xxz_u1mpo(L, 0.5, device=DEV)
    env = ta.initialenv(mps, mpo)
    AA = M.contract(mps.As[0], (1, 2, -1), mps.As[1], (-1, 3, 4))
    x = M.randn(AA.space, AA.charge, seed=9, stream=1, device=DEV)
    y = M.randn(AA.space, AA.charge, seed=9, stream=2, device=DEV)
    H = lambda t: M._applymps2site(t, env[0], env[3], mpo.Ws[0], mpo.Ws[1], path=path)
    a, b = M.dot(x, H(y)), M.dot(H(x), y)
    assert abs(a - b) < 1e-12 * max(abs(a), 1.0)


def test_lanczos_lowest_matches_dense_eigh():
    """Device Lanczos vs the lowest eigenvalue of the explicit H_eff matrix (built column by column)."""
    L = 8
    mps = mt.random_mps(L, 4, 6, seed=6, device=DEV)
    mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
    env = ta.initialenv(mps, mpo)
    AA = M.contract(mps.As[0], (1, 2, -1), mps.As[1], (-1, 3, 4))
    Hop = ta.HeffOperator(env[0], env[3], mpo.Ws[0], mpo.Ws[1])
    n = AA.struct.numel
    cols = []
    for i in range(n):
        e = M.zeros(AA.space, AA.charge, device=DEV)
        e.data[i] = 1.0
        cols.append(Hop(e).data.cpu().numpy())
    Hm = np.stack(cols, axis=1)
    assert np.allclose(Hm, Hm.T, atol=1e-12)
    lam = np.linalg.eigvalsh(0.5 * (Hm + Hm.T))[0]
    theta, x, info = ta.eigsolve(Hop, AA, tol=1e-12, krylovdim=n if n < 30 else 30, maxiter=20)
    assert abs(theta - lam) < 1e-10 * max(1, abs(lam))


def test_dmrg_xxz_reaches_ed():
    L, N, delta = 12, 6, 1.0
    mps = mt.random_mps(L, N, 16, seed=7, device=DEV)
    mpo = mt.xxz_u1mpo(L, delta, device=DEV)
    env = ta.initialenv(mps, mpo)
    e_ed = ota.u1_ground_energy(L, _chain_bonds(L), N, jz=delta)
    for sweep in range(4):
        es = ta.dmrg2sitesweep(mps, mpo, env, maxdim=64, tol=1e-12, lanczostol=1e-10, krylovdim=8,
                               krylovmaxiter=8)
        print(f"sweep {sweep}: E = {es[-1]:.12f} (ED {e_ed:.12f})", flush=True)
    assert abs(es[-1] - e_ed) < 1e-9 * abs(e_ed)
    assert abs(mt.measure(mps, mpo) - e_ed) < 1e-9 * abs(e_ed)


def test_dmrg_sweep_matches_oracle_sweep():
    """One device sweep and one oracle sweep from the same initial state: same energy list."""
    L, N = 10, 5
    mps = mt.random_mps(L, N, 8, seed=8, device=DEV)
    mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
    omps = ota.MPState([o_from_product(A) for A in mps.As], 1)
    ompo = ota.MPOperator([o_from_product(W) for W in mpo.Ws])
    env = ta.initialenv(mps, mpo)
    oenv = ota.initialenv(omps, ompo)
    kw = dict(maxdim=16, tol=1e-10, lanczostol=1e-9, krylovdim=6, krylovmaxiter=6)
    es = ta.dmrg2sitesweep(mps, mpo, env, **kw)
    eo = ota.dmrg2sitesweep(omps, ompo, oenv, **kw)
    assert len(es) == len(eo) == 2 * L - 3 + (L - 2)
    np.testing.assert_allclose(es, eo, rtol=1e-9, atol=1e-10)
    assert [A.space for A in mps.As] == [tuple(M.U1Space(*V.items(), isdual=V.isdual) for V in A.space)
                                         for A in omps.As]


def test_dmrg_mixed_precision_refined_to_fp64():
    """FP32 / BF16 H_eff matvecs with FP64 refinement cycles reach the FP64 energy to 1e-9."""
    L, N = 12, 6
    e_ed = ota.u1_ground_energy(L, _chain_bonds(L), N)
    for prec in ("f32", "bf16"):
        mps = mt.random_mps(L, N, 16, seed=7, device=DEV)
        mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
        env = ta.initialenv(mps, mpo)
        for _ in range(4):
            es = ta.dmrg2sitesweep(mps, mpo, env, maxdim=64, tol=1e-12, lanczostol=1e-6, krylovdim=8,
                                   krylovmaxiter=8, prec=prec, refine_tol=1e-10)
        print(f"{prec}: E = {es[-1]:.12f} (ED {e_ed:.12f})", flush=True)
        assert abs(es[-1] - e_ed) < 1e-9 * abs(e_ed)


def test_dmrg_cylinder_reaches_ed():
    W, Lx = 3, 4
    L, N = W * Lx, W * Lx // 2
    mps = mt.random_mps(L, N, 16, seed=11, device=DEV)
    mpo = mt.heisenberg_cylinder_mpo(W, Lx, device=DEV)
    env = ta.initialenv(mps, mpo)
    e_ed = ota.u1_ground_energy(L, mt.heisenberg_cylinder_bonds(W, Lx), N)
    for _ in range(5):
        es = ta.dmrg2sitesweep(mps, mpo, env, maxdim=64, tol=1e-12, lanczostol=1e-10, krylovdim=8,
                               krylovmaxiter=10)
    print(f"cylinder {W}x{Lx}: E = {es[-1]:.12f} (ED {e_ed:.12f})", flush=True)
    assert abs(es[-1] - e_ed) < 1e-8 * abs(e_ed)


def test_dmrg_window_equals_sweep_left_to_right():
    """dmrg2site_window over bonds 1..L-2 from center 1 = the L->R half of dmrg2sitesweep."""
    L, N = 10, 5
    mps = mt.random_mps(L, N, 8, seed=12, device=DEV)
    mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
    env = ta.initialenv(mps, mpo)
    mps2 = mt.MPState(list(mps.As), mps.center)
    env2 = list(env)
    kw = dict(maxdim=16, tol=1e-10, lanczostol=1e-9, krylovdim=6, krylovmaxiter=6)
    es = ta.dmrg2sitesweep(mps, mpo, env, **kw)
    timing = []
    ew = ta.dmrg2site_window(mps2, mpo, env2, L - 2, timing=timing, **kw)
    np.testing.assert_allclose(ew, es[:L - 2], rtol=1e-12, atol=1e-12)
    assert mps2.center == L - 1 and len(timing) == L - 2
    assert all(t["ms_total"] > 0 for t in timing)


def _tv_sector_energy(L, N, t, V):
    """Exact ground energy of the open t-V chain in the N-particle sector (JW: nearest-
    neighbour hopping without strings), from the sector's sparse matrix."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    states = [s for s in range(1 << L) if bin(s).count("1") == N]
    index = {s: i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for s in states:
        i = index[s]
        diag = 0.0
        for k in range(L - 1):
            b1, b2 = (s >> (L - 1 - k)) & 1, (s >> (L - 2 - k)) & 1
            diag += V * b1 * b2
            if b1 != b2:
                s2 = s ^ (1 << (L - 1 - k)) ^ (1 << (L - 2 - k))
                rows.append(index[s2])
                cols.append(i)
                vals.append(-t)
        rows.append(i)
        cols.append(i)
        vals.append(diag)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(len(states), len(states)))
    return float(sla.eigsh(H, k=1, which="SA", tol=1e-14)[0][0])


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
def test_dmrg_tv_chain_reaches_ed(prec):
    """Spinless t-V chain (t=1, V=2, half filling, Jordan-Wigner U(1) MPO): device DMRG
    with FP64 / FP32 / BF16 matvecs (reduced precisions refined in FP64) reaches the
    exact sector energy to 1e-9 relative."""
    L, N = 12, 6
    e_ed = _tv_sector_energy(L, N, 1.0, 2.0)
    mps = mt.random_mps(L, N, 16, seed=17, device=DEV)
    mpo = mt.tv_u1mpo(L, 1.0, 2.0, device=DEV)
    env = ta.initialenv(mps, mpo)
    kw = dict(maxdim=64, tol=1e-12, krylovdim=8, krylovmaxiter=8)
    if prec != "f64":
        kw.update(prec=prec, refine_tol=1e-10, lanczostol=1e-6)
    else:
        kw.update(lanczostol=1e-10)
    for _ in range(5):
        es = ta.dmrg2sitesweep(mps, mpo, env, **kw)
    print(f"t-V {prec}: E = {es[-1]:.12f} (ED {e_ed:.12f})", flush=True)
    assert abs(es[-1] - e_ed) < 1e-9 * abs(e_ed)


def test_dmrg_sweep_with_preconditioned_svd():
    """XXZ chain L=20, N=10 from a seeded random MPS at chi=256: the two-site
    blocks reach 100+ columns, so the truncations run the Gram-preconditioned
    Jacobi SVD.  One sweep with it and one with plain Jacobi sweeps
    (MPSH_SVD_PRECOND=0, fresh SVD plans) from the same start give the same
    energies to 1e-10 relative (maxdim above every bond dimension, so only the
    absolute tol truncates and no cut can fall inside a degenerate multiplet),
    and each sweep's last energy equals <psi|H|psi> of its MPS to 1e-12."""
    import os
    L, N = 20, 10
    runs, widest = [], [0]
    old = os.environ.get("MPSH_SVD_PRECOND")
    svd_blocks = mt._svd_blocks

    def recording(M):  # the widest block (columns of W = min(m, n)) any truncation factorises
        facts = svd_blocks(M)
        widest[0] = max([widest[0]] + [min(U.shape[0], Vh.shape[1]) for _, U, _, Vh in facts])
        return facts
    mt._svd_blocks = recording
    try:
        for flag in ("1", "0"):
            os.environ["MPSH_SVD_PRECOND"] = flag
            mt._SVD_PLANS.clear()
            mps = mt.random_mps(L, N, 256, seed=29, device=DEV)
            mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
            env = ta.initialenv(mps, mpo)
            es = ta.dmrg2sitesweep(mps, mpo, env, maxdim=4096, tol=1e-13, lanczostol=1e-10, krylovdim=8,
                                   krylovmaxiter=4)
            assert abs(es[-1] - mt.measure(mps, mpo)) < 1e-12 * abs(es[-1])
            runs.append(np.asarray(es))
    finally:
        mt._svd_blocks = svd_blocks
        mt._SVD_PLANS.clear()
        if old is None:
            os.environ.pop("MPSH_SVD_PRECOND", None)
        else:
            os.environ["MPSH_SVD_PRECOND"] = old
    assert widest[0] >= 96, widest[0]  # the preconditioned class was exercised
    np.testing.assert_allclose(runs[0], runs[1], rtol=1e-10, atol=0)


def test_sweeps_under_an_environment_budget_equal_unbounded_sweeps():
    """Two sweeps (XXZ L=24, chi=128) with all environments resident and with an
    EnvStore holding about a quarter of them (the rest dropped and recomputed by
    the fused updates when read): the energies are bitwise equal, so recomputed
    environments are the ones dropped; the store stayed near its budget."""
    L, N, chi = 24, 12, 128
    kw = dict(maxdim=chi, tol=1e-10, lanczostol=1e-9, krylovdim=6, krylovmaxiter=4)
    runs, store = [], None
    full = None
    for budgeted in (False, True):
        mps = mt.random_mps(L, N, chi, seed=41, device=DEV)
        mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
        if not budgeted:
            env = ta.initialenv(mps, mpo)
            full = sum(t.data.numel() * t.data.element_size() for t in env)
            largest = max(t.data.numel() * t.data.element_size() for t in env)
        else:
            env = store = ta.initialenv(mps, mpo, budget=full // 4)
        es = []
        for _ in range(2):
            es += ta.dmrg2sitesweep(mps, mpo, env, **kw)
        runs.append(es)
    print(f"env budget {full // 4} of {full} bytes: {store.dropped} dropped, {store.recomputed} recomputed, "
          f"peak {store.peak}", flush=True)
    assert runs[0] == runs[1]
    assert store.dropped > 0 and store.recomputed > 0
    assert store.peak <= full // 4 + (store.protect + 2) * largest
