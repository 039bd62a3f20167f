"""The oracle's H_eff / environments / DMRG against closed forms, dense einsum and ED."""
import numpy as np
import pytest

from oracle import symtensors as ost
from oracle import tensornetalgs as ota


def test_ed_closed_forms():
    """Two-site singlet -3/4; open 4-site Heisenberg -(3 + 2 sqrt 3)/4."""
    assert abs(ota.u1_ground_energy(2, [(1, 2)], 1) + 0.75) < 1e-14
    e4 = ota.u1_ground_energy(4, [(1, 2), (2, 3), (3, 4)], 2)
    assert abs(e4 + (3 + 2 * np.sqrt(3)) / 4) < 1e-12


def test_ed_block_matches_dense_xxz():
    """U(1) block ED == dense xxz_explicit restricted to the magnetisation block (xxz.jl:13-38 vs :56-113)."""
    L, delta = 6, 0.7
    H = ota.xxz_dense_hamiltonian(L, delta)
    full = np.linalg.eigvalsh(H)
    blk = ota.u1_block_hamiltonian(L, [(i, i + 1) for i in range(1, L)], 3, jz=delta).toarray()
    ev = np.linalg.eigvalsh(blk)
    # every eigenvalue of the M=3 block is an eigenvalue of the full H
    for e in ev:
        assert np.min(np.abs(full - e)) < 1e-10


def _rand_env_case(seed=0, nsec=4):
    rng = np.random.default_rng(seed)
    Vl = ost.Space([(q, int(rng.integers(1, 4))) for q in range(nsec)])
    Vr = ost.Space([(q + 1, int(rng.integers(1, 4))) for q in range(nsec)])
    Vw = ost.Space([(-1, 2), (0, 3), (1, 2)])
    Vd = ost.U1Space((0, 1), (1, 1))
    R = lambda sp: ost.from_function(lambda T, z: rng.standard_normal(z), 0, sp)
    return (R((Vl, Vd, Vd, Vr.dual())), R((Vl.dual(), Vw.dual(), Vl)), R((Vr, Vw, Vr.dual())),
            R((Vw, Vd, Vw.dual(), Vd.dual())), R((Vw, Vd, Vw.dual(), Vd.dual())))


@pytest.mark.parametrize("seed", range(3))
def test_heff_equals_dense_einsum(seed):
    """apply.jl:93-105 (block sparse) == apply.jl:80-91 (dense) on array() embeddings."""
    v, L, R, WL, WR = _rand_env_case(seed)
    got = ost.array(ota.applymps2site(v, L, R, WL, WR))
    want = ota.dense_applymps2site(*[ost.array(t) for t in (v, L, R, WL, WR)])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(3))
def test_env_updates_equal_dense_einsum(seed):
    v, L, R, WL, WR = _rand_env_case(seed)
    rng = np.random.default_rng(50 + seed)
    Vl, Vd, Vr = v.space[0], v.space[1], v.space[3].dual()
    Al = ost.from_function(lambda T, z: rng.standard_normal(z), 0, (Vl, Vd, Vl.dual()))
    Ar = ost.from_function(lambda T, z: rng.standard_normal(z), 0, (Vr, Vd, Vr.dual()))
    np.testing.assert_allclose(ost.array(ota.mpsupdateleft(L, Al, WL)),
                               ota.dense_mpsupdateleft(ost.array(L), ost.array(Al), ost.array(WL)), atol=1e-12)
    np.testing.assert_allclose(ost.array(ota.mpsupdateright(R, Ar, WR)),
                               ota.dense_mpsupdateright(ost.array(R), ost.array(Ar), ost.array(WR)), atol=1e-12)


def _oracle_mps(L, N, chi, seed=1911):
    from mps_amd.workloads import binom_capped_profile, gaussian
    bonds = [[(0, 1)]] + [binom_capped_profile(l, L, N, chi) for l in range(1, L)] + [[(N, 1)]]
    Vd = ost.U1Space((0, 1), (1, 1))
    flats = []
    for l in range(1, L + 1):
        sp = (ost.Space(bonds[l - 1]), Vd, ost.Space(bonds[l]).dual())
        n = sum(int(np.prod(z)) for z in ost.allsectorsandsizes(0, sp)[1])
        flats.append(gaussian(seed, 1000 + l, n))
    mps = ota.random_u1mps_oracle(L, N, bonds, flats)
    ota.center_at(mps, 1)
    ota.normalize_mps(mps)
    return mps


def test_oracle_dmrg_reaches_ed_energy():
    """Two-site sweeps (dmrg.jl:58-179) converge to the U(1) ED ground energy; <H> = measure()."""
    L, N, delta = 10, 5, 1.0
    mps = _oracle_mps(L, N, 8)
    mpo = ota.xxz_u1mpo(L, delta)
    env = ota.initialenv(mps, mpo)
    e_meas0 = ota.measure(mps, mpo)
    # H_eff expectation at the centre equals the full measure (canonical form)
    AA = ost.contract(mps.As[0], (1, 2, -1), mps.As[1], (-1, 3, 4))
    e_heff = ost.dot(AA, ota.applymps2site(AA, env[0], env[3], mpo.Ws[0], mpo.Ws[1]))
    assert abs(e_heff - e_meas0) < 1e-10 * max(1, abs(e_meas0))
    for _ in range(3):
        es = ota.dmrg2sitesweep(mps, mpo, env, maxdim=32, tol=1e-12, lanczostol=1e-10, krylovdim=8,
                                krylovmaxiter=8)
    e_ed = ota.u1_ground_energy(L, [(i, i + 1) for i in range(1, L)], N, jz=delta)
    assert abs(es[-1] - e_ed) < 1e-9 * abs(e_ed)
    assert abs(ota.measure(mps, mpo) - e_ed) < 1e-9 * abs(e_ed)
