"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle still reproduces them (a regression pin on the oracle itself), and
the fixtures are internally consistent with the mathematics they encode."""
import numpy as np

from oracle import symtensors as ost
from oracle import tensornetalgs as ota
from tests.helpers import load_fixture


def _rel(a, b):
    fa = np.concatenate([np.ravel(x) for x in a.blocks.values()])
    fb = np.concatenate([np.ravel(x) for x in b.blocks.values()])
    assert list(a.blocks) == list(b.blocks)
    return np.linalg.norm(fa - fb) / max(np.linalg.norm(fb), 1e-300)


def test_golden_heff():
    T, _, _ = load_fixture("heff_xxz_L32_chi64.npz")
    got = ota.applymps2site(T["v"], T["envL"], T["envR"], T["WL"], T["WR"])
    assert _rel(got, T["out"]) < 1e-14
    # the dense @tensor formula (src/apply.jl:86-89) on the array() embeddings
    dense = ota.dense_applymps2site(*(ost.array(T[k]) for k in ("v", "envL", "envR", "WL", "WR")))
    assert np.linalg.norm(dense - ost.array(T["out"])) < 1e-12 * np.linalg.norm(dense)


def test_golden_env_updates():
    T, _, _ = load_fixture("envupdate_xxz_L32_chi64.npz")
    assert _rel(ota.mpsupdateleft(T["Lenv"], T["A"], T["W"]), T["left"]) < 1e-14
    assert _rel(ota.mpsupdateright(T["Renv"], T["A"], T["W"]), T["right"]) < 1e-14
    dl = ota.dense_mpsupdateleft(ost.array(T["Lenv"]), ost.array(T["A"]), ost.array(T["W"]))
    assert np.linalg.norm(dl - ost.array(T["left"])) < 1e-12 * np.linalg.norm(dl)


def test_golden_sweep_trace():
    T, extra, meta = load_fixture("sweep_xxz_L8.npz")
    L = 8
    mps = ota.MPState([T[f"A{l}"] for l in range(1, L + 1)], 1)
    mpo = ota.xxz_u1mpo(L, 1.0)
    env = ota.initialenv(mps, mpo)
    es = ota.dmrg2sitesweep(mps, mpo, env, **meta["settings"])
    np.testing.assert_allclose(es, extra["energies"], rtol=1e-10, atol=1e-10)
    # variational: every energy lies above the exact ground energy
    assert np.all(extra["energies"] >= float(extra["e_ed"]) - 1e-9)


def test_golden_sweep_trace_xxz_L32_chi64():
    """XXZ L=32, Delta=1, chi=64 (13 sectors per bond), krylovdim 8: the fixture's start
    is right-canonical at site 1 and normalised, the oracle sweep from it reproduces
    the stored energy list, and the list has the reference's layout (quirk Q6)."""
    T, extra, meta = load_fixture("sweep_xxz_L32_chi64.npz")
    L = 32
    As = [T[f"A{l}"] for l in range(1, L + 1)]
    for A in As[1:]:
        I = ost.contract(A, (1, -1, -2), ost.dual(A), (2, -1, -2))
        for b in I.blocks.values():
            assert np.allclose(b, np.eye(b.shape[0]), atol=1e-12)
    assert abs(ost.norm(As[0]) - 1.0) < 1e-12
    mps, mpo = ota.MPState(list(As), 1), ota.xxz_u1mpo(L, 1.0)
    es = ota.dmrg2sitesweep(mps, mpo, ota.initialenv(mps, mpo), **meta["settings"])
    assert len(es) == len(extra["energies"]) == 2 * L - 3 + (L - 2)
    np.testing.assert_allclose(es, extra["energies"], rtol=1e-10, atol=1e-10)
