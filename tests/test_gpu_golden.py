"""Device path against the committed golden fixtures (tests/golden/*.npz):
_applymps2site (fused FP64 / plain FP64 / fused FP32), both environment
updates, and one full device sweep from the fixture's canonical MPS.
Bars: FP64 1e-12 relative, FP32 1e-5, sweep energies 1e-9."""
import numpy as np
import pytest
import torch

from tests.helpers import load_fixture, p_space

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import mps_amd as M  # noqa: E402
from mps_amd import mpstools as mt  # noqa: E402
from mps_amd import tensornetalgs as ta  # noqa: E402

DEV = torch.device("cuda", 0)


def dev(O, dtype=torch.float64):
    flat = np.concatenate([np.asarray(b, np.float64).reshape(-1, order="F") for b in O.blocks.values()])
    return M.SymTensor.from_flat(O.charge, tuple(p_space(V) for V in O.space), flat, dtype, DEV)


def flat_of(O):
    return np.concatenate([np.asarray(b, np.float64).reshape(-1, order="F") for b in O.blocks.values()])


@pytest.mark.parametrize("path,prec,bar", [("fused", "f64", 1e-12), ("plain", None, 1e-12), ("fused", "f32", 1e-5)])
def test_heff_vs_golden(path, prec, bar):
    T, _, _ = load_fixture("heff_xxz_L32_chi64.npz")
    dt = torch.float32 if prec == "f32" else torch.float64
    ins = [dev(T[k], dt) for k in ("v", "envL", "envR", "WL", "WR")]
    kw = {"path": path} if prec is None else {"path": path, "prec": prec}
    out = M._applymps2site(*ins, **kw)
    want = flat_of(T["out"])
    got = out.data.double().cpu().numpy()
    assert got.shape == want.shape
    assert np.linalg.norm(got - want) <= bar * np.linalg.norm(want)


def test_env_updates_vs_golden():
    T, _, _ = load_fixture("envupdate_xxz_L32_chi64.npz")
    A, W, Le, Re = (dev(T[k]) for k in ("A", "W", "Lenv", "Renv"))
    for got, key in ((M._mpsupdateleft(Le, A, W), "left"), (M._mpsupdateright(Re, A, W), "right")):
        want = flat_of(T[key])
        g = got.data.cpu().numpy()
        assert np.linalg.norm(g - want) <= 1e-12 * np.linalg.norm(want)


def test_sweep_vs_golden():
    T, extra, meta = load_fixture("sweep_xxz_L8.npz")
    L = 8
    mps = mt.MPState([dev(T[f"A{l}"]) for l in range(1, L + 1)], 1)
    mpo = mt.xxz_u1mpo(L, 1.0, device=DEV)
    env = ta.initialenv(mps, mpo)
    es = ta.dmrg2sitesweep(mps, mpo, env, **meta["settings"])
    np.testing.assert_allclose(es, extra["energies"], rtol=1e-9, atol=1e-9)
