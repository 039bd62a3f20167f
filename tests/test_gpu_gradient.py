"""Gradient of the Rayleigh quotient <v|H_eff v>/<v|v> w.r.t. the two-site tensor
and both environments (the differentiation configuration).

Checked against central finite differences (h=1e-6, 20 seeded directions per
variable) and against the same gradient built from the oracle's (reference
restatement) contractions on the host.  Environments are recomputed from the
MPS (initialenv), so H_eff is the Hermitian DMRG operator."""
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


def _bond(L, chi, delta, seed, dtype=torch.float64):
    mps = mt.random_mps(L, L // 2, chi, seed=seed, device=DEV)
    mpo = mt.xxz_u1mpo(L, delta, device=DEV)
    c = L // 2
    mt.center_at(mps, c)
    env = ta.initialenv(mps, mpo, at=c)
    v = M.contract(mps.As[c - 1], (1, 2, -1), mps.As[c], (-1, 3, 4))
    ops = (v, env[c - 1], env[c + 2], mpo.Ws[c - 1], mpo.Ws[c])
    return tuple(t.to(dtype) for t in ops) if dtype != torch.float64 else ops


def _E(v, L, R, WL, WR):
    return M.dot(v, M._applymps2site(v, L, R, WL, WR)) / M.dot(v, v)


def _perturbed(T, d, h):
    out = T.copy()
    M.axpby_(h, d, 1.0, out)
    return out


def _dense_dE_denvL(v, WL, WR, R):
    """sum v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] R[c,k,e] v[b,q,r,e] -> [a,m,b], pairwise."""
    t = np.einsum("apsc,cke->apske", v, R, optimize=True)
    t = np.einsum("apske,nrks->apenr", t, WR, optimize=True)
    t = np.einsum("apenr,mqnp->aermq", t, WL, optimize=True)
    return np.einsum("aermq,bqre->amb", t, v, optimize=True)


def _dense_dE_denvR(L, v, WL, WR):
    """sum L[a,m,b] v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] v[b,q,r,e] -> [c,k,e], pairwise."""
    t = np.einsum("amb,apsc->mbpsc", L, v, optimize=True)
    t = np.einsum("mbpsc,mqnp->bscqn", t, WL, optimize=True)
    t = np.einsum("bscqn,nrks->bcqrk", t, WR, optimize=True)
    return np.einsum("bcqrk,bqre->cke", t, v, optimize=True)


@pytest.mark.parametrize("L,chi", [(12, 16), (64, 256)])
def test_gradient_fd_and_oracle(L, chi):
    v, eL, eR, WL, WR = _bond(L, chi, 0.5, seed=31)
    E, gv, gL, gR = ta.rayleigh_gradient(v, eL, eR, WL, WR)
    assert abs(E - _E(v, eL, eR, WL, WR)) < 1e-12 * abs(E)
    h = 1e-6
    for name, var, grad in (("v", 0, gv), ("envL", 1, gL), ("envR", 2, gR)):
        base = [v, eL, eR, WL, WR]
        for k in range(20):
            d = M.randn(base[var].space, base[var].charge, seed=100 + k, stream=var + 1, device=DEV)
            M.rmul_(d, 1.0 / M.norm(d))
            plus, minus = list(base), list(base)
            plus[var] = _perturbed(base[var], d, h)
            minus[var] = _perturbed(base[var], d, -h)
            fd = (_E(*plus) - _E(*minus)) / (2 * h)
            an = M.dot(grad, d)
            assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (name, k, fd, an)
    # oracle gradient (host contractions in the reference order)
    ov, oL, oR, oWL, oWR = (o_from_product(t) for t in (v, eL, eR, WL, WR))
    nn = ost.dot(ov, ov)
    Hv = ota.applymps2site(ov, oL, oR, oWL, oWR)
    Eo = ost.dot(ov, Hv) / nn
    assert abs(E - Eo) < 1e-12 * abs(Eo)
    want_v = {k: 2 * (Hv.blocks[k] - Eo * ov.blocks[k]) / nn for k in ov.blocks}
    got_v = gv.blocks()
    num = sum(np.sum((got_v[k] - want_v[k]) ** 2) for k in want_v) ** 0.5
    den = sum(np.sum(want_v[k] ** 2) for k in want_v) ** 0.5
    assert num <= 1e-10 * den
    # dE/denvL from the dense @tensor form of <v|H v> differentiated in envL
    dv, dL, dR, dWL, dWR = (ost.array(t) for t in (ov, oL, oR, oWL, oWR))
    XL = _dense_dE_denvL(dv, dWL, dWR, dR) / nn
    YR = _dense_dE_denvR(dL, dv, dWL, dWR) / nn
    assert np.linalg.norm(M.array(gL) - XL) <= 1e-11 * np.linalg.norm(XL)
    assert np.linalg.norm(M.array(gR) - YR) <= 1e-11 * np.linalg.norm(YR)


def test_gradient_fp32_close_to_fp64():
    ops64 = _bond(12, 16, 0.5, seed=32)
    ops32 = tuple(t.to(torch.float32) for t in ops64)
    E64, gv64, gL64, gR64 = ta.rayleigh_gradient(*ops64)
    E32, gv32, gL32, gR32 = ta.rayleigh_gradient(*ops32)
    assert abs(E32 - E64) < 1e-5 * abs(E64)
    for a, b in ((gv32, gv64), (gL32, gL64), (gR32, gR64)):
        x, y = a.data.double().cpu().numpy(), b.data.cpu().numpy()
        assert np.linalg.norm(x - y) < 1e-4 * np.linalg.norm(y)


@pytest.mark.timeout(600)
def test_gradient_fp32_L64_chi256_vs_fd_and_oracle():
    """The FP32 gradient at the differentiation configuration (XXZ L=64, Delta=0.5,
    chi=256): directional derivatives against central finite differences of the FP64
    Rayleigh quotient (h=1e-6, 20 seeded directions per variable) and the gradients
    against the oracle's, each to 1e-4 relative (FP32 rounding of the contractions)."""
    ops64 = _bond(64, 256, 0.5, seed=31)
    ops32 = tuple(t.to(torch.float32) for t in ops64)
    E32, gv, gL, gR = ta.rayleigh_gradient(*ops32)
    assert gv.dtype == gL.dtype == gR.dtype == torch.float32
    E64 = _E(*ops64)
    assert abs(E32 - E64) < 1e-5 * abs(E64)
    h = 1e-6
    for name, var, grad in (("v", 0, gv), ("envL", 1, gL), ("envR", 2, gR)):
        g64 = grad.to(torch.float64)
        for k in range(20):
            d = M.randn(ops64[var].space, ops64[var].charge, seed=200 + k, stream=var + 1, device=DEV)
            M.rmul_(d, 1.0 / M.norm(d))
            plus, minus = list(ops64), list(ops64)
            plus[var] = _perturbed(ops64[var], d, h)
            minus[var] = _perturbed(ops64[var], d, -h)
            fd = (_E(*plus) - _E(*minus)) / (2 * h)
            an = M.dot(g64, d)
            assert abs(fd - an) < 1e-4 * max(1.0, abs(fd)), (name, k, fd, an)
    ov, oL, oR, oWL, oWR = (o_from_product(t) for t in ops64)
    nn = ost.dot(ov, ov)
    Hv = ota.applymps2site(ov, oL, oR, oWL, oWR)
    Eo = ost.dot(ov, Hv) / nn
    want_v = np.concatenate([(2 * (Hv.blocks[k] - Eo * ov.blocks[k]) / nn).reshape(-1, order="F") for k in ov.blocks])
    got_v = gv.data.double().cpu().numpy()
    assert np.linalg.norm(got_v - want_v) <= 1e-4 * np.linalg.norm(want_v)
    dv, dL, dR, dWL, dWR = (ost.array(t) for t in (ov, oL, oR, oWL, oWR))
    XL = _dense_dE_denvL(dv, dWL, dWR, dR) / nn
    YR = _dense_dE_denvR(dL, dv, dWL, dWR) / nn
    assert np.linalg.norm(M.array(gL) - XL) <= 1e-4 * np.linalg.norm(XL)
    assert np.linalg.norm(M.array(gR) - YR) <= 1e-4 * np.linalg.norm(YR)
