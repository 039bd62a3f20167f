"""Device Lanczos: basis kernels (mpsh_lanczos_*), hipGraph-captured iterations.

Graph replays are checked bit for bit against the eager iteration, several
times in a row with eager work and synchronisations in between; the
graph-captured eigensolve must give the eager eigenpair."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import mps_amd as M  # noqa: E402
from mps_amd import _lib as L  # noqa: E402
from mps_amd import mpstools as mt  # noqa: E402
from mps_amd import tensornetalgs as ta  # noqa: E402

DEV = torch.device("cuda", 0)


def _operator(prec="f64", L_=12, chi=16, seed=41):
    mps = mt.random_mps(L_, L_ // 2, chi, seed=seed, device=DEV)
    mpo = mt.xxz_u1mpo(L_, 1.0, device=DEV)
    c = L_ // 2
    mt.center_at(mps, c)
    env = ta.initialenv(mps, mpo, at=c)
    v = M.contract(mps.As[c - 1], (1, 2, -1), mps.As[c], (-1, 3, 4))
    return ta.HeffOperator(env[c - 1], env[c + 2], mpo.Ws[c - 1], mpo.Ws[c], prec), v


def test_basis_kernels_match_torch():
    n, k = 100_003, 6
    g = torch.Generator(device="cpu").manual_seed(3)
    V = torch.randn(k + 1, n, generator=g, dtype=torch.float64).to(DEV)
    w = torch.randn(n, generator=g, dtype=torch.float64).to(DEV)
    c = torch.empty(k, dtype=torch.float64, device=DEV)
    part = torch.empty(k * L.LZ_PARTS, dtype=torch.float64, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    L.check(L.lib().mpsh_lanczos_dots(k, n, V.data_ptr(), n, w.data_ptr(), c.data_ptr(), part.data_ptr(), st))
    want = V[:k] @ w
    assert torch.allclose(c, want, rtol=1e-12, atol=1e-10)
    w2 = w.clone()
    L.check(L.lib().mpsh_lanczos_update(k, n, V.data_ptr(), n, c.data_ptr(), -1.0, w2.data_ptr(), st))
    assert torch.allclose(w2, w - V[:k].T @ c, rtol=1e-12, atol=1e-10)
    beta = torch.empty(1, dtype=torch.float64, device=DEV)
    out = torch.empty_like(w)
    L.check(L.lib().mpsh_lanczos_normalize(n, w.data_ptr(), out.data_ptr(), beta.data_ptr(), part.data_ptr(), st))
    assert abs(float(beta) - float(torch.linalg.norm(w))) < 1e-10 * float(beta)
    assert torch.allclose(out, w / torch.linalg.norm(w), rtol=1e-12, atol=1e-14)


def test_graph_replay_equals_eager_iteration():
    H, v = _operator()
    B = ta._LanczosBasis(v, 4)
    B.V[0, :B.n] = v.data / torch.linalg.norm(v.data)
    H(B.vec(0))  # prepare outside any capture
    B.step(H, 0, graph=False)
    torch.cuda.synchronize()
    want_v1, want_sc = B.V[1].clone(), B.sc.clone()
    v0 = B.V[0].clone()
    for rep in range(3):
        B.V[1].zero_()
        B.sc.zero_()
        B.V[0].copy_(v0)
        torch.cuda.synchronize()
        B.step(H, 0, graph=True)  # captured on the first pass, replayed afterwards
        torch.cuda.synchronize()
        assert torch.equal(B.V[1], want_v1), rep
        assert torch.equal(B.sc, want_sc), rep
        _ = M.dot(v, v)  # eager work between replays


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_graph_eigsolve_equals_eager(prec):
    H, v = _operator(prec)
    e0, x0, i0 = ta.eigsolve(H, v, tol=1e-10, krylovdim=6, maxiter=6, graphs=False,
                             refine_tol=1e-10 if prec != "f64" else None)
    H2, _ = _operator(prec)
    e1, x1, i1 = ta.eigsolve(H2, v, tol=1e-10, krylovdim=6, maxiter=6, graphs=True,
                             refine_tol=1e-10 if prec != "f64" else None)
    assert abs(e0 - e1) < 1e-11 * abs(e0)
    assert abs(abs(M.dot(x0, x1)) - 1.0) < 1e-9
    assert i0.numops == i1.numops
