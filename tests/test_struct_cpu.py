"""libmpsh integer bookkeeping against the oracle, bit-exact, on the CPU.

No kernel runs here: only structure creation and plan building, which are
host-side C++ behind the C ABI.
"""
import os
import re

import numpy as np
import pytest

from oracle import symtensors as ost
from oracle import tensornetalgs as ota
from tests.helpers import o_space, p_space, random_space

mps = pytest.importorskip("mps_amd")
from mps_amd import _lib  # noqa: E402
from mps_amd import symtensors as pst  # noqa: E402
from mps_amd.symtensors import U1Space  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mpsh.h")).read()
    declared = sorted(set(re.findall(r"\b(mpsh_[a-z0-9_]+)\s*\(", hdr)))
    assert declared, "no declarations found"
    h = _lib.lib()
    for name in declared:
        assert hasattr(h, name), f"libmpsh.so does not export {name}"
    assert sorted(_lib.exported_symbols()) == declared


def _struct_tables(charge, ospaces):
    st = pst.get_struct(charge, tuple(p_space(V) for V in ospaces))
    s, z, o = st.tables()
    return [tuple(r) for r in s.tolist()], [tuple(r) for r in z.tolist()], o.tolist()


@pytest.mark.parametrize("seed", range(40))
def test_allsectorsandsizes_bitexact(seed):
    rng = np.random.default_rng(seed)
    rank = int(rng.integers(1, 6))
    spaces = tuple(random_space(rng) for _ in range(rank))
    charge = int(rng.integers(-3, 4))
    os_, oz = ost.allsectorsandsizes(charge, spaces)
    ls, lz, lo = _struct_tables(charge, spaces)
    assert ls == os_
    assert lz == oz
    assert lo == list(np.cumsum([0] + [int(np.prod(z)) for z in oz]))


def _oracle_zero(charge, spaces):
    return ost.fill(0.0, spaces, charge)


def _same_structure(P, O):
    """product Struct vs oracle SymTensor: charge, legs, sector list, sizes."""
    assert P.charge == O.charge
    assert tuple(o_space(V) for V in P.spaces) == O.space
    s, z, _ = P.tables()
    assert [tuple(r) for r in s.tolist()] == list(O.blocks.keys())
    assert [tuple(r) for r in z.tolist()] == [tuple(b.shape) for b in O.blocks.values()]


def _plan_contract(OA, idxA, OB, idxB):
    sA = pst.get_struct(OA.charge, tuple(p_space(V) for V in OA.space))
    sB = pst.get_struct(OB.charge, tuple(p_space(V) for V in OB.space))
    return pst.contract_plan(sA, idxA, sB, idxB)


def test_contract_structures_kat_spaces():
    vlist = (ost.U1Space((0, 2), (1, 2), (2, 2)), ost.U1Space((0, 2), (-1, 2), (-2, 2), (-3, 2)),
             ost.U1Space((0, 2), (1, 2), (2, 2)), ost.U1Space((0, 2), (1, 2)).dual(), ost.U1Space((0, 2), (1, 2)),
             ost.U1Space((0, 2), (1, 2), (2, 2), (3, 2), (4, 2)), ost.U1Space((0, 2), (1, 2)),
             ost.U1Space((0, 2), (-1, 2), (-2, 2)))
    A = _oracle_zero(0, vlist[0:4])
    B = _oracle_zero(0, vlist[4:8])
    C = ost.contract(A, (1, 2, 3, -1), B, (-1, 4, 5, 6))
    _, st = _plan_contract(A, (1, 2, 3, -1), B, (-1, 4, 5, 6))
    _same_structure(st, C)


@pytest.mark.parametrize("seed", range(30))
def test_contract_structures_random(seed):
    """Random tensors with one or two contracted legs and a random output order."""
    rng = np.random.default_rng(100 + seed)
    for _ in range(50):
        case = _random_contract_case(rng)
        if case is not None:
            break
    A, idxA, B, idxB = case
    C = ost.contract(A, tuple(idxA), B, tuple(idxB))
    _, st = _plan_contract(A, tuple(idxA), B, tuple(idxB))
    _same_structure(st, C)


def _random_contract_case(rng):
    nA, nB = int(rng.integers(2, 5)), int(rng.integers(2, 5))
    ncon = int(rng.integers(1, min(nA, nB)))
    sA = [random_space(rng, dmax=3) for _ in range(nA)]
    sB = [random_space(rng, dmax=3) for _ in range(nB)]
    conA = sorted(rng.choice(nA, ncon, replace=False).tolist())
    conB = sorted(rng.choice(nB, ncon, replace=False).tolist())
    for i, (a, b) in enumerate(zip(conA, conB)):
        sB[b] = sA[a].dual()
    nout = nA + nB - 2 * ncon
    outs = (rng.permutation(nout) + 1).tolist()
    idxA, idxB = [0] * nA, [0] * nB
    labels = (rng.permutation(ncon) + 1).tolist()
    for i, (a, b) in enumerate(zip(conA, conB)):
        idxA[a] = -labels[i]
        idxB[b] = -labels[i]
    k = 0
    for i in range(nA):
        if idxA[i] == 0:
            idxA[i] = outs[k]
            k += 1
    for i in range(nB):
        if idxB[i] == 0:
            idxB[i] = outs[k]
            k += 1
    cA, cB = int(rng.integers(-1, 2)), int(rng.integers(-1, 2))
    A = _oracle_zero(cA, tuple(sA))
    B = _oracle_zero(cB, tuple(sB))
    if not A.blocks or not B.blocks:
        return None
    return A, idxA, B, idxB


def test_not_contractible_raises_space_mismatch():
    V = ost.U1Space((0, 2), (1, 3))
    W = ost.U1Space((0, 2), (1, 4))
    A = _oracle_zero(0, (V, V.dual()))
    B = _oracle_zero(0, (W, W.dual()))
    with pytest.raises(ValueError):
        ost.contract(A, (1, -1), B, (-1, 2))
    with pytest.raises(_lib.SpaceMismatch):
        _plan_contract(A, (1, -1), B, (-1, 2))


def test_rank1_output_unsupported_like_reference():
    V = ost.U1Space((0, 2), (1, 3))
    A = _oracle_zero(0, (V, V.dual()))
    B = _oracle_zero(0, (V.dual(), V, V))
    with pytest.raises(_lib.Unsupported):
        _plan_contract(A, (-1, -2), B, (-1, -2, 1))


def test_releg_structures():
    vlist = (ost.U1Space((0, 2), (1, 3), (2, 4)), ost.U1Space((-2, 4), (-1, 3), (0, 2), (1, 1)),
             ost.U1Space((-1, 1), (0, 2), (1, 3), (2, 4)), ost.U1Space((-2, 3), (-1, 2), (0, 1)))
    A = ost.fill_linearindex(vlist)
    sA = pst.get_struct(0, tuple(p_space(V) for V in vlist))
    import mps_amd._lib as L
    for (l, n, d) in [(1, 2, False), (3, 2, True), (2, 3, False), (1, 4, True), (2, 1, True)]:
        O = ost.fuselegs(A, l, n, d)
        plan, out = L._p(), L._p()
        L.check(L.lib().mpsh_fuselegs_plan(sA.handle, l, n, 1 if d else 0, plan, out))
        _same_structure(pst._struct_from_handle(out), O)
        L.lib().mpsh_releg_destroy(plan)


def test_heff_plan_structure_and_flops_cfg1():
    """Fused H_eff plan: output structure == the oracle chain's; FLOPs == sum 2MNK of its GEMMs."""
    from mps_amd.workloads import heff_spaces
    left, right, mpo = heff_spaces("xxz_L32_chi64")
    Vl, Vr, Vw = ost.Space(left), ost.Space(right), ost.Space(mpo)
    Vd = ost.U1Space((0, 1), (1, 1))
    sp = dict(v=(Vl, Vd, Vd, Vr.dual()), L=(Vl.dual(), Vw.dual(), Vl), R=(Vr, Vw, Vr.dual()),
              W=(Vw, Vd, Vw.dual(), Vd.dual()))
    O = {k: _oracle_zero(0, s) for k, s in sp.items()}
    T1, T2, T3, out = ota.applymps2site_steps(O["v"], O["L"], O["R"], O["W"], O["W"])
    S = {k: pst.get_struct(0, tuple(p_space(V) for V in s)) for k, s in sp.items()}
    from mps_amd.tensornetalgs import HeffPlan
    hp = HeffPlan(S["v"], S["L"], S["R"], S["W"], S["W"])
    _same_structure(hp.out, out)
    # the reference-order FLOPs: 2*M*N*K over the fused blocks each step multiplies
    def step_flops(A, idxA, B, idxB):
        ra, ca, _ = ost.contract_index_perm(idxA)
        rb, cb, _ = ost.contract_index_perm(idxB)
        FA, FB = ost.symmatrix(A, ra, ca), ost.symmatrix(B, cb, rb)
        f = 0
        for (c1, c2), blk in ost.matmul(FA, FB).blocks.items():
            _c1 = -c1 if FA.space[0].isdual else c1
            sA = (c1, -FA.charge + _c1) if FA.space[1].isdual else (c1, FA.charge - _c1)
            if sA in FA.blocks:
                _c2 = -c2 if FB.space[1].isdual else c2
                sB = (FB.charge - _c2, c2) if FA.space[1].isdual else (-FB.charge + _c2, c2)
                if sB in FB.blocks:
                    f += 2 * blk.shape[0] * blk.shape[1] * FA.blocks[sA].shape[1]
        return f
    want = [step_flops(O["L"], (-1, 2, 1), O["v"], (-1, 3, 4, 5)),
            step_flops(T1, (1, -1, -2, 4, 5), O["W"], (-1, 2, 3, -2)),
            step_flops(T2, (1, 2, -1, -2, 5), O["W"], (-1, 3, 4, -2)),
            step_flops(T3, (1, 2, 3, -1, -2), O["R"], (-2, -1, 4))]
    assert hp.step_flops == want
    assert hp.ref_flops == sum(want)



def test_gather_blocks_rejects_bad_arguments_on_the_host():
    """mpsh_gather_blocks checks its arguments before any device call: a negative
    block count, missing buffers and a table not aligned to 8 bytes are
    MPSH_ERR_INVALID with a message; an empty gather is a no-op."""
    lb = _lib.lib()
    assert lb.mpsh_gather_blocks(-1, 0, None, None, None, None, None) == 1
    assert b"gather" in lb.mpsh_last_error()
    assert lb.mpsh_gather_blocks(1, 1, None, None, None, None, None) == 1
    assert lb.mpsh_gather_blocks(1, 1, 4099, 4096, None, 8192, None) == 1
    assert b"aligned" in lb.mpsh_last_error()
    assert lb.mpsh_gather_blocks(0, 0, None, None, None, None, None) == 0
