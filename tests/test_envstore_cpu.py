"""EnvStore (environments under a device-memory budget, recomputed on demand)
and the bounded plan caches, on the host: the store is driven through the
access pattern of dmrg2sitesweep (src/dmrg.jl:58-179) with stand-in update
functions, and every read must equal the value an unbounded list holds."""
import numpy as np
import pytest
import torch

from mps_amd.symtensors import PlanLRU, SymTensor, U1Space
from mps_amd.tensornetalgs import EnvStore

V = U1Space((0, 2), (1, 3))
W = U1Space((-1, 1), (0, 2), (1, 1))
LEFT = (V.dual(), W.dual(), V)    # left environments: first leg dual
RIGHT = (V, W, V.dual())          # right environments: first leg not dual


def _t(space, x):
    from mps_amd.symtensors import get_struct
    st = get_struct(0, space)
    return SymTensor(0, space, torch.full((st.numel,), float(x), dtype=torch.float64), st)


class _Site:
    def __init__(self, x):
        self.data = torch.tensor([float(x)], dtype=torch.float64)


# stand-in updates with the dependency structure of the real ones
def _left(prev, A, Wt):
    return _t(LEFT, 0.5 * float(prev.data[0]) + float(A.data[0]) + 0.25 * Wt)


def _right(nxt, A, Wt):
    return _t(RIGHT, 0.75 * float(nxt.data[0]) - float(A.data[0]) + 0.125 * Wt)


class _MPS:
    def __init__(self, L, rng):
        self.As = [_Site(rng.normal()) for _ in range(L)]


class _MPO:
    def __init__(self, L):
        self.Ws = [float(i + 1) for i in range(L)]


@pytest.mark.parametrize("budget_envs", [3, 5, 100])
def test_envstore_follows_an_unbounded_list_through_a_sweep(budget_envs):
    L = 24
    rng = np.random.default_rng(5)
    mps, mpo = _MPS(L, rng), _MPO(L)
    nbytes = _t(LEFT, 0).data.numel() * 8
    store = EnvStore(L + 2, mps, mpo, budget_envs * nbytes, update_left=_left, update_right=_right)
    ref = [None] * (L + 2)
    store.set(0, _t(LEFT, 1.0), "L")
    store.set(L + 1, _t(RIGHT, 1.0), "R")
    ref[0], ref[L + 1] = _t(LEFT, 1.0), _t(RIGHT, 1.0)
    for l in range(L, 0, -1):  # initialenv(at=1)
        store.set(l, _right(store[l + 1], mps.As[l - 1], mpo.Ws[l - 1]), "R")
        ref[l] = _right(ref[l + 1], mps.As[l - 1], mpo.Ws[l - 1])

    def read(i):
        a, b = store[i], ref[i]
        assert float(a.data[0]) == float(b.data[0]), i
        assert store.resident_bytes() <= max(store.budget, (store.protect + 2) * nbytes)

    for l in range(1, L - 1):  # left to right
        read(l - 1)
        read(l + 2)
        mps.As[l - 1] = _Site(rng.normal())
        store.set(l, _left(store[l - 1], mps.As[l - 1], mpo.Ws[l - 1]), "L")
        ref[l] = _left(ref[l - 1], mps.As[l - 1], mpo.Ws[l - 1])
    read(L - 2)
    read(L + 1)
    for l in range(L - 1, 1, -1):  # right to left
        read(l - 1)
        read(l + 2)
        mps.As[l] = _Site(rng.normal())
        store.set(l + 1, _right(store[l + 2], mps.As[l], mpo.Ws[l]), "R")
        ref[l + 1] = _right(ref[l + 2], mps.As[l], mpo.Ws[l])
        read(l - 2)
        read(l + 1)
    if budget_envs < L:
        assert store.dropped > 0 and store.recomputed > 0
    else:
        assert store.dropped == 0 and store.recomputed == 0
    assert store.peak <= max(store.budget, (store.protect + 2) * nbytes)


class _Bonded:
    """Stand-in tensor that carries bond labels: site tensors (left bond, right
    bond), environments the bond they close.  Truncations create new labels, so
    an update from a stale neighbour or site tensor fails its label check, the
    role the real updates' leg checks play."""

    def __init__(self, lab, side=None, left=None, right=None):
        self.lab, self.side, self.l, self.r = lab, side, left, right

        class _D:
            def numel(self):
                return 8

            def element_size(self):
                return 1
        self.data = _D()


def _bl(prev, A, W):
    assert prev.lab == A.l, ("left update from a stale pair", prev.lab, A.l)
    return _Bonded(A.r, "L")


def _br(nxt, A, W):
    assert nxt.lab == A.r, ("right update from a stale pair", nxt.lab, A.r)
    return _Bonded(A.l, "R")


@pytest.mark.parametrize("budget_envs", [6, 10, 1000])
def test_envstore_recomputes_from_current_tensors_over_two_sweeps(budget_envs):
    """The index and site-tensor bookkeeping of dmrg2sitesweep (src/dmrg.jl:58-179),
    with a new bond label at every truncation: every environment the store hands
    out, recomputed or not, closes the bond its reader expects."""
    import itertools
    L = 24
    fresh = itertools.count(1000)

    class _M:
        pass
    mps, mpo = _M(), _M()
    mps.As = [_Bonded(None, left=i, right=i + 1) for i in range(L)]
    mpo.Ws = [0] * L
    env = EnvStore(L + 2, mps, mpo, budget_envs * 8, update_left=_bl, update_right=_br)
    env.set(0, _Bonded(0, "L"), "L")
    env.set(L + 1, _Bonded(L, "R"), "R")
    for l in range(L, 0, -1):
        env.set(l, _br(env[l + 1], mps.As[l - 1], 0), "R")

    def expect(i, lab):
        assert env[i].lab == lab, (i, env[i].lab, lab)

    for _ in range(2):
        A = mps.As[0]
        for l in range(1, L - 1):
            aa = (A.l, mps.As[l].r)
            expect(l - 1, aa[0])
            expect(l + 2, aa[1])
            nb = next(fresh)
            mps.As[l - 1] = _Bonded(None, left=aa[0], right=nb)
            env.set(l, _bl(env[l - 1], mps.As[l - 1], 0), "L")
            A = _Bonded(None, left=nb, right=aa[1])
        aa = (A.l, mps.As[L - 1].r)
        expect(L - 2, aa[0])
        for l in range(L - 1, 1, -1):
            nb = next(fresh)
            expect(l + 2, aa[1])
            mps.As[l] = _Bonded(None, left=nb, right=aa[1])
            env.set(l + 1, _br(env[l + 2], mps.As[l], 0), "R")
            A = _Bonded(None, left=aa[0], right=nb)
            aa = (mps.As[l - 2].l, A.r)
            expect(l - 2, aa[0])
            expect(l + 1, aa[1])
        nb = next(fresh)
        mps.As[1] = _Bonded(None, left=nb, right=aa[1])
        env.set(2, _br(env[3], mps.As[1], 0), "R")
        mps.As[0] = _Bonded(None, left=aa[0], right=nb)
    if budget_envs < L:
        assert env.recomputed > 0
    assert env.peak <= max(env.budget, (env.protect + 2) * 8)


def test_envstore_assignment_reads_the_side_from_the_legs():
    """``env[i] = T`` (no side given): a left environment starts with A_i's right bond
    leg, a right one with A_i's left bond leg."""
    Vl, Vr = U1Space((0, 2), (1, 1)), U1Space((0, 1), (1, 2), (2, 1))

    class _A:
        space = (Vl, U1Space((0, 1), (1, 1)), Vr.dual())

    class _M:
        As = [_A(), _A()]
    store = EnvStore(4, _M(), None, 1 << 20)
    store[1] = _t((Vr.dual(), W.dual(), Vr), 0.0)
    store[2] = _t((Vl, W, Vl.dual()), 0.0)
    assert store.kind[1] == "L" and store.kind[2] == "R"
    with pytest.raises(ValueError):
        store[1] = _t((W, W, W), 0.0)


def test_envstore_rejects_reads_of_unset_entries():
    store = EnvStore(6, _MPS(4, np.random.default_rng(1)), _MPO(4), 1 << 20, update_left=_left,
                     update_right=_right)
    with pytest.raises(KeyError):
        store[2]


def test_plan_lru_is_bounded_and_keeps_recent_entries():
    c = PlanLRU(3)
    for k in range(10):
        c[k] = k
        assert len(c) == min(k + 1, 3)
    assert list(c) == [7, 8, 9] and c.get(0) is None
    c.get(7)
    c[10] = 10
    assert list(c) == [9, 7, 10]
    assert len(c) <= 3


def test_plan_caches_stay_bounded_over_many_structures():
    """The library's relegging plans are LRU-cached: many distinct bond structures
    (as successive truncations produce) leave at most the cache size alive."""
    from mps_amd import symtensors as pst
    pst._RELEG.clear()
    old = pst._RELEG.maxsize
    pst._RELEG.maxsize = 8
    try:
        for d in range(1, 40):
            Vd = U1Space((0, 1), (1, 1))
            Vb = U1Space((0, d), (1, 2))
            space = (Vb, Vd, Vb.dual())
            st = pst.get_struct(0, space)
            A = SymTensor(0, space, torch.zeros(st.numel, dtype=torch.float64), st)
            key = ("perm", st.key, (3, 2, 1))
            pst._RELEG[key] = (pst.PlanHandle(None, "mpsh_releg_destroy"), st)
            assert len(pst._RELEG) <= 8
        assert len(pst._RELEG) == 8
    finally:
        pst._RELEG.maxsize = old
        pst._RELEG.clear()
