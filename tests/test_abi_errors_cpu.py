"""Host-side argument checks of the C ABI (include/mpsh.h), on the CPU.

Every call here fails, or answers, before any device work: malformed
structures, mismatched H_eff / environment-update operands, bad SVD block
lists and missing buffers come back as MPSH_ERR_* codes with a message from
mpsh_last_error, never as a crash.  The reference raises on the same
conditions (src/symtensor.jl:20-26, src/contract.jl:237-238, src/apply.jl:93-105
via contract)."""
import ctypes as C

import pytest

mps = pytest.importorskip("mps_amd")
from mps_amd import _lib  # noqa: E402
from mps_amd.symtensors import U1Space, get_struct  # noqa: E402

OK, INVALID, SPACE_MISMATCH = 0, 1, 2


def _lib_h():
    return _lib.lib()


def _create(charge, isdual, legs):
    """mpsh_struct_create over legs given as [(charge, dim), ...] lists."""
    lb = _lib_h()
    nch = [len(l) for l in legs]
    ch = [q for l in legs for q, _ in l]
    dm = [d for l in legs for _, d in l]
    out = C.c_void_p()
    st = lb.mpsh_struct_create(charge, len(legs), (C.c_uint8 * max(1, len(legs)))(*isdual),
                               (C.c_int32 * max(1, len(legs)))(*nch), (C.c_int64 * max(1, len(ch)))(*ch),
                               (C.c_int64 * max(1, len(dm)))(*dm), C.byref(out))
    return st, out


def test_struct_create_rejects_malformed_legs():
    st, out = _create(0, [0, 1], [[(0, 2), (1, 3)], [(0, 2), (1, 3)]])
    assert st == OK and out.value
    _lib_h().mpsh_struct_destroy(out)
    st, _ = _create(0, [0, 1], [[(0, 2), (1, -3)], [(0, 2), (1, 3)]])  # negative dimension
    assert st == INVALID, _lib_h().mpsh_last_error()
    assert _lib_h().mpsh_last_error()
    st, _ = _create(0, [0, 1], [[(0, 2), (0, 3)], [(0, 2), (1, 3)]])  # repeated charge on one leg
    assert st == INVALID, _lib_h().mpsh_last_error()
    lb = _lib_h()
    out = C.c_void_p()
    assert lb.mpsh_struct_create(0, -1, None, None, None, None, C.byref(out)) == INVALID
    assert lb.mpsh_struct_create(0, 2, None, None, None, None, C.byref(out)) == INVALID


def test_struct_queries_reject_bad_handles_and_indices():
    lb = _lib_h()
    s = get_struct(0, (U1Space((0, 2), (1, 3)), U1Space((0, 2), (1, 3)).dual()))
    assert lb.mpsh_struct_find(None, None) == -1
    assert lb.mpsh_struct_find(s.handle, None) == -1
    assert lb.mpsh_struct_find(s.handle, (C.c_int64 * 2)(1, 1)) >= 0
    assert lb.mpsh_struct_find(s.handle, (C.c_int64 * 2)(0, 1)) == -1  # charge 0 - 1 != 0: no such block
    assert lb.mpsh_struct_leg(s.handle, 2, None, None, None) == INVALID
    assert b"leg" in lb.mpsh_last_error()
    assert lb.mpsh_struct_leg(s.handle, -1, None, None, None) == INVALID


def _heff_structs(left, right, mpo):
    Vl, Vr, Vw = U1Space(*left), U1Space(*right), U1Space(*mpo)
    Vd = U1Space((0, 1), (1, 1))
    v = get_struct(0, (Vl, Vd, Vd, Vr.dual()))
    L = get_struct(0, (Vl.dual(), Vw.dual(), Vl))
    R = get_struct(0, (Vr, Vw, Vr.dual()))
    W = get_struct(0, (Vw, Vd, Vw.dual(), Vd.dual()))
    return v, L, R, W


def test_heff_plan_rejects_mismatched_operands():
    lb = _lib_h()
    v, L, R, W = _heff_structs(((0, 3), (1, 4), (2, 3)), ((1, 3), (2, 4), (3, 3)), ((-1, 1), (0, 3), (1, 1)))
    plan, out = C.c_void_p(), C.c_void_p()
    assert lb.mpsh_heff_plan(v.handle, L.handle, R.handle, W.handle, W.handle, C.byref(plan), C.byref(out)) == OK
    # a missing workspace is refused before any device call
    assert lb.mpsh_heff_apply(plan, 0, None, None, None, None, None, None, None, None) == INVALID
    lb.mpsh_heff_destroy(plan)
    lb.mpsh_struct_destroy(out)
    # envR in place of envL: ranks match, legs do not
    st = lb.mpsh_heff_plan(v.handle, R.handle, R.handle, W.handle, W.handle, C.byref(plan), C.byref(out))
    assert st in (INVALID, SPACE_MISMATCH), st
    assert lb.mpsh_last_error()
    # a right environment of another bond (other dims on the shared leg)
    _, _, R2, _ = _heff_structs(((0, 3), (1, 4), (2, 3)), ((1, 3), (2, 5), (3, 3)), ((-1, 1), (0, 3), (1, 1)))
    st = lb.mpsh_heff_plan(v.handle, L.handle, R2.handle, W.handle, W.handle, C.byref(plan), C.byref(out))
    assert st in (INVALID, SPACE_MISMATCH), st
    # wrong ranks (the MPO tensor as the two-site tensor's environment)
    st = lb.mpsh_heff_plan(v.handle, W.handle, R.handle, W.handle, W.handle, C.byref(plan), C.byref(out))
    assert st == INVALID, st
    assert b"rank" in lb.mpsh_last_error()
    assert lb.mpsh_heff_plan(None, L.handle, R.handle, W.handle, W.handle, C.byref(plan), C.byref(out)) == INVALID


def test_envupdate_plan_rejects_mismatched_operands():
    lb = _lib_h()
    Vl, Vr, Vw = U1Space((0, 3), (1, 4)), U1Space((1, 3), (2, 4)), U1Space((-1, 1), (0, 3), (1, 1))
    Vd = U1Space((0, 1), (1, 1))
    A = get_struct(0, (Vl, Vd, Vr.dual()))
    L = get_struct(0, (Vl.dual(), Vw.dual(), Vl))
    W = get_struct(0, (Vw, Vd, Vw.dual(), Vd.dual()))
    plan, out = C.c_void_p(), C.c_void_p()
    assert lb.mpsh_envupdate_left_plan(L.handle, A.handle, W.handle, C.byref(plan), C.byref(out)) == OK
    lb.mpsh_envupdate_destroy(plan)
    lb.mpsh_struct_destroy(out)
    L2 = get_struct(0, (U1Space((0, 3), (1, 5)).dual(), Vw.dual(), U1Space((0, 3), (1, 5))))
    st = lb.mpsh_envupdate_left_plan(L2.handle, A.handle, W.handle, C.byref(plan), C.byref(out))
    assert st in (INVALID, SPACE_MISMATCH), st
    assert lb.mpsh_last_error()
    st = lb.mpsh_envupdate_left_plan(A.handle, A.handle, W.handle, C.byref(plan), C.byref(out))
    assert st in (INVALID, SPACE_MISMATCH), st


def test_svd_plan_rejects_bad_block_lists():
    lb = _lib_h()
    plan = C.c_void_p()
    i64 = lambda *x: (C.c_int64 * max(1, len(x)))(*x)  # noqa: E731
    assert lb.mpsh_svd_plan(2, i64(3, 4), i64(2, 5), i64(0, 6), C.byref(plan)) == OK
    lb.mpsh_svd_destroy(plan)
    assert lb.mpsh_svd_plan(-1, None, None, None, C.byref(plan)) == INVALID
    assert lb.mpsh_svd_plan(1, i64(-3), i64(2), i64(0), C.byref(plan)) == INVALID
    assert lb.mpsh_last_error()
    assert lb.mpsh_svd_plan(1, None, None, None, C.byref(plan)) == INVALID
