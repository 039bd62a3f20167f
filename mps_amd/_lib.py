"""ctypes binding of libmpsh.so (declared in include/mpsh.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpsh.so")

MPSH_F64, MPSH_F32, MPSH_BF16 = 0, 1, 2
DOT_PARTIALS = 1024
LZ_PARTS = 256  # MPSH_LZ_PARTS
GATHER_FIELDS = 8  # MPSH_GATHER_FIELDS
GATHER_TILE = 4096  # elements per work group of mpsh_gather_blocks
NOALLOC_SCRATCH_BYTES = 64 << 20  # MPSH_NOALLOC_SCRATCH_BYTES

STATUS_NAMES = {
    0: "OK", 1: "INVALID", 2: "SPACE_MISMATCH", 3: "SECTOR_MISMATCH", 4: "NOT_SIMILAR",
    5: "NONFINITE", 6: "DEVICE", 7: "OOM", 8: "UNSUPPORTED",
}


class MpshError(RuntimeError):
    """A non-OK mpsh_status; ``code`` is the status, the message is mpsh_last_error()."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{STATUS_NAMES.get(code, code)}] {msg}")
        self.code = code


class SpaceMismatch(MpshError, ValueError):
    pass


class SectorMismatch(MpshError, ValueError):
    pass


class Unsupported(MpshError, NotImplementedError):
    pass


_p = C.c_void_p
_i32, _i64, _u8, _dbl, _sz = C.c_int32, C.c_int64, C.c_uint8, C.c_double, C.c_size_t
_pi32, _pi64, _pu8 = C.POINTER(_i32), C.POINTER(_i64), C.POINTER(_u8)

_SIGS = {
    "mpsh_last_error": (C.c_char_p, []),
    "mpsh_version": (C.c_char_p, []),
    "mpsh_struct_create": (C.c_int, [_i64, _i32, _pu8, _pi32, _pi64, _pi64, C.POINTER(_p)]),
    "mpsh_struct_destroy": (None, [_p]),
    "mpsh_struct_rank": (_i32, [_p]),
    "mpsh_struct_charge": (_i64, [_p]),
    "mpsh_struct_nblocks": (_i64, [_p]),
    "mpsh_struct_numel": (_i64, [_p]),
    "mpsh_struct_leg_ncharges": (_i32, [_p, _i32]),
    "mpsh_struct_leg": (C.c_int, [_p, _i32, _pu8, _pi64, _pi64]),
    "mpsh_struct_blocks": (C.c_int, [_p, _pi64, _pi64, _pi64]),
    "mpsh_struct_find": (_i64, [_p, _pi64]),
    "mpsh_permutelegs_plan": (C.c_int, [_p, _pi32, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_fuselegs_plan": (C.c_int, [_p, _i32, _i32, _i32, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_splitleg_plan": (C.c_int, [_p, _i32, _i32, _pu8, _pi32, _pi64, _pi64, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_symmatrix_plan": (C.c_int, [_p, _i32, _pi32, _i32, _pi32, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_releg_execute": (C.c_int, [_p, C.c_int, _p, _p, _p]),
    "mpsh_releg_destroy": (None, [_p]),
    "mpsh_contract_plan": (C.c_int, [_p, _pi32, _p, _pi32, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_contract_workspace": (_sz, [_p, C.c_int]),
    "mpsh_contract_flops": (_i64, [_p]),
    "mpsh_contract_execute": (C.c_int, [_p, C.c_int, _p, _p, _p, _p, _p]),
    "mpsh_contract_destroy": (None, [_p]),
    "mpsh_heff_plan": (C.c_int, [_p, _p, _p, _p, _p, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_heff_workspace": (_sz, [_p, C.c_int]),
    "mpsh_heff_stats": (C.c_int, [_p, _pi64]),
    "mpsh_heff_engine_stats": (C.c_int, [_p, _pi64]),
    "mpsh_heff_apply": (C.c_int, [_p, C.c_int, _p, _p, _p, _p, _p, _p, _p, _p]),
    "mpsh_heff_prepared_size": (_sz, [_p, C.c_int]),
    "mpsh_heff_apply_workspace": (_sz, [_p, C.c_int]),
    "mpsh_heff_prepare": (C.c_int, [_p, C.c_int, _p, _p, _p, _p, _p, _p]),
    "mpsh_heff_apply_prepared": (C.c_int, [_p, C.c_int, _p, _p, _p, _p, _p]),
    "mpsh_heff_apply_prepared_noalloc": (C.c_int, [_p, C.c_int, _p, _p, _p, _p]),
    "mpsh_heff_apply_prepared_stages": (C.c_int, [_p, C.c_int, _i32, _p, _p, _p, _p, _p]),
    "mpsh_heff_tune": (C.c_int, [_p, C.c_int, _p, _p, _p, _p, _i32, _i32, _pi32, _p]),
    "mpsh_heff_destroy": (None, [_p]),
    "mpsh_envupdate_left_plan": (C.c_int, [_p, _p, _p, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_envupdate_right_plan": (C.c_int, [_p, _p, _p, C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_envupdate_workspace": (_sz, [_p, C.c_int]),
    "mpsh_envupdate_stats": (C.c_int, [_p, _pi64]),
    "mpsh_envupdate_execute": (C.c_int, [_p, C.c_int, _p, _p, _p, _p, _p, _p]),
    "mpsh_envupdate_destroy": (None, [_p]),
    "mpsh_heff_adjoint_plan": (C.c_int, [_p, _p, _p, _p, _p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p)]),
    "mpsh_heff_adjoint_workspace": (_sz, [_p, C.c_int]),
    "mpsh_heff_adjoint_stats": (C.c_int, [_p, _pi64]),
    "mpsh_heff_adjoint_execute": (C.c_int, [_p, C.c_int, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "mpsh_heff_adjoint_destroy": (None, [_p]),
    "mpsh_svd_plan": (C.c_int, [_i64, _pi64, _pi64, _pi64, C.POINTER(_p)]),
    "mpsh_svd_sizes": (C.c_int, [_p, _pi64, _pi64, _pi64, _pi64, _pi64]),
    "mpsh_svd_workspace": (C.c_size_t, [_p]),
    "mpsh_svd_execute": (C.c_int, [_p, _p, _p, _p, _p, _p, _p]),
    "mpsh_svd_destroy": (None, [_p]),
    "mpsh_dot": (C.c_int, [C.c_int, _i64, _p, _p, _p, _p, _p]),
    "mpsh_lanczos_dots": (C.c_int, [_i32, _i64, _p, _i64, _p, _p, _p, _p]),
    "mpsh_lanczos_update": (C.c_int, [_i32, _i64, _p, _i64, _p, _dbl, _p, _p]),
    "mpsh_lanczos_normalize": (C.c_int, [_i64, _p, _p, _p, _p, _p]),
    "mpsh_axpby": (C.c_int, [C.c_int, _i64, _dbl, _p, _dbl, _p, _p]),
    "mpsh_scal": (C.c_int, [C.c_int, _i64, _dbl, _p, _p]),
    "mpsh_convert": (C.c_int, [C.c_int, _p, C.c_int, _p, _i64, _p]),
    "mpsh_gather_blocks": (C.c_int, [C.c_int32, _i64, _p, _p, _p, _p, _p]),
}

_lib = None


def lib():
    """The loaded library (raises ImportError if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mps_amd.build` "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(status: int):
    if status != 0:
        msg = lib().mpsh_last_error().decode(errors="replace")
        cls = {2: SpaceMismatch, 3: SectorMismatch, 8: Unsupported}.get(status, MpshError)
        raise cls(status, msg)


def i32arr(xs):
    xs = list(xs)
    return (_i32 * max(1, len(xs)))(*xs)


def i64arr(xs):
    xs = list(xs)
    return (_i64 * max(1, len(xs)))(*xs)


def u8arr(xs):
    xs = list(xs)
    return (_u8 * max(1, len(xs)))(*xs)
