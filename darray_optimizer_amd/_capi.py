"""ctypes bindings of include/darray_hip.h (the C ABI of libdarray_hip.so).

Only plain C types cross this boundary.  The library is loaded from the
package directory; there is no fallback: a missing or stale library raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libdarray_hip.so")

_i, _c, _d, _f, _p = C.c_int, C.c_char, C.c_double, C.c_float, C.c_void_p
_u32, _u64, _sz = C.c_uint32, C.c_uint64, C.c_size_t


class LraProblem(C.Structure):
    """dh_lra_problem of include/darray_hip.h."""
    _fields_ = [("n", _i), ("d", _i), ("l", _i), ("q", _i), ("X", _p), ("ldx", _i), ("y", _p),
                ("gamma", _d), ("Omega", _p), ("ldo", _i), ("Q", _p), ("ldq", _i), ("W", _p),
                ("ldw", _i)]


# name -> (restype, argtypes); mirrors include/darray_hip.h one to one.
SIGNATURES: dict[str, tuple] = {
    "dh_create": (_i, [_i, _p, C.POINTER(_p)]),
    "dh_destroy": (_i, [_p]),
    "dh_synchronize": (_i, [_p]),
    "dh_last_error_string": (C.c_char_p, []),
    "dh_workspace_bytes": (_sz, [_p]),
    "dh_timing_enable": (_i, [_p, _i]),
    "dh_timing_read": (_i, [_p, C.POINTER(_d), C.POINTER(_i)]),
    "dh_malloc": (_i, [_p, _sz, C.POINTER(_p)]),
    "dh_free": (_i, [_p, _p]),
    "dh_copy_2d": (_i, [_p, _i, _i, _i, _sz, _p, _i, _p, _i]),
    "dh_convert_f32_f64": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "dh_convert_f64_f32": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "dh_sqnorms_f64": (_i, [_p, _i, _i, _p, _i, _p]),
    "dh_rbf_tile_f64": (_i, [_p, _i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _p, _i]),
    "dh_rbf_tile_f32": (_i, [_p, _i, _i, _i, _p, _i, _p, _i, _p, _p, _f, _p, _i]),
    "dh_rbf_project_f64": (_i, [_p, _i, _i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _p, _i, _d, _p,
                                _i]),
    "dh_rbf_project_f64_noalloc": (_i, [_p, _i, _i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _p, _i,
                                        _d, _p, _i]),
    "dh_rbf_project_bf16": (_i, [_p, _i, _i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _p, _i, _p,
                                 _i]),
    "dh_rbf_project_f16": (_i, [_p, _i, _i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _p, _i, _p,
                                _i]),
    "dh_rbf_project_fp8": (_i, [_p, _i, _i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _p, _i, _p,
                                _i]),
    "dh_gemm_f64": (_i, [_p, _c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
    "dh_gemm_f32": (_i, [_p, _c, _c, _i, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i]),
    "dh_syrk_ut_f64": (_i, [_p, _i, _i, _d, _p, _i, _d, _p, _i]),
    "dh_syrk_ut_f32": (_i, [_p, _i, _i, _f, _p, _i, _f, _p, _i]),
    "dh_trsm_f64": (_i, [_p, _c, _c, _c, _c, _i, _i, _d, _p, _i, _p, _i]),
    "dh_trsm_f32": (_i, [_p, _c, _c, _c, _c, _i, _i, _f, _p, _i, _p, _i]),
    "dh_potrf_f64": (_i, [_p, _c, _i, _p, _i, _i]),
    "dh_potrf_f32": (_i, [_p, _c, _i, _p, _i, _i]),
    "dh_posv_f64": (_i, [_p, _c, _i, _i, _p, _i, _p, _i, _i]),
    "dh_posv_mixed": (_i, [_p, _c, _i, _i, _p, _i, _p, _i, _p, _i, _i, _d, _i, C.POINTER(_i),
                           C.POINTER(_d)]),
    "dh_cholqr2_f64": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "dh_lra": (_i, [_p, _i, _i, _i, _i, _i, _p, _i, _p, _d, _p, _i, _p, _i, _p, _i]),
    "dh_lra_batched": (_i, [_p, _i, _i, C.POINTER(LraProblem), C.POINTER(_i)]),
    "dh_lra_error": (_i, [_p, _i, _i, _i, _p, _i, _p, _d, _p, _i, _p, _i, C.POINTER(_d)]),
    "dh_gesvj_f64": (_i, [_p, _i, _i, _p, _i, _p, _p, _i, C.POINTER(_i)]),
    "dh_nystrom_f64": (_i, [_p, _i, _i, _i, _p, _i, _p, _i, _p, _i, _p]),
    "dh_svm_pcg": (_i, [_p, _i, _i, _p, _i, _p, _p, _p, _d, _i, C.POINTER(_i), C.POINTER(_d)]),
    "dh_svm_barrier": (_i, [_p, _i, _i, _p, _i, _p, _d, _d, _i, _i, _i, _p, C.POINTER(_i),
                            C.POINTER(_d)]),
    "dh_read_libsvm": (_i, [C.c_char_p, _i, _i, C.POINTER(_d), _i, C.POINTER(_d)]),
    "dh_write_libsvm_model": (_i, [C.c_char_p, _i, _i, _i, C.POINTER(_d), _d, C.POINTER(_d), _i,
                                   C.POINTER(_d), C.POINTER(_d), _i, _d, C.POINTER(_d)]),
    "dh_mt19937_raw": (_i, [_u32, _i, C.POINTER(_u32)]),
    "dh_fill_uniform": (_i, [_u32, _d, _d, _i, _i, C.POINTER(_d), _i]),
    "dh_fill_normal": (_i, [_u32, _d, _d, _i, _i, C.POINTER(_d), _i]),
    "dh_fill_mixture": (_i, [_u64, _i, _i, _i, _d, C.POINTER(_d), _i, C.POINTER(_d),
                             C.POINTER(_d)]),
}

DH_PREC_F64 = 0
DH_PREC_BF16 = 1
DH_PREC_F16 = 2
DH_PREC_FP8 = 3
DH_ERR_HIP = -1000
DH_ERR_ALLOC = -1001
DH_ERR_UNSUPPORTED = -1002
DH_ERR_NOT_CONVERGED = -1003

_lib = None


def lib() -> C.CDLL:
    """Load libdarray_hip.so and declare every entry point."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m darray_optimizer_amd.build`")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


class DHError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} returned {status}: {msg}")
        self.status = status


def call(name: str, *args) -> int:
    """Call an entry point; raise DHError on a negative status, return info >= 0."""
    st = getattr(lib(), name)(*args)
    if st < 0:
        raise DHError(name, st, (lib().dh_last_error_string() or b"").decode())
    return st


def last_error() -> str:
    return (lib().dh_last_error_string() or b"").decode()
