"""ctypes binding of libddgx.so (C ABI declared in include/ddgx.h).

The library is built in-tree by ``python -m ddg_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libddgx.so"
_lib = None

DDGX_F64 = 0
DDGX_F32 = 1
REC_WORDS = 16

EXPORTS = (
    "ddgx_last_error", "ddgx_table_size", "ddgx_op_create", "ddgx_op_destroy",
    "ddgx_op_info", "ddgx_sipg_apply", "ddgx_sipg_rhs", "ddgx_block_jacobi_create",
    "ddgx_block_jacobi_blocks", "ddgx_pc_apply", "ddgx_pc_destroy", "ddgx_pcg_solve",
    "ddgx_dot", "ddgx_mg_create", "ddgx_ir_solve", "ddgx_sipg_param_grad",
    "ddgx_solver_create", "ddgx_solver_bytes", "ddgx_solver_pcg", "ddgx_solver_ir",
    "ddgx_solver_destroy", "ddgx_pcg_solve_noalloc", "ddgx_op_create_curved",
    "ddgx_batch_create", "ddgx_batch_bytes", "ddgx_batch_pcg", "ddgx_batch_destroy",
)


class CGConfig(C.Structure):
    _fields_ = [("tol", C.c_double), ("atol", C.c_double),
                ("maxiter", C.c_int), ("check_every", C.c_int)]


class CGInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int),
                ("residual2", C.c_double), ("atol2", C.c_double)]


class MGConfig(C.Structure):
    _fields_ = [("nu_smooth", C.c_int), ("omega", C.c_double), ("alpha", C.c_double),
                ("regularize", C.c_int), ("eps", C.c_double)]


class IRConfig(C.Structure):
    _fields_ = [("tol", C.c_double), ("atol", C.c_double), ("max_outer", C.c_int),
                ("inner_tol", C.c_double), ("inner_maxiter", C.c_int), ("check_every", C.c_int)]


class IRInfo(C.Structure):
    _fields_ = [("outer", C.c_int), ("inner_total", C.c_int), ("converged", C.c_int),
                ("residual2", C.c_double), ("atol2", C.c_double), ("at_floor", C.c_int)]


class DdgxError(RuntimeError):
    pass


class DdgxBusyError(DdgxError):
    """DDGX_ERR_BUSY: the call's shared buffers (a handle's default solve
    state, a solver or batch context) are held by a concurrent call."""


def lib():
    """Load libddgx.so once and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise DdgxError(f"{_LIB_PATH} is missing: build it with `python -m ddg_amd.build` "
                        "(the HIP path has no CPU fallback)")
    L = C.CDLL(str(_LIB_PATH))
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "ddgx_last_error": (C.c_char_p, []),
        "ddgx_table_size": (i32, [i32]),
        "ddgx_op_create": (i32, [i32, i64, i32, vp, vp, vp, vp, vp, C.POINTER(vp)]),
        "ddgx_op_create_curved": (i32, [i32, i64, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp,
                                         C.POINTER(vp)]),
        "ddgx_op_destroy": (i32, [vp]),
        "ddgx_op_info": (i32, [vp, C.POINTER(i32), C.POINTER(i64), C.POINTER(i32)]),
        "ddgx_sipg_apply": (i32, [vp, vp, vp, vp, dbl, dbl, vp]),
        "ddgx_sipg_rhs": (i32, [vp, vp, vp, vp, vp, vp]),
        "ddgx_block_jacobi_create": (i32, [vp, dbl, i32, vp, C.POINTER(vp)]),
        "ddgx_block_jacobi_blocks": (i32, [vp, vp, vp]),
        "ddgx_pc_apply": (i32, [vp, vp, vp, vp]),
        "ddgx_pc_destroy": (i32, [vp]),
        "ddgx_pcg_solve": (i32, [vp, dbl, dbl, vp, vp, vp, C.POINTER(CGConfig),
                                 C.POINTER(CGInfo), vp]),
        "ddgx_pcg_solve_noalloc": (i32, [vp, dbl, dbl, vp, vp, vp, C.POINTER(CGConfig),
                                         C.POINTER(CGInfo), vp]),
        "ddgx_dot": (i32, [vp, vp, vp, C.POINTER(dbl), vp]),
        "ddgx_sipg_param_grad": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "ddgx_ir_solve": (i32, [vp, vp, dbl, dbl, vp, vp, vp, C.POINTER(IRConfig),
                                C.POINTER(IRInfo), vp]),
        "ddgx_mg_create": (i32, [i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_int32),
                                 C.POINTER(C.POINTER(dbl)), C.POINTER(MGConfig), vp,
                                 C.POINTER(vp)]),
        "ddgx_solver_create": (i32, [vp, vp, vp, C.POINTER(vp)]),
        "ddgx_solver_bytes": (i32, [vp, C.POINTER(i64)]),
        "ddgx_solver_pcg": (i32, [vp, dbl, dbl, vp, vp, C.POINTER(CGConfig),
                                  C.POINTER(CGInfo), vp]),
        "ddgx_solver_ir": (i32, [vp, dbl, dbl, vp, vp, C.POINTER(IRConfig),
                                 C.POINTER(IRInfo), vp]),
        "ddgx_solver_destroy": (i32, [vp]),
        "ddgx_batch_create": (i32, [vp, vp, i32, vp, C.POINTER(vp)]),
        "ddgx_batch_bytes": (i32, [vp, C.POINTER(i64)]),
        "ddgx_batch_pcg": (i32, [vp, C.POINTER(dbl), vp, vp, C.POINTER(CGConfig),
                                 C.POINTER(CGInfo), vp]),
        "ddgx_batch_destroy": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().ddgx_last_error().decode(errors="replace")
        kind = {1: ValueError, 2: NotImplementedError, 5: DdgxBusyError}.get(rc, DdgxError)
        raise kind(f"{what}: {msg}" if what else msg)
