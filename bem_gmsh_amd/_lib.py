"""ctypes binding of libbemhip.so (the C ABI declared in include/bemhip.h).

The shared library is built in-tree by ``python -m bem_gmsh_amd.build`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there is
no CPU fallback for any compute entry point.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbemhip.so")

_lib = None

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_int_p = C.POINTER(C.c_int)
c_int64_p = C.POINTER(C.c_int64)

# (name, restype, argtypes) for every symbol of include/bemhip.h
SIGNATURES = [
    ("bemhip_version", C.c_int, []),
    ("bemhip_strerror", C.c_char_p, [C.c_int]),
    ("bemhip_last_error", C.c_char_p, [C.c_void_p]),
    ("bemhip_gauss_legendre", C.c_int, [C.c_int, c_double_p, c_double_p]),
    ("bemhip_device_count", C.c_int, [c_int_p]),
    ("bemhip_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("bemhip_destroy", None, [C.c_void_p]),
    ("bemhip_set_geometry", C.c_int,
     [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p, c_double_p,
      c_double_p, c_double_p, c_double_p, C.c_double]),
    ("bemhip_set_elasticity", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("bemhip_assemble", C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int]),
    ("bemhip_get_HG", C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    ("bemhip_get_HG_f32", C.c_int, [C.c_void_p, c_float_p, c_float_p]),
    ("bemhip_near_pairs", C.c_int, [C.c_void_p, c_int64_p]),
    ("bemhip_get_near_pairs", C.c_int, [C.c_void_p, c_int_p, c_int_p, c_double_p, c_double_p]),
    ("bemhip_apply_bc", C.c_int, [C.c_void_p, c_int64_p, c_double_p, c_double_p, c_double_p]),
    ("bemhip_solve", C.c_int,
     [C.c_void_p, C.c_int, C.c_double, C.c_int, c_double_p, c_int_p, c_double_p]),
    ("bemhip_apply_operator", C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    ("bemhip_set_lu_pivoting", C.c_int, [C.c_void_p, C.c_int]),
    ("bemhip_set_lu_gemm", C.c_int, [C.c_void_p, C.c_int]),
    ("bemhip_set_lu_timers", C.c_int, [C.c_void_p, C.c_int]),
    ("bemhip_lu_factor", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, c_int_p]),
    ("bemhip_lu_solve", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, c_int_p,
                                  c_double_p, C.c_int]),
    ("bemhip_adjoint", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p]),
    ("bemhip_solve_rhs_noalloc", C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    ("bemhip_solve_gmres", C.c_int,
     [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, c_double_p, c_int_p, c_double_p]),
    ("bemhip_interior_laplace", C.c_int,
     [C.c_void_p, C.c_int, c_double_p, c_double_p, c_double_p, C.c_int, C.c_double,
      c_double_p]),
    ("bemhip_counters", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ("bemhip_synchronize", C.c_int, [C.c_void_p]),
]


def lib():
    """Load libbemhip.so once and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first with `python -m bem_gmsh_amd.build` "
            "(there is no CPU fallback for the HIP path)")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(c_double_p)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32
    return a.ctypes.data_as(c_float_p)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(c_int_p)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64
    return a.ctypes.data_as(c_int64_p)


class BemHipError(RuntimeError):
    pass


STATUS = {0: "ok", 1: "bad argument", 2: "HIP runtime error", 3: "invalid state",
          4: "singular matrix", 5: "no convergence", 6: "out of memory", 7: "no device"}


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = lib().bemhip_last_error(ctx)
        text = msg.decode() if msg else ""
        kind = STATUS.get(status, f"status {status}")
        if status == 1:
            raise ValueError(f"bemhip: {kind}: {text}")
        raise BemHipError(f"bemhip: {kind}: {text}")


def gauss_legendre(n: int):
    """Gauss-Legendre nodes and weights on [-1, 1], ascending (host-only routine)."""
    x = np.empty(n, dtype=np.float64)
    w = np.empty(n, dtype=np.float64)
    check(lib().bemhip_gauss_legendre(int(n), dptr(x), dptr(w)))
    return x, w
