"""``assemble`` / ``apply_bc`` / ``solve`` on the HIP library.

Python counterparts of the reference's student-facing calls:

* ``assemble(dad, npg=20, method="hip", precision="fp64")`` —
  ``assemble!(dad; method, npg)`` (src/Core/Assembly_full.jl:107-118). The
  method symbol is the plugin seam; ``"hip"`` is this library
  (``"dense"`` is an alias). Stores ``dad.H``, ``dad.G``, ``qsi``, ``w``,
  ``near_factor`` in the cache like H_G_full_direct (:61-85).
* ``apply_bc(dad)`` — ``applyBC`` (src/Core/Boundary_conditions.jl:20-75).
* ``solve(dad, mode="fp64")`` — ``solve`` (src/Core/Solver.jl:24-64) with
  ``split_sol!`` (:88-98); sets ``dad.T`` (nt) and ``dad.q`` (n).
* Elasticity (``format2d(mesh, Elasticity(...))``): the same calls build the
  2-D Kelvin system (H_G_full_direct for a Vectorial problem, :120-175; the
  per-dof applyBC, src/Core/Boundary_conditions.jl:153-175; solve and
  split_sol!, src/Core/Solver.jl:115-146) and set ``dad.u``, ``dad.traction``
  (2n, node j's component b at 2j + b) and ``dad.T = dad.u``.
* ``solve_Hmat(dad; atol, rtol, itmax)`` — GMRES on the device
  (src/Core/Solver.jl:66-86), the solve of a ``precision="matrix_free"``
  assembly (no stored H, G), which ``solve`` dispatches to.

Matrices are returned to the host lazily: the device copies stay resident in
the context (``dad.hip``); ``dad.H`` / ``dad.G`` are fetched on first access
unless ``fetch=False``.
"""

from __future__ import annotations

import atexit
import ctypes as C
import json
import weakref
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, dptr, fptr, i64ptr, iptr, lib
from .problem import BEMData, Elasticity

PRECISIONS = {"fp64": 0, "mixed": 1, "matrix_free": 2, "fmm": 3}
MODES = {"fp64": 0, "mixed": 1, "mixed_gmres": 2}


# Every open context, closed by an atexit hook: Python runs atexit callbacks
# before it tears modules down and before the C runtime's own exit handlers
# (HIP's among them), so no bemhip_destroy is left to a finalizer that runs in
# an arbitrary order at shutdown.
_LIVE = weakref.WeakSet()


def _close_all_contexts():
    for ctx in list(_LIVE):
        ctx.close()


atexit.register(_close_all_contexts)


class HipContext:
    """Owns one bemhip_ctx (device buffers, factors) bound to a GPU."""

    def __init__(self, device: int = 0):
        L = lib()
        cnt = C.c_int(0)
        if L.bemhip_device_count(C.byref(cnt)) != 0 or cnt.value == 0:
            raise ArgumentError("assemble(method='hip') needs a visible GPU; none was found")
        h = C.c_void_p()
        check(L.bemhip_create(C.byref(h), int(device)))
        self.h = h
        self.device = device
        _LIVE.add(self)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().bemhip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, status: int):
        check(status, self.h)

    def counters(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        self.check(lib().bemhip_counters(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())


class ArgumentError(ValueError):
    """Invalid option (the reference throws ArgumentError)."""


def _context(dad: BEMData, device: int = 0) -> HipContext:
    if dad.hip is None or dad.hip.device != device:
        dad.hip = HipContext(device)
    return dad.hip


def upload_geometry(dad: BEMData, ctx: HipContext) -> None:
    nn = dad.p + 1
    coll = np.ascontiguousarray(dad.collocation, dtype=np.float64)
    nrm = np.ascontiguousarray(dad.Normal, dtype=np.float64)
    J = np.ascontiguousarray(dad.node_J(), dtype=np.float64)
    ew = np.ascontiguousarray(dad.elem_weight, dtype=np.float64)
    geo = np.ascontiguousarray(dad.elem_geo.reshape(-1, 2), dtype=np.float64)
    elen = np.ascontiguousarray(dad.elem_L, dtype=np.float64)
    elastic = isinstance(dad.properties, Elasticity)
    k = 1.0 if elastic else float(dad.properties.k)
    ctx.check(lib().bemhip_set_geometry(ctx.h, dad.p, dad.ne, dad.n, dad.ni, dptr(coll),
                                        dptr(nrm), dptr(J), dptr(ew), dptr(geo), dptr(elen), k))
    if elastic:
        pr = dad.properties
        ctx.check(lib().bemhip_set_elasticity(ctx.h, float(pr.mu), float(pr.effective_nu())))


def assemble(dad: BEMData, npg: int = 20, method: str = "hip", precision: str = "fp64",
             near_factor: Optional[float] = 1.5, device: int = 0, fetch: bool = True):
    """Assemble H (nt x nt) and G (nt x n) on the GPU; returns (H, G)."""
    if method not in ("hip", "dense", "gpu"):
        raise ArgumentError(f"assemble method must be :hip (aliases :dense, :gpu); got {method}")
    if precision not in PRECISIONS:
        raise ArgumentError(f"precision must be 'fp64' or 'mixed'; got {precision!r}")
    if near_factor is None or near_factor == "auto":
        near_factor = 1.5
    if not isinstance(near_factor, (int, float)):
        raise ArgumentError(f"near_factor must be a Real, :auto, or nothing; got {near_factor!r}")
    if precision in ("matrix_free", "fmm"):
        fetch = False
    if isinstance(dad.properties, Elasticity) and precision == "fmm":
        raise ArgumentError("elasticity has no FMM far field (precision 'fp64', 'mixed' or "
                            "'matrix_free')")
    nf = float(near_factor)
    if nf == float("inf"):
        nf = 1e300
    ctx = _context(dad, device)
    upload_geometry(dad, ctx)
    ctx.check(lib().bemhip_assemble(ctx.h, int(npg), nf, PRECISIONS[precision]))
    qsi, w = _lib.gauss_legendre(int(npg))
    for key in ("H", "G", "A", "B", "b", "T", "q", "u", "traction"):
        dad.cache.pop(key, None)
    dad.set_cache(qsi=qsi, w=w, near_factor=float(near_factor),
                  assembly_info={"precision": precision, "npg": int(npg)})
    if fetch:
        H, G = get_HG(dad)
        return H, G
    return None


def get_HG(dad: BEMData):
    """Copy H, G from the device into the cache (column-major nt x nt, nt x n;
    elasticity 2nt x 2nt, 2nt x 2n)."""
    ctx = dad.hip
    prec = dad.assembly_info["precision"]
    nt, n = dad.dof * dad.nt, dad.dof * dad.n
    if prec == "fp64":
        H = np.empty((nt, nt), dtype=np.float64, order="F")
        G = np.empty((nt, n), dtype=np.float64, order="F")
        ctx.check(lib().bemhip_get_HG(ctx.h, dptr(H), dptr(G)))
    else:
        H = np.empty((nt, nt), dtype=np.float32, order="F")
        G = np.empty((nt, n), dtype=np.float32, order="F")
        ctx.check(lib().bemhip_get_HG_f32(ctx.h, fptr(H), fptr(G)))
    dad.set_cache(H=H, G=G)
    return H, G


def near_pairs(dad: BEMData):
    """(rows, elems, h, g) of the near / on-element pairs of the last assembly."""
    ctx = dad.hip
    cnt = C.c_int64(0)
    ctx.check(lib().bemhip_near_pairs(ctx.h, C.byref(cnt)))
    m = cnt.value
    nn = dad.p + 1
    shape = (m, nn) if dad.dof == 1 else (m, nn, 2, 2)  # elasticity: [pair][node][a][b]
    rows = np.empty(m, dtype=np.int32)
    elems = np.empty(m, dtype=np.int32)
    h = np.empty(shape, dtype=np.float64)
    g = np.empty(shape, dtype=np.float64)
    if m:
        ctx.check(lib().bemhip_get_near_pairs(ctx.h, iptr(rows), iptr(elems), dptr(h), dptr(g)))
    return rows, elems, h, g


def apply_bc(dad: BEMData, fetch: bool = False) -> None:
    """Mixed-BC column swap on the device; b (and optionally A) to the cache."""
    ctx = dad.hip
    if ctx is None:
        raise RuntimeError("cache.H is not set — assemble/solve first?")
    BC = np.ascontiguousarray(dad.BC, dtype=np.int64)
    BV = np.ascontiguousarray(dad.BV, dtype=np.float64)
    ns = dad.dof * dad.nt
    b = np.empty(ns, dtype=np.float64)
    if fetch:
        A = np.empty((ns, ns), dtype=np.float64, order="F")
        ctx.check(lib().bemhip_apply_bc(ctx.h, i64ptr(BC), dptr(BV), dptr(A), dptr(b)))
        dad.set_cache(A=A)
    else:
        ctx.check(lib().bemhip_apply_bc(ctx.h, i64ptr(BC), dptr(BV), None, dptr(b)))
    dad.set_cache(b=b)


def split_sol(dad: BEMData, x: np.ndarray):
    """split_sol! (src/Core/Solver.jl:88-98): x -> (T over nt, q over n);
    elasticity (:137-146): x -> (u, traction) over the 2n boundary dofs."""
    if isinstance(dad.properties, Elasticity):
        nb = 2 * dad.n
        d = dad.BC == 0
        u = np.where(d, dad.BV, x[:nb])
        t = np.where(d, x[:nb], dad.BV)
        return u, t
    T = x[: dad.nt].copy()
    q = np.zeros(dad.n)
    dirichlet = dad.BC == 0
    q[dirichlet] = T[: dad.n][dirichlet]
    T[: dad.n][dirichlet] = dad.BV[dirichlet]
    q[~dirichlet] = dad.BV[~dirichlet]
    return T, q


def solve(dad: BEMData, mode: Optional[str] = None, tol: float = 1e-12, maxit: int = 30):
    """Apply BCs and solve on the GPU. ``mode`` 'fp64' (FP64 LU), 'mixed' (FP32
    LU + FP64 refinement, GMRES-IR when refinement stops contracting) or
    'mixed_gmres' (FP32 LU + GMRES-IR from the start); default follows the
    assembly precision."""
    if dad.hip is None or "assembly_info" not in dad.cache:
        raise RuntimeError("cache.H is not set — assemble/solve first?")
    if dad.assembly_info["precision"] in ("matrix_free", "fmm") and mode in (None, "gmres"):
        return solve_Hmat(dad)
    if mode == "gmres":
        return solve_Hmat(dad)
    if mode is None:
        mode = "fp64" if dad.assembly_info["precision"] == "fp64" else "mixed"
    if mode not in MODES:
        raise ArgumentError(f"mode must be 'fp64', 'mixed' or 'mixed_gmres'; got {mode!r}")
    apply_bc(dad)
    ctx = dad.hip
    x = np.empty(dad.dof * dad.nt, dtype=np.float64)
    it = C.c_int(0)
    rel = C.c_double(0.0)
    ctx.check(lib().bemhip_solve(ctx.h, MODES[mode], float(tol), int(maxit), dptr(x),
                                 C.byref(it), C.byref(rel)))
    if isinstance(dad.properties, Elasticity):
        u, t = split_sol(dad, x)
        # internal-point displacements (2 ni), the dad.uint of the reference's
        # thermoelastic solve (src/Elasticity/Thermoelasticity.jl:222-225)
        dad.set_cache(u=u, traction=t, T=u, uint=x[2 * dad.n:].copy(),
                      solve_info={"mode": mode, "iters": it.value, "relres": rel.value, "x": x})
        return dad.u
    T, q = split_sol(dad, x)
    dad.set_cache(T=T, q=q, solve_info={"mode": mode, "iters": it.value, "relres": rel.value,
                                        "x": x})
    return dad.T


def solve_Hmat(dad: BEMData, Pl=None, atol: float = 1e-10, rtol: float = 1e-8, itmax: int = 0,
               restart: int = 50):
    """GMRES on the device for the applied (not stored) system, as
    ``solve_Hmat`` (src/Core/Solver.jl:66-86): stop when ||b - A x|| <= atol +
    rtol ||b||; ``itmax = 0`` means max(4 nt, 200) products. Right Jacobi
    preconditioning is built in (``Pl`` is not supported). Sets dad.T, dad.q
    and ``gmres_stats``."""
    if Pl is not None:
        raise ArgumentError("solve_Hmat: a user preconditioner is not supported "
                            "(diagonal scaling is built in)")
    if dad.hip is None or "assembly_info" not in dad.cache:
        raise RuntimeError("cache.H is not set — assemble/solve first?")
    apply_bc(dad)
    ctx = dad.hip
    ns = dad.dof * dad.nt
    itm = int(itmax) if itmax > 0 else max(4 * ns, 200)
    x = np.empty(ns, dtype=np.float64)
    it = C.c_int(0)
    rel = C.c_double(0.0)
    status = lib().bemhip_solve_gmres(ctx.h, float(atol), float(rtol), int(restart), itm,
                                      dptr(x), C.byref(it), C.byref(rel))
    converged = status == 0
    if status not in (0, 5):          # 5: BEMHIP_ERR_NOCONV, x is still the last iterate
        ctx.check(status)
    stats = {"niter": it.value, "relres": rel.value, "solved": converged}
    info = {"mode": "gmres", "iters": it.value, "relres": rel.value, "x": x}
    if isinstance(dad.properties, Elasticity):
        u, t = split_sol(dad, x)
        dad.set_cache(u=u, traction=t, T=u, uint=x[2 * dad.n:].copy(), gmres_stats=stats,
                      solve_info=info)
        return dad.u
    T, q = split_sol(dad, x)
    dad.set_cache(T=T, q=q, gmres_stats={"niter": it.value, "relres": rel.value,
                                         "solved": converged},
                  solve_info={"mode": "gmres", "iters": it.value, "relres": rel.value, "x": x})
    return dad.T


def apply_operator(dad: BEMData, x: np.ndarray) -> np.ndarray:
    """y = A x in FP64 for the current BC system (verification helper)."""
    ctx = dad.hip
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    ctx.check(lib().bemhip_apply_operator(ctx.h, dptr(x), dptr(y)))
    return y


PIVOTING = {"partial": 0, "tournament": 1}


def set_lu_pivoting(ctx: HipContext, pivoting: str) -> None:
    """FP32 LU pivoting of a context: 'partial' (default; GEPP) or 'tournament'
    (TSLU panels). The FP64 LU always pivots partially."""
    if pivoting not in PIVOTING:
        raise ArgumentError(f"pivoting must be 'tournament' or 'partial'; got {pivoting!r}")
    ctx.check(lib().bemhip_set_lu_pivoting(ctx.h, PIVOTING[pivoting]))


LU_GEMM = {"fp32": 0, "split_bf16": 1}


def set_lu_gemm(ctx: HipContext, mode: str) -> None:
    """Arithmetic of the FP32 LU's large trailing updates: 'fp32' (FP32 matrix
    cores) or 'split_bf16' (the default: each FP32 operand split exactly into
    three bf16 parts, the six partial products down to 2^-16 |a b| on the bf16
    matrix cores with FP32 sums)."""
    if mode not in LU_GEMM:
        raise ArgumentError(f"mode must be 'fp32' or 'split_bf16'; got {mode!r}")
    ctx.check(lib().bemhip_set_lu_gemm(ctx.h, LU_GEMM[mode]))


def lu_factor(A: np.ndarray, ctx: Optional[HipContext] = None, pivoting: Optional[str] = None):
    """Dense LU on the GPU of a host matrix (float64/float32): partial pivoting in
    FP64; FP32 pivots by ``pivoting`` ('tournament' or 'partial'; default: the
    context's setting, partial for a new one). Returns (LU, ipiv) in the
    LAPACK convention (0-based ipiv)."""
    ctx = ctx or HipContext(0)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ArgumentError("lu_factor: square matrix required")
    if pivoting is not None:
        set_lu_pivoting(ctx, pivoting)
    dtype = 0 if A.dtype == np.float64 else 1
    LU = np.array(A, dtype=np.float64 if dtype == 0 else np.float32, order="F", copy=True)
    ipiv = np.empty(A.shape[0], dtype=np.int32)
    ctx.check(lib().bemhip_lu_factor(ctx.h, A.shape[0], dtype, LU.ctypes.data_as(C.c_void_p),
                                     iptr(ipiv)))
    return LU, ipiv


def lu_solve(LU: np.ndarray, ipiv: np.ndarray, b: np.ndarray,
             ctx: Optional[HipContext] = None, trans: int = 0) -> np.ndarray:
    """Solve A x = b (trans=0) or A^T x = b (trans=1) with the factors of lu_factor."""
    ctx = ctx or HipContext(0)
    dtype = 0 if LU.dtype == np.float64 else 1
    LUc = np.asfortranarray(LU)
    x = np.array(b, dtype=np.float64, copy=True)
    ctx.check(lib().bemhip_lu_solve(ctx.h, LU.shape[0], dtype, LUc.ctypes.data_as(C.c_void_p),
                                    iptr(np.ascontiguousarray(ipiv, dtype=np.int32)), dptr(x),
                                    int(trans)))
    return x


def adjoint(dad: BEMData, q_target: np.ndarray, w: Optional[np.ndarray] = None):
    """Adjoint shape gradient of J = 1/2 sum_j w_j (q_j - q_target_j)^2 for the
    system of the last ``solve`` (no counterpart in the reference).

    ``w`` defaults to the boundary quadrature weights J_j w_k, which are then
    differentiated too; BC values and internal points are held fixed.
    Returns (J, dJ_dgeo) with dJ_dgeo shaped like ``dad.elem_geo``
    (ne, p+1, 2): the derivative with respect to every element's copy of its
    geometric nodes (see ``unique_node_gradient`` for shared end nodes).
    """
    ctx = dad.hip
    if ctx is None or "solve_info" not in dad.cache:
        raise RuntimeError("cache.T is not set — assemble/solve first?")
    qt = np.ascontiguousarray(q_target, dtype=np.float64)
    if qt.shape != (dad.n,):
        raise ArgumentError(f"q_target must have n = {dad.n} entries")
    wv = None
    if w is not None:
        wv = np.ascontiguousarray(w, dtype=np.float64)
        if wv.shape != (dad.n,):
            raise ArgumentError(f"w must have n = {dad.n} entries")
    out = np.empty((dad.ne, dad.p + 1, 2), dtype=np.float64)
    J = C.c_double(0.0)
    ctx.check(lib().bemhip_adjoint(ctx.h, dptr(qt), dptr(wv) if wv is not None else None,
                                   dptr(out), C.byref(J)))
    return J.value, out


def unique_node_gradient(dad: BEMData, dgeo: np.ndarray):
    """Sum per-element geometric-node gradients over coincident nodes.
    Returns (nodes (m, 2), grad (m, 2), index (ne, p+1)) with
    dgeo[e, k] contributing to grad[index[e, k]]."""
    pts = np.ascontiguousarray(dad.elem_geo.reshape(-1, 2))
    nodes, index = np.unique(pts, axis=0, return_inverse=True)
    grad = np.zeros_like(nodes)
    np.add.at(grad, index.reshape(-1), dgeo.reshape(-1, 2))
    return nodes, grad, index.reshape(dad.ne, dad.p + 1)
