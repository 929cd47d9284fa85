"""ctypes wrapper of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

The FP64 CPU restatement of the reference algorithm (oracle/bem_oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
as the checker; the product (bem_gmsh_amd/) never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_L = None

dp = C.POINTER(C.c_double)
ip64 = C.POINTER(C.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_gausslegendre.argtypes = [C.c_int, dp, dp]
        L.oracle_poly.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.oracle_shapefun.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
        L.oracle_fundamental.argtypes = [C.c_double] * 5 + [dp, dp]
        L.oracle_fundamental.restype = None
        L.oracle_closest_point.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double, dp, dp]
        L.oracle_complex_pole.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double, dp, dp]
        L.oracle_p3c_inv_pole.argtypes = [C.c_double, C.c_double, dp, dp]
        L.oracle_p3c_inv_pole.restype = None
        L.oracle_map.argtypes = [C.c_int, dp, dp, C.c_int, C.c_double, C.c_double, dp, dp]
        L.oracle_interp_rule.argtypes = [C.c_double, C.c_int, dp, dp, dp, dp]
        L.oracle_interp_integral.argtypes = [C.c_double, C.c_int, dp, C.c_int, dp]
        L.oracle_interp_integral.restype = C.c_double
        L.oracle_near_element.argtypes = [C.c_double, C.c_double, dp, C.c_int, C.c_double,
                                          C.c_double]
        L.oracle_assemble.argtypes = [C.c_int] * 4 + [dp] * 6 + [C.c_double, C.c_int,
                                                                 C.c_double, C.c_int, C.c_int,
                                                                 C.c_int, dp, dp]
        L.oracle_pair.argtypes = [C.c_int] * 4 + [dp] * 6 + [C.c_double, C.c_int, C.c_int,
                                                             C.c_int, dp, dp]
        L.oracle_pair_plan.argtypes = [C.c_int] * 4 + [dp, C.c_int, C.c_int, C.c_int, dp]
        L.oracle_apply_bc.argtypes = [C.c_int, C.c_int, dp, dp, ip64, dp, dp, dp]
        L.oracle_kelvin.argtypes = [C.c_double] * 6 + [dp, dp]
        L.oracle_kelvin.restype = None
        L.oracle_assemble_kelvin.argtypes = [C.c_int] * 4 + [dp] * 6 + [
            C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]
        L.oracle_pair_kelvin.argtypes = [C.c_int] * 4 + [dp] * 6 + [
            C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]
        _L = L
    return _L


def _p(a):
    return a.ctypes.data_as(dp)


def gausslegendre(n):
    x = np.empty(n)
    w = np.empty(n)
    lib().oracle_gausslegendre(n, _p(x), _p(w))
    return x, w


def poly(kind, p):
    x = np.empty(p + 1)
    bw = np.empty(p + 1)
    D = np.empty((p + 1) * (p + 1))
    lib().oracle_poly(kind, p, _p(x), _p(bw), _p(D))
    return x, bw, D.reshape(p + 1, p + 1)


def shapefun(kind, p, xi):
    N = np.empty(p + 1)
    dN = np.empty(p + 1)
    lib().oracle_shapefun(kind, p, float(xi), _p(N), _p(dN))
    return N, dN


def fundamental(r, n, k=1.0):
    U = C.c_double()
    T = C.c_double()
    lib().oracle_fundamental(float(r[0]), float(r[1]), float(n[0]), float(n[1]), float(k),
                             C.byref(U), C.byref(T))
    return U.value, T.value


def closest_point(kind, p, nodes, pf):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    xi = C.c_double()
    d = C.c_double()
    lib().oracle_closest_point(kind, p, _p(nodes), float(pf[0]), float(pf[1]), C.byref(xi),
                               C.byref(d))
    return xi.value, d.value


def complex_pole(kind, p, nodes, pf):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    z = C.c_double()
    e = C.c_double()
    lib().oracle_complex_pole(kind, p, _p(nodes), float(pf[0]), float(pf[1]), C.byref(z),
                              C.byref(e))
    return z.value, e.value


def p3c_inv_pole(zeta, eta):
    a = C.c_double()
    b = C.c_double()
    lib().oracle_p3c_inv_pole(float(zeta), float(eta), C.byref(a), C.byref(b))
    return a.value, b.value


MAPS = {"tanp3c": 0, "tangent": 1, "p3c": 2, "eq41": 3}


def nearmap(which, u, w, zeta, eta):
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.empty(2 * u.size)
    wo = np.empty(2 * u.size)
    m = lib().oracle_map(MAPS[which], _p(u), _p(w), u.size, float(zeta), float(eta), _p(x),
                         _p(wo))
    return x[:m].copy(), wo[:m].copy()


def interp_rule(a, n=20):
    xi, w, Ni, d = (np.empty(n) for _ in range(4))
    m = lib().oracle_interp_rule(float(a), n, _p(xi), _p(w), _p(Ni), _p(d))
    if m < 0:
        raise ValueError("interp_rule: fewer than 2 nodes")
    return xi[:m].copy(), w[:m].copy(), Ni[:m].copy(), d[:m].copy()


def interp_integral(f, a, order, n=20):
    """Guiggiani integral of f on [-1,1] with Laurent order 0 (log) or -1 (CPV) at a."""
    xi, _, _, _ = interp_rule(a, n)
    fi = np.ascontiguousarray(np.array([f(x) for x in xi], dtype=np.float64))
    Cf = np.empty(3)
    val = lib().oracle_interp_integral(float(a), n, _p(fi), int(order), _p(Cf))
    return val, Cf


def near_element(pf, xg, L, factor=1.5):
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    return bool(lib().oracle_near_element(float(pf[0]), float(pf[1]), _p(xg), xg.shape[0],
                                          float(L), float(factor)))


def _geom_args(dad):
    coll = np.ascontiguousarray(dad.collocation, dtype=np.float64)
    nrm = np.ascontiguousarray(dad.Normal, dtype=np.float64)
    J = np.ascontiguousarray(dad.node_J(), dtype=np.float64)
    ew = np.ascontiguousarray(dad.elem_weight, dtype=np.float64)
    geo = np.ascontiguousarray(dad.elem_geo.reshape(-1, 2), dtype=np.float64)
    L = np.ascontiguousarray(dad.elem_L, dtype=np.float64)
    return coll, nrm, J, ew, geo, L


def assemble(dad, npg=20, near_factor=1.5, nthreads=0, rows=None):
    """H (nt x nt), G (nt x n), column-major, as H_G_full_direct."""
    coll, nrm, J, ew, geo, L = _geom_args(dad)
    nt, n = dad.nt, dad.n
    H = np.zeros((nt, nt), order="F")
    G = np.zeros((nt, n), order="F")
    lo, hi = (0, nt) if rows is None else rows
    st = lib().oracle_assemble(dad.p, dad.ne, n, dad.ni, _p(coll), _p(nrm), _p(J), _p(ew),
                               _p(geo), _p(L), float(dad.properties.k), int(npg),
                               float(near_factor), int(nthreads), int(lo), int(hi),
                               H.ctypes.data_as(dp), G.ctypes.data_as(dp))
    if st != 0:
        raise ValueError("oracle_assemble: bad arguments")
    return H, G


def interior_laplace(dad, pts, T, q, npg=20, near_factor=1e300):
    """u(p) = (sum_j g_j q_j - sum_j h_j T_j) / c, c = -sum_j h_j, at points p
    (src/Laplace/Lubrication.jl:399-417; test/laplace.jl:36-53): the rows of
    an assembly with the points as internal nodes (far lumping beyond
    near_factor * L, every pair through the near rule when near_factor > 1e20)."""
    coll, nrm, J, ew, geo, L = _geom_args(dad)
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 2))
    n, m = dad.n, pts.shape[0]
    coll = np.ascontiguousarray(np.concatenate([coll[:n], pts], axis=0))
    nt = n + m
    H = np.zeros((nt, nt), order="F")
    G = np.zeros((nt, n), order="F")
    st = lib().oracle_assemble(dad.p, dad.ne, n, m, _p(coll), _p(nrm), _p(J), _p(ew), _p(geo),
                               _p(L), float(dad.properties.k), int(npg), float(near_factor), 0,
                               n, nt, H.ctypes.data_as(dp), G.ctypes.data_as(dp))
    if st != 0:
        raise ValueError("oracle_assemble: bad arguments")
    h = H[n:, :n]
    g = G[n:, :]
    c = -h.sum(axis=1)
    T = np.asarray(T, dtype=np.float64)[:n]
    q = np.asarray(q, dtype=np.float64)[:n]
    return (g @ q - h @ T) / c


def pair(dad, i, e, npg=20):
    coll, nrm, J, ew, geo, L = _geom_args(dad)
    nn = dad.p + 1
    h = np.empty(nn)
    g = np.empty(nn)
    lib().oracle_pair(dad.p, dad.ne, dad.n, dad.ni, _p(coll), _p(nrm), _p(J), _p(ew), _p(geo),
                      _p(L), float(dad.properties.k), int(npg), int(i), int(e), _p(h), _p(g))
    return h, g


def pair_plan(dad, i, e, npg=20):
    """(zeta0, eta0, branch, clamped) of pair (i, e): branch -1 on-element,
    0 Mobius, 1 p3c only, 2 compose, 3 split; clamped = rule nodes whose
    tangent angle hit the +-(pi/2 - 1e-8) clamp."""
    coll = np.ascontiguousarray(dad.collocation, dtype=np.float64)
    out = np.zeros(4)
    lib().oracle_pair_plan(dad.p, dad.ne, dad.n, dad.ni, _p(coll), int(npg), int(i), int(e),
                           _p(out))
    return float(out[0]), float(out[1]), int(out[2]), int(out[3])


def apply_bc(dad, H, G):
    nt, n = dad.nt, dad.n
    H = np.asfortranarray(H, dtype=np.float64)
    G = np.asfortranarray(G, dtype=np.float64)
    A = np.empty((nt, nt), order="F")
    b = np.empty(nt)
    BC = np.ascontiguousarray(dad.BC, dtype=np.int64)
    BV = np.ascontiguousarray(dad.BV, dtype=np.float64)
    lib().oracle_apply_bc(nt, n, H.ctypes.data_as(dp), G.ctypes.data_as(dp),
                          BC.ctypes.data_as(ip64), _p(BV), A.ctypes.data_as(dp), _p(b))
    return A, b


def solve(dad, H, G):
    """applyBC + dense LU (scipy LAPACK getrf/getrs) + split_sol!."""
    import scipy.linalg as sl
    A, b = apply_bc(dad, H, G)
    x = sl.lu_solve(sl.lu_factor(A), b)
    T = x[: dad.nt].copy()
    q = np.zeros(dad.n)
    d = dad.BC == 0
    q[d] = T[: dad.n][d]
    T[: dad.n][d] = dad.BV[d]
    q[~d] = dad.BV[~d]
    return T, q, x


# ---------------------------------------------------------------------------
# 2-D isotropic elasticity (Kelvin): H_G_full_direct for a Vectorial problem.
# ---------------------------------------------------------------------------

def _kelvin_consts(dad):
    pr = dad.properties
    return float(pr.mu), float(pr.effective_nu())


def kelvin(r, n, mu, nu):
    """U, T (2x2) of fundamental(props::Elasticity, r, n)
    (src/Elasticity/Fundamental.jl:28-57); r = x_field - x_source."""
    U = np.empty(4)
    T = np.empty(4)
    lib().oracle_kelvin(float(r[0]), float(r[1]), float(n[0]), float(n[1]), float(mu),
                        float(nu), _p(U), _p(T))
    return U.reshape(2, 2), T.reshape(2, 2)


def assemble_kelvin(dad, npg=20, near_factor=1.5, nthreads=0, rows=None):
    """H (2nt x 2nt), G (2nt x 2n), column-major, dof 2j + b for node j
    component b (src/Core/Assembly_full.jl:120-175)."""
    coll, nrm, J, ew, geo, L = _geom_args(dad)
    nt, n = dad.nt, dad.n
    mu, nu = _kelvin_consts(dad)
    H = np.zeros((2 * nt, 2 * nt), order="F")
    G = np.zeros((2 * nt, 2 * n), order="F")
    lo, hi = (0, nt) if rows is None else rows
    st = lib().oracle_assemble_kelvin(dad.p, dad.ne, n, dad.ni, _p(coll), _p(nrm), _p(J), _p(ew),
                                      _p(geo), _p(L), mu, nu, int(npg), float(near_factor),
                                      int(nthreads), int(lo), int(hi), H.ctypes.data_as(dp),
                                      G.ctypes.data_as(dp))
    if st != 0:
        raise ValueError("oracle_assemble_kelvin: bad arguments")
    return H, G


def pair_kelvin(dad, i, e, npg=20):
    """Near / on-element Kelvin pair (i, e): h, g of shape (p+1, 2, 2)."""
    coll, nrm, J, ew, geo, L = _geom_args(dad)
    mu, nu = _kelvin_consts(dad)
    nn = dad.p + 1
    h = np.empty(nn * 4)
    g = np.empty(nn * 4)
    lib().oracle_pair_kelvin(dad.p, dad.ne, dad.n, dad.ni, _p(coll), _p(nrm), _p(J), _p(ew),
                             _p(geo), _p(L), mu, nu, int(npg), int(i), int(e), _p(h), _p(g))
    return h.reshape(nn, 2, 2), g.reshape(nn, 2, 2)


def apply_bc_kelvin(dad, H, G):
    """applyBC for elasticity (src/Core/Boundary_conditions.jl:153-175): per dof."""
    nd, nb = 2 * dad.nt, 2 * dad.n
    H = np.asfortranarray(H, dtype=np.float64)
    G = np.asfortranarray(G, dtype=np.float64)
    A = np.empty((nd, nd), order="F")
    b = np.empty(nd)
    BC = np.ascontiguousarray(dad.BC, dtype=np.int64)
    BV = np.ascontiguousarray(dad.BV, dtype=np.float64)
    lib().oracle_apply_bc(nd, nb, H.ctypes.data_as(dp), G.ctypes.data_as(dp),
                          BC.ctypes.data_as(ip64), _p(BV), A.ctypes.data_as(dp), _p(b))
    return A, b


def solve_kelvin(dad, H, G):
    """applyBC + dense LU + split_sol! for elasticity (src/Core/Solver.jl:115-146):
    returns u (2n), traction (2n) and the solution x (2nt)."""
    import scipy.linalg as sl
    A, b = apply_bc_kelvin(dad, H, G)
    x = sl.lu_solve(sl.lu_factor(A), b)
    nb = 2 * dad.n
    d = dad.BC == 0
    u = np.where(d, dad.BV, x[:nb])
    t = np.where(d, x[:nb], dad.BV)
    return u, t, x
