"""The near-singular rule templates (bem_gmsh_amd/csrc/near_rule.hpp) built for
the host: the one source of the tan-p3c and interpolated-Guiggiani steps that
the assembly kernels instantiate in FP64 and the adjoint in dual numbers.

- In FP64 they reproduce the oracle's pair values (src/Core/Integration.jl
  :1117-1365, src/Core/Assembly_full.jl:397-424, :649-695) exactly.
- In dual numbers, the derivative along a geometric-node move (the node's
  element(s) and every source point it carries move together, as the adjoint
  chains them) matches the oracle's central differences, Richardson-
  extrapolated from (h, 2h) with h = 3e-6 or 1e-5, on the config-4 star (1365 quadratic
  elements, npg = 12), around the nodes where the adjoint once erred.
  Pair (2253, 750) there is a compose-branch pair (|zeta0| = 1.25 > 1, eta0 =
  1.05e-7): At, Bt ~ 1/eta0 ~ 1e7, and a derivative of uL = Bt - At pi/2
  taken through the reference's own formula loses 14 digits; the rule takes
  it from the cancellation-free form, and this test fails without it.
"""

import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "bem_gmsh_amd", "csrc")
NPG = 12
STEPS = (3e-6, 1e-5)   # central-difference steps (Richardson pairs h, 2h)

dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(dp)


@pytest.fixture(scope="module")
def rule(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("rule") / "libnear_rule_host.so")
    subprocess.run([cxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I", CSRC, os.path.join(HERE, "rule_host", "near_rule_host.cpp"), "-o", out],
                   check=True)
    L = C.CDLL(out)
    for fn in (L.rule_pair_dual, L.rule_pair):
        fn.restype = None
    return L


@pytest.fixture(scope="module")
def star():
    mesh = B.star_mesh(1365, seed=1)
    dad = B.format2d(mesh, B.Laplace(1.0))
    nodes, index = np.unique(dad.elem_geo.reshape(-1, 2), axis=0, return_inverse=True)
    return mesh, dad, nodes, index.reshape(dad.ne, 3)


def _init(rule, dad):
    rule.rule_init(NPG, C.c_double(dad.properties.k),
                   _p(np.ascontiguousarray(dad.elem_weight, dtype=np.float64)))


def _args(dad, i, e):
    gx = np.ascontiguousarray(dad.elem_geo[e, :, 0])
    gy = np.ascontiguousarray(dad.elem_geo[e, :, 1])
    cx = np.ascontiguousarray(dad.collocation[3 * e:3 * e + 3, 0])
    cy = np.ascontiguousarray(dad.collocation[3 * e:3 * e + 3, 1])
    kloc = i - 3 * e if 3 * e <= i < 3 * e + 3 else -1
    return gx, gy, cx, cy, kloc


def _near_pairs_of_node(dad, index, u):
    """(row, element) near pairs whose values move with geometric node u."""
    els = sorted(set(np.where(index == u)[0].tolist()))
    rows = sorted({r for e in els for r in range(3 * e, 3 * e + 3)})
    coll = dad.collocation
    pairs = []
    for e in range(dad.ne):
        for i in (range(dad.nt) if e in els else rows):
            if 3 * e <= i < 3 * e + 3 or O.near_element(coll[i], coll[3 * e:3 * e + 3],
                                                        dad.elem_L[e]):
                pairs.append((i, e))
    return els, pairs


def test_fp64_rule_equals_oracle_pair_values(rule, star):
    _, dad, nodes, index = star
    _init(rule, dad)
    coll, nrm = dad.collocation, dad.Normal
    n = 0
    for u in (77, 402, 349):
        _, pairs = _near_pairs_of_node(dad, index, u)
        for i, e in pairs:
            gx, gy, cx, cy, kloc = _args(dad, i, e)
            h, g = np.empty(3), np.empty(3)
            rule.rule_pair(_p(gx), _p(gy), _p(cx), _p(cy), C.c_double(coll[i, 0]),
                           C.c_double(coll[i, 1]), C.c_double(nrm[3 * e, 0]),
                           C.c_double(nrm[3 * e, 1]), kloc, _p(h), _p(g))
            hr, gr = O.pair(dad, i, e, NPG)
            if kloc < 0:   # near pairs: the same operations in the same order
                assert np.array_equal(h, hr) and np.array_equal(g, gr), (i, e)
            else:          # on-element: the 20-node sums in another order
                sc = max(np.abs(hr).max(), np.abs(gr).max())
                assert np.abs(h - hr).max() <= 1e-14 * sc and np.abs(g - gr).max() <= 1e-14 * sc
            n += 1
    print(f"\n{n} pairs equal to the oracle's")


@pytest.mark.parametrize("u", [77, 402, 349, 949])
def test_dual_derivatives_match_oracle_differences(rule, star, u):
    mesh, dad, nodes, index = star
    _init(rule, dad)
    coll, nrm = dad.collocation, dad.Normal
    lx = np.polynomial.legendre.leggauss(3)[0]
    Ngeo = np.array([[x * (x - 1) / 2, 1 - x * x, x * (x + 1) / 2] for x in lx])  # [k][m]
    els, pairs = _near_pairs_of_node(dad, index, u)
    worst = 0.0
    for axis in (0, 1):
        moved = {s: B.format2d(B.move_geometric_node(mesh, nodes[u], axis, s), B.Laplace(1.0))
                 for h in STEPS for s in (h, -h, 2 * h, -2 * h)}
        for i, e in pairs:
            gx, gy, cx, cy, kloc = _args(dad, i, e)
            h, g = np.empty(3), np.empty(3)
            dh, dg = np.empty(24), np.empty(24)
            rule.rule_pair_dual(_p(gx), _p(gy), _p(cx), _p(cy), C.c_double(coll[i, 0]),
                                C.c_double(coll[i, 1]), C.c_double(nrm[3 * e, 0]),
                                C.c_double(nrm[3 * e, 1]), kloc, _p(h), _p(g), _p(dh), _p(dg))
            dv = np.concatenate([dh.reshape(3, 8), dg.reshape(3, 8)])
            d = np.zeros(6)
            for m in range(3):           # the element's own copies of node u
                if index[e, m] == u:
                    d += dv[:, 2 + 2 * m + axis]
            ei, ki = divmod(i, 3)        # the source point, when u shapes its element
            if ei != e:
                for m in range(3):
                    if index[ei, m] == u:
                        d += dv[:, axis] * Ngeo[ki, m]
            vals = {s: np.concatenate(O.pair(md, i, e, NPG)) for s, md in moved.items()}
            errs = []
            for hs in STEPS:   # rounding noise ~1e-12 / h, truncation ~h^4 after Richardson
                f1 = (vals[hs] - vals[-hs]) / (2 * hs)
                f2 = (vals[2 * hs] - vals[-2 * hs]) / (4 * hs)
                scale = max(np.abs(f1).max(), 1e-3)
                assert np.abs(f1 - f2).max() <= 1e-2 * scale, (i, e, "not smooth at this h")
                errs.append(np.abs(d - (4 * f1 - f2) / 3).max() / scale)
            err = min(errs)
            worst = max(worst, err)
            assert err <= 1e-6, (u, axis, i, e, errs)
    print(f"\nnode {u} (elements {els}): {len(pairs)} pairs, worst |dual - fd| / scale {worst:.1e}")
