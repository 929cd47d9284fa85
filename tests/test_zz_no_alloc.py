"""bemhip_solve_rhs_noalloc: the re-solve with resident FP64 factors that
allocates nothing (device scratch in a fixed per-GPU array), through the C ABI,
against bemhip_solve on the same system: two calls in a row on two systems,
the second smaller, each bit for bit equal to bemhip_solve's solution for the
assembled right-hand side, and a random right-hand side against numpy's solve
of the same FP64 A (normwise backward error <= n eps)."""

import ctypes as C

import numpy as np
import pytest

import bem_gmsh_amd as B
from bem_gmsh_amd._lib import dptr, lib
from bem_gmsh_amd.api import upload_geometry

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def _star(ne, seed):
    d = B.format2d(B.star_mesh(ne, seed=seed), B.Laplace(1.0))
    B.apply_bc_from_physical(d, B.ana_exp_cos())
    return d


def _noalloc(d, b):
    x = np.empty(d.nt)
    st = lib().bemhip_solve_rhs_noalloc(d.hip.h, dptr(np.ascontiguousarray(b)), dptr(x))
    return st, x


def test_noalloc_resolve_matches_solve_two_calls():
    for ne, seed in ((1200, 3), (500, 4)):           # second call: a smaller system
        d = _star(ne, seed)
        B.assemble(d, 16, fetch=False)
        B.apply_bc(d, fetch=True)
        A, b = d.A.copy(), d.b.copy()
        B.solve(d, mode="fp64")
        x_ref = d.solve_info["x"].copy()
        st, x = _noalloc(d, b)
        assert st == 0, lib().bemhip_last_error(d.hip.h)
        same = np.array_equal(x, x_ref)
        print(f"\n[N={d.nt}] noalloc vs bemhip_solve: {'bit-identical' if same else 'differs'}, "
              f"max diff {np.abs(x - x_ref).max():.1e}")
        assert same
        rng = np.random.default_rng(seed)
        br = rng.standard_normal(d.nt)
        st, xr = _noalloc(d, br)
        assert st == 0
        back = np.abs(br - A @ xr).max() / (np.abs(A).sum(axis=1).max() * np.abs(xr).max())
        xn = np.linalg.solve(A, br)
        print(f"  random rhs: backward error {back:.1e}, vs numpy "
              f"{np.linalg.norm(xr - xn) / np.linalg.norm(xn):.1e}")
        assert back <= d.nt * EPS


def test_noalloc_errors():
    d = _star(200, 5)
    B.assemble(d, 16, precision="mixed", fetch=False)
    B.solve(d, mode="mixed")
    st, _ = _noalloc(d, np.ones(d.nt))
    assert st == 3                                    # BEMHIP_ERR_STATE: no FP64 factors
    assert b"FP64 factors" in lib().bemhip_last_error(d.hip.h)
    big = B.format2d(B.star_mesh(21846, seed=1), B.Laplace(1.0))   # 65538 unknowns
    big.hip = B.HipContext(0)
    upload_geometry(big, big.hip)
    st = lib().bemhip_solve_rhs_noalloc(big.hip.h, dptr(np.ones(big.nt)), dptr(np.empty(big.nt)))
    assert st == 1 and b"maximum of 65536" in lib().bemhip_last_error(big.hip.h)
