"""The C-ABI boundary and the host side, without a GPU.

* libbemhip.so loads and exports every function include/bemhip.h declares,
  and the ctypes table binds each of them;
* host-only entry points (version, status names, Gauss-Legendre rule) and
  the argument checks that run before any device work;
* the Python mirror fails loudly (no CPU fallback) when no GPU is visible;
* Gmsh .msh input (a Gmsh 4.1 ASCII fixture, and write -> read round trips).
"""

import ctypes as C
import os
import re

import numpy as np
import pytest

import bem_gmsh_amd as B
from bem_gmsh_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bemhip.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bemhip_\w+)\s*\(", src)))


def _gpu_visible():
    n = C.c_int(0)
    return _lib.lib().bemhip_device_count(C.byref(n)) == 0 and n.value > 0


def test_header_symbols_exported_and_bound():
    names = _declared()
    assert len(names) >= 20
    L = _lib.lib()
    bound = {s[0] for s in _lib.SIGNATURES}
    for name in names:
        assert hasattr(L, name), f"{name} declared in bemhip.h but not exported"
        assert name in bound, f"{name} has no ctypes signature"
    assert bound <= set(names)


def test_host_entry_points():
    L = _lib.lib()
    assert L.bemhip_version() >= 10000
    assert L.bemhip_strerror(0) == b"ok"
    assert L.bemhip_strerror(1) == b"bad argument"
    assert L.bemhip_strerror(99) == b"unknown status"
    for n in (1, 2, 5, 20, 64):
        x, w = _lib.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xr).max() <= 4e-16 and np.abs(w - wr).max() <= 4e-15
    with pytest.raises(ValueError):
        _lib.gauss_legendre(0)


def test_argument_errors_before_device_work():
    L = _lib.lib()
    dp = _lib.c_double_p
    assert L.bemhip_set_geometry(None, 2, 1, 3, 0, None, None, None, None, None, None, 1.0) == 1
    assert b"NULL context" in L.bemhip_last_error(None)
    assert L.bemhip_assemble(None, 16, 1.5, 0) == 1
    assert L.bemhip_apply_bc(None, None, None, None, None) == 1
    assert L.bemhip_solve(None, 0, 1e-12, 30, None, None, None) == 1
    assert L.bemhip_adjoint(None, None, None, None, None) == 1
    assert L.bemhip_lu_factor(None, 4, 0, None, None) == 1
    assert L.bemhip_set_lu_gemm(None, 0) == 1
    assert L.bemhip_set_lu_timers(None, 1) == 1
    assert L.bemhip_set_lu_pivoting(None, 0) == 1
    assert L.bemhip_create(None, 0) == 1
    del dp


@pytest.mark.skipif("_gpu_visible()")
def test_no_gpu_fails_loudly():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.bemhip_create(C.byref(h), 0) == 7          # BEMHIP_ERR_NODEV
    assert L.bemhip_last_error(None)
    dad = B.format2d(B.quadrado(ndiv=5), B.Laplace(1.0))
    with pytest.raises(ValueError, match="GPU"):
        B.assemble(dad, 12)
    with pytest.raises(RuntimeError):
        B.solve(dad)


def test_read_msh_v41_fixture():
    """Unit square, 2 linear elements per side, one quad cell; physical names
    "t;v" give the BCs (T = 0 on x = 0, T = 1 on x = 1, q = 0 elsewhere)."""
    m = B.read_msh(os.path.join(GOLDEN, "square_2x2_v41.msh"))
    assert m.order == 1 and m.lines.shape == (8, 2, 2) and len(m.cells) == 1
    dad = B.format2d(m, B.Laplace(1.0))
    assert (dad.ne, dad.n, dad.ni) == (8, 16, 1)
    assert np.allclose(dad.collocation[-1], [0.5, 0.5])
    x = dad.Nodes
    left, right = np.abs(x[:, 0]) < 1e-12, np.abs(x[:, 0] - 1) < 1e-12
    assert np.all(dad.BC[left | right] == 0) and np.all(dad.BC[~(left | right)] == 1)
    assert np.all(dad.BV[right] == 1.0) and np.all(dad.BV[~right] == 0.0)
    # same boundary collocation as the built-in quadrado(ndiv=3) (2 elements per side)
    ref = B.format2d(B.quadrado(ndiv=3), B.Laplace(1.0))
    a = np.array(sorted(map(tuple, np.round(dad.Nodes, 14))))
    b = np.array(sorted(map(tuple, np.round(ref.Nodes, 14))))
    assert np.allclose(a, b, atol=1e-14)
    # outward normals
    c = np.array([0.5, 0.5])
    assert np.all(np.sum((dad.Nodes - c) * dad.Normal, axis=1) > 0)


@pytest.mark.parametrize("builder", ["square_p1", "square_p2", "star_p2"])
def test_write_read_msh_roundtrip(builder, tmp_path):
    mesh = {"square_p1": lambda: B.quadrado(ndiv=6),
            "square_p2": lambda: B.quadrado(ndiv=5, ordem=2),
            "star_p2": lambda: B.star_mesh(40, seed=4)}[builder]()
    path = str(tmp_path / "m.msh")
    B.write_msh(mesh, path)
    back = B.read_msh(path)
    d0 = B.format2d(mesh, B.Laplace(1.0))
    d1 = B.format2d(back, B.Laplace(1.0))
    assert (d0.ne, d0.n, d0.ni) == (d1.ne, d1.n, d1.ni)
    assert np.allclose(d0.collocation, d1.collocation, rtol=0, atol=1e-13)
    assert np.allclose(d0.Normal, d1.Normal, rtol=0, atol=1e-13)
    assert np.array_equal(d0.BC, d1.BC) and np.allclose(d0.BV, d1.BV)


def test_geometry_helpers_for_the_adjoint():
    mesh = B.star_mesh(30, seed=2)
    dad = B.format2d(mesh, B.Laplace(1.0))
    nodes, grad, idx = B.unique_node_gradient(dad, np.ones_like(dad.elem_geo))
    # closed curve of quadratic elements: ne shared end nodes + ne mid nodes
    assert nodes.shape[0] == 2 * dad.ne
    assert sorted(np.unique(grad[:, 0]).tolist()) == [1.0, 2.0]
    moved = B.move_geometric_node(mesh, nodes[0], 1, 1e-3)
    d2 = B.format2d(moved, B.Laplace(1.0))
    diff = np.abs(d2.elem_geo - dad.elem_geo).max(axis=-1)
    assert np.count_nonzero(diff) == np.count_nonzero(idx == 0)


def test_gmres_argument_errors():
    L = _lib.lib()
    x = np.zeros(4)
    assert L.bemhip_solve_gmres(None, 0.0, 1e-8, 20, 10, _lib.dptr(x), None, None) == 1
    assert b"NULL context" in L.bemhip_last_error(None)
