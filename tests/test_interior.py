"""Interior points on the CPU: grids, inside test, and the oracle's potential.

* internal_grid / internal_layer / point_in_domain (src/Core/InternalPoints.jl)
  against closed forms on the unit square;
* the oracle's interior potential (src/Laplace/Lubrication.jl:399-417) pinned
  three ways: the internal rows of its own assembled system (same numbers),
  the reference's near-boundary approach test (test/laplace.jl:55-85: square,
  T = x, |u(0.5, d) - 0.5| < 5e-3 for d = 1e-1 .. 1e-6), and Green's
  representation of a harmonic field from its exact boundary values (a sign,
  normal or term mistake fails it by orders of magnitude);
* the C ABI rejects a NULL context before any device work.
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
from bem_gmsh_amd import _lib
import oracle as O
from helpers import ellipse, square, star


def test_internal_grid_layouts_on_unit_square():
    dad = square(ndiv=10, pontointerno=False)
    pts = B.internal_grid(dad, 4, 3)
    xs = np.arange(1, 5) / 5.0
    ys = np.arange(1, 4) / 4.0
    ref = np.array([(x, y) for y in ys for x in xs])        # x fastest
    assert pts.shape == (12, 2) and np.abs(pts - ref).max() <= 1e-15
    cell = B.internal_grid(dad, 4, 4, layout="cell")
    c = (np.arange(1, 5) - 0.5) / 4
    assert np.abs(cell - np.array([(x, y) for y in c for x in c])).max() <= 1e-15
    cheb = B.internal_grid(dad, 3, 3, layout="cheb")
    ch = (1 - np.cos(np.pi * (np.arange(1, 4) - 0.5) / 3)) / 2
    assert np.abs(cheb - np.array([(x, y) for y in ch for x in ch])).max() <= 1e-15
    # clearance: element length 1/9; d_min = 1.0001 drops the ring at 0.1
    assert B.internal_grid(dad, 9, 9).shape[0] == 81
    assert B.internal_grid(dad, 9, 9, d_min=1.0001).shape[0] == 49
    with pytest.raises(ValueError):
        B.internal_grid(dad, 0, 3)
    with pytest.raises(ValueError):
        B.internal_grid(dad, 3, 3, layout="hex")


def test_point_in_domain_cells_and_chords():
    sq = square(ndiv=10)                     # domain cells cached by format2d
    assert sq.has_cache("cells")
    el = ellipse(ne=64)                      # no cells: chords of the boundary
    assert not el.has_cache("cells")
    for dad in (sq, el):
        assert B.point_in_domain(dad, (0.5 if dad is sq else 0.0, 0.5 if dad is sq else 0.0))
        assert not B.point_in_domain(dad, (1.5, 0.25))
        assert not B.point_in_domain(dad, (-2.0, 0.0))
    assert B.point_in_domain(el, (0.9, 0.0)) and not B.point_in_domain(el, (0.0, 0.51))


def test_internal_layer_square():
    dad = square(ndiv=10, pontointerno=False)
    m = dad.ne // 4                           # elements per side, length h = 1/m
    delta = 1.25 / m                          # 1.25 x typical h = 1.25 x 1 / (ne / 4)
    pts = B.internal_layer(dad)
    # the element at each corner has its mid-point h/2 < 0.75 delta from the other edge
    assert pts.shape == (4 * (m - 2), 2)
    d = np.minimum.reduce([pts[:, 0], 1 - pts[:, 0], pts[:, 1], 1 - pts[:, 1]])
    assert np.abs(d - delta).max() <= 1e-14
    more = B.internal_layer(dad, fill=4)
    assert more.shape[0] == 4 * (m - 2) + 4   # the 4 core cells farther than 2 delta
    with pytest.raises(ValueError):
        B.internal_layer(dad, every=0)


def test_set_internal_nodes():
    dad = square(ndiv=10)
    n = dad.n
    pts = B.internal_grid(dad, 3, 3)
    B.set_internal_nodes(dad, pts)
    assert dad.ni == 9 and dad.nt == n + 9
    assert np.array_equal(dad.internalNodes, pts)


def test_oracle_interior_equals_internal_rows():
    dad = square(ndiv=10)                     # 81 internal cell centroids
    H, G = O.assemble(dad, npg=16)
    T, q, _ = O.solve(dad, H, G)
    u = O.interior_laplace(dad, dad.internalNodes, T, q, npg=16, near_factor=1.5)
    assert np.abs(u - T[dad.n:]).max() <= 1e-14


def test_oracle_interior_reference_approach():
    """test/laplace.jl:55-85 with the default tan-p3c map."""
    dad = B.format2d(B.quadrado(ndiv=20), B.Laplace(1.0), pontointerno=False)
    B.attach_analytical(dad, B.ana_laplace_linear())
    H, G = O.assemble(dad, npg=16)
    T, q, _ = O.solve(dad, H, G)
    pts = np.array([(0.5, d) for d in (1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6)])
    u = O.interior_laplace(dad, pts, T, q, npg=16)
    assert np.abs(u - 0.5).max() < 5e-3


@pytest.mark.parametrize("case", ["ellipse", "star"])
def test_oracle_interior_green_identity(case):
    dad = ellipse(ne=200) if case == "ellipse" else star(ne=300)
    ana = B.ana_exp_cos()
    Tb = ana.u(dad.Nodes)
    qb = ana.q(dad.Nodes, dad.Normal)
    pts = B.internal_grid(dad, 7, 7, d_min=0.05)
    ue = ana.u(pts)
    u_near = O.interior_laplace(dad, pts, Tb, qb, npg=16)
    u_far = O.interior_laplace(dad, pts, Tb, qb, npg=16, near_factor=1.5)
    bar_near, bar_far = (1e-11, 1e-7) if case == "ellipse" else (1e-9, 1e-6)
    assert np.abs(u_near - ue).max() <= bar_near
    assert np.abs(u_far - ue).max() <= bar_far
    # a planted mistake (the T term with the wrong sign) fails by orders of magnitude
    bad = O.interior_laplace(dad, pts, -Tb, qb, npg=16)
    assert np.abs(bad - ue).max() > 0.1


def test_interior_abi_argument_errors():
    L = _lib.lib()
    assert L.bemhip_interior_laplace(None, 1, None, None, None, 16, 1.5, None) == 1
    assert b"NULL context" in L.bemhip_last_error(None)
