"""GPU interior potential (bemhip_interior_laplace) against the oracle.

* the same boundary data through the GPU pass and the oracle's interior rows,
  every pair through the near rule (as the reference's interior_laplace,
  src/Laplace/Lubrication.jl:399-417) and with the assembly's far lumping
  (near_factor 1.5): within 1e-11 of max |u| (FP64 both, summation orders
  differ);
* at the internal nodes of an assembled and solved system the GPU pass
  reproduces the system's own internal unknowns (the same rows);
* the reference's near-boundary approach test (test/laplace.jl:55-85) on the
  GPU solve;
* Green's representation of a harmonic field from exact boundary data;
* many points: results do not depend on how the points are split into calls
  or into the library's internal chunks (bit-identical per point).
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O
from helpers import ellipse, square, star

pytestmark = pytest.mark.gpu

CASES = {
    "square_p1": lambda: square(ndiv=20, pontointerno=False),
    "ellipse_p2": lambda: ellipse(ne=200),
    "star_p2": lambda: star(ne=300, seed=3),
}


def _exact_data(dad):
    ana = B.ana_exp_cos()
    return ana, ana.u(dad.Nodes), ana.q(dad.Nodes, dad.Normal)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("nf", [float("inf"), 1.5])
def test_interior_matches_oracle(case, nf):
    dad = CASES[case]()
    _, Tb, qb = _exact_data(dad)
    pts = B.internal_grid(dad, 9, 9, d_min=0.02)
    lay = B.internal_layer(dad, delta=1e-3 * float(dad.elem_L.max()))
    pts = np.concatenate([pts, lay])
    print(f"case={case} nf={nf} points={len(pts)}", flush=True)
    u = B.interior_laplace(dad, pts, T=Tb, q=qb, npg=16, near_factor=nf)
    ref = O.interior_laplace(dad, pts, Tb, qb, npg=16,
                             near_factor=1e300 if nf == float("inf") else nf)
    assert np.all(np.isfinite(u))
    assert np.abs(u - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_interior_equals_internal_unknowns(precision):
    dad = square(ndiv=20)                      # 361 internal cell centroids
    B.assemble(dad, 16, precision=precision, fetch=False)
    B.solve(dad, tol=1e-13)
    u = B.interior_laplace(dad, dad.internalNodes, near_factor=1.5)
    Ti = dad.T[dad.n:]
    assert np.abs(u - Ti).max() <= 1e-11 * np.abs(Ti).max()


def test_interior_reference_approach():
    dad = B.format2d(B.quadrado(ndiv=20), B.Laplace(1.0), pontointerno=False)
    B.attach_analytical(dad, B.ana_laplace_linear())
    B.assemble(dad, 16, fetch=False)
    B.solve(dad)
    assert B.rel_error(dad) < 0.03
    pts = np.array([(0.5, d) for d in (1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6)])
    u = B.interior_laplace(dad, pts)
    assert np.abs(u - 0.5).max() < 5e-3
    assert abs(B.interior_laplace(dad, (0.5, 0.5)) - 0.5) < 5e-3


def test_interior_green_identity():
    dad = ellipse(ne=200)
    ana, Tb, qb = _exact_data(dad)
    pts = B.internal_grid(dad, 15, 15, d_min=0.05)
    u = B.interior_laplace(dad, pts, T=Tb, q=qb, npg=16)
    # discretisation error of the representation: below 1e-12 in the core,
    # a few 1e-9 at the grid points closest to the boundary
    err = float(np.abs(u - ana.u(pts)).max())
    assert err <= 1e-8, err


def test_interior_split_invariance_many_points():
    """70000 points with the assembly rules (two internal chunks of 65536) and
    33000 with the near rule on 2100 elements (chunks of 2^26 / 2100 = 31956
    points): each point's value equals the one from a separate small call."""
    dad = star(ne=2100, seed=5)
    _, Tb, qb = _exact_data(dad)
    rng = np.random.default_rng(7)
    th = rng.uniform(0, 2 * np.pi, 70000)
    r = rng.uniform(0.05, 0.6, 70000)
    pts = np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
    for nf, m in ((1.5, 70000), (float("inf"), 33000)):
        print(f"nf={nf} points={m}", flush=True)
        u = B.interior_laplace(dad, pts[:m], T=Tb, q=qb, npg=12, near_factor=nf)
        for lo, hi in ((0, 500), (m - 500, m)):
            part = B.interior_laplace(dad, pts[lo:hi], T=Tb, q=qb, npg=12, near_factor=nf)
            assert np.array_equal(part, u[lo:hi])
    c = dad.hip.counters()
    assert c["stats"]["interior_points"] == 500


def test_interior_keeps_the_solved_system():
    dad = square(ndiv=12)
    B.assemble(dad, 16, fetch=False)
    B.solve(dad)
    x0 = dad.solve_info["x"].copy()
    B.interior_laplace(dad, B.internal_grid(dad, 5, 5))
    y = B.apply_operator(dad, x0)
    assert np.abs(y - dad.b).max() <= 1e-10 * np.abs(dad.b).max()
