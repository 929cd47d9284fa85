"""Where the seeded star's flux error comes from: the reference's tan-p3c clamp.

On fine star boundaries the FP64 oracle (the reference algorithm restated)
gives rel_error_flux ~ 0.1 (N = 8193) up to ~1 (N = 32769), against ~1e-2 on
coarser stars. The mechanism: for a near pair whose source lies ~1e-8
element-lengths off the element's extended curve beyond an end
(|zeta0| > 1, 1e-8 <= eta0 <~ (1 + |zeta0|) 1e-8), the compose branch of
_tanp3ctrans maps nodes through theta = (u - Bt)/At with At ~ 1/eta0, and the
nodes nearest the tangent map's pole reach the +-(pi/2 - 1e-8) clamp of
src/Core/Integration.jl:1321-1365 (_tanp3c_compose -> tangent step), where
x = zeta0 + eta0 tan(pi/2 - 1e-8) no longer follows the map. Below eta0 = 1e-8
the Mobius branch takes over (:1163-1168) and nothing is clamped.

This test lists every near pair of the N = 8193 seed-1 star (2731 quadratic
elements, npg = 16) whose rule clamps a node, checks that exactly those pairs
are wrong (against scipy adaptive quadrature of the exact integrands) while
the others are right, and that replacing exactly those pairs' entries with the
exact integrals brings rel_error_flux back to the coarse-star level.
"""

import numpy as np
import pytest
from scipy.spatial import cKDTree

import bem_gmsh_amd as B
import oracle as O
from test_oracle import _exact_pair

NPG = 16


def _near_pairs(dad):
    """Every (row, element) pair the assembly integrates with a near or
    on-element rule (src/Core/Assembly_full.jl:736-754)."""
    coll = dad.collocation
    tree = cKDTree(coll)
    pairs = []
    for e in range(dad.ne):
        c = coll[3 * e:3 * e + 3]
        ctr = c.mean(axis=0)
        rad = 1.5 * dad.elem_L[e] + 2 * np.abs(c - ctr).max()  # covers every near row
        for i in tree.query_ball_point(ctr, rad):
            if 3 * e <= i < 3 * e + 3 or O.near_element(coll[i], c, dad.elem_L[e]):
                pairs.append((i, e))
    return pairs


@pytest.fixture(scope="module")
def star8193():
    dad = B.format2d(B.star_mesh(2731, seed=1), B.Laplace(1.0))
    B.apply_bc_from_physical(dad, B.ana_exp_cos())
    pairs = _near_pairs(dad)
    plans = {p: O.pair_plan(dad, p[0], p[1], NPG) for p in pairs}
    H, G = O.assemble(dad, NPG)
    return dad, pairs, plans, H, G


def test_near_pair_enumeration_matches_assembly(star8193):
    # the pairs listed here are exactly the ones the assembly integrated: a
    # far pair's entry is the lumped (p+1)-point value, a near pair's the rule's
    dad, pairs, plans, H, G = star8193
    i = 1234
    row = sorted(e for (r, e) in pairs if r == i)
    assert row and all(np.array_equal(O.pair(dad, i, e, NPG)[1], G[i, 3 * e:3 * e + 3])
                       for e in row)


def test_flux_error_is_the_clamped_pairs(star8193):
    dad, pairs, plans, H, G = star8193
    clamped = [p for p in pairs if plans[p][3] > 0]
    print(f"\n{len(pairs)} near pairs, {len(clamped)} with clamped nodes")
    assert 10 <= len(clamped) <= 200
    for (i, e) in clamped:
        zeta, eta, branch, _ = plans[(i, e)]
        assert branch == 2 and abs(zeta) > 1.0
        assert 1e-8 <= eta <= (1.0 + abs(zeta)) * 1e-8, (i, e, zeta, eta)
    # the clamped pairs are wrong, the others right
    worst_c = []
    H2, G2 = H.copy(), G.copy()
    for (i, e) in clamped:
        hx, gx = _exact_pair(dad, i, e)
        sc = max(np.abs(hx).max(), np.abs(gx).max())
        worst_c.append(max(np.abs(H[i, 3 * e:3 * e + 3] - hx).max(),
                           np.abs(G[i, 3 * e:3 * e + 3] - gx).max()) / sc)
        H2[i, 3 * e:3 * e + 3] = hx
        G2[i, 3 * e:3 * e + 3] = gx
    rng = np.random.default_rng(0)
    others = [p for p in pairs if plans[p][3] == 0 and plans[p][2] >= 0]
    worst_u = 0.0
    for k in rng.choice(len(others), 30, replace=False):
        i, e = others[k]
        h, g = O.pair(dad, i, e, NPG)
        hx, gx = _exact_pair(dad, i, e)
        sc = max(np.abs(hx).max(), np.abs(gx).max())
        worst_u = max(worst_u, max(np.abs(h - hx).max(), np.abs(g - gx).max()) / sc)
    print(f"clamped pairs vs exact: median {np.median(worst_c):.2e}, max {max(worst_c):.2e}; "
          f"30 unclamped near pairs: max {worst_u:.2e}")
    assert np.median(worst_c) > 1e-2 and worst_u < 1e-6
    # rigid-body diagonal of the touched rows (src/Core/Assembly_full.jl:384-390)
    for i in sorted({i for i, _ in clamped}):
        H2[i, i] = 0.0
        H2[i, i] = -H2[i].sum()
    T, q, _ = O.solve(dad, H, G)
    dad.set_cache(T=T, q=q)
    f_ref = B.rel_error_flux(dad)
    T2, q2, _ = O.solve(dad, H2, G2)
    dad.set_cache(T=T2, q=q2)
    f_fix = B.rel_error_flux(dad)
    print(f"rel_error_flux: reference algorithm {f_ref:.3e}, clamped pairs exact {f_fix:.3e}")
    assert f_ref > 5e-2 and f_fix < 2e-2
