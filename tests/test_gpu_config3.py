"""The headline workload itself: the seed-1 star with 10923 quadratic
elements (N = 32769, the nearest multiple of 3 to 32768), npg = 16, T = e^x cos y
with Dirichlet/Neumann arcs of 64 elements (src/Core/Assembly_full.jl:61-85,
src/Core/Solver.jl:24-64 on the GPU).

* Mixed pipeline (FP32 far field, FP32 LU, FP64 refinement to 1e-10) against
  the FP64 pipeline per field (T, q): within 1e-10, the metric's accuracy.
* The near-pair list against the oracle's classification
  (src/Core/Assembly_full.jl:736-754) on 96 sampled rows, every element.
* Two blocks of 24 assembled FP64 rows against the oracle's rows: the near and
  on-element entries within 1e-7 of the row's scale (the bar the sweep uses,
  set by the near pairs whose FP64 values move by up to ~3e-8 under a 1-ulp
  input shift), every far entry within 1e-12 relative.
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

NE, NPG = 10923, 16


@pytest.fixture(scope="module")
def star():
    d = B.format2d(B.star_mesh(NE, seed=1), B.Laplace(1.0))
    B.apply_bc_from_physical(d, B.ana_exp_cos())
    assert d.nt == 32769
    return d


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_config3_mixed_vs_fp64(star):
    d = star
    B.assemble(d, NPG, fetch=False)
    B.solve(d, mode="fp64")
    T64, q64 = d.T.copy(), d.q.copy()
    B.assemble(d, NPG, precision="mixed", fetch=False)
    B.solve(d, mode="mixed", tol=1e-10)
    eT, eq = _rel(d.T, T64), _rel(d.q, q64)
    ms = d.hip.counters()["ms"]
    print(f"\n[star N=32769] mixed vs fp64: T {eT:.2e} q {eq:.2e}; refinement "
          f"{d.solve_info['iters']} steps; solve {ms['solve_mixed']:.0f} ms")
    assert eT <= 1e-10 and eq <= 1e-10
    assert "refine_gmres_ir" not in d.hip.counters()["stats"]


def test_config3_near_pairs_match_classification(star):
    d = star
    B.assemble(d, NPG, precision="mixed", fetch=False)
    rows, elems, _, _ = B.near_pairs(d)
    print(f"\n[star N=32769] {rows.size} near pairs")
    rng = np.random.default_rng(2)
    sample = np.sort(rng.choice(d.nt, 96, replace=False))
    coll = d.collocation
    for i in sample:
        got = set(elems[rows == i].tolist())
        want = {e for e in range(d.ne)
                if 3 * e <= i < 3 * e + 3 or O.near_element(coll[i], coll[3 * e:3 * e + 3],
                                                            d.elem_L[e])}
        assert got == want, (i, sorted(got ^ want))


def test_config3_rows_match_oracle(star):
    d = star
    H, G = B.assemble(d, NPG)                 # FP64, fetched
    rows, elems, _, _ = B.near_pairs(d)
    worst_near = worst_far = 0.0
    for lo in (0, d.nt // 2):
        Hr, Gr = O.assemble(d, NPG, rows=(lo, lo + 24))
        for r in range(lo, lo + 24):
            near_cols = np.zeros(d.nt, dtype=bool)
            for e in elems[rows == r]:
                near_cols[3 * e:3 * e + 3] = True
            near_cols[r] = False               # the diagonal needs the whole row
            h, hr, g, gr = H[r], Hr[r], G[r], Gr[r]
            sc = max(np.abs(hr).max(), np.abs(gr).max())
            worst_near = max(worst_near, np.abs(h - hr)[near_cols].max() / sc,
                             np.abs(g - gr)[near_cols].max() / sc)
            far = ~near_cols
            far[r] = False
            den_h = np.maximum(np.abs(hr[far]), 1e-3 * np.abs(hr).max())
            den_g = np.maximum(np.abs(gr[far]), 1e-3 * np.abs(gr).max())
            worst_far = max(worst_far, (np.abs(h[far] - hr[far]) / den_h).max(),
                            (np.abs(g[far] - gr[far]) / den_g).max())
        del Hr, Gr
    print(f"\n[star N=32769] 48 rows vs oracle: near entries {worst_near:.2e} of the row scale, "
          f"far entries {worst_far:.2e} relative")
    assert worst_near <= 1e-7 and worst_far <= 1e-12
