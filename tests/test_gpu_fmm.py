"""FMM far field (BEMHIP_FMM assembly) on the GPU.

The FMM evaluates the same lumped far sum as the recomputed FP64 far field,
to its truncation error (48 terms: about 1e-13 of the interacting mass), and
the near pairs and rigid-body diagonal are the same; so the applied operator
and the right-hand side match the stored FP64 system to ~1e-12, and GMRES on
the FMM operator reproduces the FP64 LU solution.
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
from helpers import ellipse, square, star

pytestmark = pytest.mark.gpu

CASES = {
    "square_p1_internal": lambda: square(ndiv=120),          # 476 elements + 14161 internal rows
    "ellipse_p2": lambda: ellipse(ne=800),
    "star_p2": lambda: star(ne=900, seed=3),
}


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", list(CASES))
def test_fmm_operator_matches_fp64(case):
    d64, df = CASES[case](), CASES[case]()
    B.assemble(d64, 16, precision="fp64", fetch=False)
    B.apply_bc(d64)
    B.assemble(df, 16, precision="fmm")
    B.apply_bc(df)
    c = df.hip.counters()["stats"]
    print(case, "nt", df.nt, "levels", c["fmm_levels"], "boxes", c["fmm_boxes"],
          "M2L pairs", c["fmm_m2l_pairs"], flush=True)
    eb = _rel(df.b, d64.b)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(d64.nt)
    ey = _rel(B.apply_operator(df, x), B.apply_operator(d64, x))
    print(case, "rhs rel err", eb, "A x rel err", ey, flush=True)
    assert eb <= 1e-11 and ey <= 1e-11


@pytest.mark.parametrize("case", list(CASES))
def test_fmm_near_pairs_equal_all_pairs_classification(case):
    """The FMM assembly classifies only the elements with a node in the target's
    3 x 3 leaf neighbourhood; its pair list (rows, elements, values) equals the
    all-pairs classification's exactly."""
    dm, df = CASES[case](), CASES[case]()
    B.assemble(dm, 16, precision="matrix_free")
    B.assemble(df, 16, precision="fmm")
    for a, b in zip(B.near_pairs(dm), B.near_pairs(df)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["square_p1_internal", "star_p2"])
def test_fmm_gmres_matches_lu(case):
    d64, df = CASES[case](), CASES[case]()
    B.assemble(d64, 16, precision="fp64", fetch=False)
    B.solve(d64, mode="fp64")
    B.assemble(df, 16, precision="fmm")
    B.solve_Hmat(df, atol=0.0, rtol=1e-12, restart=100)
    print(case, "GMRES products", df.gmres_stats["niter"], flush=True)
    assert df.gmres_stats["solved"]
    x64 = d64.solve_info["x"]
    assert _rel(df.solve_info["x"], x64) <= 1e-9


def test_fmm_needs_a_fine_boundary():
    dad = square(ndiv=5)
    with pytest.raises(ValueError):
        B.assemble(dad, 12, precision="fmm")


def test_stored_modes_refuse_what_fmm_assembles():
    """N > 2^20 nodes: H and G would take > 8 TiB, so the stored modes refuse
    the size at assembly, while the FMM mode assembles it (O(N) memory)."""
    dad = B.format2d(B.quadrado(ndiv=262146, cells=False), B.Laplace(1.0),
                     pontointerno=False)
    assert dad.nt > 1 << 20
    with pytest.raises(ValueError, match="too large"):
        B.assemble(dad, 16, precision="fp64", fetch=False)
    B.assemble(dad, 16, precision="fmm")
    st = dad.hip.counters()["stats"]
    assert st["fmm_levels"] >= 8 and st["fmm_candidates"] > 0
