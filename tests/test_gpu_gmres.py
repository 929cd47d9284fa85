"""Matrix-free assembly + GMRES (bemhip_solve_gmres) on the GPU.

* the matrix-free operator equals the stored-FP64 operator bit for bit (same
  near pairs, the same FP64 far recompute, the same rigid-body diagonal: the
  far row sums of an FP64 operator pass with unit weights are the FP64 far
  kernel's sums);
* GMRES on the applied system reproduces the FP64 LU solution of the stored
  system (within its tolerance) and the oracle's end-to-end solution;
* the reference's square test (T = x, rel_error < 0.05) through solve_Hmat;
* errors: no H, G / LU after a matrix-free assembly; itmax reached.
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O
from helpers import ellipse, square, star

pytestmark = pytest.mark.gpu

CASES = {
    "square_p1": lambda: square(ndiv=20),
    "ellipse_p2": lambda: ellipse(ne=200),
    "star_p2": lambda: star(ne=300, seed=3),
}


@pytest.mark.parametrize("case", list(CASES))
def test_matrix_free_operator_equals_stored_fp64(case):
    d64, dmf = CASES[case](), CASES[case]()
    B.assemble(d64, 16, precision="fp64", fetch=False)
    B.apply_bc(d64)
    B.assemble(dmf, 16, precision="matrix_free")
    B.apply_bc(dmf)
    assert np.array_equal(d64.b, dmf.b)
    rng = np.random.default_rng(1)
    for _ in range(2):
        x = rng.standard_normal(d64.nt)
        assert np.array_equal(B.apply_operator(d64, x), B.apply_operator(dmf, x))


def test_matrix_free_operator_with_all_far_tiles():
    """N = 9000: most (row block, column block) tiles are all-far, so the far
    kernel and the operator pass take their path without pair tests; the
    matrix-free operator still equals the stored FP64 one bit for bit."""
    d64, dmf = star(ne=3000, seed=5), star(ne=3000, seed=5)
    B.assemble(d64, 16, precision="fp64", fetch=False)
    B.apply_bc(d64)
    B.assemble(dmf, 16, precision="matrix_free")
    B.apply_bc(dmf)
    assert np.array_equal(d64.b, dmf.b)
    x = np.random.default_rng(2).standard_normal(d64.nt)
    assert np.array_equal(B.apply_operator(d64, x), B.apply_operator(dmf, x))


@pytest.mark.parametrize("case", list(CASES))
def test_gmres_matches_fp64_lu(case):
    d64, dmf = CASES[case](), CASES[case]()
    B.assemble(d64, 16, precision="fp64", fetch=False)
    B.solve(d64, mode="fp64")
    x64 = d64.solve_info["x"]
    B.assemble(dmf, 16, precision="matrix_free")
    B.solve_Hmat(dmf, atol=0.0, rtol=1e-13, restart=60)
    st = dmf.gmres_stats
    print(case, "GMRES products", st["niter"], "relres", st["relres"], flush=True)
    assert st["solved"] and st["relres"] <= 1e-13
    xg = dmf.solve_info["x"]
    assert np.linalg.norm(xg - x64) <= 1e-10 * np.linalg.norm(x64)
    T, q, _ = O.solve(dmf, *O.assemble(dmf, npg=16))
    assert np.linalg.norm(dmf.T - T) <= 1e-9 * np.linalg.norm(T)
    assert np.linalg.norm(dmf.q - q) <= 1e-8 * np.linalg.norm(q)


def test_gmres_after_stored_assembly_and_restarts():
    """GMRES(30) on the star needs restarts (81 products at m = 60 above);
    restarted GMRES with a very small basis (m = 8) stagnates on this system."""
    dad = star(ne=300, seed=3)
    B.assemble(dad, 16, precision="mixed", fetch=False)
    B.solve(dad, mode="mixed")
    xm = dad.solve_info["x"].copy()
    B.solve_Hmat(dad, atol=0.0, rtol=1e-12, restart=30)
    print("GMRES(30) products", dad.gmres_stats["niter"], flush=True)
    assert dad.gmres_stats["solved"] and dad.gmres_stats["niter"] > 30
    assert np.linalg.norm(dad.solve_info["x"] - xm) <= 1e-9 * np.linalg.norm(xm)


def test_reference_square_through_solve_Hmat():
    dad = B.format2d(B.quadrado(ndiv=10), B.Laplace(1.0), pontointerno=True)
    B.attach_analytical(dad, B.ana_laplace_linear())
    B.assemble(dad, 12, precision="matrix_free")
    B.solve(dad)                                  # dispatches to solve_Hmat
    assert dad.gmres_stats["solved"]
    assert B.rel_error(dad) < 0.05


def test_matrix_free_errors():
    dad = square(ndiv=8)
    B.assemble(dad, 12, precision="matrix_free")
    with pytest.raises(RuntimeError):
        B.get_HG(dad)
    with pytest.raises(RuntimeError):
        B.solve(dad, mode="mixed")
    with pytest.raises(RuntimeError):
        B.apply_bc(dad, fetch=True)
    B.solve_Hmat(dad, atol=0.0, rtol=1e-30, itmax=3, restart=2)
    st = dad.gmres_stats
    assert not st["solved"] and st["niter"] == 3
    with pytest.raises(ValueError):
        B.solve_Hmat(dad, Pl=np.eye(2))


def _circle_geometry(ne):
    """Discontinuous linear elements on the unit circle, packed for
    bemhip_set_geometry (numpy only, so a million elements build in seconds)."""
    t = np.linspace(0.0, 2.0 * np.pi, ne + 1)
    p0 = np.stack([np.cos(t[:-1]), np.sin(t[:-1])], axis=1)
    p1 = np.stack([np.cos(t[1:]), np.sin(t[1:])], axis=1)
    geo = np.ascontiguousarray(np.stack([p0, p1], axis=1))            # ne x 2 x 2
    xi = np.array([-1.0, 1.0]) / np.sqrt(3.0)                          # 2-point GL nodes
    coll = (0.5 * (1 - xi)[None, :, None] * p0[:, None, :] +
            0.5 * (1 + xi)[None, :, None] * p1[:, None, :]).reshape(-1, 2)
    L = np.linalg.norm(p1 - p0, axis=1)
    tang = (p1 - p0) / L[:, None]
    nrm = np.repeat(np.stack([tang[:, 1], -tang[:, 0]], axis=1), 2, axis=0)
    J = np.repeat(0.5 * L, 2)
    return (np.ascontiguousarray(coll), np.ascontiguousarray(nrm), J, np.array([1.0, 1.0]),
            geo, L)


def test_matrix_free_refuses_scratch_beyond_device_memory():
    """The dense matrix-free mode keeps 20 bytes per (row, 32-element column
    block): with 1.2e6 elements (2.4e6 rows) that is ~1.8 TB, far above one
    GPU. The assembly refuses before allocating, with BEMHIP_ERR_NOMEM and a
    message naming the O(N) FMM mode; the context stays usable."""
    from bem_gmsh_amd._lib import lib, dptr
    L = lib()
    ctx = B.HipContext(0)
    ne = 1_200_000
    coll, nrm, J, w, geo, elen = _circle_geometry(ne)
    n = 2 * ne
    assert L.bemhip_set_geometry(ctx.h, 1, ne, n, 0, dptr(coll), dptr(nrm), dptr(J), dptr(w),
                                 dptr(geo), dptr(elen), 1.0) == 0
    st = L.bemhip_assemble(ctx.h, 4, 1.5, 2)  # BEMHIP_MATRIX_FREE
    msg = L.bemhip_last_error(ctx.h).decode()
    print("\n", st, msg)
    assert st == 6 and "BEMHIP_FMM" in msg  # BEMHIP_ERR_NOMEM
    # a small problem on the same context still assembles
    coll, nrm, J, w, geo, elen = _circle_geometry(64)
    assert L.bemhip_set_geometry(ctx.h, 1, 64, 128, 0, dptr(coll), dptr(nrm), dptr(J), dptr(w),
                                 dptr(geo), dptr(elen), 1.0) == 0
    assert L.bemhip_assemble(ctx.h, 4, 1.5, 2) == 0
    ctx.close()
