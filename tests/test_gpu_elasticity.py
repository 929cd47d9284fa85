"""2-D elasticity (Kelvin) on the GPU against the FP64 oracle.

H_G_full_direct for a Vectorial problem (src/Core/Assembly_full.jl:120-175),
the per-dof applyBC (src/Core/Boundary_conditions.jl:153-175) and the solve
(src/Core/Solver.jl:115-146) through libbemhip.so, compared on the same inputs
with oracle.assemble_kelvin / apply_bc_kelvin / solve_kelvin:
  far and on-element blocks   1e-12 per entry (row-scale floor 1e-3)
  near blocks                 1e-12 per entry, or within 10x of the oracle's
                              own change under a 1-ulp move of the source or
                              element nodes (the rule's nodes come from device
                              atan/tan/log: ulp-level differences, amplified
                              by the pair's conditioning)
  rigid-body diagonal blocks  within the row's summed entry differences
  A, b of the BC swap         A bit-identical to the oracle's applyBC applied
                              to the GPU's own H, G; b within the rounding of
                              its sum taken in chunks of 256 dofs
  FP64 / mixed solutions      1e-9 relative to the oracle's LAPACK solve
"""

import ctypes as C

import numpy as np
import pytest

import bem_gmsh_amd as B
from bem_gmsh_amd._lib import lib
import oracle as O

pytestmark = pytest.mark.gpu

ANA = B.ana_elasticity_patch(1.0, 0.3, 0.01, 0.004, 0.002)


def square_internal():
    d = B.format2d(B.quadrado_elasticity(ndiv=10), B.Elasticity(1.0, 0.3, 1.0),
                   pontointerno=True)
    B.apply_bc_from_physical(d, ANA)
    return d


def ellipse_p2():
    d = B.format2d(B.ellipse_mesh(150), B.Elasticity(1.0, 0.3, 1.0), pontointerno=False)
    B.apply_analytical_bc(d, ANA, range(0, d.n, 2))
    return d


def star_p2():
    d = B.format2d(B.star_mesh(160, seed=3), B.Elasticity(2.0, 0.25, 1.0, plane_stress=True),
                   pontointerno=False)
    B.apply_analytical_bc(d, B.ana_elasticity_patch(2.0, 0.2, -0.003, 0.006, 0.001),
                          range(1, d.n, 3))
    return d


CASES = {"square_p1_internal": square_internal, "ellipse_p2": ellipse_p2, "star_p2": star_p2}


def _near_mask(d, npg=16):
    """[node row, node col] True where the oracle integrates the pair (i, elem(j))."""
    nn = d.p + 1
    M = np.zeros((d.nt, d.n), dtype=bool)
    for e in range(d.ne):
        xg = d.collocation[e * nn:(e + 1) * nn]
        for i in range(d.nt):
            if (e * nn <= i < (e + 1) * nn) or O.near_element(d.collocation[i], xg,
                                                              d.elem_L[e]):
                M[i, e * nn:(e + 1) * nn] = True
    return M


def _pair_ulp_sensitivity(d, i, e, npg):
    """Largest relative change of the oracle's Kelvin (h, g) for pair (i, e)
    when one coordinate of the source point or of the element's collocation
    nodes moves by one ulp: the pair's intrinsic FP64 conditioning."""
    import copy
    h0, g0 = O.pair_kelvin(d, i, e, npg)
    nn = d.p + 1
    worst = 0.0
    for pt in [i] + [e * nn + k for k in range(nn)]:
        for comp in (0, 1):
            for sgn in (1.0, -1.0):
                d2 = copy.copy(d)
                d2.collocation = d.collocation.copy()
                d2.collocation[pt, comp] = np.nextafter(d2.collocation[pt, comp], sgn * np.inf)
                h1, g1 = O.pair_kelvin(d2, i, e, npg)
                for a, b in ((h1, h0), (g1, g0)):
                    floor = 1e-3 * np.abs(b).max()
                    rel = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
                    worst = max(worst, rel)
    return worst


def _block_err(A, Ref, nt, ncol):
    """Per-entry error with a floor of 1e-3 * max|row of Ref|, as (nt, ncol) node blocks."""
    rowmax = np.abs(Ref).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(Ref), 1e-3 * rowmax)
    den = np.where(den > 0, den, 1.0)
    e = (np.abs(A - Ref) / den).reshape(nt, 2, ncol, 2)
    return e.max(axis=(1, 3))


@pytest.mark.parametrize("case", list(CASES))
def test_kelvin_assembly_matches_oracle(case):
    d = CASES[case]()
    print(f"\n[{case}] nt={d.nt} n={d.n} p={d.p}", flush=True)
    _check_kelvin_assembly(d, 16)


@pytest.mark.parametrize("npg", [5, 24, 40])
def test_kelvin_assembly_matches_oracle_other_npg(npg):
    """The Kelvin near rule at the other lane groups of its pair kernel: npg = 5,
    24 (32 lanes per pair) and 40 (64 lanes), on the star. Same bars."""
    d = star_p2()
    print(f"\n[star_p2 npg={npg}] nt={d.nt} n={d.n}", flush=True)
    _check_kelvin_assembly(d, npg)


def _check_kelvin_assembly(d, npg):
    H, G = B.assemble(d, npg)
    Hr, Gr = O.assemble_kelvin(d, npg)
    near = _near_mask(d, npg)
    nt, n = d.nt, d.n
    eG = _block_err(G, Gr, nt, n)
    eH = _block_err(H, Hr, nt, nt)[:, :n]
    diag = np.zeros((nt, n), dtype=bool)
    diag[np.arange(n), np.arange(n)] = True
    nn = d.p + 1
    onel = np.zeros((nt, n), dtype=bool)
    for i in range(n):
        e = i // nn
        onel[i, e * nn:(e + 1) * nn] = True
    far = ~near
    nreg = near & ~onel
    print(f"  G far {eG[far].max():.2e} onel {eG[onel].max():.2e} near {eG[nreg].max():.2e}; "
          f"H far {eH[far].max():.2e} onel {eH[onel & ~diag].max():.2e} "
          f"near {eH[nreg].max():.2e}", flush=True)
    assert eG[far].max() <= 1e-12 and eH[far].max() <= 1e-12
    assert eG[onel].max() <= 1e-12 and eH[onel & ~diag].max() <= 1e-12
    # near pairs: 1e-12, or within 10x of the oracle's own 1-ulp sensitivity
    worst = 0.0
    nsens = 0
    for e in range(d.ne):
        cols = slice(e * nn, (e + 1) * nn)
        for i in np.nonzero(nreg[:, e * nn])[0]:
            err = max(eG[i, cols].max(), eH[i, cols].max())
            if err <= 1e-12:
                continue
            kappa = _pair_ulp_sensitivity(d, int(i), e, npg)
            nsens += 1
            worst = max(worst, err / max(kappa, 1e-300))
            assert err <= 10.0 * kappa, (int(i), e, err, kappa)
    print(f"  near pairs above 1e-12: {nsens}, worst err / ulp sensitivity {worst:.2f}",
          flush=True)
    # rigid-body diagonal blocks: translations in the null space of H, and the
    # blocks within the row's summed entry differences of the oracle's
    for b in range(2):
        tr = np.zeros(2 * nt)
        tr[b::2] = 1.0
        assert np.abs(H @ tr).max() <= 1e-12 * np.abs(H).max() * nt ** 0.5
    Hb, Hrb = H.reshape(nt, 2, nt, 2), Hr.reshape(nt, 2, nt, 2)
    off = np.abs(H - Hr).reshape(nt, 2, nt, 2).copy()
    idx = np.arange(nt)
    off[idx, :, idx, :] = 0.0
    # + the rounding of the row sums taken in another order (cancellation to ~0)
    rowabs = np.abs(Hr).sum(axis=1).reshape(nt, 2)[:, :, None]
    bound = off.sum(axis=(2, 3))[:, :, None] + 1e-13 * rowabs
    assert np.all(np.abs(Hb[idx, :, idx, :] - Hrb[idx, :, idx, :]) <= bound)
    # internal-node columns: zero but the diagonal blocks
    if d.ni:
        cols = Hb[:, :, n:, :].copy()
        cols[np.arange(n, nt), :, np.arange(nt - n), :] = 0.0
        assert not cols.any()


@pytest.mark.parametrize("case", list(CASES))
def test_kelvin_bc_and_solves_match_oracle(case):
    d = CASES[case]()
    H, G = B.assemble(d, 16)
    B.apply_bc(d, fetch=True)
    A0, b0 = O.apply_bc_kelvin(d, H, G)
    assert np.array_equal(d.A, A0)
    terms = np.abs(np.where(d.BC == 0, H[:, :2 * d.n], G) * d.BV).sum(axis=1)
    assert np.all(np.abs(d.b - b0) <= 1e-14 * terms)
    Hr, Gr = O.assemble_kelvin(d, 16)
    ur, tr, xr = O.solve_kelvin(d, Hr, Gr)
    u = B.solve(d, mode="fp64")
    x64 = d.solve_info["x"]
    e64 = np.linalg.norm(x64 - xr) / np.linalg.norm(xr)
    assert e64 <= 1e-9
    assert np.array_equal(d.T, d.u) and d.traction.shape == (2 * d.n,)
    B.solve(d, mode="mixed", tol=1e-12)
    xm = d.solve_info["x"]
    em = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    print(f"\n[{case}] fp64 vs oracle {e64:.2e}, mixed vs fp64 {em:.2e}, refine "
          f"{d.solve_info['iters']} relres {d.solve_info['relres']:.1e}", flush=True)
    assert em <= 1e-9 and d.solve_info["relres"] <= 1e-12
    # the operator of the mixed refinement is the stored FP64 system
    z = np.random.default_rng(1).standard_normal(2 * d.nt)
    y = B.apply_operator(d, z)
    assert np.abs(y - A0 @ z).max() <= 1e-12 * np.abs(A0).sum(axis=1).max()


@pytest.mark.parametrize("case", list(CASES))
def test_kelvin_mixed_storage_operator_and_solve(case):
    """Mixed assembly: FP32 H, G within one FP32 ulp of the FP64 assembly's
    values rounded to FP32 (the far values computed with one division per
    pair, FP64 ulps from the FP64 mode's); the FP64
    right-hand side and operator recomputed from the geometry (no FP64 matrix
    stored) against the stored FP64 system; the mixed solve against FP64."""
    d = CASES[case]()
    H64, G64 = B.assemble(d, 16)
    B.apply_bc(d, fetch=True)
    A64, b64 = d.A.copy(), d.b.copy()
    B.solve(d, mode="fp64")
    x64 = d.solve_info["x"].copy()
    z = np.random.default_rng(2).standard_normal(2 * d.nt)
    H32, G32 = B.assemble(d, 16, precision="mixed")
    assert H32.dtype == np.float32 and G32.dtype == np.float32
    # the FP64 values rounded to FP32; far values and the diagonal come from the
    # one-division kernel (FP64 ulps apart), so within one FP32 ulp
    # (plus 1e-12 of the row's scale: entries that cancel to exactly 0 in the
    # FP64 mode, e.g. r.n on a straight side, are ~1e-18 there)
    for M32, M64 in ((H32, H64), (G32, G64)):
        ref = M64.astype(np.float32)
        floor = 1e-12 * np.abs(M64).max(axis=1, keepdims=True)
        assert np.all(np.abs(M32.astype(np.float64) - ref) <= np.spacing(np.abs(ref)) + floor)
    B.apply_bc(d)
    terms = np.abs(np.where(d.BC == 0, H64[:, :2 * d.n], G64) * d.BV).sum(axis=1)
    assert np.all(np.abs(d.b - b64) <= 1e-13 * terms)
    y = B.apply_operator(d, z)
    rowabs = np.abs(A64) @ np.abs(z)
    assert np.all(np.abs(y - A64 @ z) <= 1e-13 * rowabs)
    B.solve(d, mode="mixed", tol=1e-12)
    em = np.linalg.norm(d.solve_info["x"] - x64) / np.linalg.norm(x64)
    print(f"\n[{case}] mixed storage: solve vs fp64 {em:.2e}, refine {d.solve_info['iters']}, "
          f"relres {d.solve_info['relres']:.1e}", flush=True)
    assert em <= 1e-9 and d.solve_info["relres"] <= 1e-12
    with pytest.raises(Exception):
        B.solve(d, mode="fp64")        # the FP64 LU needs the FP64 assembly


@pytest.mark.parametrize("case", list(CASES))
def test_kelvin_matrix_free_gmres(case):
    """Matrix-free assembly (near pairs and FP64 diagonal blocks only): the
    right-hand side and A x from the operator pass against the stored FP64
    system, H, G and the LU refused, and GMRES (right Jacobi) against the FP64
    LU solution; GMRES also after the FP64 assembly."""
    d = CASES[case]()
    H64, G64 = B.assemble(d, 16)
    B.apply_bc(d, fetch=True)
    A64, b64 = d.A.copy(), d.b.copy()
    B.solve(d, mode="fp64")
    x64 = d.solve_info["x"].copy()
    B.solve_Hmat(d, atol=0.0, rtol=1e-12, restart=100)
    assert d.gmres_stats["solved"]
    e_stored = np.linalg.norm(d.solve_info["x"] - x64) / np.linalg.norm(x64)
    B.assemble(d, 16, precision="matrix_free")
    B.apply_bc(d)
    terms = np.abs(np.where(d.BC == 0, H64[:, :2 * d.n], G64) * d.BV).sum(axis=1)
    assert np.all(np.abs(d.b - b64) <= 1e-13 * terms)
    z = np.random.default_rng(3).standard_normal(2 * d.nt)
    y = B.apply_operator(d, z)
    assert np.all(np.abs(y - A64 @ z) <= 1e-13 * (np.abs(A64) @ np.abs(z)))
    with pytest.raises(Exception):
        B.get_HG(d)
    B.solve(d)                         # matrix-free: solve dispatches to solve_Hmat
    e_mf = np.linalg.norm(d.solve_info["x"] - x64) / np.linalg.norm(x64)
    print(f"\n[{case}] GMRES after FP64 assembly {e_stored:.2e}; matrix-free "
          f"{d.gmres_stats['niter']} products, vs FP64 LU {e_mf:.2e}", flush=True)
    assert d.gmres_stats["solved"]
    assert e_stored <= 1e-9 and e_mf <= 1e-6
    B.solve_Hmat(d, atol=0.0, rtol=1e-12, restart=100)
    e_mf12 = np.linalg.norm(d.solve_info["x"] - x64) / np.linalg.norm(x64)
    assert e_mf12 <= 1e-9 and d.uint.shape == (2 * d.ni,)


def test_kelvin_near_pairs_match_oracle_pairs():
    d = B.format2d(B.quadrado_elasticity(ndiv=9), B.Elasticity(1.0, 0.3), pontointerno=True)
    B.assemble(d, 16, fetch=False)
    rows, elems, h, g = B.near_pairs(d)
    assert h.shape == (rows.size, 2, 2, 2)
    rng = np.random.default_rng(4)
    doubled = 0
    pick = list(rng.choice(rows.size, 60, replace=False))
    # rows on the extension of a corner element: the rule's collapsed half
    pick += [k for k in range(rows.size) if (rows[k], elems[k]) in ((45, 24), (32, 15))]
    for k in pick:
        hr, gr = O.pair_kelvin(d, int(rows[k]), int(elems[k]), npg=16)
        s = max(np.abs(hr).max(), np.abs(gr).max())
        assert np.abs(h[k] - hr).max() <= 1e-10 * s and np.abs(g[k] - gr).max() <= 1e-10 * s
        doubled += (rows[k], elems[k]) in ((45, 24), (32, 15))
    assert doubled == 2


def test_reference_elasticity_patch_gpu():
    # test/elasticity.jl:27-36 on the GPU path: rel_error < 0.15, tractions as the oracle's
    d = B.format2d(B.quadrado_elasticity(ndiv=8), B.Elasticity(1.0, 0.3, 1.0),
                   pontointerno=False)
    ana = B.ana_elasticity_patch(E=1.0, nu=0.3, exx=0.01)
    B.apply_analytical_bc(d, ana)
    B.assemble(d, 12)
    B.solve(d)
    assert B.rel_error(d) < 0.15
    Hr, Gr = O.assemble_kelvin(d, 12)
    ur, tr, _ = O.solve_kelvin(d, Hr, Gr)
    assert np.linalg.norm(d.traction - tr) <= 1e-9 * np.linalg.norm(tr)


def test_kelvin_internal_point_displacements():
    # internal nodes of the square (pontointerno): their rows of H, G give u at
    # the cell centroids; the uniform-strain field is linear, so exact up to
    # the boundary solution's error
    d = square_internal()
    B.assemble(d, 16, fetch=False)
    B.solve(d)
    assert d.uint.shape == (2 * d.ni,)
    exact = ANA.u(d.internalNodes).reshape(-1)
    err = np.linalg.norm(d.uint - exact) / np.linalg.norm(exact)
    print(f"\n  internal displacements: rel err {err:.2e}", flush=True)
    assert err < 1e-2


def test_kelvin_patch_on_ellipse_gpu():
    d = B.format2d(B.ellipse_mesh(300), B.Elasticity(1.0, 0.3), pontointerno=False)
    B.apply_analytical_bc(d, ANA, range(0, d.n, 2))
    B.assemble(d, 16, fetch=False)
    B.solve(d, mode="mixed")
    assert B.rel_error(d) < 1e-3
    assert B.rel_error_flux(d) < 0.05


def _ctx_with_geometry(d):
    ctx = B.HipContext(0)
    B.api.upload_geometry(d, ctx)
    return ctx


def test_kelvin_argument_and_state_errors():
    L = lib()
    d = B.format2d(B.quadrado_elasticity(ndiv=6), B.Elasticity(1.0, 0.3), pontointerno=False)
    ctx = B.HipContext(0)
    assert L.bemhip_set_elasticity(ctx.h, 0.4, 0.3) == 3          # no geometry yet
    ctx.close()
    ctx = _ctx_with_geometry(d)
    assert L.bemhip_set_elasticity(ctx.h, 0.4, 0.5) == 1          # nu must be < 0.5
    assert L.bemhip_set_elasticity(ctx.h, -1.0, 0.3) == 1         # mu must be > 0
    assert L.bemhip_set_elasticity(ctx.h, 0.4, 0.3) == 0
    assert L.bemhip_assemble(ctx.h, 16, 1.5, 3) == 1              # no FMM far field
    assert L.bemhip_assemble(ctx.h, 16, 1.5, 0) == 0
    BC = np.zeros(2 * d.n, dtype=np.int64)
    BV = np.zeros(2 * d.n)
    BC[1] = 2
    assert L.bemhip_apply_bc(ctx.h, BC.ctypes.data_as(C.POINTER(C.c_int64)),
                             BV.ctypes.data_as(C.POINTER(C.c_double)), None, None) == 1
    BC[1] = 0
    assert L.bemhip_apply_bc(ctx.h, BC.ctypes.data_as(C.POINTER(C.c_int64)),
                             BV.ctypes.data_as(C.POINTER(C.c_double)), None, None) == 0
    pts = np.array([[0.5, 0.5]])
    u = np.empty(1)
    z = np.zeros(d.n)
    dp = C.POINTER(C.c_double)
    assert L.bemhip_interior_laplace(ctx.h, 1, pts.ctypes.data_as(dp), z.ctypes.data_as(dp),
                                     z.ctypes.data_as(dp), 16, 1e300,
                                     u.ctypes.data_as(dp)) == 3
    ctx.close()
    with pytest.raises(B.ArgumentError):
        B.assemble(d, 16, precision="fmm")


def test_laplace_after_elasticity_on_one_context():
    # set_geometry resets the problem: the same context then assembles Laplace
    d = B.format2d(B.quadrado_elasticity(ndiv=7), B.Elasticity(1.0, 0.3), pontointerno=True)
    B.assemble(d, 16, fetch=False)
    dl = B.format2d(B.quadrado(ndiv=7), B.Laplace(1.0), pontointerno=True)
    dl.hip = d.hip
    H, G = B.assemble(dl, 16)
    Hr, Gr = O.assemble(dl, 16)
    assert H.shape == (dl.nt, dl.nt)
    assert np.abs(G - Gr).max() <= 1e-10 * np.abs(Gr).max()


def test_kelvin_star_N8190_matches_oracle():
    """A larger Kelvin case: star of 1365 quadratic elements (4095 nodes, 8190
    dofs), plane strain. Full FP64 H, G against the oracle: entries above
    1e-10 within 10x their pair's 1-ulp sensitivity; the FP64 solve against
    LAPACK on the oracle's applyBC of the same H, G (backward error <= N eps,
    difference within the conditioning); the mixed pipeline against FP64."""
    import scipy.linalg as sl
    d = B.format2d(B.star_mesh(1365, seed=2), B.Elasticity(1.0, 0.3, 1.0))
    neu = np.nonzero(d.BC.reshape(-1, 2)[:, 0] == 1)[0]
    B.apply_analytical_bc(d, ANA, neu)
    nt, n, nn = d.nt, d.n, d.p + 1
    H, G = B.assemble(d, 16)
    Hr, Gr = O.assemble_kelvin(d, 16)
    rowH = np.abs(Hr).max(axis=1, keepdims=True)
    rowG = np.abs(Gr).max(axis=1, keepdims=True)
    EH = (np.abs(H - Hr) / np.maximum(np.abs(Hr), 1e-3 * rowH)).reshape(nt, 2, nt, 2)
    EG = (np.abs(G - Gr) / np.maximum(np.abs(Gr), 1e-3 * rowG)).reshape(nt, 2, n, 2)
    idx = np.arange(nt)
    EH[idx, :, idx, :] = 0.0                 # diagonal blocks: whole-row sums
    eh = EH.max(axis=(1, 3))[:, :n]
    eg = EG.max(axis=(1, 3))
    bad = np.argwhere((eh > 1e-10) | (eg > 1e-10))
    pairs = sorted({(int(i), int(j) // nn) for i, j in bad})
    worst = 0.0
    for i, e in pairs:
        sl_ = slice(e * nn, (e + 1) * nn)
        err = max(eh[i, sl_].max(), eg[i, sl_].max())
        kappa = _pair_ulp_sensitivity(d, i, e, 16)
        worst = max(worst, err / max(kappa, 1e-300))
        assert err <= 10.0 * kappa, (i, e, err, kappa)
    print(f"\n[star N={2 * nt}] worst entry {max(eh.max(), eg.max()):.2e}; {len(pairs)} pairs "
          f"above 1e-10, worst err / ulp sensitivity {worst:.2f}", flush=True)
    A, b = O.apply_bc_kelvin(d, H, G)
    lu = sl.lu_factor(A)
    xr = sl.lu_solve(lu, b)
    rcond, _ = sl.lapack.dgecon(lu[0], np.abs(A).sum(axis=0).max(), norm="1")
    B.solve(d, mode="fp64")
    x64 = d.solve_info["x"].copy()
    for x in (x64, xr):
        assert (np.abs(b - A @ x).max() <= 2 * nt * np.finfo(float).eps
                * np.abs(A).sum(axis=1).max() * np.abs(x).max())
    eo = np.linalg.norm(x64 - xr) / np.linalg.norm(xr)
    assert eo <= max(1e-10, 4.0 * np.finfo(float).eps / rcond)
    B.assemble(d, 16, precision="mixed", fetch=False)
    B.solve(d, mode="mixed", tol=1e-12)
    em = np.linalg.norm(d.solve_info["x"] - x64) / np.linalg.norm(x64)
    print(f"  LAPACK {eo:.2e} (kappa_1 {1 / rcond:.1e}), mixed vs fp64 {em:.2e}, u err "
          f"{B.rel_error(d):.2e}", flush=True)
    assert em <= 1e-9
