"""GPU parity of the HIP path against the FP64 oracle (through the C ABI).

Bars: FP64 H and G per entry within 1e-12 relative (entries that cancel to ~0
are measured against 1e-3 of their row's largest entry) for far, on-element and
diagonal entries; near (off-element) entries within 1e-12, or within 10x the
oracle's own change under a 1-ulp input shift where the reference algorithm is
that sensitive (adjacent elements on smooth curves: tan-p3c with a nearly real
pole). T and q within 1e-10 relative, rel_error unchanged; FP32 far field
within 1e-6 relative Frobenius of FP64 (test/laplace.jl:453-455).
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O
from helpers import (ellipse, entry_err, fp64_parity, frob_rel, solution_noise_floor,
                     square, star)

pytestmark = pytest.mark.gpu

CASES = {
    "square_p1": lambda: square(ndiv=20),                 # config 1 (513 x 513, internal nodes)
    "square_p2": lambda: square(ndiv=9, ordem=2),
    "ellipse_p2": lambda: ellipse(ne=200),
    "star_p2": lambda: star(ne=300, seed=3),
}


def _oracle_pairs(dad, near_factor=1.5):
    nn = dad.p + 1
    out = set()
    for e in range(dad.ne):
        xg = dad.collocation[e * nn:(e + 1) * nn]
        for i in range(dad.nt):
            if (e * nn <= i < (e + 1) * nn) or O.near_element(dad.collocation[i], xg,
                                                              dad.elem_L[e], near_factor):
                out.add((i, e))
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_fp64_assembly_matches_oracle(case):
    dad = CASES[case]()
    npg = 20 if case == "square_p1" else 16
    print(f"\n[{case}] nt={dad.nt} n={dad.n} npg={npg}", flush=True)
    H, G = B.assemble(dad, npg)
    Hr, Gr = O.assemble(dad, npg)
    rep = fp64_parity(dad, H, G, Hr, Gr, npg)
    print("  " + ", ".join(f"{k}={v}" for k, v in rep.items()), flush=True)
    # far lumping and on-element (Guiggiani) entries: 1e-12 per entry
    assert rep["far"] <= 1e-12 and rep["onel"] <= 1e-12
    # near entries: 1e-12, or within 10x of the oracle's own 1-ulp sensitivity
    assert not rep["near_violations"]
    # rigid-body diagonal: rows of H sum to zero; the diagonal differs from the
    # oracle's by no more than the row's summed entry differences
    assert np.abs(H.sum(axis=1)).max() <= 1e-12 * np.abs(H).max() * dad.nt ** 0.5
    off = np.abs(H - Hr)
    np.fill_diagonal(off, 0.0)
    ddiff = np.abs(np.diag(H) - np.diag(Hr))
    assert np.all(ddiff <= off.sum(axis=1) + 1e-12 * np.abs(np.diag(Hr)))


@pytest.mark.parametrize("npg", [5, 33, 64])
def test_fp64_assembly_matches_oracle_other_npg(npg):
    """The near rule at quadrature orders outside the cases above: npg = 5 (the
    smallest lane group), 33 (the first with 64 lanes per near pair: a split
    rule's 2 npg nodes in two rounds) and 64 (the largest the library takes),
    on the star with its nearly real poles. Same bars as the assembly test."""
    dad = star(ne=120, seed=5)
    H, G = B.assemble(dad, npg)
    Hr, Gr = O.assemble(dad, npg)
    rep = fp64_parity(dad, H, G, Hr, Gr, npg)
    print(f"\n[star nt={dad.nt} npg={npg}] " + ", ".join(f"{k}={v}" for k, v in rep.items()),
          flush=True)
    assert rep["far"] <= 1e-12 and rep["onel"] <= 1e-12
    assert not rep["near_violations"]
    off = np.abs(H - Hr)
    np.fill_diagonal(off, 0.0)
    ddiff = np.abs(np.diag(H) - np.diag(Hr))
    assert np.all(ddiff <= off.sum(axis=1) + 1e-12 * np.abs(np.diag(Hr)))


@pytest.mark.parametrize("case", ["square_p1", "ellipse_p2"])
def test_near_pair_list_matches_classification(case):
    dad = CASES[case]()
    B.assemble(dad, 16, fetch=False)
    rows, elems, h, g = B.near_pairs(dad)
    got = set(zip(rows.tolist(), elems.tolist()))
    assert got == _oracle_pairs(dad)
    # row-sorted CSR order
    key = rows.astype(np.int64) * dad.ne + elems
    assert np.all(np.diff(key) > 0)


@pytest.mark.parametrize("case", ["square_p1", "ellipse_p2", "star_p2"])
def test_fp64_solve_matches_oracle_on_same_matrices(case):
    """applyBC + LU + solve + split_sol: GPU vs oracle (LAPACK) on the GPU's H, G."""
    dad = CASES[case]()
    npg = 20 if case == "square_p1" else 16
    H, G = B.assemble(dad, npg)
    Tr, qr, xr = O.solve(dad, H, G)
    B.solve(dad, mode="fp64")
    eT = np.linalg.norm(dad.T - Tr) / np.linalg.norm(Tr)
    eq = np.linalg.norm(dad.q - qr) / max(np.linalg.norm(qr), 1e-300)
    print(f"\n[{case}] same-matrix solve: T err {eT:.3e} q err {eq:.3e}")
    assert eT <= 1e-10 and eq <= 1e-10
    A, b = O.apply_bc(dad, H, G)
    assert np.abs(dad.b - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("case", ["square_p1", "ellipse_p2", "star_p2"])
def test_fp64_end_to_end_matches_oracle(case):
    """GPU assemble + solve vs the oracle from scratch: T, q within 1e-10, or
    within 10x the reference algorithm's own solution-level noise floor (near
    pairs whose FP64 values move under a 1-ulp input shift)."""
    dad = CASES[case]()
    npg = 20 if case == "square_p1" else 16
    B.assemble(dad, npg)
    B.solve(dad, mode="fp64")
    Hr, Gr = O.assemble(dad, npg)
    Tr, qr, _ = O.solve(dad, Hr, Gr)
    eT = np.linalg.norm(dad.T - Tr) / np.linalg.norm(Tr)
    eq = np.linalg.norm(dad.q - qr) / max(np.linalg.norm(qr), 1e-300)
    if case.startswith("square"):
        # straight edges: every branch decision is reproduced exactly; strict bar
        nT = nq = 0.0
        changed = 0
    else:
        nT, nq, changed = solution_noise_floor(dad, npg)
    print(f"\n[{case}] end-to-end: T err {eT:.3e} q err {eq:.3e} | noise floor T {nT:.2e} "
          f"q {nq:.2e} ({changed} sensitive pairs)")
    assert eT <= max(1e-10, 10 * nT) and eq <= max(1e-10, 10 * nq)
    err_gpu, flux_gpu = B.rel_error(dad), B.rel_error_flux(dad)
    dad.set_cache(T=Tr, q=qr)
    err_ref, flux_ref = B.rel_error(dad), B.rel_error_flux(dad)
    # rel_error is itself a relative norm: unchanged to 1e-10 absolute
    assert abs(err_gpu - err_ref) <= max(1e-10, 10 * nT)
    assert abs(flux_gpu - flux_ref) <= max(1e-10, 10 * nq)


@pytest.mark.parametrize("case", ["square_p1", "ellipse_p2", "star_p2"])
def test_mixed_precision_matches_fp64(case):
    dad = CASES[case]()
    npg = 20 if case == "square_p1" else 16
    H64, G64 = B.assemble(dad, npg)
    T64 = B.solve(dad, mode="fp64").copy()
    q64 = dad.q.copy()
    H32, G32 = B.assemble(dad, npg, precision="mixed")
    print(f"\n[{case}] FP32 H frob {frob_rel(H32, H64):.3e} G frob {frob_rel(G32, G64):.3e}")
    assert frob_rel(H32, H64) < 1e-6 and frob_rel(G32, G64) < 1e-6
    # diagonal of the mixed H is the FP64 rigid-body diagonal rounded to FP32
    # (its far row sum uses a Newton reciprocal: within one FP32 ulp)
    d32, d64 = np.diag(H32).astype(np.float64), np.diag(H64)
    assert np.all(np.abs(d32 - d64) <= np.spacing(np.abs(d64).astype(np.float32)))
    B.solve(dad, mode="mixed", tol=1e-12)
    eT = np.linalg.norm(dad.T - T64) / np.linalg.norm(T64)
    eq = np.linalg.norm(dad.q - q64) / max(np.linalg.norm(q64), 1e-300)
    print(f"  mixed vs fp64: T {eT:.3e} q {eq:.3e} iters {dad.solve_info['iters']}")
    assert eT <= 1e-10 and eq <= 1e-10


def test_reference_square_T_equals_x():
    # test/laplace.jl:17-24: quadrado(ndiv=10), npg=12, reference BCs, rel_error < 0.05
    dad = square(ndiv=10, bcs="reference")
    B.assemble(dad, 12)
    B.solve(dad)
    assert B.rel_error(dad) < 0.05


def _perm_of(piv):
    perm = np.arange(len(piv))
    for i in range(len(piv)):
        perm[i], perm[piv[i]] = perm[piv[i]], perm[i]
    return perm


def _check_lu(A, LU, piv, dtype, lbound):
    """P A = L U within FP backward error, |L| <= lbound, and the factors solve A x = b."""
    n = A.shape[0]
    assert np.all((piv >= np.arange(n)) & (piv < n))
    L = np.tril(LU.astype(np.float64), -1) + np.eye(n)
    U = np.triu(LU.astype(np.float64))
    eps = np.finfo(dtype).eps
    rec = np.linalg.norm(A.astype(np.float64)[_perm_of(piv)] - L @ U) / np.linalg.norm(A)
    lmax = np.abs(np.tril(LU, -1)).max() if n > 1 else 0.0
    assert rec < 50 * n ** 0.5 * eps, rec
    assert lmax <= lbound, lmax
    rng = np.random.default_rng(n + 1)
    b = rng.standard_normal(n)
    x = B.lu_solve(LU, piv, b)
    assert np.linalg.norm(A.astype(np.float64) @ x - b) / np.linalg.norm(b) < 1e3 * n * eps
    return rec, lmax


@pytest.mark.parametrize("dtype,pivoting", [(np.float64, "partial"), (np.float32, "partial"),
                                            (np.float32, "tournament")])
# 6149: trailing updates large enough for the 256-wide GEMM tiles (>= 2 tiles per CU)
@pytest.mark.parametrize("n", [1, 7, 130, 517, 1000, 6149])
def test_lu_matches_lapack(n, dtype, pivoting):
    import scipy.linalg as sl
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)).astype(dtype)
    LU, piv = B.lu_factor(A, pivoting=pivoting)
    LUr, pivr = sl.lu_factor(A.astype(np.float64))
    # partial pivoting: |L| <= 1 (LAPACK's pivots, ties aside); tournament
    # pivoting bounds |L| less tightly (a few units in practice)
    rec, lmax = _check_lu(A, LU, piv, dtype, 1.0 + 1e-6 if pivoting == "partial" else 16.0)
    print(f"\n[lu n={n} {dtype.__name__} {pivoting}] |PA-LU|/|A| = {rec:.2e} max|L| {lmax:.2f} "
          f"pivots==LAPACK {np.array_equal(piv, pivr)}")


@pytest.mark.parametrize("n", [100, 300, 512])
def test_tournament_single_leaf_is_gepp(n):
    """Panels of at most 512 rows are one leaf: tournament pivoting is then GEPP,
    with the same arithmetic as the partial-pivoting panel, so pivots and
    factors are identical."""
    rng = np.random.default_rng(7 * n)
    A = rng.standard_normal((n, n)).astype(np.float32)
    ctx = B.HipContext(0)
    LUt, pivt = B.lu_factor(A, ctx=ctx, pivoting="tournament")
    LUp, pivp = B.lu_factor(A, ctx=ctx, pivoting="partial")
    assert np.array_equal(pivt, pivp)
    assert np.array_equal(LUt, LUp)


@pytest.mark.parametrize("n", [2600, 8200])
def test_tournament_tree(n):
    """Several tree levels (2600 rows: 6 leaves, 2 levels; 8200: 17 leaves, 3
    levels). The first panel's pivot rows are checked against a CPU replay of
    the tournament (GEPP on each leaf's original rows, then on the stacked
    candidates), and the factorization as a whole like LAPACK's."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)).astype(np.float32)
    LU, piv = B.lu_factor(A, pivoting="tournament")
    rec, lmax = _check_lu(A, LU, piv, np.float32, 16.0)
    print(f"\n[tslu n={n}] |PA-LU|/|A| = {rec:.2e} max|L| {lmax:.2f}")
    # CPU replay of the first panel's tournament in FP64 (the choices are
    # decided by clear margins, so FP32 vs FP64 rounding does not move them)
    P = A[:, :128].astype(np.float64)

    def gepp_rows(rows):
        M = P[rows].copy()
        chosen, alive = [], np.ones(len(rows), bool)
        for c in range(128):
            a = np.where(alive, np.abs(M[:, c]), -1.0)
            t = int(np.argmax(a))
            chosen.append(rows[t])
            alive[t] = False
            if M[t, c] != 0:
                l = M[:, c] / M[t, c]
                l[~alive] = 0.0
                M -= np.outer(l, M[t])
        return chosen

    sets = [gepp_rows(list(range(s, min(s + 512, n)))) for s in range(0, n, 512)]
    while len(sets) > 1:
        sets = [gepp_rows(sum(sets[i:i + 4], [])) for i in range(0, len(sets), 4)]
    winners = sets[0]
    perm = _perm_of(piv)
    assert list(perm[:128]) == winners


@pytest.mark.parametrize("n", [6149])
def test_lu_split_gemm_matches_fp32_gemm(n):
    """The FP32 LU with its large trailing updates on the bf16 matrix cores
    (each operand split exactly into three bf16 parts, six partial products)
    against the same LU with FP32 matrix-core updates, on one matrix: the split
    drops terms within 2^-23 (1 + 2^-9) |a b| per product, so the backward error stays at
    the FP32 level. A lost or doubled partial product (error 2^-16 |a b| per
    product or more) would raise it many times over."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((n, n)).astype(np.float32)
    ctx = B.HipContext(0)
    rec = {}
    for mode in ["fp32", "split_bf16"]:
        B.set_lu_gemm(ctx, mode)
        LU, piv = B.lu_factor(A, ctx=ctx)
        rec[mode], _ = _check_lu(A, LU, piv, np.float32, 1.0 + 1e-6)
    print(f"\n[lu n={n}] |PA-LU|/|A| fp32 GEMM {rec['fp32']:.3e}, split-bf16 {rec['split_bf16']:.3e}")
    assert rec["split_bf16"] < 1.25 * rec["fp32"], rec
