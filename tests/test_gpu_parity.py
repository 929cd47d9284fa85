"""GPU parity of the HIP path against the FP64 oracle (through the C ABI).

Bars (north star): FP64 H and G per entry within 1e-12 relative (entries that
cancel to ~0 are measured against 1e-3 of their row's largest entry), T and q
within 1e-10 relative, rel_error unchanged; FP32 far field within 1e-6
relative Frobenius of FP64 (test/laplace.jl:453-455).
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O
from helpers import ellipse, entry_err, frob_rel, square, star

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
    eh, eg = entry_err(H, Hr), entry_err(G, Gr)
    print(f"  H entry err {eh:.3e} frob {frob_rel(H, Hr):.3e} | "
          f"G entry err {eg:.3e} frob {frob_rel(G, Gr):.3e}", flush=True)
    assert eh <= 1e-12 and eg <= 1e-12
    # rigid-body diagonal: rows of H sum to zero
    assert np.abs(H.sum(axis=1)).max() <= 1e-12 * np.abs(H).max() * dad.nt ** 0.5


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
def test_fp64_solve_matches_oracle(case):
    dad = CASES[case]()
    npg = 20 if case == "square_p1" else 16
    H, G = B.assemble(dad, npg)
    Hr, Gr = O.assemble(dad, npg)
    Tr, qr, xr = O.solve(dad, Hr, Gr)
    B.solve(dad, mode="fp64")
    eT = np.linalg.norm(dad.T - Tr) / np.linalg.norm(Tr)
    eq = np.linalg.norm(dad.q - qr) / max(np.linalg.norm(qr), 1e-300)
    print(f"\n[{case}] T err {eT:.3e} q err {eq:.3e} relres {dad.solve_info['relres']:.3e}")
    assert eT <= 1e-10 and eq <= 1e-10
    ref = B.BEMData.__new__(B.BEMData)
    err_gpu = B.rel_error(dad)
    dad.set_cache(T=Tr, q=qr)
    err_ref = B.rel_error(dad)
    assert abs(err_gpu - err_ref) <= 1e-10 * max(err_ref, 1e-16) + 1e-14


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
    # diagonal of the mixed H is the FP64 rigid-body diagonal
    assert np.allclose(np.diag(H32), np.diag(H64).astype(np.float32), rtol=0, atol=0)
    B.solve(dad, mode="mixed", tol=1e-13)
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


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 7, 130, 517, 1000])
def test_lu_matches_lapack(n, dtype):
    import scipy.linalg as sl
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)).astype(dtype)
    LU, piv = B.lu_factor(A)
    LUr, pivr = sl.lu_factor(A.astype(np.float64))
    # same factorization up to rounding (pivot ties aside): reconstruct P A = L U
    L = np.tril(LU.astype(np.float64), -1) + np.eye(n)
    U = np.triu(LU.astype(np.float64))
    perm = np.arange(n)
    for i in range(n):
        perm[i], perm[piv[i]] = perm[piv[i]], perm[i]
    eps = np.finfo(dtype).eps
    rec = np.linalg.norm(A.astype(np.float64)[perm] - L @ U) / np.linalg.norm(A)
    print(f"\n[lu n={n} {dtype.__name__}] |PA-LU|/|A| = {rec:.2e}  pivots==LAPACK "
          f"{np.array_equal(piv, pivr)}")
    assert rec < 50 * n ** 0.5 * eps
    assert np.abs(np.tril(LU, -1)).max() <= 1.0 + 1e-6
    b = rng.standard_normal(n)
    x = B.lu_solve(LU, piv, b)
    xr = np.linalg.solve(A.astype(np.float64), b)
    assert np.linalg.norm(A.astype(np.float64) @ x - b) / np.linalg.norm(b) < 1e3 * n * eps
