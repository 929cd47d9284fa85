"""Size sweep: 64 seeded stars with N ~ U[2048, 16384] rounded to the nearest
multiple of 3, otherwise the N = 32769 star workload (quadratic elements,
+-10% angular widths, T = e^x cos y, Dirichlet/Neumann arcs of 64 elements,
npg = 16).

Every seed: the mixed pipeline (FP32 far field, FP32 LU, FP64 refinement)
against the FP64 pipeline on the GPU, per field (T, q) within 1e-10, the
relative accuracy the benchmark metric names. Every 8th seed also: the FP64
solve against LAPACK's solve of the oracle's applyBC of the same (GPU) H, G
(both with normwise backward error <= N eps; forward difference within max(1e-10,
4 kappa_1 eps), kappa_1 from LAPACK's dgecon), and two blocks
of 24 assembled rows against the oracle's rows (off-diagonal entries, relative
to the row's scale; the pairs whose FP64 values move by up to 3e-8 under a
1-ulp input shift set the 1e-7 bar).
"""

import numpy as np
import pytest
import scipy.linalg as sl

import bem_gmsh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

NSEEDS = 64


def sweep_sizes():
    out = []
    for s in range(NSEEDS):
        rng = B.SplitMix64(1000 + s)
        N = int(round(rng.uniform(2048.0, 16384.0) / 3.0)) * 3
        out.append((s, N))
    return out


def star(ne, seed):
    d = B.format2d(B.star_mesh(ne, seed=seed), B.Laplace(1.0))
    B.apply_bc_from_physical(d, B.ana_exp_cos())
    return d


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_sweep_sizes_are_the_configured_range():
    sizes = [N for _, N in sweep_sizes()]
    assert len(sizes) == NSEEDS and all(N % 3 == 0 for N in sizes)
    assert min(sizes) >= 2046 and max(sizes) <= 16386
    assert max(sizes) - min(sizes) > 8000   # spread over the range


@pytest.mark.parametrize("block", range(8))
def test_sweep_mixed_vs_fp64_and_oracle(block):
    worst_T = worst_q = 0.0
    for s, N in sweep_sizes()[block * 8:(block + 1) * 8]:
        d = star(N // 3, seed=s + 1)
        assert d.nt == N
        B.assemble(d, 16, fetch=False)
        B.solve(d, mode="fp64")
        T64, q64, x64 = d.T.copy(), d.q.copy(), d.solve_info["x"].copy()
        B.assemble(d, 16, precision="mixed", fetch=False)
        B.solve(d, mode="mixed", tol=1e-12)
        eT, eq = _rel(d.T, T64), _rel(d.q, q64)
        worst_T, worst_q = max(worst_T, eT), max(worst_q, eq)
        print(f"\n[seed {s}] N={N}: mixed vs fp64 T {eT:.2e} q {eq:.2e}, refine "
              f"{d.solve_info['iters']}", flush=True)
        assert eT <= 1e-10 and eq <= 1e-10
        if s % 8 == 0:
            H, G = B.assemble(d, 16)          # FP64, fetched
            A, b = O.apply_bc(d, H, G)
            lu = sl.lu_factor(A)
            xr = sl.lu_solve(lu, b)
            anorm = np.abs(A).sum(axis=0).max()
            rcond, _ = sl.lapack.dgecon(lu[0], anorm, norm="1")
            B.solve(d, mode="fp64")
            xg = d.solve_info["x"]
            eo = _rel(xg, xr)
            # both backward stable; forward difference within the conditioning
            for x in (xg, xr):
                assert (np.abs(b - A @ x).max() <= N * np.finfo(float).eps
                        * np.abs(A).sum(axis=1).max() * np.abs(x).max())
            bar = max(1e-10, 4.0 * np.finfo(float).eps / rcond)
            er = 0.0
            for lo in (0, N // 2):            # two blocks of 24 rows
                Hr, Gr = O.assemble(d, 16, rows=(lo, lo + 24))
                for r in range(lo, lo + 24):
                    h = H[r].copy()
                    hr = Hr[r].copy()
                    h[r] = hr[r] = 0.0        # the diagonal needs the whole row
                    sc = max(np.abs(hr).max(), np.abs(Gr[r]).max())
                    er = max(er, np.abs(h - hr).max() / sc, np.abs(G[r] - Gr[r]).max() / sc)
                del Hr, Gr
            print(f"  oracle: LAPACK solve on the same H, G {eo:.2e} (bar {bar:.1e}, "
                  f"kappa_1 {1 / rcond:.1e}), sampled rows {er:.2e}", flush=True)
            assert eo <= bar
            assert er <= 1e-7
    print(f"\n[block {block}] worst mixed vs fp64: T {worst_T:.2e} q {worst_q:.2e}")
