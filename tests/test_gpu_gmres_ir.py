"""GMRES-IR: the FP32 factors as a GMRES preconditioner for the FP64
correction equations, the mixed solve's fallback when classical refinement
stops contracting (the reference's solve falls through to an iterative method
when its factorization fails, src/Core/Solver.jl:39-50).

The ill-conditioned system: a circle of radius a0 + 1e-10 with Dirichlet data
everywhere (A = -G). The 2-D Laplace single layer has a degenerate scale: on a
circle of radius a its constant mode has eigenvalue ~ -a log a, zero at a = 1;
the discretization (200 quadratic elements, npg = 16) moves the zero to
a0 = 0.99991088895... (located on the CPU with the oracle's matrices, by
bisection on the eigenvalue nearest 0), and near it kappa(A) ~ 0.52 / (a - a0):
kappa ~ 5e9 here (checked with numpy's SVD of the oracle's A before the GPU
runs). Classical refinement with FP32 factors then contracts by anything from
0.1 to 0.9 per step depending on the FP32 rounding pattern; GMRES-IR converges
to the FP64 noise floor (kappa eps64 ~ 1e-6) in a few steps.
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

A0 = 0.9999108889531998      # discrete degenerate radius, 200 quadratic elements, npg 16
DELTA = 1e-10
EPS = np.finfo(float).eps


def _circle(radius, ne=200):
    d = B.format2d(B.ellipse_mesh(ne, a=radius, b=radius), B.Laplace(1.0))
    B.apply_analytical_bc(d, B.ana_exp_cos())           # Dirichlet everywhere
    return d


@pytest.fixture(scope="module")
def degenerate():
    d = _circle(A0 + DELTA)
    H, G = O.assemble(d, 16)
    A, b = O.apply_bc(d, H, G)
    s = np.linalg.svd(A, compute_uv=False)
    kappa = s[0] / s[-1]
    print(f"\n[degenerate circle] N={d.n}, kappa_2(A) = {kappa:.2e}")
    assert 1e9 <= kappa <= 2e10
    B.assemble(d, 16, fetch=False)
    B.solve(d, mode="fp64")
    return d, A, b, kappa, np.linalg.solve(A, b), d.solve_info["x"].copy()


def _check(d, A, b, kappa, xr, x64, xm):
    """Forward error against the FP64 LU (same operator) and numpy's solve of the
    oracle's A (whose near entries differ from the GPU's by up to ~1e-9 here,
    the pairs' own rounding sensitivity): within 20 kappa eps. Backward error
    on the library's own FP64 operator: ||b - A x||_inf (the solve's final
    FP64 residual) <= n eps ||A||_inf ||x||_inf."""
    bar = 20.0 * kappa * EPS
    e64 = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    eor = np.linalg.norm(xm - xr) / np.linalg.norm(xr)
    back = (d.solve_info["relres"] * np.abs(b).max()
            / (np.abs(A).sum(axis=1).max() * np.abs(xm).max()))
    print(f"  vs FP64 LU {e64:.2e}, vs numpy on the oracle's A {eor:.2e} (bar 20 kappa eps "
          f"{bar:.1e}); backward error {back:.1e}")
    assert e64 <= bar and eor <= bar and back <= d.n * EPS


def test_gmres_ir_mode(degenerate):
    d, A, b, kappa, xr, x64 = degenerate
    B.assemble(d, 16, precision="mixed", fetch=False)
    B.solve(d, mode="mixed_gmres", tol=1e-10)
    st = d.hip.counters()["stats"]
    steps = int(st["refine_gmres_outer"])
    print(f"  GMRES-IR: {steps} steps, {int(st['refine_gmres_products'])} preconditioned "
          "products; ||d||/||x|| (products) "
          + " ".join(f"{st[f'refine_gmres_dx_{k:02d}']:.1e} ({int(st[f'refine_gmres_its_{k:02d}'])})"
                     for k in range(steps)))
    assert st["refine_gmres_ir"] == 1.0 and steps <= 6
    _check(d, A, b, kappa, xr, x64, d.solve_info["x"])


def test_mixed_switches_when_refinement_stalls(degenerate):
    """Mixed mode on the same system: the library falls back exactly when the
    classical iteration with its own FP32 matrix contracts by more than ~1/2,
    which the CPU predicts from that matrix: rho = spectral radius of
    I - A32^-1 A64 (numpy, FP64 arithmetic on the GPU's stored values)."""
    d, A, b, kappa, xr, x64 = degenerate
    B.assemble(d, 16, precision="mixed")        # fetches the FP32 H, G
    A32 = -np.asarray(d.G, dtype=np.float64)    # every column Dirichlet
    B.solve(d, mode="mixed", tol=1e-10, maxit=30)
    st = d.hip.counters()["stats"]
    rho = np.abs(np.linalg.eigvals(np.eye(d.n) - np.linalg.solve(A32, A))).max()
    dx = [st[k] for k in sorted(st) if k.startswith("refine_dx_")]
    used = st.get("refine_gmres_ir") == 1.0
    print(f"\n  predicted classical contraction rho = {rho:.2f}; corrections "
          + " ".join(f"{v:.1e}" for v in dx) + f"; GMRES-IR {'used' if used else 'not used'}")
    if rho > 0.7:
        assert used
    if rho < 0.3:
        assert not used
    _check(d, A, b, kappa, xr, x64, d.solve_info["x"])


def test_gmres_ir_on_a_well_conditioned_star():
    d = B.format2d(B.star_mesh(300, seed=2), B.Laplace(1.0))
    B.apply_bc_from_physical(d, B.ana_exp_cos())
    B.assemble(d, 16, fetch=False)
    B.solve(d, mode="fp64")
    x64 = d.solve_info["x"].copy()
    B.assemble(d, 16, precision="mixed", fetch=False)
    B.solve(d, mode="mixed", tol=1e-10)
    st = d.hip.counters()["stats"]
    assert "refine_gmres_ir" not in st and d.solve_info["iters"] <= 5
    B.solve(d, mode="mixed_gmres", tol=1e-12)
    st = d.hip.counters()["stats"]
    err = np.linalg.norm(d.solve_info["x"] - x64) / np.linalg.norm(x64)
    print(f"\n  star N={d.n}: GMRES-IR {int(st['refine_gmres_outer'])} steps, "
          f"{int(st['refine_gmres_products'])} products, vs FP64 {err:.1e}")
    assert err <= 1e-10 and st["refine_gmres_outer"] <= 3


def test_adjoint_after_gmres_ir(degenerate):
    """The adjoint of the ill-conditioned system after the GMRES-IR solve: its
    transposed solve refines the same way (switching to GMRES-IR with the
    transposed FP32 factors) and gives the FP64 adjoint's gradient to the
    conditioning's accuracy."""
    d, A, b, kappa, xr, x64 = degenerate
    qt = B.analytical_flux(d)[:d.n]
    B.assemble(d, 16, fetch=False)
    B.solve(d, mode="fp64")
    J64, g64 = B.adjoint(d, qt)
    B.assemble(d, 16, precision="mixed", fetch=False)
    B.solve(d, mode="mixed_gmres", tol=1e-10)
    Jm, gm = B.adjoint(d, qt)
    st = d.hip.counters()["stats"]
    err = float(np.abs(gm - g64).max() / np.abs(g64).max())
    print(f"\n  adjoint: J {Jm:.6e} vs {J64:.6e}, gradient vs FP64 {err:.1e} "
          f"(transposed GMRES-IR {'used' if st.get('adjoint_gmres_ir') else 'not used'})")
    # J within 20 kappa eps; the gradient combines two kappa-amplified
    # quantities (the solution and lambda): 100 kappa eps
    assert abs(Jm - J64) <= 20 * kappa * EPS * abs(J64) and err <= 100 * kappa * EPS
