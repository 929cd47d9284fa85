"""The adjoint's reverse-mode binding (bem_gmsh_amd.autodiff): the
ChainRules-style rrule of the flux misfit and its torch.autograd.Function
form, through the C ABI, against central differences of the FP64 oracle's
path (mesh -> format2d -> assemble -> LAPACK solve -> J).

The pullback is taken with respect to mesh.lines, every element's own copy of
its nodes (Gmsh order): moving one copy of a shared end node moves that
element only, which the discontinuous elements allow, so each coordinate is
checked on its own.
"""

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O
from bem_gmsh_amd.autodiff import _with_lines

pytestmark = pytest.mark.gpu

NPG = 12


def _case(ne=60, seed=5):
    mesh = B.star_mesh(ne, seed=seed)
    base = B.format2d(mesh, B.Laplace(1.0))
    B.apply_bc_from_physical(base, B.ana_exp_cos())
    mp = B.star_mesh(ne, seed=seed, params=B.star_params(seed, perturb=True))
    dp = B.format2d(mp, B.Laplace(1.0))
    B.apply_bc_from_physical(dp, B.ana_exp_cos())
    B.assemble(dp, NPG, fetch=False)
    B.solve(dp, mode="fp64")
    return mesh, base, dp.q.copy()


def _J_oracle(mesh, base, qt):
    d = B.format2d(mesh, B.Laplace(1.0))
    d.BC, d.BV = base.BC.copy(), base.BV.copy()
    H, G = O.assemble(d, NPG)
    T, q, _ = O.solve(d, H, G)
    d.set_cache(T=T, q=q)
    return B.objective(d, qt)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rrule_pullback_matches_oracle_differences(precision):
    mesh, base, qt = _case()
    J, pb = B.rrule_flux_misfit(mesh, qt, base, npg=NPG, precision=precision)
    assert abs(J - _J_oracle(mesh, base, qt)) <= 1e-10 * abs(J)
    g = pb(1.0)
    assert g.shape == mesh.lines.shape
    assert np.array_equal(pb(2.5), 2.5 * g)
    rng = np.random.default_rng(4)
    scale = np.abs(g).max()
    worst = 0.0
    h = 1e-6
    for flat in rng.choice(g.size, 12, replace=False):
        e, k, a = np.unravel_index(flat, g.shape)
        L = np.array(mesh.lines, dtype=np.float64)
        Lp, Lm = L.copy(), L.copy()
        Lp[e, k, a] += h
        Lm[e, k, a] -= h
        fd = (_J_oracle(_with_lines(mesh, Lp), base, qt)
              - _J_oracle(_with_lines(mesh, Lm), base, qt)) / (2 * h)
        err = abs(fd - g[e, k, a]) / scale
        worst = max(worst, err)
        print(f"  [{precision}] lines[{e},{k},{a}]: pullback {g[e, k, a]: .10e} oracle fd "
              f"{fd: .10e} err/max {err:.1e}")
    print(f"  worst {worst:.1e}")
    assert worst <= 1e-6


def test_torch_function_matches_rrule():
    torch = pytest.importorskip("torch")
    mesh, base, qt = _case(ne=40, seed=3)
    J, pb = B.rrule_flux_misfit(mesh, qt, base, npg=NPG)
    lines = torch.tensor(np.array(mesh.lines, dtype=np.float64), requires_grad=True)
    Jt = B.FluxMisfit().apply(lines, mesh, qt, base, NPG, "fp64")
    (3.0 * Jt).backward()
    assert abs(float(Jt) - J) <= 1e-12 * abs(J)
    assert np.allclose(lines.grad.numpy(), 3.0 * pb(1.0), rtol=0, atol=1e-14 * np.abs(pb(1.0)).max())
