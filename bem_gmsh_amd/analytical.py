"""Analytic fields, BC imposition and error metrics.

Mirrors src/Core/Analytical.jl: ``AnalyticalSolution`` (:8-30),
``attach_analytical!`` (:73-76), ``apply_analytical_bc!`` (:84-91, :113-130),
``ana_laplace_linear`` (:171-182), ``ana_elasticity_patch`` (:546-596),
``rel_error`` (:617-630) and ``rel_error_flux`` (:667-679). Flux convention
``q = -k dT/dn``. For elasticity every field is per dof: node j component b at
``2 j + b``; ``u`` returns (m, 2) displacements, ``q`` (m, 2) tractions.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Iterable, Optional

import numpy as np

from .problem import BEMData, Elasticity


@dataclass
class AnalyticalSolution:
    name: str
    u: Callable[[np.ndarray], np.ndarray]                  # (m,2) -> (m,)
    q: Optional[Callable[[np.ndarray, np.ndarray], np.ndarray]] = None  # (x, n) -> (m,)
    description: str = ""

    def __call__(self, x):
        return self.u(np.atleast_2d(np.asarray(x, dtype=np.float64)))


def ana_laplace_linear(direction=(1.0, 0.0), k: float = 1.0) -> AnalyticalSolution:
    """T = d.x, q = -k d.n (src/Core/Analytical.jl:171-182)."""
    d = np.asarray(direction, dtype=np.float64)
    return AnalyticalSolution(
        "laplace_linear",
        lambda x: x[:, 0] * d[0] + x[:, 1] * d[1],
        lambda x, n: -k * (d[0] * n[:, 0] + d[1] * n[:, 1]),
        "T = d.x (harmonic), q = -k dT/dn")


def ana_quadratic_mixed(c: float = 0.3, k: float = 1.0) -> AnalyticalSolution:
    """T = x^2 - y^2 + c x y (harmonic), grad T = (2x + c y, -2y + c x)."""
    def u(x):
        return x[:, 0] ** 2 - x[:, 1] ** 2 + c * x[:, 0] * x[:, 1]

    def q(x, n):
        gx = 2.0 * x[:, 0] + c * x[:, 1]
        gy = -2.0 * x[:, 1] + c * x[:, 0]
        return -k * (gx * n[:, 0] + gy * n[:, 1])
    return AnalyticalSolution("quadratic_mixed", u, q, f"T = x^2 - y^2 + {c} x y")


def ana_exp_cos(k: float = 1.0) -> AnalyticalSolution:
    """T = e^x cos y (harmonic), grad T = (e^x cos y, -e^x sin y)."""
    def u(x):
        return np.exp(x[:, 0]) * np.cos(x[:, 1])

    def q(x, n):
        ex = np.exp(x[:, 0])
        return -k * (ex * np.cos(x[:, 1]) * n[:, 0] - ex * np.sin(x[:, 1]) * n[:, 1])
    return AnalyticalSolution("exp_cos", u, q, "T = e^x cos y")


def ana_elasticity_patch(E: float = 1.0, nu: float = 0.3, exx: float = 0.01, eyy: float = 0.0,
                         exy: float = 0.0) -> AnalyticalSolution:
    """Uniform strain patch u = eps . x, t = sigma n, plane strain with the
    Lame pair of the material nu (src/Core/Analytical.jl:546-596, dim = 2)."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    sxx = lam * (exx + eyy) + 2.0 * mu * exx
    syy = lam * (exx + eyy) + 2.0 * mu * eyy
    sxy = 2.0 * mu * exy

    def u(x):
        return np.stack([exx * x[:, 0] + exy * x[:, 1], exy * x[:, 0] + eyy * x[:, 1]], axis=1)

    def q(x, n):
        return np.stack([sxx * n[:, 0] + sxy * n[:, 1], sxy * n[:, 0] + syy * n[:, 1]], axis=1)
    return AnalyticalSolution("elasticity_patch", u, q, "Uniform strain patch (plane strain)")


def _is_elastic(dad: BEMData) -> bool:
    return isinstance(dad.properties, Elasticity)


def attach_analytical(dad: BEMData, ana: AnalyticalSolution) -> BEMData:
    dad.set_cache(analytical=ana)
    return dad


def apply_analytical_bc(dad: BEMData, ana: AnalyticalSolution,
                        neumann_nodes: Optional[Iterable[int]] = None) -> BEMData:
    """Dirichlet from ``ana.u`` everywhere except ``neumann_nodes`` (flux ``ana.q``).
    Elasticity: every dof of a node takes the node's type (:93-111, :132-158)."""
    x = dad.Nodes
    uval = ana.u(x)
    if _is_elastic(dad):
        mask = np.zeros(dad.n, dtype=bool)
        if neumann_nodes is not None and ana.q is not None:
            mask[np.fromiter(neumann_nodes, dtype=np.int64)] = True
            qval = ana.q(x, dad.Normal)
        else:
            qval = np.zeros_like(uval)
        m2 = np.repeat(mask, 2)
        dad.BC[:] = np.where(m2, 1, 0)
        dad.BV[:] = np.where(m2, qval.reshape(-1), uval.reshape(-1))
        return attach_analytical(dad, ana)
    if neumann_nodes is None or ana.q is None:
        dad.BC[:] = 0
        dad.BV[:] = uval
    else:
        mask = np.zeros(dad.n, dtype=bool)
        idx = np.fromiter(neumann_nodes, dtype=np.int64)
        mask[idx] = True
        qval = ana.q(x, dad.Normal)
        dad.BC[:] = np.where(mask, 1, 0)
        dad.BV[:] = np.where(mask, qval, uval)
    return attach_analytical(dad, ana)


def apply_bc_from_physical(dad: BEMData, ana: AnalyticalSolution) -> BEMData:
    """Keep each dof's BC type from the mesh, take its value from the analytic field."""
    if _is_elastic(dad):
        bc = dad.BC.copy()
        uval = ana.u(dad.Nodes).reshape(-1)
        qval = ana.q(dad.Nodes, dad.Normal).reshape(-1)
        dad.BV[:] = np.where(bc == 1, qval, uval)
        return attach_analytical(dad, ana)
    neu = np.nonzero(dad.BC == 1)[0]
    return apply_analytical_bc(dad, ana, neu)


def analytical(dad: BEMData) -> np.ndarray:
    if not dad.has_cache("analytical"):
        raise RuntimeError("No analytical solution attached. Use attach_analytical(dad, ana).")
    return np.asarray(dad.analytical.u(dad.collocation)).reshape(-1)


def analytical_flux(dad: BEMData) -> np.ndarray:
    ana = dad.analytical
    if ana.q is None:
        raise RuntimeError("analytical solution has no flux `q`")
    return np.asarray(ana.q(dad.Nodes, dad.Normal)).reshape(-1)


def rel_error(dad: BEMData) -> float:
    """||T - T_ana||_2 / ||T_ana||_2 over the collocation points."""
    Tana = analytical(dad)
    Tnum = dad.T
    n = min(Tnum.size, Tana.size)
    num = np.linalg.norm(Tnum[:n] - Tana[:n])
    den = np.linalg.norm(Tana[:n])
    return float(num / den) if den > 0 else float(num)


def rel_error_flux(dad: BEMData) -> float:
    qana = analytical_flux(dad)
    qnum = dad.traction if _is_elastic(dad) else dad.q
    n = min(qnum.size, qana.size)
    num = np.linalg.norm(qnum[:n] - qana[:n])
    den = np.linalg.norm(qana[:n])
    return float(num / den) if den > 0 else float(num)
