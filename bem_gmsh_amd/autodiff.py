"""Reverse-mode binding of the adjoint shape gradient.

The reference has no adjoint; its differentiable stack is ChainRulesCore
(ForwardDiff is a dependency, Project.toml:17). These are the two forms a
caller of this library binds it in:

* ``rrule_flux_misfit(mesh, q_target, bc_from, ...)`` -> ``(J, pullback)``, the
  ChainRulesCore ``rrule`` convention: the forward pass is the reference's
  path ``format2d`` (src/Core/Input.jl:136-197) -> ``assemble!``
  (src/Core/Assembly_full.jl:107-118) -> ``solve`` (src/Core/Solver.jl:24-64),
  on the GPU, followed by the flux misfit

      J = 1/2 sum_j w_j (q_j - q_target_j)^2,   w_j = J_j w_k (or given);

  ``pullback(Jbar)`` returns ``Jbar * dJ/d(mesh.lines)`` (the mesh's element
  nodes, Gmsh order; a node shared by two elements has an entry per element
  copy) from one ``bemhip_adjoint`` call: a transposed solve with the factors
  of the forward solve, still resident, and one gradient pass.
* ``FluxMisfit``, the same as a ``torch.autograd.Function`` of the element
  nodes (float64 CPU tensor, shaped like ``mesh.lines``), so torch optimizers
  can drive the boundary shape.

Boundary-condition types and values are taken from ``bc_from`` and held fixed
(they are data, not functions of the shape), as ``bemhip_adjoint`` assumes.
"""

from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np

from .api import adjoint, assemble, solve
from .mesh import Mesh2D, _gmsh_line_reorder, format2d
from .problem import BEMData, Laplace


def _with_lines(mesh: Mesh2D, lines: np.ndarray) -> Mesh2D:
    lines = np.ascontiguousarray(lines, dtype=np.float64)
    if lines.shape != np.shape(mesh.lines):
        raise ValueError(f"lines must have shape {np.shape(mesh.lines)}, got {lines.shape}")
    return Mesh2D(mesh.name, mesh.order, lines, list(mesh.physical), np.array(mesh.region),
                  [np.array(c) for c in mesh.cells], dict(mesh.extras))


def objective(dad: BEMData, q_target: np.ndarray, w: Optional[np.ndarray] = None) -> float:
    """J = 1/2 sum_j w_j (q_j - q_target_j)^2 of a solved problem (w defaults to
    the boundary quadrature weights J_j w_k)."""
    if w is None:
        w = dad.node_J() * np.tile(dad.elem_weight, dad.ne)
    return 0.5 * float(np.sum(w * (dad.q - q_target) ** 2))


def rrule_flux_misfit(mesh: Mesh2D, q_target: np.ndarray, bc_from: BEMData, npg: int = 20,
                      precision: str = "fp64", w: Optional[np.ndarray] = None,
                      properties=None) -> Tuple[float, Callable[[float], np.ndarray]]:
    """(J, pullback) of the flux misfit of the boundary mesh ``mesh`` (see the
    module docstring). ``precision`` 'fp64' (FP64 assembly and LU) or 'mixed'
    (FP32 far field and LU, FP64 refinement; the transposed solve is refined
    the same way). The returned pullback may be called any number of times
    while the problem it closes over is alive."""
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed'; got {precision!r}")
    dad = format2d(mesh, properties if properties is not None else Laplace(1.0))
    if dad.n != bc_from.n:
        raise ValueError("bc_from has another number of boundary nodes")
    dad.BC = np.array(bc_from.BC, copy=True)
    dad.BV = np.array(bc_from.BV, copy=True)
    qt = np.ascontiguousarray(q_target, dtype=np.float64)
    assemble(dad, npg, precision=precision, fetch=False)
    solve(dad, mode=precision, tol=1e-12)
    J, dgeo = adjoint(dad, qt, w)
    inv = np.argsort(np.asarray(_gmsh_line_reorder(dad.p)))   # parametric -> Gmsh order
    grad_lines = np.ascontiguousarray(dgeo[:, inv, :])

    def pullback(Jbar: float = 1.0) -> np.ndarray:
        return float(Jbar) * grad_lines

    return J, pullback


def FluxMisfit():
    """The ``torch.autograd.Function`` form (torch imported on first use):

        J = FluxMisfit().apply(lines, mesh, q_target, bc_from, npg, precision)

    with ``lines`` a float64 CPU tensor shaped like ``mesh.lines``; only
    ``lines`` receives a gradient."""
    import torch

    class _FluxMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, lines, mesh, q_target, bc_from, npg=20, precision="fp64"):
            m = _with_lines(mesh, lines.detach().cpu().numpy())
            J, pb = rrule_flux_misfit(m, q_target, bc_from, npg=npg, precision=precision)
            ctx.pullback = pb
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            g = torch.from_numpy(ctx.pullback(float(gJ)))
            return g, None, None, None, None, None

    return _FluxMisfit
