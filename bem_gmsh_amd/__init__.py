"""MI355X-native dense 2-D Laplace BEM path of BEM_gmsh.

    dad = format2d(quadrado(ndiv=20), Laplace(1.0))
    attach_analytical(dad, ana_laplace_linear())
    assemble(dad, 20)          # H, G on the GPU (libbemhip.so)
    solve(dad)                 # FP64 LU on the GPU -> dad.T, dad.q
    rel_error(dad)
"""

from .problem import BEMData, Elasticity, Element, Laplace, n_dof
from .mesh import (Mesh2D, SplitMix64, ellipse_mesh, format2d, formatdata, move_geometric_node,
                   n_eff, parse_bc_pairs, quadrado, quadrado_elasticity, read_msh, star_mesh,
                   star_param_jacobian, star_params, write_msh)
from .analytical import (AnalyticalSolution, ana_elasticity_patch, ana_exp_cos, ana_laplace_linear,
                         ana_quadratic_mixed, analytical, analytical_flux, apply_analytical_bc,
                         apply_bc_from_physical, attach_analytical, rel_error, rel_error_flux)
from .api import (ArgumentError, HipContext, adjoint, apply_bc, apply_operator, assemble,
                  get_HG, lu_factor, lu_solve, near_pairs, set_lu_gemm, set_lu_pivoting, solve, solve_Hmat,
                  split_sol, unique_node_gradient)
from .autodiff import FluxMisfit, objective, rrule_flux_misfit
from .interior import (gera_p_in, interior_laplace, internal_grid, internal_grid_, internal_layer,
                       internal_layer_, point_in_domain, set_internal_nodes)

__all__ = [
    "BEMData", "Elasticity", "Element", "Laplace", "n_dof", "Mesh2D", "SplitMix64",
    "ellipse_mesh", "format2d", "formatdata", "move_geometric_node", "n_eff", "parse_bc_pairs",
    "quadrado", "quadrado_elasticity", "read_msh", "star_mesh",
    "star_param_jacobian", "star_params", "write_msh",
    "AnalyticalSolution", "ana_elasticity_patch", "ana_exp_cos", "ana_laplace_linear", "ana_quadratic_mixed",
    "analytical", "analytical_flux", "apply_analytical_bc", "apply_bc_from_physical",
    "attach_analytical", "rel_error", "rel_error_flux", "ArgumentError", "HipContext",
    "adjoint", "apply_bc", "apply_operator", "assemble", "get_HG", "lu_factor", "lu_solve",
    "near_pairs", "set_lu_gemm", "set_lu_pivoting", "solve", "solve_Hmat", "split_sol", "unique_node_gradient", "gera_p_in",
    "interior_laplace", "internal_grid", "internal_grid_", "internal_layer", "internal_layer_",
    "point_in_domain", "set_internal_nodes", "FluxMisfit", "objective", "rrule_flux_misfit",
]
