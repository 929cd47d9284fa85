"""Shared builders for the tests (configurations of the path at test sizes)."""

import numpy as np

import bem_gmsh_amd as B


def square(ndiv=20, ordem=1, pontointerno=True, bcs="config1"):
    """Unit square of quadrado(); config-1 BCs: Dirichlet T=x on x=0,1, Neumann on y=0,1."""
    dad = B.format2d(B.quadrado(ndiv=ndiv, ordem=ordem), B.Laplace(1.0),
                     pontointerno=pontointerno)
    ana = B.ana_laplace_linear()
    if bcs == "config1":
        x = dad.Nodes
        neu = np.nonzero((np.abs(x[:, 1]) < 1e-12) | (np.abs(x[:, 1] - 1.0) < 1e-12))[0]
        B.apply_analytical_bc(dad, ana, neu)
    else:
        B.attach_analytical(dad, ana)
    return dad


def ellipse(ne=200):
    dad = B.format2d(B.ellipse_mesh(ne), B.Laplace(1.0))
    B.apply_bc_from_physical(dad, B.ana_quadratic_mixed(0.3))
    return dad


def star(ne=300, seed=1):
    dad = B.format2d(B.star_mesh(ne, seed=seed), B.Laplace(1.0))
    B.apply_bc_from_physical(dad, B.ana_exp_cos())
    return dad


def entry_err(A, Ref):
    """Per-entry relative error with an absolute floor of 1e-3 * max|row of Ref|
    (entries that cancel to ~0 are compared against their row's scale)."""
    A = np.asarray(A, dtype=np.float64)
    Ref = np.asarray(Ref, dtype=np.float64)
    rowmax = np.abs(Ref).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(Ref), 1e-3 * rowmax)
    den = np.where(den > 0, den, 1.0)
    return float(np.max(np.abs(A - Ref) / den))


def frob_rel(A, Ref):
    return float(np.linalg.norm(np.asarray(A, np.float64) - Ref) / np.linalg.norm(Ref))
