"""Problem description and the collocation container.

Mirrors the reference's data model for the 2-D path:
``Laplace`` (src/Core/Structures.jl:101-103), ``Elasticity``
(src/Core/Structures.jl:149-263), ``Element``
(src/Core/Structures.jl:317-324) and ``BEMdata`` (src/Core/Structures.jl:504-519)
with its cache forwarding (:522-553). Collocation points are boundary nodes
``0:n`` followed by internal nodes ``n:nt``; element ``e`` owns the contiguous
boundary nodes ``e*(p+1) : (e+1)*(p+1)`` (src/Core/Input.jl:170).

Arrays are stored structure-of-arrays in numpy (float64 / int64) so they can be
handed to the C ABI of ``libbemhip.so`` without conversion.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, Optional

import numpy as np


@dataclass
class Laplace:
    """Steady heat conduction, conductivity ``k``; flux ``q = -k dT/dn``."""

    k: float = 1.0


class Elasticity:
    """2-D isotropic linear elasticity (src/Core/Structures.jl:149-263).

    ``Elasticity(E, nu, rho; plane_strain=True)`` or ``plane_stress=True``.
    The Kelvin kernels use the effective Poisson ratio (plane stress maps to
    the equivalent plane-strain ``nu / (1 + nu)``, :260-263); the shear
    modulus is ``mu = E / (2 (1 + nu))`` with the material ``nu`` and the first
    Lame constant uses the effective ratio (``lame_constants``, :156-165).
    """

    def __init__(self, E: float = 1.0, nu: float = 0.3, rho: float = 1.0,
                 plane_strain: bool = True, plane_stress: bool = False):
        self.E = float(E)
        self.nu = float(nu)
        self.rho = float(rho)
        self.plane_strain = False if plane_stress else bool(plane_strain)

    def effective_nu(self) -> float:
        return self.nu if self.plane_strain else self.nu / (1.0 + self.nu)

    @property
    def mu(self) -> float:
        return self.E / (2.0 * (1.0 + self.nu))

    @property
    def lambda_(self) -> float:
        ne = self.effective_nu()
        den = (1.0 + ne) * (1.0 - 2.0 * ne)
        return float("inf") if abs(den) < np.finfo(float).eps else self.E * ne / den

    def __repr__(self) -> str:
        mode = "plane_strain" if self.plane_strain else "plane_stress"
        return f"Elasticity(E={self.E}, nu={self.nu}, rho={self.rho}, {mode})"


def n_dof(properties) -> int:
    """Unknowns per node: 1 (Laplace) or 2 (2-D elasticity)."""
    return 2 if isinstance(properties, Elasticity) else 1


@dataclass
class Element:
    """One boundary element (view onto the SoA arrays of a BEMData)."""

    index: np.ndarray      # global node ids (p+1,)
    Jacobian: np.ndarray   # |dx/dxi| at the collocation nodes (p+1,)
    Length: float
    Region: int
    geo: np.ndarray        # (p+1, 2) geometric nodes, xi = -1..1 equispaced


class BEMData:
    """Collocation container for a 2-D boundary-element problem.

    Attributes follow the reference names: ``collocation`` (nt, 2),
    ``Normal`` (n, 2), ``BC`` (dof*n,) with 0 = Dirichlet (unknown q, or the
    traction component) and 1 = Neumann (unknown T, or the displacement
    component), ``BV`` (dof*n,), ``n``, ``ni``, ``nt``, ``elem_weight``
    (p+1,) and the per-element arrays ``elem_J`` (ne, p+1), ``elem_L`` (ne,),
    ``elem_geo`` (ne, p+1, 2), ``elem_region`` (ne,).

    Results of ``assemble``/``solve`` live in ``cache`` and are forwarded as
    attributes (``dad.H``, ``dad.G``, ``dad.A``, ``dad.b``, ``dad.T``, ``dad.q``),
    raising an error when missing, like src/Core/Structures.jl:539-553.
    """

    _CACHE_KEYS = ("H", "G", "A", "B", "b", "T", "q", "u", "traction", "uint", "qsi", "w", "near_factor",
                   "analytical", "cells", "solve_info", "assembly_info", "gmres_stats")

    def __init__(self, name: str, p: int, collocation: np.ndarray, normal: np.ndarray,
                 elem_J: np.ndarray, elem_L: np.ndarray, elem_geo: np.ndarray,
                 elem_region: np.ndarray, elem_weight: np.ndarray, properties,
                 BC: np.ndarray, BV: np.ndarray, n: int, ni: int):
        object.__setattr__(self, "cache", {})
        self.name = name
        self.dimension = 2
        self.p = int(p)
        self.collocation = np.ascontiguousarray(collocation, dtype=np.float64)
        self.Normal = np.ascontiguousarray(normal, dtype=np.float64)
        self.elem_J = np.ascontiguousarray(elem_J, dtype=np.float64)
        self.elem_L = np.ascontiguousarray(elem_L, dtype=np.float64)
        self.elem_geo = np.ascontiguousarray(elem_geo, dtype=np.float64)
        self.elem_region = np.ascontiguousarray(elem_region, dtype=np.int64)
        self.elem_weight = np.ascontiguousarray(elem_weight, dtype=np.float64)
        self.properties = properties
        self.BC = np.ascontiguousarray(BC, dtype=np.int64)
        self.BV = np.ascontiguousarray(BV, dtype=np.float64)
        self.n = int(n)
        self.ni = int(ni)
        self.nt = int(n + ni)
        self.hip = None  # device context handle, created by assemble(method="hip")

    # -- reference-style views ------------------------------------------------
    @property
    def ne(self) -> int:
        return int(self.elem_L.shape[0])

    @property
    def Nodes(self) -> np.ndarray:
        return self.collocation[: self.n]

    @property
    def internalNodes(self) -> np.ndarray:
        return self.collocation[self.n:]

    @property
    def points(self) -> np.ndarray:
        return self.collocation

    @property
    def elements(self):
        nn = self.p + 1
        return [Element(index=np.arange(e * nn, (e + 1) * nn), Jacobian=self.elem_J[e],
                        Length=float(self.elem_L[e]), Region=int(self.elem_region[e]),
                        geo=self.elem_geo[e]) for e in range(self.ne)]

    def node_J(self) -> np.ndarray:
        """Jacobian at every boundary node, shape (n,)."""
        return self.elem_J.reshape(-1)

    def node_Jw(self) -> np.ndarray:
        """Far-field lumping weight J_k * w_k per boundary node (src/Core/Assembly_full.jl:806)."""
        ne, nn = self.elem_J.shape
        return (self.elem_J * self.elem_weight[None, :]).reshape(-1)

    # -- cache forwarding -----------------------------------------------------
    def __getattr__(self, name: str) -> Any:
        cache = object.__getattribute__(self, "cache")
        if name in cache:
            return cache[name]
        if name in BEMData._CACHE_KEYS:
            raise AttributeError(f"cache.{name} is not set — assemble/solve first?")
        raise AttributeError(f"BEMData has no property or cache field `{name}`")

    def set_cache(self, **kw) -> None:
        self.cache.update(kw)

    def has_cache(self, name: str) -> bool:
        return name in self.cache

    @property
    def dof(self) -> int:
        return n_dof(self.properties)

    def __repr__(self) -> str:
        return (f"BEMData({self.name!r}, p={self.p}, ne={self.ne}, n={self.n}, "
                f"ni={self.ni}, nt={self.nt}, {self.properties!r})")
