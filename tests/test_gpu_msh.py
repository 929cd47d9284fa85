"""Gmsh .msh input through the HIP path: read_msh -> format2d
(src/Core/Input.jl:136-197, the mesh a user's Gmsh run would give) -> HIP
assembly -> HIP solve, against the FP64 oracle on the same BEMData.

* tests/golden/square_2x2_v41.msh: the unit square as one quadrilateral cell
  with 2 linear boundary elements per side, physical names "0;0" (x = 0),
  "0;1" (x = 1) and "1;0" (y = 0, 1), so T = x exactly, with the cell centroid
  as the internal node.
* A curved boundary: the seeded star (quadratic elements) written by the
  product writer as format 4.1 binary and 2.2 ASCII and read back; the
  geometry is exact in both, so H, G equal those of the mesh built in memory
  bit for bit. That in-memory star is held to the per-entry parity bar
  elsewhere; against the oracle this test allows 1e-9, the size of the near
  pairs' own 1-ulp sensitivity on this boundary.
"""

import os

import numpy as np
import pytest

import bem_gmsh_amd as B
import oracle as O
from helpers import entry_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _vs_oracle(dad, npg):
    H, G = B.assemble(dad, npg)
    Hr, Gr = O.assemble(dad, npg)
    eh, eg = entry_err(H, Hr), entry_err(G, Gr)
    B.solve(dad, mode="fp64")
    Tr, qr, _ = O.solve(dad, Hr, Gr)
    eT = float(np.linalg.norm(dad.T - Tr) / np.linalg.norm(Tr))
    eq = float(np.linalg.norm(dad.q - qr) / max(np.linalg.norm(qr), 1e-300))
    return H, G, eh, eg, eT, eq


def test_golden_square_msh_through_hip():
    dad = B.format2d(B.read_msh(os.path.join(GOLDEN, "square_2x2_v41.msh")), B.Laplace(1.0))
    assert dad.ni == 1 and set(np.unique(dad.BC)) == {0, 1}
    H, G, eh, eg, eT, eq = _vs_oracle(dad, 20)
    B.attach_analytical(dad, B.ana_laplace_linear())
    err = B.rel_error(dad)
    print(f"\n[square_2x2_v41.msh] nt={dad.nt}: H {eh:.1e} G {eg:.1e}, T {eT:.1e} q {eq:.1e}, "
          f"rel_error vs T = x {err:.1e}")
    assert eh <= 1e-12 and eg <= 1e-12 and eT <= 1e-10 and eq <= 1e-10
    # the reference's own bar for its square (test/laplace.jl:17-24); with two
    # elements a side the corner rows' doubled split rule dominates the error
    assert err < 0.05


@pytest.mark.parametrize("fmt,binary", [("4.1", True), ("2.2", False)])
def test_curved_msh_round_trip_through_hip(tmp_path, fmt, binary):
    mesh = B.star_mesh(120, seed=4)
    path = str(tmp_path / f"star.{fmt}.msh")
    B.write_msh(mesh, path, fmt=fmt, binary=binary)
    dad = B.format2d(B.read_msh(path), B.Laplace(1.0))
    B.apply_bc_from_physical(dad, B.ana_exp_cos())
    ref = B.format2d(mesh, B.Laplace(1.0))
    B.apply_bc_from_physical(ref, B.ana_exp_cos())
    assert np.array_equal(dad.collocation, ref.collocation) and np.array_equal(dad.BC, ref.BC)
    H, G, eh, eg, eT, eq = _vs_oracle(dad, 16)
    Hm, Gm = B.assemble(ref, 16)
    print(f"\n[star .msh {fmt}{' binary' if binary else ''}] N={dad.n}: H {eh:.1e} G {eg:.1e}, "
          f"T {eT:.1e} q {eq:.1e}, rel_error {B.rel_error(dad):.2e}")
    assert np.array_equal(H, Hm) and np.array_equal(G, Gm)
    assert eh <= 1e-9 and eg <= 1e-9 and eT <= 1e-9 and eq <= 1e-9
    assert B.rel_error(dad) < 1e-3
