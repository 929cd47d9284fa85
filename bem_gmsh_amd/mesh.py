"""Boundary meshes and ``format2d`` (mesh -> collocation container).

The reference reads Gmsh models (src/Core/Input.jl:108-197). Gmsh is not part
of this path's hot loop, so meshes here are small in-memory descriptions that
hold what ``format2d_lagrange`` reads from a Gmsh model:

* 1-D line elements with their geometric nodes in Gmsh order (ends first, then
  interior nodes; src/Core/Input.jl:253-267) and a physical-group name
  ``"type;value"`` per element (src/Core/Input.jl:270-277, :564-582);
* optional 2-D domain cells (corner polygons) used to orient normals outward
  and to place internal collocation points at the cell centroids
  (src/Core/Input.jl:376-413).

Builders:

* ``quadrado(ndiv, ordem)`` — the unit square of data/Laplace/Laplace_dad.jl:238-280.
  Gmsh's ``setTransfiniteCurve(l, ndiv)`` takes a NODE count, so each side has
  ``ndiv - 1`` elements; curves run bottom (->), right (^), top (<-), left (v).
* ``ellipse_mesh`` and ``star_mesh`` — curved quadratic boundaries for the
  large workloads (seeded with SplitMix64 so the inputs are reproducible).
* ``read_msh`` — Gmsh 2.2 / 4.1 ASCII ``.msh`` reader for the same data.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import gauss_legendre
from .problem import BEMData, Elasticity, Laplace, n_dof

# ---------------------------------------------------------------------------
# Mesh description
# ---------------------------------------------------------------------------


@dataclass
class Mesh2D:
    """Line elements + physical names + domain cells (a Gmsh-model stand-in)."""

    name: str
    order: int                       # geometric order p of the line elements
    lines: np.ndarray                # (ne, p+1, 2) nodes in Gmsh order
    physical: List[str]              # per-element physical name ("t;v" or label)
    region: np.ndarray               # per-element entity tag
    cells: List[np.ndarray] = field(default_factory=list)  # corner polygons (k, 2)
    extras: dict = field(default_factory=dict)


def _gmsh_line_reorder(p: int) -> Sequence[int]:
    """Gmsh 1-D node order -> parametric order (src/Core/Input.jl:253-267)."""
    if p == 1:
        return (0, 1)
    return (0,) + tuple(range(2, p + 1)) + (1,)


def _to_gmsh_order(X: np.ndarray) -> np.ndarray:
    """Parametric order (xi = -1..1) -> Gmsh order (ends, then interior)."""
    p = X.shape[1] - 1
    if p == 1:
        return X
    return np.concatenate([X[:, :1], X[:, -1:], X[:, 1:-1]], axis=1)


def parse_pairs(s: str) -> Tuple[int, float]:
    """First BC type and value of a physical name (src/Core/Input.jl:564-582)."""
    t, v = parse_bc_pairs(s)
    return t[0], v[0]


def parse_bc_pairs(s: str) -> Tuple[List[int], List[float]]:
    """Every (type, value) pair of a physical name, ``"t;v"`` (scalar) or
    ``"tx;vx;ty;vy"`` (elasticity), as parse_pairs (src/Core/Input.jl:564-582):
    bare labels give ``[1], [0.0]``; a malformed pair raises ValueError."""
    parts = [t.strip() for t in s.split(";") if t.strip() != ""]
    try:
        int(parts[0])
    except (ValueError, IndexError):
        return [1], [0.0]
    if len(parts) < 2:
        return [1], [0.0]
    ts, vs = [], []
    for k in range(len(parts) // 2):
        try:
            ts.append(int(parts[2 * k]))
            vs.append(float(parts[2 * k + 1]))
        except ValueError:
            raise ValueError(f"Bad BC pair in '{s}'")
    return ts, vs


# ---------------------------------------------------------------------------
# Barycentric Lagrange bases (src/Core/Interpolation.jl) for host geometry.
# ---------------------------------------------------------------------------

def legendre_basis(p: int):
    """Nodes, barycentric weights and Dmat of Legendre(p) (Interpolation.jl:62-71)."""
    x, w = gauss_legendre(p + 1)
    sgn = np.array([1.0 if (i + 1) % 2 == 1 else -1.0 for i in range(p + 1)])
    bw = sgn * np.sqrt((1.0 - x * x) * w)
    return x, bw, _diff_matrix(x, bw)


def equispaced_basis(p: int):
    """Nodes, barycentric weights (eq. 5.1) and Dmat of Equispaced(p)."""
    x = np.linspace(-1.0, 1.0, p + 1)
    bw = np.array([(1.0 if ((p % 2 == 1) != (j % 2 == 0)) else -1.0) * math.comb(p, j)
                   for j in range(p + 1)], dtype=np.float64)
    return x, bw, _diff_matrix(x, bw)


def _diff_matrix(x: np.ndarray, bw: np.ndarray) -> np.ndarray:
    n = x.size
    D = np.zeros((n, n))
    for i in range(n):
        s = 0.0
        for j in range(n):
            if j == i:
                continue
            t = (bw[j] / bw[i]) / (x[i] - x[j])
            D[i, j] = t
            s += t
        D[i, i] = -s
    return D


def interpolation_matrix(x0: np.ndarray, bw: np.ndarray, xs: np.ndarray) -> np.ndarray:
    """Rows of the barycentric interpolation matrix (Interpolation.jl:148-178)."""
    xs = np.atleast_1d(np.asarray(xs, dtype=np.float64))
    M = np.zeros((xs.size, x0.size))
    for r, xx in enumerate(xs):
        exact = -1
        msum = 0.0
        for i in range(x0.size):
            if xx == x0[i]:
                exact = i
            with np.errstate(divide="ignore"):
                M[r, i] = bw[i] / (xx - x0[i])
            msum += M[r, i]
        if msum == 0.0:
            M[r, :] = 0.0
        elif exact >= 0:
            M[r, :] = 0.0
            M[r, exact] = 1.0
        else:
            M[r, :] = M[r, :] / msum
    return M


def shapefun(basis, xs):
    x0, bw, D = basis
    L = interpolation_matrix(x0, bw, xs)
    return L, L @ D


# ---------------------------------------------------------------------------
# format2d
# ---------------------------------------------------------------------------

def _point_in_polygon(px: float, py: float, verts: np.ndarray) -> bool:
    """Even-odd ray test, boundary counts as outside (src/Core/Input.jl:324-338)."""
    nv = verts.shape[0]
    if nv < 3:
        return False
    inside = False
    j = nv - 1
    for i in range(nv):
        yi, yj = verts[i, 1], verts[j, 1]
        if (yi > py) != (yj > py):
            xint = (verts[j, 0] - verts[i, 0]) * (py - yi) / (yj - yi + 1e-30) + verts[i, 0]
            if px < xint:
                inside = not inside
        j = i
    return inside


def format2d(mesh: Mesh2D, properties, tipo: Optional[int] = None,
             pontointerno: bool = True) -> BEMData:
    """Build the collocation container from a boundary mesh.

    Follows ``format2d_lagrange`` (src/Core/Input.jl:136-197): discontinuous
    Lagrange elements with (p+1) Gauss-Legendre collocation nodes, geometry from
    the equispaced Lagrange interpolant of the element's geometric nodes,
    ``J = |dx/dxi|``, ``n = (dx_2, -dx_1)/J`` (tan2normal, :584),
    ``L = |sum_k J_k w_k|`` and BCs from the physical names (``"t;v"`` for
    Laplace, ``"tx;vx;ty;vy"`` per dof for Elasticity, ``_assign_bc!``
    :537-559); internal nodes at the domain-cell centroids when ``pontointerno``.
    """
    if not isinstance(properties, (Laplace, Elasticity)):
        raise TypeError("format2d: properties must be Laplace or Elasticity")
    dof = n_dof(properties)
    p = int(mesh.order)
    if tipo is not None and int(tipo) != p:
        raise ValueError(f"format2d: tipo={tipo} differs from the mesh order {p}; "
                         "re-mesh at the requested order")
    if p < 1:
        raise ValueError(f"degree must be >= 1, got {p}")
    lines = np.asarray(mesh.lines, dtype=np.float64)
    ne = lines.shape[0]
    if ne == 0:
        raise ValueError("format2d: no 1D mesh elements")
    nn = p + 1
    reorder = list(_gmsh_line_reorder(p))
    X = lines[:, reorder, :]                        # (ne, p+1, 2) parametric order
    qsi, wi = gauss_legendre(p + 1)
    Ngeo, dNgeo = shapefun(equispaced_basis(p), qsi)  # (p+1, p+1)
    # nodes = Ngeo * X ; dx = dNgeo * X (sequential sums over geometric nodes)
    nodes = np.zeros((ne, nn, 2))
    dx = np.zeros((ne, nn, 2))
    for k in range(nn):
        for m in range(nn):
            nodes[:, k, :] += Ngeo[k, m] * X[:, m, :]
            dx[:, k, :] += dNgeo[k, m] * X[:, m, :]
    J = np.sqrt(dx[..., 0] * dx[..., 0] + dx[..., 1] * dx[..., 1])
    normal = np.stack([dx[..., 1] / J, -(dx[..., 0] / J)], axis=-1)
    Lsum = np.zeros(ne)
    for k in range(nn):
        Lsum = Lsum + J[:, k] * wi[k]
    L = np.abs(Lsum)
    BC = np.ones(dof * ne * nn, dtype=np.int64)
    BV = np.zeros(dof * ne * nn, dtype=np.float64)
    for e in range(ne):
        ts, vs = parse_bc_pairs(mesh.physical[e])
        tc, vc = [1] * dof, [0.0] * dof
        m = min(len(ts), dof)
        tc[:m], vc[:m] = ts[:m], vs[:m]
        for d in range(dof):
            BC[dof * e * nn + d:dof * (e + 1) * nn:dof] = tc[d]
            BV[dof * e * nn + d:dof * (e + 1) * nn:dof] = vc[d]
    cells = [np.asarray(c, dtype=np.float64) for c in mesh.cells]
    cells = [c for c in cells if abs(_polygon_area(c)) >= 1e-18]
    # _orient_normals_outward! (Input.jl:376-389)
    if cells:
        for e in range(ne):
            delta = 0.1 * float(L[e])
            if delta < 1e-16:
                continue
            tx = nodes[e, 0, 0] - delta * normal[e, 0, 0]
            ty = nodes[e, 0, 1] - delta * normal[e, 0, 1]
            if any(_point_in_polygon(tx, ty, c) for c in cells):
                continue
            normal[e] = -normal[e]
    internal = (np.array([c.mean(axis=0) for c in cells]).reshape(-1, 2)
                if (pontointerno and cells) else np.zeros((0, 2)))
    coll = np.concatenate([nodes.reshape(-1, 2), internal], axis=0)
    dad = BEMData(mesh.name, p, coll, normal.reshape(-1, 2), J, L, X,
                  np.asarray(mesh.region, dtype=np.int64), wi, properties, BC, BV,
                  n=ne * nn, ni=internal.shape[0])
    if cells:
        dad.set_cache(cells=cells)
    return dad


formatdata = format2d


def _polygon_area(v: np.ndarray) -> float:
    n = v.shape[0]
    if n < 3:
        return 0.0
    a = 0.0
    for i in range(n):
        j = 0 if i == n - 1 else i + 1
        a += v[i, 0] * v[j, 1] - v[j, 0] * v[i, 1]
    return 0.5 * a


# ---------------------------------------------------------------------------
# Builders
# ---------------------------------------------------------------------------

def quadrado(ndiv: int = 10, ordem: int = 1, Lx: float = 1.0, Ly: float = 1.0,
             nome: str = "quadrado", cells: bool = True) -> Mesh2D:
    """Unit-square mesh of data/Laplace/Laplace_dad.jl:238-280 (no Gmsh).

    ``ndiv`` nodes per side (ndiv-1 elements), curve loop bottom, right, top,
    left (counter-clockwise, so the element normals point outward); BCs "1;0"
    (bottom, top), "1;-1" (right), "0;0" (left); a recombined transfinite
    (ndiv-1)^2 quad grid as domain cells (``cells=False``: boundary only, for
    large ndiv).
    """
    return _square(ndiv, ordem, Lx, Ly, nome, cells, ("1;0", "1;-1", "1;0", "0;0"))


def quadrado_elasticity(ndiv: int = 10, ordem: int = 1, Lx: float = 1.0, Ly: float = 1.0,
                        nome: str = "quadrado_elast", cells: bool = True) -> Mesh2D:
    """Square of data/Laplace/Laplace_dad.jl:289-320 for 2-D elasticity (no
    Gmsh): the mesh of ``quadrado`` with 4-token BC names ``"tx;vx;ty;vy"``,
    left edge clamped ``"0;0;0;0"``, the others traction-free ``"1;0;1;0"``."""
    return _square(ndiv, ordem, Lx, Ly, nome, cells, ("1;0;1;0", "1;0;1;0", "1;0;1;0",
                                                      "0;0;0;0"))


def _square(ndiv, ordem, Lx, Ly, nome, cells, names) -> Mesh2D:
    if ndiv < 2:
        raise ValueError("ndiv must be >= 2")
    if ordem < 1:
        raise ValueError("ordem must be >= 1")
    m = ndiv - 1
    corners = np.array([[0.0, 0.0], [Lx, 0.0], [Lx, Ly], [0.0, Ly]])
    sides = [(0, 1, names[0], 1), (1, 2, names[1], 2), (2, 3, names[2], 3), (3, 0, names[3], 4)]
    t = np.linspace(0.0, 1.0, m + 1)
    lines, phys, reg = [], [], []
    for a, b, name, tag in sides:
        A, B = corners[a], corners[b]
        for s in range(m):
            ts = np.linspace(t[s], t[s + 1], ordem + 1)
            X = A[None, :] + ts[:, None] * (B - A)[None, :]
            X[0] = A + t[s] * (B - A)
            X[-1] = A + t[s + 1] * (B - A)
            lines.append(X)
            phys.append(name)
            reg.append(tag)
    lines = _to_gmsh_order(np.array(lines))
    xs = np.linspace(0.0, Lx, m + 1)
    ys = np.linspace(0.0, Ly, m + 1)
    quads = [np.array([[xs[i], ys[j]], [xs[i + 1], ys[j]], [xs[i + 1], ys[j + 1]],
                       [xs[i], ys[j + 1]]]) for j in range(m) for i in range(m)] if cells else []
    return Mesh2D(nome, ordem, lines, phys, np.array(reg), quads)


class SplitMix64:
    """Portable seeded PRNG (Steele, Lea & Flood 2014) for synthetic geometry."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))


def _curve_mesh(name: str, theta_ends: np.ndarray, curve, bc_of_element,
                order: int = 2) -> Mesh2D:
    ne = theta_ends.size - 1
    lines = np.zeros((ne, order + 1, 2))
    for e in range(ne):
        th = np.linspace(theta_ends[e], theta_ends[e + 1], order + 1)
        th[0], th[-1] = theta_ends[e], theta_ends[e + 1]
        lines[e] = curve(th)
    if abs(theta_ends[-1] - theta_ends[0] - 2.0 * np.pi) < 1e-12:
        lines[-1, -1] = lines[0, 0]   # closed curve: the last end node is the first one
    phys = [bc_of_element(e, 0.5 * (theta_ends[e] + theta_ends[e + 1])) for e in range(ne)]
    return Mesh2D(name, order, _to_gmsh_order(lines), phys, np.ones(ne, dtype=np.int64))


def ellipse_mesh(ne: int, a: float = 1.0, b: float = 0.5, order: int = 2) -> Mesh2D:
    """Ellipse (a cos t, b sin t), CCW, ``ne`` elements uniform in t.

    Physical names mark elements whose mid-angle lies in (0, pi) Dirichlet
    ("0;0") and the others Neumann ("1;0"); values come from the analytic field.
    """
    th = 2.0 * np.pi * np.arange(ne + 1) / ne
    th[-1] = 2.0 * np.pi

    def curve(t):
        return np.stack([a * np.cos(t), b * np.sin(t)], axis=-1)

    def bc(e, tm):
        return "0;0" if 0.0 < tm < np.pi else "1;0"

    m = _curve_mesh(f"ellipse_{ne}", th, curve, bc, order)
    m.extras.update(a=a, b=b)
    return m


def star_params(seed: int, perturb: bool = False):
    """(a_m, phi_m) for m = 2..8 of the star r = 1 + sum a_m cos(m t + phi_m)."""
    rng = SplitMix64(seed)
    am = np.array([rng.uniform(0.0, 0.05) for _ in range(7)])
    ph = np.array([rng.uniform(0.0, 2.0 * np.pi) for _ in range(7)])
    if perturb:
        prng = SplitMix64(seed ^ 0x5DEECE66D)
        am = am + np.array([prng.uniform(-0.01, 0.01) for _ in range(7)])
    return am, ph


def star_radius(t, am, ph):
    r = np.ones_like(t)
    for idx, m in enumerate(range(2, 9)):
        r = r + am[idx] * np.cos(m * t + ph[idx])
    return r


def star_mesh(ne: int, seed: int = 1, jitter: float = 0.1, arc: int = 64,
              order: int = 2, params=None) -> Mesh2D:
    """Random star-shaped domain, ``ne`` elements with +-``jitter`` angular widths.

    Arcs of ``arc`` elements alternate Dirichlet ("0;0") / Neumann ("1;0"),
    starting Dirichlet. ``params`` overrides (a_m, phi_m) (shape perturbation).
    """
    am, ph = star_params(seed) if params is None else params
    rng = SplitMix64(seed * 7919 + 17)
    widths = np.array([1.0 + jitter * rng.uniform(-1.0, 1.0) for _ in range(ne)])
    th = np.concatenate([[0.0], np.cumsum(widths)])
    th = 2.0 * np.pi * th / th[-1]
    th[-1] = 2.0 * np.pi

    def curve(t):
        r = star_radius(t, am, ph)
        return np.stack([r * np.cos(t), r * np.sin(t)], axis=-1)

    def bc(e, tm):
        return "0;0" if (e // arc) % 2 == 0 else "1;0"

    m = _curve_mesh(f"star_{ne}_{seed}", th, curve, bc, order)
    m.extras.update(am=am, ph=ph, theta=th)
    return m


def star_param_jacobian(mesh: Mesh2D) -> np.ndarray:
    """d(geometric nodes)/d(a_m, phi_m) of a star mesh, shaped (ne, p+1, 2, 14):
    columns 0..6 are a_2..a_8, 7..13 phi_2..phi_8. A node at angle t sits at
    r(t) (cos t, sin t), so dX/da_m = cos(m t + phi_m) (cos t, sin t) and
    dX/dphi_m = -a_m sin(m t + phi_m) (cos t, sin t)."""
    am, ph = mesh.extras["am"], mesh.extras["ph"]
    th = mesh.extras["theta"]
    p = int(mesh.order)
    ne = th.size - 1
    t = np.zeros((ne, p + 1))
    for e in range(ne):
        te = np.linspace(th[e], th[e + 1], p + 1)
        te[0], te[-1] = th[e], th[e + 1]
        t[e] = te                                   # parametric order (xi = -1..1)
    dirv = np.stack([np.cos(t), np.sin(t)], axis=-1)  # (ne, p+1, 2)
    Jac = np.zeros((ne, p + 1, 2, 14))
    for idx, m in enumerate(range(2, 9)):
        Jac[..., idx] = (np.cos(m * t + ph[idx]))[..., None] * dirv
        Jac[..., 7 + idx] = (-am[idx] * np.sin(m * t + ph[idx]))[..., None] * dirv
    return Jac


def move_geometric_node(mesh: Mesh2D, node, axis: int, delta: float) -> Mesh2D:
    """Copy of ``mesh`` with every copy of the geometric node ``node`` (x, y)
    moved by ``delta`` along ``axis`` (a shared end node moves in both elements)."""
    lines = np.array(mesh.lines, dtype=np.float64, copy=True)
    hit = np.all(lines == np.asarray(node, dtype=np.float64), axis=-1)
    if not hit.any():
        raise ValueError("move_geometric_node: node not found")
    lines[hit, axis] += delta
    return Mesh2D(mesh.name, mesh.order, lines, list(mesh.physical), np.array(mesh.region),
                  [np.array(c) for c in mesh.cells], dict(mesh.extras))


def n_eff(N: int, p: int = 2) -> int:
    """Nearest multiple of (p+1) to N: discontinuous elements carry p+1 nodes."""
    nn = p + 1
    return max(nn, int(round(N / nn)) * nn)


# ---------------------------------------------------------------------------
# Gmsh .msh (ASCII 2.2 and 4.1) reader
# ---------------------------------------------------------------------------

_LINE_TYPES = {1: 1, 8: 2, 26: 3, 27: 4, 28: 5}    # Gmsh element type -> order
_CELL_CORNERS = {2: 3, 3: 4, 9: 3, 10: 4, 16: 4, 21: 3, 23: 3}


def read_msh(path: str, name: Optional[str] = None) -> Mesh2D:
    """Read a Gmsh mesh (format 2.2 or 4.1, ASCII or binary) into a Mesh2D.

    Line elements carry the physical-group name of their entity (first group);
    2-D elements become domain cells (corners only, as src/Core/Input.jl:307-322).
    """
    with open(path, "rb") as f:
        data = f.read()
    version, binary, dsize = _msh_format(data)
    if binary:
        parsed = _parse_msh_binary(data, version, dsize)
    else:
        parsed = _parse_msh_ascii(data.decode("utf-8", errors="replace").split("\n"))
    names, nodes, lines_raw, cells_raw = parsed
    nm = name or path.rsplit("/", 1)[-1].rsplit(".", 1)[0]
    return _mesh_from_raw(nm, names, nodes, lines_raw, cells_raw)


def _msh_format(data: bytes):
    """(version, binary, data size) from the $MeshFormat header."""
    k = data.find(b"$MeshFormat")
    if k < 0:
        raise ValueError("read_msh: no $MeshFormat section")
    line = data[k:].split(b"\n", 2)[1].split()
    version, ftype, dsize = float(line[0]), int(line[1]), int(line[2])
    if ftype not in (0, 1):
        raise ValueError(f"read_msh: unknown file type {ftype}")
    if ftype == 1 and dsize != 8:
        raise ValueError(f"read_msh: binary data size {dsize} (only 8 is defined)")
    return version, ftype == 1, dsize


def _parse_msh_ascii(tok):
    """Sections of an ASCII .msh: (names, nodes, lines, cells)."""
    i = 0
    version = 2.2
    names = {}
    nodes = {}
    lines_raw: List[Tuple[int, List[int], int]] = []   # (type, node tags, phys or entity)
    cells_raw: List[Tuple[int, List[int]]] = []
    ent_phys = {}     # (dim, entity tag) -> physical tag (4.1)
    while i < len(tok):
        t = tok[i].strip()
        if t == "$MeshFormat":
            version = float(tok[i + 1].split()[0])
            i += 2
        elif t == "$PhysicalNames":
            cnt = int(tok[i + 1])
            for r in range(cnt):
                parts = tok[i + 2 + r].split(maxsplit=2)
                names[(int(parts[0]), int(parts[1]))] = parts[2].strip().strip('"')
            i += 2 + cnt
        elif t == "$Entities" and version >= 4:
            hdr = [int(v) for v in tok[i + 1].split()]
            r = i + 2
            r += hdr[0]                       # points
            for dim, count in ((1, hdr[1]), (2, hdr[2]), (3, hdr[3])):
                for _ in range(count):
                    parts = tok[r].split()
                    tag = int(parts[0])
                    nph = int(parts[7])
                    if nph > 0:
                        ent_phys[(dim, tag)] = abs(int(parts[8]))
                    r += 1
            i = r
        elif t == "$Nodes":
            if version >= 4:
                hdr = [int(v) for v in tok[i + 1].split()]
                r = i + 2
                for _ in range(hdr[0]):
                    eh = [int(v) for v in tok[r].split()]
                    cnt = eh[3]
                    tags = [int(tok[r + 1 + q]) for q in range(cnt)]
                    for q in range(cnt):
                        xyz = [float(v) for v in tok[r + 1 + cnt + q].split()]
                        nodes[tags[q]] = (xyz[0], xyz[1])
                    r += 1 + 2 * cnt
                i = r
            else:
                cnt = int(tok[i + 1])
                for r in range(cnt):
                    parts = tok[i + 2 + r].split()
                    nodes[int(parts[0])] = (float(parts[1]), float(parts[2]))
                i += 2 + cnt
        elif t == "$Elements":
            if version >= 4:
                hdr = [int(v) for v in tok[i + 1].split()]
                r = i + 2
                for _ in range(hdr[0]):
                    eh = [int(v) for v in tok[r].split()]
                    dim, etag, etype, cnt = eh
                    for q in range(cnt):
                        parts = [int(v) for v in tok[r + 1 + q].split()]
                        if dim == 1 and etype in _LINE_TYPES:
                            lines_raw.append((etype, parts[1:], ent_phys.get((1, etag), -1),
                                              etag))
                        elif dim == 2:
                            cells_raw.append((etype, parts[1:]))
                    r += 1 + cnt
                i = r
            else:
                cnt = int(tok[i + 1])
                for r in range(cnt):
                    parts = [int(v) for v in tok[i + 2 + r].split()]
                    etype, ntag = parts[1], parts[2]
                    tags = parts[3:3 + ntag]
                    conn = parts[3 + ntag:]
                    phys = tags[0] if ntag > 0 else -1
                    ent = tags[1] if ntag > 1 else -1
                    if etype in _LINE_TYPES:
                        lines_raw.append((etype, conn, phys, ent))
                    elif etype in _CELL_CORNERS or len(conn) >= 3:
                        if etype not in (15,):
                            cells_raw.append((etype, conn))
                i += 2 + cnt
        else:
            i += 1
    return names, nodes, lines_raw, cells_raw


# Nodes per element of the Gmsh element types a binary file may hold (the
# binary element records have no length field).
_MSH_NODES = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6, 10: 9, 11: 10, 12: 27,
              13: 18, 14: 14, 15: 1, 16: 8, 17: 20, 18: 15, 19: 13, 20: 9, 21: 10, 22: 12,
              23: 15, 24: 15, 25: 21, 26: 4, 27: 5, 28: 6, 29: 20, 30: 35, 31: 56, 36: 16,
              37: 25, 38: 36}


class _Cursor:
    """Byte cursor over a binary .msh: text lines and little/big-endian arrays."""

    def __init__(self, data: bytes, pos: int, endian: str):
        self.d, self.p, self.e = data, pos, endian

    def line(self) -> str:
        k = self.d.index(b"\n", self.p)
        s = self.d[self.p:k].decode("utf-8", errors="replace").strip()
        self.p = k + 1
        return s

    def arr(self, code: str, count: int) -> np.ndarray:
        dt = np.dtype(self.e + code)
        a = np.frombuffer(self.d, dtype=dt, count=count, offset=self.p)
        self.p += dt.itemsize * count
        return a

    def one(self, code: str):
        return self.arr(code, 1)[0].item()


def _parse_msh_binary(data: bytes, version: float, dsize: int):
    """Sections of a binary .msh (Gmsh file format 2.2 and 4.1, data size 8):
    ints are 4 bytes, size_t 8 bytes, reals 8-byte IEEE; the endianness comes
    from the integer 1 written after the format line."""
    k = data.find(b"$MeshFormat")
    k = data.index(b"\n", data.index(b"\n", k) + 1) + 1     # after the format line
    one = data[k:k + 4]
    endian = "<" if np.frombuffer(one, "<i4")[0] == 1 else ">"
    if np.frombuffer(one, endian + "i4")[0] != 1:
        raise ValueError("read_msh: bad endianness marker")
    cur = _Cursor(data, k + 4, endian)
    names, nodes = {}, {}
    lines_raw, cells_raw = [], []
    ent_phys = {}
    v4 = version >= 4
    st = "u8"                                               # size_t
    while cur.p < len(data):
        t = cur.line()
        if t == "$PhysicalNames":                           # ASCII in both versions
            for _ in range(int(cur.line())):
                parts = cur.line().split(maxsplit=2)
                names[(int(parts[0]), int(parts[1]))] = parts[2].strip().strip('"')
        elif t == "$Entities" and v4:
            npt, ncu, nsu, nvo = (int(x) for x in cur.arr(st, 4))
            for _ in range(npt):
                cur.arr("i4", 1); cur.arr("f8", 3)
                cur.arr("i4", cur.one(st))
            for dim, count in ((1, ncu), (2, nsu), (3, nvo)):
                for _ in range(count):
                    tag = cur.one("i4")
                    cur.arr("f8", 6)
                    ph = cur.arr("i4", cur.one(st))
                    if ph.size:
                        ent_phys[(dim, tag)] = abs(int(ph[0]))
                    cur.arr("i4", cur.one(st))
        elif t == "$Nodes":
            if v4:
                nblk, _, _, _ = (int(x) for x in cur.arr(st, 4))
                for _ in range(nblk):
                    cur.arr("i4", 2)
                    par = cur.one("i4")
                    cnt = cur.one(st)
                    tags = cur.arr(st, cnt)
                    xyz = cur.arr("f8", cnt * (3 + 3 * par)).reshape(cnt, -1)
                    for q in range(cnt):
                        nodes[int(tags[q])] = (float(xyz[q, 0]), float(xyz[q, 1]))
            else:
                cnt = int(cur.line())
                rec = np.dtype([("tag", endian + "i4"), ("xyz", endian + "f8", 3)])
                a = np.frombuffer(data, dtype=rec, count=cnt, offset=cur.p)
                cur.p += rec.itemsize * cnt
                for r in a:
                    nodes[int(r["tag"])] = (float(r["xyz"][0]), float(r["xyz"][1]))
        elif t == "$Elements":
            if v4:
                nblk, _, _, _ = (int(x) for x in cur.arr(st, 4))
                for _ in range(nblk):
                    dim, etag, etype = (int(x) for x in cur.arr("i4", 3))
                    cnt = cur.one(st)
                    if etype not in _MSH_NODES:
                        raise ValueError(f"read_msh: unknown element type {etype}")
                    recs = cur.arr(st, cnt * (1 + _MSH_NODES[etype])).reshape(cnt, -1)
                    for q in range(cnt):
                        conn = [int(x) for x in recs[q, 1:]]
                        if dim == 1 and etype in _LINE_TYPES:
                            lines_raw.append((etype, conn, ent_phys.get((1, etag), -1), etag))
                        elif dim == 2:
                            cells_raw.append((etype, conn))
            else:
                total = int(cur.line())
                seen = 0
                while seen < total:
                    etype, follow, ntag = (int(x) for x in cur.arr("i4", 3))
                    if etype not in _MSH_NODES:
                        raise ValueError(f"read_msh: unknown element type {etype}")
                    width = 1 + ntag + _MSH_NODES[etype]
                    recs = cur.arr("i4", follow * width).reshape(follow, width)
                    for q in range(follow):
                        tags = [int(x) for x in recs[q, 1:1 + ntag]]
                        conn = [int(x) for x in recs[q, 1 + ntag:]]
                        phys = tags[0] if ntag > 0 else -1
                        ent = tags[1] if ntag > 1 else -1
                        if etype in _LINE_TYPES:
                            lines_raw.append((etype, conn, phys, ent))
                        elif etype in _CELL_CORNERS:
                            cells_raw.append((etype, conn))
                    seen += follow
        elif t.startswith("$") and not t.startswith("$End"):
            # a section this reader does not use: skip to its end marker
            cur.p = data.index(b"$End" + t[1:].encode(), cur.p)
    return names, nodes, lines_raw, cells_raw


def _mesh_from_raw(nm, names, nodes, lines_raw, cells_raw) -> Mesh2D:
    if not lines_raw:
        raise ValueError("read_msh: no 1D mesh elements")
    order = _LINE_TYPES[lines_raw[0][0]]
    if any(_LINE_TYPES[l[0]] != order for l in lines_raw):
        raise ValueError("read_msh: mixed line orders are not supported")
    lines = np.array([[nodes[t] for t in conn] for (_, conn, _, _) in lines_raw])
    phys = [names.get((1, ph), "") if ph >= 0 else "" for (_, _, ph, _) in lines_raw]
    region = np.array([ent for (_, _, _, ent) in lines_raw], dtype=np.int64)
    cells = []
    for etype, conn in cells_raw:
        nc = _CELL_CORNERS.get(etype, min(len(conn), 4))
        cells.append(np.array([nodes[t] for t in conn[:nc]]))
    return Mesh2D(nm, order, lines, phys, region, cells)




def write_msh(mesh: Mesh2D, path: str, fmt: str = "2.2", binary: bool = False) -> None:
    """Write a Mesh2D as a Gmsh mesh: lines with their physical names, quad /
    triangle domain cells. fmt "2.2" (ASCII) or "4.1" (ASCII or binary)."""
    if fmt not in ("2.2", "4.1"):
        raise ValueError(f"write_msh: fmt must be '2.2' or '4.1', got {fmt!r}")
    if binary and fmt != "4.1":
        raise ValueError("write_msh: binary output is written in format 4.1")
    names = {}
    for nm in mesh.physical:
        if nm and nm not in names:
            names[nm] = len(names) + 1
    pts = []
    index = {}

    def tag(xy):
        key = (float(xy[0]), float(xy[1]))
        if key not in index:
            index[key] = len(pts) + 1
            pts.append(key)
        return index[key]

    etype = {1: 1, 2: 8, 3: 26, 4: 27, 5: 28}[mesh.order]
    lconn = [[tag(x) for x in mesh.lines[e]] for e in range(mesh.lines.shape[0])]
    cconn = [[tag(x) for x in c] for c in mesh.cells]
    dom = len(names) + 1
    if fmt == "2.2":
        elines = []
        for e, conn in enumerate(lconn):
            ph = names.get(mesh.physical[e], 0)
            elines.append(f"{etype} 2 {ph} {int(mesh.region[e])} " + " ".join(map(str, conn)))
        for conn in cconn:
            ct = 3 if len(conn) == 4 else 2
            elines.append(f"{ct} 2 {dom} 1 " + " ".join(map(str, conn)))
        with open(path, "w") as f:
            f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
            f.write(f"$PhysicalNames\n{len(names) + 1}\n")
            for nm, t in names.items():
                f.write(f'1 {t} "{nm}"\n')
            f.write(f'2 {dom} "Domain"\n$EndPhysicalNames\n')
            f.write(f"$Nodes\n{len(pts)}\n")
            for q, (x, y) in enumerate(pts):
                f.write(f"{q + 1} {x!r} {y!r} 0\n")
            f.write("$EndNodes\n")
            f.write(f"$Elements\n{len(elines)}\n")
            for q, s_ in enumerate(elines):
                f.write(f"{q + 1} {s_}\n")
            f.write("$EndElements\n")
        return
    _write_msh41(path, binary, names, dom, pts, etype, lconn, cconn, mesh)


def _write_msh41(path, binary, names, dom, pts, etype, lconn, cconn, mesh):
    """Format 4.1: one curve entity per (region, physical name), one surface for
    the cells, all nodes in one block on the surface."""
    curves = {}                                  # (region, name) -> entity tag
    used = set()
    for e in range(len(lconn)):
        key = (int(mesh.region[e]), mesh.physical[e])
        if key not in curves:
            t = key[0] if key[0] > 0 and key[0] not in used else max(used | {0}) + 1
            used.add(t)
            curves[key] = t
    blocks = {}                                  # entity tag -> element indices
    for e in range(len(lconn)):
        blocks.setdefault(curves[(int(mesh.region[e]), mesh.physical[e])], []).append(e)
    cell_blocks = {}
    for q, conn in enumerate(cconn):
        cell_blocks.setdefault(3 if len(conn) == 4 else 2, []).append(q)
    nel = len(lconn) + len(cconn)
    xy = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
    with open(path, "wb") as f:
        w = f.write
        w(b"$MeshFormat\n")
        w(b"4.1 1 8\n" if binary else b"4.1 0 8\n")
        if binary:
            w(np.array([1], dtype="<i4").tobytes() + b"\n")
        w(b"$EndMeshFormat\n")
        w(f"$PhysicalNames\n{len(names) + 1}\n".encode())
        for nm, t in names.items():
            w(f'1 {t} "{nm}"\n'.encode())
        w(f'2 {dom} "Domain"\n$EndPhysicalNames\n'.encode())
        # entities: curves with their physical group, one surface
        w(b"$Entities\n")
        lo, hi = xy.min(axis=0) if len(xy) else (0, 0), xy.max(axis=0) if len(xy) else (0, 0)
        box = [float(lo[0]), float(lo[1]), 0.0, float(hi[0]), float(hi[1]), 0.0]
        if binary:
            w(np.array([0, len(curves), 1, 0], dtype="<u8").tobytes())
            for (reg, nm), t in curves.items():
                ph = [names[nm]] if nm in names else []
                w(np.array([t], "<i4").tobytes() + np.array(box, "<f8").tobytes())
                w(np.array([len(ph)], "<u8").tobytes() + np.array(ph, "<i4").tobytes())
                w(np.array([0], "<u8").tobytes())
            w(np.array([1], "<i4").tobytes() + np.array(box, "<f8").tobytes())
            w(np.array([1], "<u8").tobytes() + np.array([dom], "<i4").tobytes())
            w(np.array([0], "<u8").tobytes())
            w(b"\n")
        else:
            w(f"0 {len(curves)} 1 0\n".encode())
            bs = " ".join(repr(v) for v in box)
            for (reg, nm), t in curves.items():
                ph = f"1 {names[nm]}" if nm in names else "0"
                w(f"{t} {bs} {ph} 0\n".encode())
            w(f"1 {bs} 1 {dom} 0\n".encode())
        w(b"$EndEntities\n")
        # nodes: one block on the surface
        w(b"$Nodes\n")
        n = len(pts)
        if binary:
            w(np.array([1, n, 1, n], "<u8").tobytes())
            w(np.array([2, 1, 0], "<i4").tobytes() + np.array([n], "<u8").tobytes())
            w(np.arange(1, n + 1, dtype="<u8").tobytes())
            xyz = np.zeros((n, 3), "<f8")
            xyz[:, :2] = xy
            w(xyz.tobytes())
            w(b"\n")
        else:
            w(f"1 {n} 1 {n}\n2 1 0 {n}\n".encode())
            w("".join(f"{q + 1}\n" for q in range(n)).encode())
            w("".join(f"{x!r} {y!r} 0\n" for x, y in pts).encode())
        w(b"$EndNodes\n")
        # elements: one block per curve entity, one per cell type
        w(b"$Elements\n")
        nblk = len(blocks) + len(cell_blocks)
        tagc = 1
        if binary:
            w(np.array([nblk, nel, 1, nel], "<u8").tobytes())
        else:
            w(f"{nblk} {nel} 1 {nel}\n".encode())
        groups = [(1, t, etype, [lconn[e] for e in es]) for t, es in blocks.items()]
        groups += [(2, 1, ct, [cconn[q] for q in qs]) for ct, qs in cell_blocks.items()]
        for dim, ent, et, conns in groups:
            if binary:
                w(np.array([dim, ent, et], "<i4").tobytes() + np.array([len(conns)], "<u8").tobytes())
                recs = np.array([[tagc + q] + c for q, c in enumerate(conns)], dtype="<u8")
                w(recs.tobytes())
            else:
                w(f"{dim} {ent} {et} {len(conns)}\n".encode())
                w("".join(f"{tagc + q} " + " ".join(map(str, c)) + "\n"
                          for q, c in enumerate(conns)).encode())
            tagc += len(conns)
        if binary:
            w(b"\n")
        w(b"$EndElements\n")
