"""Points inside the domain: grids, inside test, and the potential there.

Host-side counterparts of ``src/Core/InternalPoints.jl`` (``internal_grid``,
``internal_layer``, ``point_in_domain``; 2-D only) and of
``set_internal_nodes!`` (src/Core/Structures.jl:623-634), plus
``interior_laplace`` (src/Laplace/Lubrication.jl:399-417), whose values come
from the device (``bemhip_interior_laplace``).

The grids are small host computations (one closest-chord distance per
candidate point and element, vectorised over the elements with numpy); the
potential is the device pass over every (point, element) pair.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from ._lib import dptr, lib
from .api import ArgumentError, _context, upload_geometry
from .problem import BEMData


# ---------------------------------------------------------------------------
# Geometry helpers (src/Core/InternalPoints.jl:331-395, src/Core/Integration.jl:228-312)
# ---------------------------------------------------------------------------

def _segments(dad: BEMData):
    """Chords between consecutive geometric nodes of every element
    (``_boundary_segments``, InternalPoints.jl:209-225): arrays a, b of (m, 2)."""
    g = dad.elem_geo
    return g[:, :-1, :].reshape(-1, 2), g[:, 1:, :].reshape(-1, 2)


def _boundary_bbox(dad: BEMData):
    """Bounding box of the boundary nodes and geometric nodes (InternalPoints.jl:368-395)."""
    if dad.n < 1:
        raise ArgumentError("BEMdata has no boundary nodes")
    P = np.concatenate([dad.Nodes, dad.elem_geo.reshape(-1, 2)], axis=0)
    return P[:, 0].min(), P[:, 0].max(), P[:, 1].min(), P[:, 1].max()


def _max_el_length(dad: BEMData) -> float:
    """Largest element length, 1.0 when all are zero (InternalPoints.jl:360-366)."""
    L = float(np.max(dad.elem_L)) if dad.ne else 0.0
    return L if L > 0 else 1.0


def _dist_to_boundary(dad: BEMData, p) -> float:
    """Distance from p to the union of the element chords (``_dist_to_boundary``,
    InternalPoints.jl:348-358, with ``_dist_element`` / ``_dist_chord``,
    Integration.jl:253-257, :295-305). The element bounding-box pruning of the
    reference changes no result and is omitted."""
    a, b = _segments(dad)
    px, py = float(p[0]), float(p[1])
    abx, aby = b[:, 0] - a[:, 0], b[:, 1] - a[:, 1]
    apx, apy = px - a[:, 0], py - a[:, 1]
    t = np.clip((apx * abx + apy * aby) / ((abx * abx + aby * aby) + np.finfo(float).eps),
                0.0, 1.0)
    dx = px - (a[:, 0] + t * abx)
    dy = py - (a[:, 1] + t * aby)
    return float(np.sqrt(dx * dx + dy * dy).min()) if dx.size else float("inf")


def _even_odd(px: float, py: float, a: np.ndarray, b: np.ndarray) -> bool:
    """Even-odd ray test over chords a->b (``_inside_segments``, InternalPoints.jl:227-237)."""
    yi, yj = a[:, 1], b[:, 1]
    cross = (yi > py) != (yj > py)
    if not cross.any():
        return False
    ai, bi = a[cross], b[cross]
    xint = (bi[:, 0] - ai[:, 0]) * (py - ai[:, 1]) / (bi[:, 1] - ai[:, 1] + 1e-30) + ai[:, 0]
    return bool(np.count_nonzero(px < xint) % 2)


def point_in_domain(dad: BEMData, p) -> bool:
    """``point_in_domain`` (InternalPoints.jl:193-200): even-odd over the domain
    cells when they are cached (``_point_in_cells``, src/Core/Input.jl:339-344),
    otherwise even-odd over the boundary chords."""
    p = np.asarray(p, dtype=np.float64)
    if p.shape != (2,):
        raise ArgumentError(f"2-D point vs {dad.dimension}-D mesh")
    px, py = float(p[0]), float(p[1])
    cells = dad.cache.get("cells")
    if cells:
        for c in cells:
            v = np.asarray(c, dtype=np.float64)
            if v.shape[0] < 3:
                continue
            # edge (v[i], v[i-1]) from vertex i, as point_in_polygon (Input.jl:322-336)
            if _even_odd(px, py, v, np.roll(v, 1, axis=0)):
                return True
        return False
    a, b = _segments(dad)
    return _even_odd(px, py, a, b)


# ---------------------------------------------------------------------------
# Grids (InternalPoints.jl:23-99, :101-190)
# ---------------------------------------------------------------------------

def _axis_nodes(n: int, a: float, b: float, layout: str) -> np.ndarray:
    """Candidate coordinates along one axis (``_axis_nodes``, InternalPoints.jl:70-80)."""
    if n < 1:
        raise ArgumentError("n ≥ 1")
    i = np.arange(1, n + 1, dtype=np.float64)
    if layout == "gera":
        return a + (b - a) * i / (n + 1)
    if layout == "cell":
        return a + (b - a) * (i - 0.5) / n
    if layout == "cheb":
        return a + (b - a) * (1 - np.cos(np.pi * (i - 0.5) / n)) / 2
    raise ArgumentError(f"layout must be :gera, :cell, or :cheb (got {layout})")


def internal_grid(dad: BEMData, nx: int, ny: int, d_min: float = 0.01,
                  layout: str = "gera") -> np.ndarray:
    """Cartesian candidates in the bounding box of the boundary, kept when inside
    and at least ``d_min * L_max`` from the boundary (``internal_grid``,
    InternalPoints.jl:23-40). Returns (m, 2), x fastest as the reference loops."""
    if dad.dimension != 2:
        raise ArgumentError("internal_grid(dad, nx, ny) is 2-D; pass nz for 3-D")
    if nx < 1 or ny < 1:
        raise ArgumentError("nx, ny must be ≥ 1")
    xmin, xmax, ymin, ymax = _boundary_bbox(dad)
    xs = _axis_nodes(int(nx), xmin, xmax, layout)
    ys = _axis_nodes(int(ny), ymin, ymax, layout)
    clear = float(d_min) * _max_el_length(dad)
    pts = []
    for y in ys:
        for x in xs:
            p = (float(x), float(y))
            if not point_in_domain(dad, p):
                continue
            if clear > 0 and _dist_to_boundary(dad, p) < clear:
                continue
            pts.append(p)
    return np.asarray(pts, dtype=np.float64).reshape(-1, 2)


gera_p_in = internal_grid


def _typical_h(dad: BEMData) -> float:
    """In-plane mesh size L / (ne / 4) (``_typical_h``, InternalPoints.jl:118-125)."""
    xmin, xmax, ymin, ymax = _boundary_bbox(dad)
    L = max(xmax - xmin, ymax - ymin, 0.0, 1e-12)
    return L / (max(dad.ne, 1) / 4)


def _dedup_points(pts, tol: float):
    """Drop points closer than tol to an earlier kept one (InternalPoints.jl:168-185)."""
    if not pts:
        return pts
    tol2 = tol * tol
    kept = [pts[0]]
    for p in pts[1:]:
        if not any((p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 < tol2 for q in kept):
            kept.append(p)
    return kept


def _offset_layer(dad: BEMData, delta: float, every: int):
    """One inward point per element, mid-point moved by delta along -n
    (``_offset_layer``, InternalPoints.jl:127-145)."""
    nn = dad.p + 1
    pts = []
    for e in range(dad.ne):
        if e % every != 0:
            continue
        idx = slice(e * nn, (e + 1) * nn)
        mid = dad.collocation[idx].sum(axis=0) / nn
        nrm = dad.Normal[idx].sum(axis=0)
        nnrm = float(np.hypot(nrm[0], nrm[1]))
        if nnrm < 1e-16:
            continue
        p = (float(mid[0] - (delta / nnrm) * nrm[0]), float(mid[1] - (delta / nnrm) * nrm[1]))
        if not point_in_domain(dad, p):
            continue
        if _dist_to_boundary(dad, p) < 0.75 * delta:
            continue
        pts.append(p)
    return _dedup_points(pts, 0.25 * delta)


def _core_fill(dad: BEMData, nfill: int, clear: float):
    """Cell-centred nfill x nfill lattice, kept when inside and clear of the
    boundary (``_core_fill``, InternalPoints.jl:147-166)."""
    xmin, xmax, ymin, ymax = _boundary_bbox(dad)
    xs = _axis_nodes(nfill, xmin, xmax, "cell")
    ys = _axis_nodes(nfill, ymin, ymax, "cell")
    pts = []
    for y in ys:
        for x in xs:
            p = (float(x), float(y))
            if not point_in_domain(dad, p):
                continue
            if _dist_to_boundary(dad, p) < clear:
                continue
            pts.append(p)
    return pts


def internal_layer(dad: BEMData, delta: Optional[float] = None, every: int = 1,
                   fill: int = 0) -> np.ndarray:
    """Few-pole layout: one inward point per element (every ``every``-th), plus
    an optional ``fill x fill`` core lattice (``internal_layer``,
    InternalPoints.jl:101-112). ``delta`` defaults to 1.25 x the typical h."""
    if every < 1:
        raise ArgumentError("every must be ≥ 1")
    if fill < 0:
        raise ArgumentError("fill must be ≥ 0")
    d = 1.25 * _typical_h(dad) if delta is None else float(delta)
    if not d > 0:
        raise ArgumentError("δ must be positive")
    pts = _offset_layer(dad, d, int(every))
    if fill > 0:
        pts = _dedup_points(pts + _core_fill(dad, int(fill), 2 * d), 0.25 * d)
    return np.asarray(pts, dtype=np.float64).reshape(-1, 2)


def set_internal_nodes(dad: BEMData, pts) -> BEMData:
    """Replace the internal nodes: collocation = boundary nodes then pts
    (``set_internal_nodes!``, src/Core/Structures.jl:623-634). Drops the cached
    matrices and solution, which no longer match nt."""
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
    dad.collocation = np.ascontiguousarray(np.concatenate([dad.Nodes, pts], axis=0))
    dad.ni = int(pts.shape[0])
    dad.nt = dad.n + dad.ni
    for key in ("H", "G", "A", "B", "b", "T", "q", "solve_info", "assembly_info"):
        dad.cache.pop(key, None)
    if dad.hip is not None:
        dad.hip.close()
        dad.hip = None
    return dad


def internal_grid_(dad: BEMData, nx: int, ny: int, **kw) -> BEMData:
    """``internal_grid!``: set_internal_nodes(dad, internal_grid(dad, nx, ny))."""
    return set_internal_nodes(dad, internal_grid(dad, nx, ny, **kw))


def internal_layer_(dad: BEMData, **kw) -> BEMData:
    """``internal_layer!``: set_internal_nodes(dad, internal_layer(dad, ...))."""
    return set_internal_nodes(dad, internal_layer(dad, **kw))


# ---------------------------------------------------------------------------
# Potential at interior points (src/Laplace/Lubrication.jl:399-417)
# ---------------------------------------------------------------------------

def interior_laplace(dad: BEMData, pf, T: Optional[np.ndarray] = None,
                     q: Optional[np.ndarray] = None, npg: Optional[int] = None,
                     near_factor: float = float("inf"), device: int = 0):
    """``c u = sum g q - sum h T`` at points inside the domain, c = -sum h.

    ``pf`` is one point (2,) -> float, or points (m, 2) -> (m,). ``T``, ``q``
    default to the solution of the last ``solve`` (their first n entries are
    used). ``npg`` defaults to the near rule of the last assembly. With the
    default ``near_factor=inf`` every (point, element) pair goes through the
    tan-p3c near rule, as the reference's ``integrate_element`` does; a finite
    cutoff (e.g. 1.5) applies the assembly's far nodal lumping beyond
    ``near_factor * L`` instead, the value an internal row of H, G gives.
    Points must not lie on the boundary.
    """
    if T is None or q is None:
        if "T" not in dad.cache or "q" not in dad.cache:
            raise RuntimeError("cache.T is not set — assemble/solve first?")
        T = dad.T if T is None else T
        q = dad.q if q is None else q
    T = np.ascontiguousarray(np.asarray(T, dtype=np.float64)[: dad.n])
    q = np.ascontiguousarray(np.asarray(q, dtype=np.float64)[: dad.n])
    if T.shape != (dad.n,) or q.shape != (dad.n,):
        raise ArgumentError(f"T and q need at least n = {dad.n} entries")
    if npg is None:
        info = dad.cache.get("assembly_info")
        if info is None:
            raise RuntimeError("cache.qsi is not set — assemble/solve first?")
        npg = int(info["npg"])
    P = np.asarray(pf, dtype=np.float64)
    single = P.ndim == 1
    P = np.ascontiguousarray(P.reshape(-1, 2))
    nf = float(near_factor)
    if nf == float("inf"):
        nf = 1e300
    fresh = dad.hip is None
    ctx = _context(dad, device)
    if fresh:
        upload_geometry(dad, ctx)
    u = np.empty(P.shape[0], dtype=np.float64)
    ctx.check(lib().bemhip_interior_laplace(ctx.h, int(P.shape[0]), dptr(P), dptr(T), dptr(q),
                                            int(npg), nf, dptr(u)))
    return float(u[0]) if single else u
