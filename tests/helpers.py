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


def oracle_near_pairs(dad, near_factor=1.5):
    """(row, element) pairs the oracle integrates (near or on-element)."""
    import oracle as O
    nn = dad.p + 1
    out = set()
    for e in range(dad.ne):
        xg = dad.collocation[e * nn:(e + 1) * nn]
        for i in range(dad.nt):
            if (e * nn <= i < (e + 1) * nn) or O.near_element(dad.collocation[i], xg,
                                                              dad.elem_L[e], near_factor):
                out.add((i, e))
    return out


def pair_ulp_sensitivity(dad, i, e, npg, return_values=False):
    """Largest relative change of the oracle's (h, g) for pair (i, e) when one
    coordinate of the source point or of the element's collocation nodes moves
    by one ulp: the pair's intrinsic FP64 conditioning."""
    import copy
    import oracle as O
    h0, g0 = O.pair(dad, i, e, npg)
    nn = dad.p + 1
    worst = 0.0
    points = [i] + [e * nn + k for k in range(nn)]  # source, element collocation nodes
    for pt in points:
        for comp in (0, 1):
            for sgn in (1.0, -1.0):
                d2 = copy.copy(dad)
                d2.collocation = dad.collocation.copy()
                d2.collocation[pt, comp] = np.nextafter(d2.collocation[pt, comp], sgn * np.inf)
                h1, g1 = O.pair(d2, i, e, npg)
                for a, b in ((h1, h0), (g1, g0)):
                    scale = np.maximum(np.abs(b), 1e-300)
                    rel = float(np.max(np.abs(a - b) / scale))
                    if rel > worst:
                        worst, hw, gw = rel, h1, g1
    if return_values:
        return (worst, hw, gw) if worst > 0 else (0.0, h0, g0)
    return worst


def fp64_parity(dad, H, G, Hr, Gr, npg, tol=1e-12):
    """Per-entry parity report of an FP64 assembly against the oracle.

    Far, on-element and diagonal entries: relative error <= tol (entries that
    cancel to ~0 measured against 1e-3 of their row's largest entry). Near
    (off-element) pairs: <= tol, or <= 10x the oracle's own change under a
    1-ulp shift of the source point for pairs whose FP64 result is that
    sensitive. Returns a dict of the worst errors per class."""
    nn = dad.p + 1
    pairs = oracle_near_pairs(dad)
    rowmax_h = np.abs(Hr).max(axis=1, keepdims=True)
    rowmax_g = np.abs(Gr).max(axis=1, keepdims=True)
    EH = np.abs(H - Hr) / np.maximum(np.abs(Hr), 1e-3 * rowmax_h)
    EG = np.abs(G - Gr) / np.maximum(np.abs(Gr), 1e-3 * rowmax_g)
    near_mask = np.zeros(G.shape, dtype=bool)
    onel_mask = np.zeros(G.shape, dtype=bool)
    for (i, e) in pairs:
        sl = slice(e * nn, (e + 1) * nn)
        if e * nn <= i < (e + 1) * nn:
            onel_mask[i, sl] = True
        else:
            near_mask[i, sl] = True
    far_mask = ~(near_mask | onel_mask)
    diag = np.zeros(H.shape, dtype=bool)
    np.fill_diagonal(diag, True)
    EHb = EH[:, :dad.n]
    rep = {
        "far": max(float(EHb[far_mask & ~diag[:, :dad.n]].max(initial=0)),
                   float(EG[far_mask].max(initial=0))),
        "onel": max(float(EHb[onel_mask & ~diag[:, :dad.n]].max(initial=0)),
                    float(EG[onel_mask].max(initial=0))),
        "diag": float(np.abs(np.diag(H) - np.diag(Hr)).max() / np.abs(np.diag(Hr)).max()),
        "near": max(float(EHb[near_mask].max(initial=0)), float(EG[near_mask].max(initial=0))),
        "frob_H": frob_rel(H, Hr), "frob_G": frob_rel(G, Gr),
    }
    # near pairs above tol: compare with their own 1-ulp sensitivity
    bad = []
    worst_ratio = 0.0
    n_sens = 0
    for (i, e) in pairs:
        if e * nn <= i < (e + 1) * nn:
            continue
        err = max(EHb[i, e * nn:(e + 1) * nn].max(), EG[i, e * nn:(e + 1) * nn].max())
        if err <= tol:
            continue
        kappa = pair_ulp_sensitivity(dad, i, e, npg)
        n_sens += 1
        worst_ratio = max(worst_ratio, err / max(kappa, 1e-300))
        if err > 10.0 * kappa:
            bad.append((i, e, err, kappa))
    rep.update({"near_pairs_above_tol": n_sens, "near_err_over_ulp_sensitivity": worst_ratio,
                "near_violations": bad[:10]})
    return rep


def solution_noise_floor(dad, npg, tol=1e-12):
    """Relative change of the oracle's (T, q) when each near pair whose FP64
    entries move by more than `tol` under a 1-ulp input shift takes its
    worst-case perturbed values (classification unchanged): the solution-level
    noise floor of the reference algorithm on this mesh."""
    import oracle as O
    nn = dad.p + 1
    Hr, Gr = O.assemble(dad, npg)
    T0, q0, _ = O.solve(dad, Hr, Gr)
    Hp, Gp = Hr.copy(order="F"), Gr.copy(order="F")
    changed = 0
    for (i, e) in oracle_near_pairs(dad):
        if e * nn <= i < (e + 1) * nn:
            continue
        kappa, h1, g1 = pair_ulp_sensitivity(dad, i, e, npg, return_values=True)
        if kappa > tol:
            sl = slice(e * nn, (e + 1) * nn)
            Hp[i, sl] = h1
            Gp[i, sl] = g1
            changed += 1
    if changed:  # rigid-body diagonal of the perturbed H
        np.fill_diagonal(Hp, 0.0)
        np.fill_diagonal(Hp, -Hp.sum(axis=1))
    T1, q1, _ = O.solve(dad, Hp, Gp)
    return (float(np.linalg.norm(T1 - T0) / np.linalg.norm(T0)),
            float(np.linalg.norm(q1 - q0) / max(np.linalg.norm(q0), 1e-300)), changed)
