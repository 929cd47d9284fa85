"""2-D elasticity (Kelvin) on the CPU: the oracle's restatement pinned by the
mathematics, and the host logic of the Elasticity path.

Pins of the oracle (oracle/bem_oracle.cpp, Kelvin section):
  kernel      Somigliana's identity on a finely sampled circle (interior source:
              int T = -I and u(xi) = int (U t - T u) for a uniform strain field;
              exterior source: int T = 0), symmetry of U;
  on-element  straight linear elements: the node sums of h and g and the first
              moment of h against closed forms (CPV and log integrals), each g
              against the interpolation rule applied to exact samples;
  near        off-element pairs against adaptive quadrature of the exact
              integrand;
  assembly    rigid translations in the null space of H, G's 2x2 blocks
              symmetric, the far blocks equal to the lumped kernel, the
              uniform-strain patch residual on a smooth boundary shrinking with
              h, and the reference's patch test (test/elasticity.jl:27-36).
"""

import math

import numpy as np
import pytest
from scipy import integrate

import bem_gmsh_amd as B
import oracle as O


def _circle(M=4000, R0=1.0):
    th = (np.arange(M) + 0.5) * 2 * np.pi / M
    xs = np.stack([R0 * np.cos(th), R0 * np.sin(th)], 1)
    return xs, xs / R0, 2 * np.pi * R0 / M


def _lame(E, nu):
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return mu, lam


@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_kelvin_somigliana_identity(nu):
    E = 2.5
    mu, lam = _lame(E, nu)
    eps = np.array([[0.01, 0.003], [0.003, -0.002]])
    sig = lam * np.trace(eps) * np.eye(2) + 2 * mu * eps
    xs, ns, ds = _circle()
    for xi, inside in ((np.array([0.2, -0.1]), True), (np.array([1.7, 0.4]), False)):
        IT = np.zeros((2, 2))
        Iu = np.zeros(2)
        for x, n in zip(xs, ns):
            U, T = O.kelvin(x - xi, n, mu, nu)
            assert U[0, 1] == U[1, 0]
            IT += T * ds
            Iu += (U @ (sig @ n) - T @ (eps @ x)) * ds
        if inside:
            assert np.abs(IT + np.eye(2)).max() < 1e-10
            assert np.abs(Iu - eps @ xi).max() < 1e-10
        else:
            assert np.abs(IT).max() < 1e-10
            assert np.abs(Iu).max() < 1e-10


def test_elasticity_properties_and_plane_stress():
    e = B.Elasticity(2.0, 0.25, 1.0)
    assert e.plane_strain and e.effective_nu() == 0.25
    assert e.mu == pytest.approx(2.0 / 2.5)
    assert e.lambda_ == pytest.approx(2.0 * 0.25 / (1.25 * 0.5))
    s = B.Elasticity(2.0, 0.25, 1.0, plane_stress=True)
    assert not s.plane_strain and s.effective_nu() == pytest.approx(0.2)
    assert s.mu == e.mu
    # lame_constants (src/Core/Structures.jl:156-165): E with the effective ratio
    assert s.lambda_ == pytest.approx(2.0 * 0.2 / (1.2 * 0.6))
    assert B.n_dof(e) == 2 and B.n_dof(B.Laplace(1.0)) == 1


def test_format2d_elasticity_bcs_and_pairs():
    assert B.parse_bc_pairs("0;0;1;2.5") == ([0, 1], [0.0, 2.5])
    assert B.parse_bc_pairs("Domain") == ([1], [0.0])
    with pytest.raises(ValueError):
        B.parse_bc_pairs("0;x;1;0")
    d = B.format2d(B.quadrado_elasticity(ndiv=5), B.Elasticity(1.0, 0.3), pontointerno=True)
    assert d.BC.shape == (2 * d.n,) and d.BV.shape == (2 * d.n,)
    left = np.abs(d.Nodes[:, 0]) < 1e-12
    bc = d.BC.reshape(-1, 2)
    assert np.all(bc[left] == 0) and np.all(bc[~left] == 1)
    assert d.ni == 16 and d.dof == 2
    # scalar names on an elasticity mesh: the missing second pair is (1, 0)
    m = B.quadrado(ndiv=4)
    d2 = B.format2d(m, B.Elasticity(1.0, 0.3), pontointerno=False)
    assert np.all(d2.BC.reshape(-1, 2)[:, 1] == 1)


def test_apply_analytical_bc_and_split_sol_per_dof():
    d = B.format2d(B.quadrado_elasticity(ndiv=4), B.Elasticity(1.0, 0.3), pontointerno=False)
    ana = B.ana_elasticity_patch(1.0, 0.3, 0.01, 0.002, 0.001)
    neu = [0, 5]
    B.apply_analytical_bc(d, ana, neu)
    u = ana.u(d.Nodes).reshape(-1)
    t = ana.q(d.Nodes, d.Normal).reshape(-1)
    m = np.zeros(d.n, bool)
    m[neu] = True
    m2 = np.repeat(m, 2)
    assert np.array_equal(d.BC, np.where(m2, 1, 0))
    assert np.allclose(d.BV, np.where(m2, t, u))
    x = np.arange(2 * d.nt, dtype=float)
    uu, tt = B.split_sol(d, x)
    assert np.array_equal(uu, np.where(m2, x[:2 * d.n], d.BV))
    assert np.array_equal(tt, np.where(m2, d.BV, x[:2 * d.n]))


def _straight_element_closed_forms(L, a, t, n, mu, nu):
    prod1 = 4 * math.pi * (1 - nu)
    base = 2 * prod1 * mu
    c34, c12 = 3 - 4 * nu, 1 - 2 * nu
    s1, s2 = (a + 1) * L / 2, (1 - a) * L / 2
    Ilog = -(s2 * math.log(s2) - s2 + s1 * math.log(s1) - s1)
    Gs = (c34 * Ilog * np.eye(2) + (s1 + s2) * np.outer(t, t)) / base
    A = np.outer(t, n) - np.outer(n, t)
    C = c12 / (4 * math.pi * (1 - nu))
    Hs = C * A * math.log(s2 / s1)
    H1 = C * A * (a * math.log(s2 / s1) + (2 / L) * (s1 + s2))
    return Gs, Hs, H1


@pytest.mark.parametrize("nu", [0.2, 0.3])
def test_kelvin_on_element_straight_closed_forms(nu):
    d = B.format2d(B.quadrado_elasticity(ndiv=6, ordem=1, Lx=1.3, Ly=0.7),
                   B.Elasticity(1.7, nu), pontointerno=False)
    mu = d.properties.mu
    xk = np.array([-1 / math.sqrt(3), 1 / math.sqrt(3)])
    for e in (0, 6, 13):
        X = d.elem_geo[e]
        L = np.linalg.norm(X[1] - X[0])
        t = (X[1] - X[0]) / L
        n = d.Normal[2 * e]
        for k in range(2):
            i = 2 * e + k
            h, g = O.pair_kelvin(d, i, e, npg=12)
            Gs, Hs, H1 = _straight_element_closed_forms(L, xk[k], t, n, mu, nu)
            assert np.abs(h.sum(0) - Hs).max() <= 1e-11 * np.abs(Hs).max()
            assert np.abs((xk[:, None, None] * h).sum(0) - H1).max() <= 1e-11 * np.abs(Hs).max()
            assert np.abs(g.sum(0) - Gs).max() <= 1e-4 * np.abs(Gs).max()
            # each g: the interpolation rule on the exact integrand's samples
            gi = _interp_rule_kelvin(d, i, e, mu, nu)
            assert np.abs(g - gi).max() <= 1e-14 * np.abs(Gs).max()


def _interp_rule_kelvin(d, i, e, mu, nu):
    p = d.p
    nn = p + 1
    geo = d.elem_geo[e]
    src = d.collocation[i]
    xn = d.collocation[e * nn:(e + 1) * nn]
    a, _ = O.closest_point(0, p, xn, src)
    a = min(max(a, np.nextafter(-1.0, 0.0)), np.nextafter(1.0, 0.0))
    nref = d.Normal[e * nn]
    out = np.empty((nn, 2, 2))
    for k in range(nn):
        for c in range(4):
            def f(xi, k=k, c=c):
                Ng, dNg = O.shapefun(1, p, xi)
                pg, dx = Ng @ geo, dNg @ geo
                J = math.hypot(dx[0], dx[1])
                nrm = np.array([dx[1] / J, -dx[0] / J])
                if nrm @ nref < 0:
                    nrm = -nrm
                U, _ = O.kelvin(pg - src, nrm, mu, nu)
                N, _ = O.shapefun(0, p, xi)
                return U.reshape(-1)[c] * N[k] * J
            out[k, c // 2, c % 2] = O.interp_integral(f, a, 0)[0]
    return out


def _exact_pair_kelvin(d, i, e, mu, nu):
    p = d.p
    nn = p + 1
    geo = d.elem_geo[e]
    src = d.collocation[i]
    nref = d.Normal[e * nn]
    ts = np.linspace(-1, 1, 4001)
    pts = np.array([O.shapefun(1, p, t)[0] @ geo for t in ts])
    tb = ts[np.argmin(np.hypot(pts[:, 0] - src[0], pts[:, 1] - src[1]))]

    def integrand(xi, k, c, which):
        Ng, dNg = O.shapefun(1, p, xi)
        pg, dx = Ng @ geo, dNg @ geo
        J = math.hypot(dx[0], dx[1])
        nrm = np.array([dx[1] / J, -dx[0] / J])
        if nrm @ nref < 0:
            nrm = -nrm
        U, T = O.kelvin(pg - src, nrm, mu, nu)
        N, _ = O.shapefun(0, p, xi)
        K = U if which == "g" else T
        return K.reshape(-1)[c] * N[k] * J

    h = np.empty((nn, 2, 2))
    g = np.empty((nn, 2, 2))
    for k in range(nn):
        for c in range(4):
            kw = dict(points=[tb], limit=400, epsabs=1e-15, epsrel=1e-13)
            g[k, c // 2, c % 2] = integrate.quad(integrand, -1, 1, args=(k, c, "g"), **kw)[0]
            h[k, c // 2, c % 2] = integrate.quad(integrand, -1, 1, args=(k, c, "h"), **kw)[0]
    return h, g


def _collapsed_half(d, i, e):
    """True when the tan-p3c rule of pair (i, e) splits at a tangent parent point
    within 1e-14 of an end: the reference's _tanp3c_half then returns the full
    tangent rule for the empty half (src/Core/Integration.jl:1344-1349), so the
    element is integrated twice. This happens for rows on the extension of an
    element that ends at a corner (complex pole within rounding of +-1)."""
    nn = d.p + 1
    zeta, eta = O.complex_pole(0, d.p, d.collocation[e * nn:(e + 1) * nn], d.collocation[i])
    eta = max(eta, 1e-14)
    ap, am = math.atan((1 - zeta) / eta), math.atan((1 + zeta) / eta)
    At = 2.0 / (ap + am)
    Bt = -At * (ap - am) / 2.0
    if not (abs(zeta) < 1 and -1 < Bt < 1):
        return False
    return (Bt + 1) / 2 <= 1e-14 or (1 - Bt) / 2 <= 1e-14


@pytest.mark.parametrize("builder", ["square", "ellipse"])
def test_kelvin_near_pairs_match_exact_integrals(builder):
    if builder == "square":
        d = B.format2d(B.quadrado_elasticity(ndiv=9), B.Elasticity(1.0, 0.3), pontointerno=True)
    else:
        d = B.format2d(B.ellipse_mesh(40), B.Elasticity(1.0, 0.3))
    mu, nu = d.properties.mu, d.properties.effective_nu()
    nn = d.p + 1
    rng = np.random.default_rng(0)
    checked = doubled = 0
    for e in rng.choice(d.ne, 4, replace=False):
        xg = d.collocation[e * nn:(e + 1) * nn]
        rows = [i for i in range(d.nt) if not (e * nn <= i < (e + 1) * nn)
                and O.near_element(d.collocation[i], xg, d.elem_L[e])]
        for i in rows[:4]:
            h, g = O.pair_kelvin(d, i, e, npg=16)
            he, ge = _exact_pair_kelvin(d, i, e, mu, nu)
            f = 2.0 if _collapsed_half(d, i, e) else 1.0
            doubled += f == 2.0
            assert np.abs(h - f * he).max() <= 2e-7 * max(np.abs(he).max(), 1e-3)
            assert np.abs(g - f * ge).max() <= 2e-8 * np.abs(ge).max()
            checked += 1
    assert checked >= 6
    if builder == "square":
        assert doubled >= 1  # the square's corners exercise the collapsed half


def test_kelvin_assembly_structure():
    d = B.format2d(B.quadrado_elasticity(ndiv=7, ordem=2), B.Elasticity(1.0, 0.3),
                   pontointerno=True)
    H, G = O.assemble_kelvin(d, 12)
    nt, n = d.nt, d.n
    assert H.shape == (2 * nt, 2 * nt) and G.shape == (2 * nt, 2 * n)
    # rigid translations: H tx = H ty = 0 (the free term)
    for b in range(2):
        tr = np.zeros(2 * nt)
        tr[b::2] = 1.0
        assert np.abs(H @ tr).max() <= 1e-13 * np.abs(H).max()
    # U is symmetric, so every 2x2 block of G is
    Gb = G.reshape(nt, 2, n, 2)
    assert np.array_equal(Gb[:, 0, :, 1], Gb[:, 1, :, 0])
    # internal-node columns of H: zero but the diagonal blocks
    Hb = H.reshape(nt, 2, nt, 2)
    for j in range(n, nt):
        col = Hb[:, :, j, :].copy()
        col[j] = 0.0
        assert not col.any()
    # a far pair: the block is the lumped kernel T * J_k w_k (index layout)
    nn = d.p + 1
    i, e = 0, d.ne // 2
    xg = d.collocation[e * nn:(e + 1) * nn]
    assert not O.near_element(d.collocation[i], xg, d.elem_L[e])
    for k in range(nn):
        j = e * nn + k
        U, T = O.kelvin(d.collocation[j] - d.collocation[i], d.Normal[j], d.properties.mu,
                        d.properties.effective_nu())
        w = d.node_J()[j] * d.elem_weight[k]
        assert np.array_equal(Hb[i, :, j, :], T * w)
        assert np.array_equal(Gb[i, :, j, :], U * w)


def test_kelvin_patch_residual_on_smooth_boundary():
    ana = B.ana_elasticity_patch(1.0, 0.3, 0.01, 0.004, 0.002)
    res = []
    for ne in (100, 200):
        d = B.format2d(B.ellipse_mesh(ne), B.Elasticity(1.0, 0.3), pontointerno=False)
        H, G = O.assemble_kelvin(d, 16)
        u = ana.u(d.Nodes).reshape(-1)
        t = ana.q(d.Nodes, d.Normal).reshape(-1)
        res.append(np.abs(H @ u - G @ t).max() / np.abs(G @ t).max())
        # mixed BCs: every other node traction-loaded
        B.apply_analytical_bc(d, ana, range(0, d.n, 2))
        uu, tt, _ = O.solve_kelvin(d, H, G)
        assert np.linalg.norm(uu - u) / np.linalg.norm(u) < 1e-3
        assert np.linalg.norm(tt - t) / np.linalg.norm(t) < 0.05
    assert res[0] < 2e-3 and res[1] < 0.7 * res[0]


def test_reference_elasticity_patch():
    # test/elasticity.jl:27-36: quadrado_elasticity(ndiv=8), Elasticity(1, 0.3, 1),
    # patch with exx = 0.01, Dirichlet everywhere, assemble!(dad, 12), rel_error < 0.15
    d = B.format2d(B.quadrado_elasticity(ndiv=8), B.Elasticity(1.0, 0.3, 1.0),
                   pontointerno=False)
    ana = B.ana_elasticity_patch(E=1.0, nu=0.3, exx=0.01)
    B.apply_analytical_bc(d, ana)
    H, G = O.assemble_kelvin(d, 12)
    u, t, x = O.solve_kelvin(d, H, G)
    d.set_cache(u=u, traction=t, T=u)
    assert B.rel_error(d) < 0.15
    assert np.isfinite(B.rel_error_flux(d))
