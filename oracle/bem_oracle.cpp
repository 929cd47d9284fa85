// FP64 CPU oracle for the dense 2-D Laplace BEM path of BEM_gmsh.
//
// TEST INFRASTRUCTURE ONLY. This file restates, in plain C++ on the CPU, the
// reference's FP64 algorithm for
//   assemble!(dad, npg)  (H_G_full_direct)  and  solve(dad)  (applyBC + LU)
// so that the HIP library can be checked against it. Nothing in the product
// (bem_gmsh_amd/) links, loads or calls this code; only tests/, the smoke check
// and the CPU-baseline leg of bench.py do.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root). Arithmetic is FP64 with the operation order
// of the Julia source; the build uses -ffp-contract=off so that a*b+c is never
// fused (Julia/LLVM does not fuse either without @fastmath).
//
// Build: see oracle/Makefile (produces oracle/liboracle.so).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.141592653589793;
const double kEps = 2.220446049250313e-16;  // eps(Float64)

// ---------------------------------------------------------------------------
// Gauss-Legendre rule (FastGaussQuadrature.gausslegendre, called from
// src/Core/Assembly_full.jl:720, src/Core/Input.jl:88,
// src/Core/Interpolation.jl:67, src/Core/Integration.jl:1599).
// Restated from the published definition: the n roots of P_n found by Newton
// on the three-term recurrence, weights 2/((1-x^2) P_n'(x)^2), ascending.
// ---------------------------------------------------------------------------
void legendre_eval(int n, double x, double* pn, double* dpn) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { *pn = 1.0; *dpn = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

void gausslegendre(int n, double* x, double* w) {
  if (n == 1) { x[0] = 0.0; w[0] = 2.0; return; }
  int m = (n + 1) / 2;
  for (int i = 0; i < m; ++i) {
    // i-th root from the right end, initial guess (Tricomi / Francesca).
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double pn = 0.0, dpn = 0.0;
    for (int it = 0; it < 100; ++it) {
      legendre_eval(n, z, &pn, &dpn);
      double dz = pn / dpn;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    legendre_eval(n, z, &pn, &dpn);
    double wi = 2.0 / ((1.0 - z * z) * dpn * dpn);
    x[n - 1 - i] = z;
    x[i] = -z;
    w[n - 1 - i] = wi;
    w[i] = wi;
  }
  if (n % 2 == 1) {
    double pn, dpn;
    x[m - 1] = 0.0;
    legendre_eval(n, 0.0, &pn, &dpn);
    w[m - 1] = 2.0 / (dpn * dpn);
  }
}

// ---------------------------------------------------------------------------
// Barycentric Lagrange bases (src/Core/Interpolation.jl).
// ---------------------------------------------------------------------------
struct Poly {
  int N = 0;                 // degree
  std::vector<double> x;     // nodes (N+1)
  std::vector<double> bw;    // barycentric weights (N+1)
  std::vector<double> D;     // differentiation matrix, row-major (N+1)^2
};

// diff_matrix (src/Core/Interpolation.jl:193-213)
void build_diff_matrix(Poly& P) {
  int n = P.N + 1;
  P.D.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double dsum = 0.0;
    for (int j = 0; j < i; ++j) {
      double t = (P.bw[j] / P.bw[i]) / (P.x[i] - P.x[j]);
      P.D[i * n + j] = t;
      dsum += t;
    }
    for (int j = i + 1; j < n; ++j) {
      double t = (P.bw[j] / P.bw[i]) / (P.x[i] - P.x[j]);
      P.D[i * n + j] = t;
      dsum += t;
    }
    P.D[i * n + i] = -dsum;
  }
}

// Legendre(p): nodes gausslegendre(p+1), weights (2*isodd(i)-1)*sqrt((1-x^2) w)
// with 1-based i (src/Core/Interpolation.jl:62-71).
Poly legendre_poly(int p) {
  Poly P;
  P.N = p;
  P.x.resize(p + 1);
  std::vector<double> w(p + 1);
  gausslegendre(p + 1, P.x.data(), w.data());
  P.bw.resize(p + 1);
  for (int i = 0; i <= p; ++i) {
    double s = ((i + 1) % 2 == 1) ? 1.0 : -1.0;
    P.bw[i] = s * std::sqrt((1.0 - P.x[i] * P.x[i]) * w[i]);
  }
  build_diff_matrix(P);
  return P;
}

double binomial(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return std::round(r);
}

// Equispaced(p): nodes range(-1,1,length=p+1), weights eq. (5.1)
// (src/Core/Interpolation.jl:98, :127).
Poly equispaced_poly(int p) {
  Poly P;
  P.N = p;
  P.x.resize(p + 1);
  P.bw.resize(p + 1);
  for (int j = 0; j <= p; ++j) {
    // Julia range(-1, 1, length=p+1) is exact at the ends and the centre.
    P.x[j] = (p == 0) ? -1.0 : (-1.0 * (p - j) + 1.0 * j) / p;
    bool odd_n = (p % 2) == 1, even_j = (j % 2) == 0;
    double s = (odd_n != even_j) ? 1.0 : -1.0;
    P.bw[j] = s * binomial(p, j);
  }
  build_diff_matrix(P);
  return P;
}

// ArbitraryPolynomial weights 1/prod_{j!=i}(x_i - x_j)
// (src/Core/Interpolation.jl:104-116).
std::vector<double> arbitrary_weights(const std::vector<double>& x) {
  std::vector<double> w(x.size());
  for (size_t i = 0; i < x.size(); ++i) {
    double wi = 1.0;
    for (size_t j = 0; j < i; ++j) wi *= x[i] - x[j];
    for (size_t j = i + 1; j < x.size(); ++j) wi *= x[i] - x[j];
    w[i] = 1.0 / wi;
  }
  return w;
}

// interpolation_matrix row (src/Core/Interpolation.jl:148-178).
void interp_row(const Poly& P, double xx, double* M) {
  int n = P.N + 1;
  double msum = 0.0;
  int exact = -1;
  for (int i = 0; i < n; ++i) {
    if (xx == P.x[i]) exact = i;
    M[i] = P.bw[i] / (xx - P.x[i]);
    msum += M[i];
  }
  if (msum == 0.0) {
    for (int i = 0; i < n; ++i) M[i] = 0.0;
  } else if (exact >= 0) {
    for (int i = 0; i < n; ++i) M[i] = 0.0;
    M[exact] = 1.0;
  } else {
    for (int i = 0; i < n; ++i) M[i] /= msum;
  }
}

// shapefun: (L, L*Dmat) (src/Core/Interpolation.jl:215-218).
void shapefun(const Poly& P, double xx, double* N, double* dN) {
  int n = P.N + 1;
  interp_row(P, xx, N);
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += N[i] * P.D[i * n + j];
    dN[j] = s;
  }
}

// ---------------------------------------------------------------------------
// Laplace kernel (src/Laplace/Fundamental.jl:19-25): r = x_field - x_source.
// ---------------------------------------------------------------------------
inline void fundamental(double rx, double ry, double nx, double ny, double k,
                        double* U, double* T) {
  double R2 = rx * rx + ry * ry;
  double inv2pi = 1.0 / (2.0 * kPi);
  *U = -std::log(R2) * ((0.5 * inv2pi) / k);
  *T = (rx * nx + ry * ny) * (inv2pi / R2);
}

inline double norm2d(double x, double y) { return std::sqrt(x * x + y * y); }

// ---------------------------------------------------------------------------
// Near/far classification (src/Core/Assembly_full.jl:736-754,
// src/Core/Integration.jl:238-312).
// ---------------------------------------------------------------------------
double dist_chord(double px, double py, double ax, double ay, double bx,
                  double by) {
  double abx = bx - ax, aby = by - ay;
  double t = ((px - ax) * abx + (py - ay) * aby) / ((abx * abx + aby * aby) + kEps);
  t = std::min(std::max(t, 0.0), 1.0);
  double qx = ax + t * abx, qy = ay + t * aby;
  return norm2d(px - qx, py - qy);
}

bool near_element(double px, double py, const double* xg, int nn, double L,
                  double factor) {
  if (factor == INFINITY || factor > 1e20) return true;
  if (nn == 0) return true;
  double lim = factor * L;
  double lim2 = lim * lim;
  double dn2 = INFINITY;
  double lox = xg[0], loy = xg[1], hix = xg[0], hiy = xg[1];
  for (int k = 0; k < nn; ++k) {
    double qx = xg[2 * k], qy = xg[2 * k + 1];
    lox = std::min(lox, qx); loy = std::min(loy, qy);
    hix = std::max(hix, qx); hiy = std::max(hiy, qy);
    double dx = px - qx, dy = py - qy;
    double d2 = dx * dx + dy * dy;
    if (d2 < dn2) dn2 = d2;
  }
  if (dn2 < lim2) return true;
  // _dist_aabb
  double d2 = 0.0;
  if (px < lox) { double d = lox - px; d2 += d * d; }
  else if (px > hix) { double d = px - hix; d2 += d * d; }
  if (py < loy) { double d = loy - py; d2 += d * d; }
  else if (py > hiy) { double d = py - hiy; d2 += d * d; }
  if (std::sqrt(d2) >= lim) return false;
  // _dist_element (2-D: polyline through the nodes)
  double d;
  if (nn == 1) d = norm2d(px - xg[0], py - xg[1]);
  else {
    d = INFINITY;
    for (int k = 0; k + 1 < nn; ++k)
      d = std::min(d, dist_chord(px, py, xg[2 * k], xg[2 * k + 1], xg[2 * k + 2], xg[2 * k + 3]));
  }
  return d < lim;
}

// ---------------------------------------------------------------------------
// Closest point (src/Core/Integration.jl:323-349, seed :1005-1010,
// residual :1023-1033).
// ---------------------------------------------------------------------------
double seed_1d(const Poly& P, const double* nodes, int nn, double px, double py) {
  double dx = nodes[2 * (nn - 1)] - nodes[0], dy = nodes[2 * (nn - 1) + 1] - nodes[1];
  double nd = norm2d(dx, dy);
  double xi0 = (P.x[nn - 1] - P.x[0]) * (dx * (px - nodes[0]) + dy * (py - nodes[1])) /
                   (nd * nd + kEps) + P.x[0];
  return std::min(std::max(xi0, -1.0), 1.0);
}

void curve_point(const Poly& P, const double* nodes, double xi, double* x, double* y,
                 double* dx, double* dy, double* d2x, double* d2y) {
  int n = P.N + 1;
  double N[16], dN[16], d2N[16];
  shapefun(P, xi, N, dN);
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += dN[i] * P.D[i * n + j];
    d2N[j] = s;
  }
  double ax = 0, ay = 0, bx = 0, by = 0, cx = 0, cy = 0;
  for (int k = 0; k < n; ++k) {
    ax += N[k] * nodes[2 * k]; ay += N[k] * nodes[2 * k + 1];
    bx += dN[k] * nodes[2 * k]; by += dN[k] * nodes[2 * k + 1];
    cx += d2N[k] * nodes[2 * k]; cy += d2N[k] * nodes[2 * k + 1];
  }
  *x = ax; *y = ay;
  if (dx) { *dx = bx; *dy = by; }
  if (d2x) { *d2x = cx; *d2y = cy; }
}

void closest_point_1d(const Poly& P, const double* nodes, double px, double py,
                      double xi0, double* xi_out, double* dist_out) {
  double xi = xi0;
  for (int it = 0; it < 25; ++it) {
    double x, y, dx, dy, d2x, d2y;
    curve_point(P, nodes, xi, &x, &y, &dx, &dy, &d2x, &d2y);
    double rx = x - px, ry = y - py;
    double res = rx * dx + ry * dy;
    double J = (dx * dx + dy * dy) + (rx * d2x + ry * d2y);
    if (std::fabs(J) < 1e-30) break;
    double d = res / J;
    xi -= d;
    if (std::fabs(d) < 1e-14) break;
  }
  double x, y;
  if (xi < -1.0 || xi > 1.0) {
    double x0, y0, x1, y1;
    curve_point(P, nodes, -1.0, &x0, &y0, nullptr, nullptr, nullptr, nullptr);
    curve_point(P, nodes, 1.0, &x1, &y1, nullptr, nullptr, nullptr, nullptr);
    if (norm2d(x0 - px, y0 - py) <= norm2d(x1 - px, y1 - py)) { xi = -1.0; x = x0; y = y0; }
    else { xi = 1.0; x = x1; y = y1; }
  } else {
    curve_point(P, nodes, xi, &x, &y, nullptr, nullptr, nullptr, nullptr);
  }
  *xi_out = xi;
  *dist_out = norm2d(x - px, y - py);
}

// ---------------------------------------------------------------------------
// Complex pole of z(xi) = x1 + i x2 (src/Core/Integration.jl:1117-1160).
// ---------------------------------------------------------------------------
using cplx = std::complex<double>;

void complex_pole_1d(const Poly& P, const double* nodes, int nn, double px, double py,
                     double* zeta, double* eta) {
  cplx y(px, py);
  if (nn == 2 || nn == 3) {
    const double* xs = P.x.data();
    cplx Z[3];
    for (int k = 0; k < nn; ++k) Z[k] = cplx(nodes[2 * k], nodes[2 * k + 1]);
    if (nn == 2) {
      double d = xs[1] - xs[0] + 1e-30;
      cplx B((Z[1] - Z[0]).real() / d, (Z[1] - Z[0]).imag() / d);
      cplx C = Z[0] - B * xs[0];
      if (std::abs(B) < 1e-30) { *zeta = 0.0; *eta = 1.0; return; }
      cplx r = (y - C) / B;
      *zeta = r.real();
      *eta = std::fabs(r.imag());
      return;
    }
    // _fit_quad_complex: Vandermonde [xi^2 xi 1] c = Z, solved by Cramer's rule.
    double a11 = xs[0] * xs[0], a12 = xs[0], a21 = xs[1] * xs[1], a22 = xs[1],
           a31 = xs[2] * xs[2], a32 = xs[2];
    double det = a11 * (a22 - a32) - a12 * (a21 - a31) + (a21 * a32 - a22 * a31);
    cplx A = (Z[0] * (a22 - a32) - a12 * (Z[1] - Z[2]) + (Z[1] * a32 - a22 * Z[2])) / det;
    cplx B = (a11 * (Z[1] - Z[2]) - Z[0] * (a21 - a31) + (a21 * Z[2] - Z[1] * a31)) / det;
    cplx C = (a11 * (a22 * Z[2] - Z[1] * a32) - a12 * (a21 * Z[2] - Z[1] * a31) +
              Z[0] * (a21 * a32 - a22 * a31)) / det;
    if (std::abs(A) < 1e-14) {
      if (std::abs(B) < 1e-30) { *zeta = 0.0; *eta = 1.0; return; }
      cplx r = (y - C) / B;
      *zeta = r.real();
      *eta = std::fabs(r.imag());
      return;
    }
    cplx disc = std::sqrt(B * B - 4.0 * A * (C - y));
    cplx r1 = (-B + disc) / (2.0 * A);
    cplx r2 = (-B - disc) / (2.0 * A);
    cplx r = std::fabs(r1.imag()) <= std::fabs(r2.imag()) ? r1 : r2;
    *zeta = r.real();
    *eta = std::fabs(r.imag());
    return;
  }
  double a, dist;
  closest_point_1d(P, nodes, px, py, seed_1d(P, nodes, nn, px, py), &a, &dist);
  double N[16], dN[16];
  shapefun(P, a, N, dN);
  double dx = 0, dy = 0;
  for (int k = 0; k < nn; ++k) { dx += dN[k] * nodes[2 * k]; dy += dN[k] * nodes[2 * k + 1]; }
  *zeta = a;
  *eta = dist / std::max(norm2d(dx, dy), 1e-14);
}

// ---------------------------------------------------------------------------
// Granados near-singular maps (src/Core/Integration.jl:1163-1365).
// A rule is a list of (node, weight) pairs on [-1, 1].
// ---------------------------------------------------------------------------
struct Rule { std::vector<double> x, w; };

// _eq41_collinear (:1163-1168)
bool eq41_collinear(double zeta, double eta, bool has_den, double den) {
  if (std::fabs(zeta) <= 1.0) return false;
  if (std::fabs(eta) < 1e-8) return true;
  if (has_den && std::fabs(den) < 1e-10) return true;
  return false;
}

// _eq41_real_pole (:1203-1217)
Rule eq41_real_pole(const Rule& u, double zeta) {
  if (std::fabs(zeta) <= 1.0) return u;
  double A = zeta * zeta - 1.0;
  double B = -zeta;
  Rule r;
  r.x.resize(u.x.size()); r.w.resize(u.x.size());
  for (size_t i = 0; i < u.x.size(); ++i) {
    double d = u.x[i] - B;
    r.x[i] = zeta - A / d;
    r.w[i] = u.w[i] * (A / (d * d));
  }
  return r;
}

// Mapped nodes whose tangent angle hit the +-(pi/2 - 1e-8) clamp of
// _tangenttrans / _tanp3c_compose / _tanp3c_half since the last reset (a
// diagnostic of oracle_pair_plan; the arithmetic does not depend on it).
thread_local long g_clamped = 0;

// _tangenttrans (:1177-1200)
Rule tangenttrans(const Rule& u, double zeta, double eta) {
  if (eq41_collinear(zeta, eta, false, 0.0)) return eq41_real_pole(u, zeta);
  eta = std::max(eta, 1e-14);
  double tp = std::atan((1.0 - zeta) / eta);
  double tm = std::atan((-1.0 - zeta) / eta);
  double den = tp - tm;
  if (eq41_collinear(zeta, eta, true, den)) return eq41_real_pole(u, zeta);
  double A = (2.0 * eta) / den;
  double B = -(tp + tm) / den;
  double halfpi = kPi / 2.0 - 1e-8;
  Rule r;
  r.x.resize(u.x.size()); r.w.resize(u.x.size());
  for (size_t i = 0; i < u.x.size(); ++i) {
    const double raw = (eta * (u.x[i] - B)) / A;
    double phi = std::min(std::max(raw, -halfpi), halfpi);
    if (phi != raw) ++g_clamped;
    double tn = std::tan(phi);
    r.x[i] = zeta + eta * tn;
    r.w[i] = u.w[i] * (((eta * eta) / A) * (1.0 + tn * tn));
  }
  return r;
}

// _p3c_inv_pole (:1220-1266)
void p3c_inv_pole(double zeta, double eta, double* zt_out, double* et_out) {
  if (std::fabs(eta) < 1e-14) {
    double x = zeta;
    for (int it = 0; it < 25; ++it) {
      double g = (((x * x * x) - (3.0 * zeta) * (x * x)) + 3.0 * x) - zeta;
      double gp = ((3.0 * (x * x)) - (6.0 * zeta) * x) + 3.0;
      if (std::fabs(gp) < 1e-30) break;
      x -= g / gp;
      if (std::fabs(g) < 1e-14) break;
    }
    *zt_out = x;
    *et_out = 0.0;
    return;
  }
  double sgn = (eta > 0) ? 1.0 : ((eta < 0) ? -1.0 : 0.0);
  double et = std::cbrt(std::fabs(eta)) * sgn;
  for (int it = 0; it < 20; ++it) {
    double g = ((2.0 * (et * et * et)) - (3.0 * eta) * (et * et)) - eta;
    double gp = (6.0 * (et * et)) - (6.0 * eta) * et;
    if (std::fabs(gp) < 1e-30) break;
    et -= g / gp;
  }
  et += std::cbrt(eta) * std::sqrt(zeta * zeta + eta * eta);  // eq. (86)
  double zt = zeta;
  for (int it = 0; it < 30; ++it) {
    double den = (1.0 + 3.0 * (zt * zt)) + 3.0 * (et * et);
    double nz = ((3.0 + zt * zt) + 3.0 * (et * et)) * zt;
    double ne = 2.0 * (et * et * et);
    double f1 = nz / den - zeta;
    double f2 = ne / den - eta;
    if (std::fabs(f1) + std::fabs(f2) < 1e-14) break;
    double denz = 6.0 * zt;
    double dene = 6.0 * et;
    double nz_z = (3.0 + 3.0 * (zt * zt)) + 3.0 * (et * et);
    double nz_e = (6.0 * et) * zt;
    double ne_z = 0.0;
    double ne_e = 6.0 * (et * et);
    double den2 = den * den;
    double j11 = (nz_z * den - nz * denz) / den2;
    double j12 = (nz_e * den - nz * dene) / den2;
    double j21 = (ne_z * den - ne * denz) / den2;
    double j22 = (ne_e * den - ne * dene) / den2;
    double detJ = j11 * j22 - j12 * j21;
    if (std::fabs(detJ) < 1e-30) break;
    zt -= (j22 * f1 - j12 * f2) / detJ;
    et -= (-j21 * f1 + j11 * f2) / detJ;
  }
  *zt_out = zt;
  *et_out = et;
}

// _p3ctrans (:1269-1288)
Rule p3ctrans(const Rule& u, double zeta, double eta) {
  double zt, et;
  p3c_inv_pole(zeta, eta, &zt, &et);
  double den = (1.0 + 3.0 * (zt * zt)) + 3.0 * (et * et);
  double A = 3.0 / den;
  double B = ((3.0 + zt * zt) * zt) / den;
  Rule r;
  r.x.resize(u.x.size()); r.w.resize(u.x.size());
  for (size_t i = 0; i < u.x.size(); ++i) {
    double s = u.x[i];
    double ds = s - zt;
    r.x[i] = A * ((ds * ds * ds) / 3.0 + (et * et) * s) + B;
    r.w[i] = u.w[i] * (A * (ds * ds + et * et));
  }
  return r;
}

// tangent step shared by _tanp3c_compose (:1321-1335) and _tanp3c_half (:1344-1365)
void tan_back(const Rule& ut, double zeta, double eta, double At, double Bt, Rule* out) {
  double halfpi = kPi / 2.0 - 1e-8;
  out->x.resize(ut.x.size()); out->w.resize(ut.x.size());
  for (size_t i = 0; i < ut.x.size(); ++i) {
    const double raw = (ut.x[i] - Bt) / At;
    double th = std::min(std::max(raw, -halfpi), halfpi);
    if (th != raw) ++g_clamped;
    double tn = std::tan(th);
    out->x[i] = zeta + eta * tn;
    out->w[i] = ut.w[i] * ((eta / At) * (1.0 + tn * tn));
  }
}

Rule tanp3c_compose(const Rule& u, double zeta, double eta, double At, double Bt,
                    double upole) {
  Rule ut = p3ctrans(u, upole, 1e-14);
  Rule r;
  tan_back(ut, zeta, eta, At, Bt, &r);
  return r;
}

Rule tanp3c_half(const Rule& u, double zeta, double eta, double At, double Bt,
                 double lo, double hi, double upole) {
  double mid = (hi + lo) / 2.0;
  double h = (hi - lo) / 2.0;
  if (h <= 1e-14) return tangenttrans(u, zeta, eta);
  double spole = (upole - mid) / h;
  Rule s = p3ctrans(u, spole, 1e-14);
  Rule ut;
  ut.x.resize(s.x.size()); ut.w.resize(s.x.size());
  for (size_t i = 0; i < s.x.size(); ++i) {
    ut.x[i] = mid + h * s.x[i];
    ut.w[i] = s.w[i] * h;
  }
  Rule r;
  tan_back(ut, zeta, eta, At, Bt, &r);
  return r;
}

// _tanp3ctrans (:1291-1319)
Rule tanp3ctrans(const Rule& u, double zeta, double eta) {
  if (eq41_collinear(zeta, eta, false, 0.0)) return eq41_real_pole(u, zeta);
  eta = std::max(eta, 1e-14);
  double ap = std::atan((1.0 - zeta) / eta);
  double am = std::atan((1.0 + zeta) / eta);
  double den = ap + am;
  if (eq41_collinear(zeta, eta, true, den)) return eq41_real_pole(u, zeta);
  if (std::fabs(den) < 1e-30) return p3ctrans(u, zeta, eta);
  double At = 2.0 / den;
  double Bt = ((-At) * (ap - am)) / 2.0;
  double uL = Bt - (At * kPi) / 2.0;
  double uR = Bt + (At * kPi) / 2.0;
  if (std::fabs(zeta) < 1.0 && -1.0 < Bt && Bt < 1.0) {
    Rule l = tanp3c_half(u, zeta, eta, At, Bt, -1.0, Bt, uL);
    Rule r = tanp3c_half(u, zeta, eta, At, Bt, Bt, 1.0, uR);
    l.x.insert(l.x.end(), r.x.begin(), r.x.end());
    l.w.insert(l.w.end(), r.w.begin(), r.w.end());
    return l;
  } else if (zeta >= 1.0) {
    return tanp3c_compose(u, zeta, eta, At, Bt, uL);
  }
  return tanp3c_compose(u, zeta, eta, At, Bt, uR);
}

// ---------------------------------------------------------------------------
// Interpolated Guiggiani on-element integral (src/Core/Integration.jl:1597-1845).
// ---------------------------------------------------------------------------
struct InterpRule {
  std::vector<double> xi, w, Ni, delta;
};

// _barycentric_eval (:1623-1645)
std::vector<double> barycentric_eval(const std::vector<double>& xi,
                                     const std::vector<double>& bw, double a) {
  size_t n = xi.size();
  std::vector<double> Ni(n);
  double msum = 0.0;
  int exact = -1;
  for (size_t i = 0; i < n; ++i) {
    double d = a - xi[i];
    if (d == 0.0) exact = (int)i;
    Ni[i] = bw[i] / d;
    msum += Ni[i];
  }
  if (exact >= 0) {
    std::fill(Ni.begin(), Ni.end(), 0.0);
    Ni[exact] = 1.0;
    return Ni;
  }
  double invs = 1.0 / msum;
  for (size_t i = 0; i < n; ++i) Ni[i] *= invs;
  return Ni;
}

// _interp_rule (:1654-1670); nodes of |xi - a| <= 1e-12 are dropped.
bool interp_rule(double a, int n, InterpRule* R) {
  std::vector<double> x(n), w(n);
  gausslegendre(n, x.data(), w.data());
  std::vector<double> bw = arbitrary_weights(x);
  int ndrop = 0;
  for (int i = 0; i < n; ++i) if (std::fabs(x[i] - a) <= 1e-12) ++ndrop;
  if (ndrop > 0) {
    std::vector<double> xk, wk;
    for (int i = 0; i < n; ++i)
      if (std::fabs(x[i] - a) > 1e-12) { xk.push_back(x[i]); wk.push_back(w[i]); }
    if (xk.size() < 2) return false;
    x = xk; w = wk;
    bw = arbitrary_weights(x);
  }
  R->xi = x;
  R->w = w;
  R->Ni = barycentric_eval(x, bw, a);
  R->delta.resize(x.size());
  for (size_t i = 0; i < x.size(); ++i) R->delta[i] = x[i] - a;
  return true;
}

// Laurent tuple (F_-2, F_-1, F_0) for orders 0 and -1.
struct Laurent { double F2, F1, F0; };

// _interp_log_coeff (:1797-1821)
Laurent interp_log_coeff(const double* fi, const InterpRule& R) {
  size_t n = R.xi.size();
  double fm1 = R.Ni[0] * (fi[0] * R.delta[0]);
  for (size_t i = 1; i < n; ++i) fm1 += R.Ni[i] * (fi[i] * R.delta[i]);
  double sL = 0, sL2 = 0, sW = 0, sWL = 0;
  int nused = 0;
  for (size_t i = 0; i < n; ++i) {
    double adi = std::fabs(R.delta[i]);
    if (adi < 1e-14) continue;
    double lg = std::log(adi);
    double wi = fi[i] - fm1 / R.delta[i];
    sL += lg;
    sL2 += lg * lg;
    sW += wi;
    sWL += wi * lg;
    ++nused;
  }
  if (nused < 2) return {0.0, fm1, 0.0};
  double detA = sL2 * nused - sL * sL;
  if (std::fabs(detA) < 1e-30) return {0.0, fm1, 0.0};
  double F0 = (nused * sWL - sL * sW) / detA;
  return {0.0, fm1, F0};
}

// _interp_cascade for order -1 (:1686-1700): p = 1, k = 1 then k = 0.
Laurent interp_cascade_m1(const double* fi, const InterpRule& R) {
  size_t n = R.xi.size();
  std::vector<double> work(fi, fi + n);
  double acc = R.Ni[0] * (work[0] * R.delta[0]);
  for (size_t i = 1; i < n; ++i) acc += R.Ni[i] * (work[i] * R.delta[i]);
  double F1 = acc;
  for (size_t i = 0; i < n; ++i) work[i] = work[i] - acc / R.delta[i];
  double acc0 = R.Ni[0] * (work[0] * 1.0);
  for (size_t i = 1; i < n; ++i) acc0 += R.Ni[i] * (work[i] * 1.0);
  return {0.0, F1, acc0};
}

// _hfp_parent k=1 and _log_parent (:1727-1739)
double hfp1(double a) { return std::log(std::fabs((1.0 - a) / (1.0 + a))); }
double log_parent(double a) {
  return ((1.0 - a) * std::log(std::fabs(1.0 - a)) + (1.0 + a) * std::log(std::fabs(1.0 + a))) - 2.0;
}

// _interp_integrate with _interp_reg_node / _interp_analytic_parent (:1741-1784)
double interp_integrate(const double* fi, const InterpRule& R, const Laurent& C,
                        int order, double a) {
  size_t n = R.xi.size();
  auto reg = [&](size_t i) {
    if (order == 0) return (fi[i] - C.F1 / R.delta[i]) - C.F0 * std::log(std::fabs(R.delta[i]));
    return fi[i] - C.F1 / R.delta[i];  // order -1
  };
  double acc = R.w[0] * reg(0);
  for (size_t i = 1; i < n; ++i) acc += R.w[i] * reg(i);
  double ana = (order == 0) ? C.F1 * hfp1(a) + C.F0 * log_parent(a) : C.F1 * hfp1(a);
  return acc + ana;
}

// ---------------------------------------------------------------------------
// Assembly (src/Core/Assembly_full.jl).
// ---------------------------------------------------------------------------
struct Mesh {
  int p, ne, n, ni, nt;
  const double* coll;  // nt x 2
  const double* nrm;   // n x 2
  const double* J;     // n
  const double* ew;    // p+1 element GL weights (dad.elem_weight)
  const double* geo;   // ne x (p+1) x 2, Equispaced order
  const double* L;     // ne
  double k;
  Poly field;          // Legendre(p)
  Poly geom;           // Equispaced(p)
};

// _shape_geom + _sample_kernel (:607-643): returns false for "nothing".
bool sample_kernel(const Mesh& M, int e, double px, double py, double xi, double nrefx,
                   double nrefy, double* U, double* T, double* Nrow, double* Jout) {
  int nn = M.p + 1;
  double dN[16], Ng[16], dNg[16];
  shapefun(M.field, xi, Nrow, dN);
  shapefun(M.geom, xi, Ng, dNg);
  const double* g = M.geo + (size_t)e * nn * 2;
  double pgx = 0, pgy = 0, dxx = 0, dxy = 0;
  for (int k = 0; k < nn; ++k) {
    pgx += Ng[k] * g[2 * k]; pgy += Ng[k] * g[2 * k + 1];
    dxx += dNg[k] * g[2 * k]; dxy += dNg[k] * g[2 * k + 1];
  }
  double J = norm2d(dxx, dxy);
  if (J < 1e-30) return false;
  double rx = pgx - px, ry = pgy - py;
  if (norm2d(rx, ry) < 1e-30) return false;
  double nx = dxy / J, ny = -(dxx / J);  // tan2normal(dx / J)
  if (nx * nrefx + ny * nrefy < 0) { nx = -nx; ny = -ny; }
  fundamental(rx, ry, nx, ny, M.k, U, T);
  *Jout = J;
  return true;
}

// _integrate_singular_guiggiani! (:397-424) -> guiggiani_GH :interp
// (Integration.jl:439-481, :1833-1845) with orders (0, -1), ninterp = 20.
void integrate_singular(const Mesh& M, int e, double px, double py, double a, double* h,
                        double* g) {
  int nn = M.p + 1;
  a = std::min(std::max(a, std::nextafter(-1.0, 0.0)), std::nextafter(1.0, 0.0));
  InterpRule R;
  interp_rule(a, 20, &R);
  size_t m = R.xi.size();
  const double* nref = M.nrm + 2 * (size_t)e * nn;  // dad.Normal[elem.index[1]]
  std::vector<double> Fg(m * nn, 0.0), Fh(m * nn, 0.0);
  for (size_t i = 0; i < m; ++i) {
    double U, T, Nrow[16], J;
    if (!sample_kernel(M, e, px, py, R.xi[i], nref[0], nref[1], &U, &T, Nrow, &J)) continue;
    for (int j = 0; j < nn; ++j) {
      double NjJ = Nrow[j] * J;
      Fg[j * m + i] = U * NjJ;
      Fh[j * m + i] = T * NjJ;
    }
  }
  for (int j = 0; j < nn; ++j) {
    Laurent Cg = interp_log_coeff(&Fg[j * m], R);
    Laurent Ch = interp_cascade_m1(&Fh[j * m], R);
    g[j] = interp_integrate(&Fg[j * m], R, Cg, 0, a);
    h[j] = interp_integrate(&Fh[j * m], R, Ch, -1, a);
  }
}

// _quad_geom (:649-669) + _accumulate_kernels! (:683-695) with transform
// (Integration.jl:68-93, :tanp3c default).
void integrate_near(const Mesh& M, const Rule& gl, int e, double px, double py, double* h,
                    double* g) {
  int nn = M.p + 1;
  const double* xn = M.coll + 2 * (size_t)e * nn;
  double zeta, eta;
  complex_pole_1d(M.field, xn, nn, px, py, &zeta, &eta);
  Rule q = tanp3ctrans(gl, zeta, eta);
  const double* gg = M.geo + (size_t)e * nn * 2;
  const double* nref = M.nrm + 2 * (size_t)e * nn;
  size_t nq = q.x.size();
  std::vector<double> N(nq * nn), rx(nq), ry(nq), nx(nq), ny(nq), wJ(nq);
  for (size_t i = 0; i < nq; ++i) {
    double dN[16], Ng[16], dNg[16];
    shapefun(M.field, q.x[i], &N[i * nn], dN);
    shapefun(M.geom, q.x[i], Ng, dNg);
    double pgx = 0, pgy = 0, dxx = 0, dxy = 0;
    for (int k = 0; k < nn; ++k) {
      pgx += Ng[k] * gg[2 * k]; pgy += Ng[k] * gg[2 * k + 1];
      dxx += dNg[k] * gg[2 * k]; dxy += dNg[k] * gg[2 * k + 1];
    }
    rx[i] = pgx - px; ry[i] = pgy - py;
    double J = norm2d(dxx, dxy);
    nx[i] = dxy / J; ny[i] = -(dxx / J);
    wJ[i] = J * q.w[i];
  }
  if (nq > 0 && (nx[0] * nref[0] + ny[0] * nref[1]) < 0) {
    for (size_t i = 0; i < nq; ++i) { nx[i] = -nx[i]; ny[i] = -ny[i]; }
  }
  for (int j = 0; j < nn; ++j) { h[j] = 0.0; g[j] = 0.0; }
  for (size_t i = 0; i < nq; ++i) {
    double U, T;
    fundamental(rx[i], ry[i], nx[i], ny[i], M.k, &U, &T);
    double wi = wJ[i];
    for (int j = 0; j < nn; ++j) {
      double Nj = N[i * nn + j] * wi;
      h[j] += T * Nj;
      g[j] += U * Nj;
    }
  }
}

// integrate_element (:262-313) for a 2-D scalar pair with source = i.
void integrate_element(const Mesh& M, const Rule& gl, int e, int i, double* h, double* g) {
  int nn = M.p + 1;
  double px = M.coll[2 * (size_t)i], py = M.coll[2 * (size_t)i + 1];
  bool singular = (i >= e * nn && i < (e + 1) * nn);
  if (singular) {
    const double* xn = M.coll + 2 * (size_t)e * nn;
    double a, dist;
    closest_point_1d(M.field, xn, px, py, seed_1d(M.field, xn, nn, px, py), &a, &dist);
    integrate_singular(M, e, px, py, a, h, g);
    return;
  }
  integrate_near(M, gl, e, px, py, h, g);
}


// ---------------------------------------------------------------------------
// 2-D isotropic elasticity (Kelvin), H_G_full_direct for a Vectorial problem
// (src/Core/Assembly_full.jl:120-175). Blocks are 2x2, component c = 2a + b
// (row component a, column component b); dof of node j component b is 2j + b.
// ---------------------------------------------------------------------------
// fundamental(props::Elasticity, r, n) (src/Elasticity/Fundamental.jl:28-57),
// r = x_field - x_source, n the field normal; mu and nu are the shear modulus
// and the effective Poisson ratio (effective_nu, src/Core/Structures.jl:260-263).
inline void kelvin(double rx, double ry, double nx, double ny, double mu, double nu,
                   double* U, double* T) {
  double R2 = rx * rx + ry * ry;
  double R = std::sqrt(R2);
  double dr1 = rx / R, dr2 = ry / R;
  double drdn = dr1 * nx + dr2 * ny;
  double prod1 = (4.0 * kPi) * (1.0 - nu);
  double prod2 = (3.0 - 4.0 * nu) * std::log(1.0 / R);
  double base = (2.0 * prod1) * mu;
  U[0] = (prod2 + dr1 * dr1) / base;
  U[1] = (dr1 * dr2) / base;
  U[2] = U[1];
  U[3] = (prod2 + dr2 * dr2) / base;
  double fat = 1.0 / (prod1 * R);
  double c12 = 1.0 - 2.0 * nu;
  T[0] = ((-drdn) * (c12 + 2.0 * (dr1 * dr1))) * fat;
  T[3] = ((-drdn) * (c12 + 2.0 * (dr2 * dr2))) * fat;
  T[1] = (-(((drdn * 2.0) * dr1) * dr2 - c12 * (dr1 * ny - dr2 * nx))) * fat;
  T[2] = (-(((drdn * 2.0) * dr1) * dr2 - c12 * (dr2 * nx - dr1 * ny))) * fat;
}

struct KMesh {
  const Mesh* M;
  double mu, nu;
};

// _sample_kernel (src/Core/Assembly_full.jl:607-643) with the Kelvin kernel.
bool sample_kelvin(const KMesh& K, int e, double px, double py, double xi, double nrefx,
                   double nrefy, double* U, double* T, double* Nrow, double* Jout) {
  const Mesh& M = *K.M;
  int nn = M.p + 1;
  double dN[16], Ng[16], dNg[16];
  shapefun(M.field, xi, Nrow, dN);
  shapefun(M.geom, xi, Ng, dNg);
  const double* g = M.geo + (size_t)e * nn * 2;
  double pgx = 0, pgy = 0, dxx = 0, dxy = 0;
  for (int k = 0; k < nn; ++k) {
    pgx += Ng[k] * g[2 * k]; pgy += Ng[k] * g[2 * k + 1];
    dxx += dNg[k] * g[2 * k]; dxy += dNg[k] * g[2 * k + 1];
  }
  double J = norm2d(dxx, dxy);
  if (J < 1e-30) return false;
  double rx = pgx - px, ry = pgy - py;
  if (norm2d(rx, ry) < 1e-30) return false;
  double nx = dxy / J, ny = -(dxx / J);
  if (nx * nrefx + ny * nrefy < 0) { nx = -nx; ny = -ny; }
  kelvin(rx, ry, nx, ny, K.mu, K.nu, U, T);
  *Jout = J;
  return true;
}

// Vectorial _integrate_singular_guiggiani! (src/Core/Assembly_full.jl:426-457):
// the interpolated Laurent rule of the scalar case applied to every component,
// orders (0, -1) (singularity_orders, src/Core/Structures.jl:84).
// h, g: [node k][component c] (nn x 4).
void integrate_singular_kelvin(const KMesh& K, int e, double px, double py, double a,
                               double* h, double* g) {
  const Mesh& M = *K.M;
  int nn = M.p + 1;
  a = std::min(std::max(a, std::nextafter(-1.0, 0.0)), std::nextafter(1.0, 0.0));
  InterpRule R;
  interp_rule(a, 20, &R);
  size_t m = R.xi.size();
  const double* nref = M.nrm + 2 * (size_t)e * nn;
  std::vector<double> Fg(m * nn * 4, 0.0), Fh(m * nn * 4, 0.0);  // [(k*4+c)][i]
  for (size_t i = 0; i < m; ++i) {
    double U[4], T[4], Nrow[16], J;
    if (!sample_kelvin(K, e, px, py, R.xi[i], nref[0], nref[1], U, T, Nrow, &J)) continue;
    for (int k = 0; k < nn; ++k) {
      double NjJ = Nrow[k] * J;
      for (int c = 0; c < 4; ++c) {
        Fg[(k * 4 + c) * m + i] = U[c] * NjJ;
        Fh[(k * 4 + c) * m + i] = T[c] * NjJ;
      }
    }
  }
  for (int kc = 0; kc < nn * 4; ++kc) {
    Laurent Cg = interp_log_coeff(&Fg[kc * m], R);
    Laurent Ch = interp_cascade_m1(&Fh[kc * m], R);
    g[kc] = interp_integrate(&Fg[kc * m], R, Cg, 0, a);
    h[kc] = interp_integrate(&Fh[kc * m], R, Ch, -1, a);
  }
}

// _quad_geom (:649-669) + vectorial _accumulate_kernels! (:697-713).
void integrate_near_kelvin(const KMesh& K, const Rule& gl, int e, double px, double py,
                           double* h, double* g) {
  const Mesh& M = *K.M;
  int nn = M.p + 1;
  const double* xn = M.coll + 2 * (size_t)e * nn;
  double zeta, eta;
  complex_pole_1d(M.field, xn, nn, px, py, &zeta, &eta);
  Rule q = tanp3ctrans(gl, zeta, eta);
  const double* gg = M.geo + (size_t)e * nn * 2;
  const double* nref = M.nrm + 2 * (size_t)e * nn;
  size_t nq = q.x.size();
  std::vector<double> N(nq * nn), rx(nq), ry(nq), nx(nq), ny(nq), wJ(nq);
  for (size_t i = 0; i < nq; ++i) {
    double dN[16], Ng[16], dNg[16];
    shapefun(M.field, q.x[i], &N[i * nn], dN);
    shapefun(M.geom, q.x[i], Ng, dNg);
    double pgx = 0, pgy = 0, dxx = 0, dxy = 0;
    for (int k = 0; k < nn; ++k) {
      pgx += Ng[k] * gg[2 * k]; pgy += Ng[k] * gg[2 * k + 1];
      dxx += dNg[k] * gg[2 * k]; dxy += dNg[k] * gg[2 * k + 1];
    }
    rx[i] = pgx - px; ry[i] = pgy - py;
    double J = norm2d(dxx, dxy);
    nx[i] = dxy / J; ny[i] = -(dxx / J);
    wJ[i] = J * q.w[i];
  }
  if (nq > 0 && (nx[0] * nref[0] + ny[0] * nref[1]) < 0) {
    for (size_t i = 0; i < nq; ++i) { nx[i] = -nx[i]; ny[i] = -ny[i]; }
  }
  for (int kc = 0; kc < nn * 4; ++kc) { h[kc] = 0.0; g[kc] = 0.0; }
  for (size_t i = 0; i < nq; ++i) {
    double U[4], T[4];
    kelvin(rx[i], ry[i], nx[i], ny[i], K.mu, K.nu, U, T);
    double wi = wJ[i];
    for (int k = 0; k < nn; ++k) {
      double Nj = N[i * nn + k] * wi;
      for (int c = 0; c < 4; ++c) {
        h[k * 4 + c] += T[c] * Nj;
        g[k * 4 + c] += U[c] * Nj;
      }
    }
  }
}

// integrate_element (:262-313) for a 2-D Kelvin pair with source = i.
void integrate_element_kelvin(const KMesh& K, const Rule& gl, int e, int i, double* h,
                              double* g) {
  const Mesh& M = *K.M;
  int nn = M.p + 1;
  double px = M.coll[2 * (size_t)i], py = M.coll[2 * (size_t)i + 1];
  bool singular = (i >= e * nn && i < (e + 1) * nn);
  if (singular) {
    const double* xn = M.coll + 2 * (size_t)e * nn;
    double a, dist;
    closest_point_1d(M.field, xn, px, py, seed_1d(M.field, xn, nn, px, py), &a, &dist);
    integrate_singular_kelvin(K, e, px, py, a, h, g);
    return;
  }
  integrate_near_kelvin(K, gl, e, px, py, h, g);
}

}  // namespace

// ===========================================================================
// C ABI for the tests (ctypes).
// ===========================================================================
extern "C" {

int oracle_gausslegendre(int n, double* x, double* w) {
  if (n < 1) return -1;
  gausslegendre(n, x, w);
  return 0;
}

// Legendre(p) / Equispaced(p) nodes, barycentric weights and Dmat.
int oracle_poly(int kind, int p, double* x, double* bw, double* D) {
  if (p < 0 || p > 14) return -1;
  Poly P = (kind == 0) ? legendre_poly(p) : equispaced_poly(p);
  std::memcpy(x, P.x.data(), sizeof(double) * (p + 1));
  std::memcpy(bw, P.bw.data(), sizeof(double) * (p + 1));
  std::memcpy(D, P.D.data(), sizeof(double) * (p + 1) * (p + 1));
  return 0;
}

int oracle_shapefun(int kind, int p, double xi, double* N, double* dN) {
  if (p < 0 || p > 14) return -1;
  Poly P = (kind == 0) ? legendre_poly(p) : equispaced_poly(p);
  shapefun(P, xi, N, dN);
  return 0;
}

void oracle_fundamental(double rx, double ry, double nx, double ny, double k, double* U,
                        double* T) {
  fundamental(rx, ry, nx, ny, k, U, T);
}

// closest_point_1d on a curve given by nodes with basis kind (0 Legendre, 1 Equispaced).
int oracle_closest_point(int kind, int p, const double* nodes, double px, double py,
                         double* xi, double* dist) {
  Poly P = (kind == 0) ? legendre_poly(p) : equispaced_poly(p);
  closest_point_1d(P, nodes, px, py, seed_1d(P, nodes, p + 1, px, py), xi, dist);
  return 0;
}

int oracle_complex_pole(int kind, int p, const double* nodes, double px, double py,
                        double* zeta, double* eta) {
  Poly P = (kind == 0) ? legendre_poly(p) : equispaced_poly(p);
  complex_pole_1d(P, nodes, p + 1, px, py, zeta, eta);
  return 0;
}

void oracle_p3c_inv_pole(double zeta, double eta, double* zt, double* et) {
  p3c_inv_pole(zeta, eta, zt, et);
}

// Maps: which = 0 tanp3c, 1 tangent, 2 p3c, 3 eq41 real pole.
// Output arrays must hold 2*n entries; returns the number of nodes.
int oracle_map(int which, const double* u, const double* w, int n, double zeta, double eta,
               double* x_out, double* w_out) {
  Rule r0;
  r0.x.assign(u, u + n);
  r0.w.assign(w, w + n);
  Rule r;
  switch (which) {
    case 0: r = tanp3ctrans(r0, zeta, eta); break;
    case 1: r = tangenttrans(r0, zeta, eta); break;
    case 2: r = p3ctrans(r0, zeta, eta); break;
    case 3: r = eq41_real_pole(r0, zeta); break;
    default: return -1;
  }
  std::memcpy(x_out, r.x.data(), sizeof(double) * r.x.size());
  std::memcpy(w_out, r.w.data(), sizeof(double) * r.w.size());
  return (int)r.x.size();
}

// Interpolated-Guiggiani rule at collocation a with n nodes: returns m (<= n).
int oracle_interp_rule(double a, int n, double* xi, double* w, double* Ni, double* delta) {
  InterpRule R;
  if (!interp_rule(a, n, &R)) return -1;
  size_t m = R.xi.size();
  std::memcpy(xi, R.xi.data(), sizeof(double) * m);
  std::memcpy(w, R.w.data(), sizeof(double) * m);
  std::memcpy(Ni, R.Ni.data(), sizeof(double) * m);
  std::memcpy(delta, R.delta.data(), sizeof(double) * m);
  return (int)m;
}

// Guiggiani integral of samples fi (taken at the oracle_interp_rule nodes) with
// Laurent order 0 (log) or -1 (CPV). Coefficients returned in C[3].
double oracle_interp_integral(double a, int n, const double* fi, int order, double* C) {
  InterpRule R;
  if (!interp_rule(a, n, &R)) return NAN;
  Laurent L = (order == 0) ? interp_log_coeff(fi, R) : interp_cascade_m1(fi, R);
  if (C) { C[0] = L.F2; C[1] = L.F1; C[2] = L.F0; }
  return interp_integrate(fi, R, L, order, a);
}

int oracle_near_element(double px, double py, const double* xg, int nn, double L,
                        double factor) {
  return near_element(px, py, xg, nn, L, factor) ? 1 : 0;
}

// Dense H (nt x nt) and G (nt x n), column-major, as H_G_full_direct
// (src/Core/Assembly_full.jl:61-85): rows are threaded statically; the
// rigid-body diagonal (:384-390) is applied last.
// rows: if row_lo/row_hi bound a sub-range, only those rows are assembled
// (the diagonal is then skipped; used for bounded CPU timing samples).
int oracle_assemble(int p, int ne, int n, int ni, const double* coll, const double* nrm,
                    const double* J, const double* ew, const double* geo, const double* L,
                    double k, int npg, double near_factor, int nthreads, int row_lo,
                    int row_hi, double* H, double* G) {
  if (p < 1 || p > 14 || n != ne * (p + 1) || npg < 1) return -1;
  Mesh M;
  M.p = p; M.ne = ne; M.n = n; M.ni = ni; M.nt = n + ni;
  M.coll = coll; M.nrm = nrm; M.J = J; M.ew = ew; M.geo = geo; M.L = L; M.k = k;
  M.field = legendre_poly(p);
  M.geom = equispaced_poly(p);
  const int nt = M.nt, nn = p + 1;
  Rule gl;
  gl.x.resize(npg); gl.w.resize(npg);
  gausslegendre(npg, gl.x.data(), gl.w.data());
  bool full = (row_lo <= 0 && row_hi >= nt);
  if (row_lo < 0) row_lo = 0;
  if (row_hi > nt) row_hi = nt;
  const size_t ldh = (size_t)nt;
  if (full) {
    std::memset(H, 0, sizeof(double) * ldh * nt);
    std::memset(G, 0, sizeof(double) * ldh * n);
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int i = row_lo; i < row_hi; ++i) {
    double px = coll[2 * (size_t)i], py = coll[2 * (size_t)i + 1];
    double h[16], g[16];
    for (int e = 0; e < ne; ++e) {
      const double* xn = coll + 2 * (size_t)e * nn;
      bool onel = (i >= e * nn && i < (e + 1) * nn);
      bool nr = near_element(px, py, xn, nn, L[e], near_factor) || onel;
      if (nr) {
        integrate_element(M, gl, e, i, h, g);
        for (int kk = 0; kk < nn; ++kk) {
          size_t j = (size_t)e * nn + kk;
          H[i + j * ldh] += h[kk];
          G[i + j * ldh] += g[kk];
        }
      } else {
        // _far_nodal_scalar! (:801-810)
        for (int kk = 0; kk < nn; ++kk) {
          size_t node = (size_t)e * nn + kk;
          double U, T;
          fundamental(coll[2 * node] - px, coll[2 * node + 1] - py, nrm[2 * node],
                      nrm[2 * node + 1], k, &U, &T);
          double wjk = J[node] * ew[kk];
          H[i + node * ldh] += T * wjk;
          G[i + node * ldh] += U * wjk;
        }
      }
    }
  }
  if (full) {
    // _rowsum_diag! (:384-390)
    for (int i = 0; i < nt; ++i) {
      H[i + i * ldh] = 0.0;
      double s = 0.0;
      for (int j = 0; j < nt; ++j) s += H[i + (size_t)j * ldh];
      H[i + i * ldh] = -s;
    }
  }
  return 0;
}

// One near or on-element pair (i, e): the (p+1) entries h, g added to row i.
int oracle_pair(int p, int ne, int n, int ni, const double* coll, const double* nrm,
                const double* J, const double* ew, const double* geo, const double* L,
                double k, int npg, int i, int e, double* h, double* g) {
  Mesh M;
  M.p = p; M.ne = ne; M.n = n; M.ni = ni; M.nt = n + ni;
  M.coll = coll; M.nrm = nrm; M.J = J; M.ew = ew; M.geo = geo; M.L = L; M.k = k;
  M.field = legendre_poly(p);
  M.geom = equispaced_poly(p);
  Rule gl;
  gl.x.resize(npg); gl.w.resize(npg);
  gausslegendre(npg, gl.x.data(), gl.w.data());
  integrate_element(M, gl, e, i, h, g);
  return 0;
}

// Diagnostics of the rule a pair (i, e) takes (integrate_element, :262-313):
// out = {zeta0, eta0, branch, clamped}. branch: -1 on-element (interpolated
// Guiggiani), 0 Mobius (_eq41_real_pole), 1 p3c only, 2 compose
// (_tanp3c_compose), 3 split (_tanp3c_half twice); clamped = how many of the
// rule's nodes had their tangent angle clamped to +-(pi/2 - 1e-8)
// (src/Core/Integration.jl:1177-1200, :1321-1365).
int oracle_pair_plan(int p, int ne, int n, int ni, const double* coll, int npg, int i, int e,
                     double* out) {
  const int nn = p + 1;
  if (e < 0 || e >= ne || i < 0 || i >= n + ni) return 1;
  out[0] = out[1] = 0.0;
  out[3] = 0.0;
  if (i >= e * nn && i < (e + 1) * nn) { out[2] = -1.0; return 0; }
  const Poly field = legendre_poly(p);
  double zeta, eta;
  complex_pole_1d(field, coll + 2 * (size_t)e * nn, nn, coll[2 * (size_t)i],
                  coll[2 * (size_t)i + 1], &zeta, &eta);
  out[0] = zeta;
  out[1] = eta;
  int branch;
  if (eq41_collinear(zeta, eta, false, 0.0)) {
    branch = 0;
  } else {
    const double et = std::max(eta, 1e-14);
    const double ap = std::atan((1.0 - zeta) / et), am = std::atan((1.0 + zeta) / et);
    const double den = ap + am;
    if (eq41_collinear(zeta, et, true, den)) branch = 0;
    else if (std::fabs(den) < 1e-30) branch = 1;
    else {
      const double At = 2.0 / den, Bt = ((-At) * (ap - am)) / 2.0;
      branch = (std::fabs(zeta) < 1.0 && -1.0 < Bt && Bt < 1.0) ? 3 : 2;
    }
  }
  out[2] = branch;
  Rule gl;
  gl.x.resize(npg); gl.w.resize(npg);
  gausslegendre(npg, gl.x.data(), gl.w.data());
  g_clamped = 0;
  tanp3ctrans(gl, zeta, eta);
  out[3] = (double)g_clamped;
  return 0;
}

// applyBC (src/Core/Boundary_conditions.jl:54-75): A = H with Dirichlet
// columns -G, b = -sum_D H[:,j] BV_j + sum_N G[:,j] BV_j. Column-major.
int oracle_apply_bc(int nt, int n, const double* H, const double* G, const int64_t* BC,
                    const double* BV, double* A, double* b) {
  const size_t ld = (size_t)nt;
  std::memcpy(A, H, sizeof(double) * ld * nt);
  for (int j = 0; j < n; ++j)
    if (BC[j] == 0)
      for (int i = 0; i < nt; ++i) A[i + j * ld] = -G[i + j * ld];
  for (int i = 0; i < nt; ++i) b[i] = 0.0;
  for (int j = 0; j < n; ++j) {
    if (BC[j] == 0)
      for (int i = 0; i < nt; ++i) b[i] -= H[i + j * ld] * BV[j];
    else
      for (int i = 0; i < nt; ++i) b[i] += G[i + j * ld] * BV[j];
  }
  return 0;
}

void oracle_kelvin(double rx, double ry, double nx, double ny, double mu, double nu, double* U,
                   double* T) {
  kelvin(rx, ry, nx, ny, mu, nu, U, T);
}

// Dense Kelvin H (2nt x 2nt) and G (2nt x 2n), column-major, as H_G_full_direct
// for a Vectorial problem (src/Core/Assembly_full.jl:120-175): near pairs by
// integrate_element, far pairs by _far_nodal_vec! (:812-822), then the
// rigid-body free term _vectorial_free_term! (:777-786). row_lo/row_hi as in
// oracle_assemble (node rows; the free term is skipped for a sub-range).
int oracle_assemble_kelvin(int p, int ne, int n, int ni, const double* coll, const double* nrm,
                           const double* J, const double* ew, const double* geo,
                           const double* L, double mu, double nu, int npg, double near_factor,
                           int nthreads, int row_lo, int row_hi, double* H, double* G) {
  if (p < 1 || p > 14 || n != ne * (p + 1) || npg < 1) return -1;
  Mesh M;
  M.p = p; M.ne = ne; M.n = n; M.ni = ni; M.nt = n + ni;
  M.coll = coll; M.nrm = nrm; M.J = J; M.ew = ew; M.geo = geo; M.L = L; M.k = 1.0;
  M.field = legendre_poly(p);
  M.geom = equispaced_poly(p);
  KMesh K{&M, mu, nu};
  const int nt = M.nt, nn = p + 1;
  Rule gl;
  gl.x.resize(npg); gl.w.resize(npg);
  gausslegendre(npg, gl.x.data(), gl.w.data());
  bool full = (row_lo <= 0 && row_hi >= nt);
  if (row_lo < 0) row_lo = 0;
  if (row_hi > nt) row_hi = nt;
  const size_t ld = 2 * (size_t)nt;
  if (full) {
    std::memset(H, 0, sizeof(double) * ld * ld);
    std::memset(G, 0, sizeof(double) * ld * 2 * (size_t)n);
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int i = row_lo; i < row_hi; ++i) {
    double px = coll[2 * (size_t)i], py = coll[2 * (size_t)i + 1];
    double h[64], g[64];
    for (int e = 0; e < ne; ++e) {
      const double* xn = coll + 2 * (size_t)e * nn;
      if (near_element(px, py, xn, nn, L[e], near_factor)) {
        integrate_element_kelvin(K, gl, e, i, h, g);
        for (int k = 0; k < nn; ++k) {
          size_t j = (size_t)e * nn + k;
          for (int c = 0; c < 4; ++c) {
            size_t r = 2 * (size_t)i + (c >> 1), col = 2 * j + (c & 1);
            H[r + col * ld] += h[k * 4 + c];
            G[r + col * ld] += g[k * 4 + c];
          }
        }
      } else {
        for (int k = 0; k < nn; ++k) {
          size_t node = (size_t)e * nn + k;
          double U[4], T[4];
          kelvin(coll[2 * node] - px, coll[2 * node + 1] - py, nrm[2 * node],
                 nrm[2 * node + 1], mu, nu, U, T);
          double wjk = J[node] * ew[k];
          for (int c = 0; c < 4; ++c) {
            size_t r = 2 * (size_t)i + (c >> 1), col = 2 * node + (c & 1);
            H[r + col * ld] += T[c] * wjk;
            G[r + col * ld] += U[c] * wjk;
          }
        }
      }
    }
  }
  if (full) {
    // _vectorial_free_term!: H[ii, ii] = 0, then H[ii, ii[b]] = -sum over the
    // columns of component b (the zeroed block included)
    for (int i = 0; i < nt; ++i) {
      for (int c = 0; c < 4; ++c)
        H[2 * (size_t)i + (c >> 1) + (2 * (size_t)i + (c & 1)) * ld] = 0.0;
      for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
          size_t r = 2 * (size_t)i + a;
          double s = 0.0;
          for (int j = 0; j < nt; ++j) s += H[r + (2 * (size_t)j + b) * ld];
          H[r + (2 * (size_t)i + b) * ld] = -s;
        }
    }
  }
  return 0;
}

// One near or on-element Kelvin pair (i, e): [node k][component c] blocks.
int oracle_pair_kelvin(int p, int ne, int n, int ni, const double* coll, const double* nrm,
                       const double* J, const double* ew, const double* geo, const double* L,
                       double mu, double nu, int npg, int i, int e, double* h, double* g) {
  Mesh M;
  M.p = p; M.ne = ne; M.n = n; M.ni = ni; M.nt = n + ni;
  M.coll = coll; M.nrm = nrm; M.J = J; M.ew = ew; M.geo = geo; M.L = L; M.k = 1.0;
  M.field = legendre_poly(p);
  M.geom = equispaced_poly(p);
  KMesh K{&M, mu, nu};
  Rule gl;
  gl.x.resize(npg); gl.w.resize(npg);
  gausslegendre(npg, gl.x.data(), gl.w.data());
  integrate_element_kelvin(K, gl, e, i, h, g);
  return 0;
}

}  // extern "C"
