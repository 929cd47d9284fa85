// Near-singular (tan-p3c) and on-element (interpolated Guiggiani) element
// integrals of the Laplace kernel for ONE (source, element) pair, written for
// a generic scalar S: double, or ad::Dual<D> to carry derivatives with
// respect to the source point and the element's geometric nodes.
//
// Same steps, in the same order and with the same branch thresholds, as the
// reference rules:
//   complex pole            src/Core/Integration.jl:1117-1160
//   p3c inverse pole        src/Core/Integration.jl:1220-1266
//   tangent / Mobius map    src/Core/Integration.jl:1163-1217
//   tan-p3c plan + nodes    src/Core/Integration.jl:1291-1365
//   near quadrature         src/Core/Assembly_full.jl:649-695
//   closest point           src/Core/Integration.jl:323-349, :1005-1033
//   interpolated Guiggiani  src/Core/Integration.jl:1597-1845,
//                           src/Core/Assembly_full.jl:397-424, :607-643
// Branches compare values only, so a Dual evaluation takes the branch the
// double evaluation takes and yields the derivative of that branch.
//
// Serial over the quadrature points (one thread per pair): the adjoint runs
// this once per near pair.
#pragma once

#include <cmath>

#include "dual.hpp"
#include "quadconst.hpp"

namespace bemhip {
namespace rule {

constexpr double kPi = 3.141592653589793;
constexpr double kEps = 2.220446049250313e-16;

template <typename S>
struct Cx {
  S re, im;
};
template <typename S>
__host__ __device__ inline Cx<S> cadd(const Cx<S>& a, const Cx<S>& b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename S>
__host__ __device__ inline Cx<S> csub(const Cx<S>& a, const Cx<S>& b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename S>
__host__ __device__ inline Cx<S> cmul(const Cx<S>& a, const Cx<S>& b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename S>
__host__ __device__ inline Cx<S> cscale(const Cx<S>& a, double s) {
  return {a.re * s, a.im * s};
}
template <typename S>
__host__ __device__ inline Cx<S> cdivr(const Cx<S>& z, double d) {
  return {z.re / d, z.im / d};
}
template <typename S>
__host__ __device__ inline Cx<S> cdiv(const Cx<S>& a, const Cx<S>& b) {  // Smith
  if (fabs(b.re) >= fabs(b.im)) {
    S r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  S r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
template <typename S>
__host__ __device__ inline S cabs(const Cx<S>& a) {
  return hypot(a.re, a.im);
}
template <typename S>
__host__ __device__ inline Cx<S> csqrt_(const Cx<S>& z) {  // principal branch
  if (z.re == 0.0 && z.im == 0.0) return {S(0.0), z.im};
  S r = hypot(z.re, z.im);
  if (z.re >= 0.0) {
    S t = sqrt(0.5 * (r + z.re));
    return {t, z.im / (2.0 * t)};
  }
  S t = sqrt(0.5 * (r - z.re));
  return {fabs(z.im) / (2.0 * t), copysign(t, z.im)};
}

template <typename S>
__host__ __device__ inline S clampS(const S& x, double lo, double hi) {
  if (x > hi) return S(hi);
  if (x < lo) return S(lo);
  return x;
}

// Barycentric interpolation row (src/Core/Interpolation.jl:148-178).
template <typename S, int NN>
__host__ __device__ inline void interp_row(const double* x0, const double* bw, const S& xx,
                                           S* M) {
  S msum = 0.0;
  int exact = -1;
  for (int i = 0; i < NN; ++i) {
    if (xx == x0[i]) exact = i;
    M[i] = bw[i] / (xx - x0[i]);
    msum += M[i];
  }
  if (msum == 0.0) {
    for (int i = 0; i < NN; ++i) M[i] = 0.0;
  } else if (exact >= 0) {
    for (int i = 0; i < NN; ++i) M[i] = (i == exact) ? 1.0 : 0.0;
  } else {
    for (int i = 0; i < NN; ++i) M[i] = M[i] / msum;
  }
}

template <typename S, int NN>
__host__ __device__ inline void shapefun(const double* x0, const double* bw, const double* D,
                                         const S& xx, S* N, S* dN) {
  interp_row<S, NN>(x0, bw, xx, N);
  for (int j = 0; j < NN; ++j) {
    S s = 0.0;
    for (int i = 0; i < NN; ++i) s += N[i] * D[i * NN + j];
    dN[j] = s;
  }
}

// Complex pole zeta + i eta of z(xi) through the collocation nodes.
template <typename S, int NN>
__host__ __device__ void complex_pole(const double* lx, const S* cx, const S* cy, const S& px,
                                      const S& py, S& zeta, S& eta) {
  const Cx<S> y = {px, py};
  if constexpr (NN == 2) {
    const double d = lx[1] - lx[0] + 1e-30;
    const Cx<S> dz = {cx[1] - cx[0], cy[1] - cy[0]};
    const Cx<S> B = {dz.re / d, dz.im / d};
    const Cx<S> C = csub(Cx<S>{cx[0], cy[0]}, cscale(B, lx[0]));
    if (cabs(B) < 1e-30) { zeta = 0.0; eta = 1.0; return; }
    const Cx<S> r = cdiv(csub(y, C), B);
    zeta = r.re;
    eta = fabs(r.im);
  } else {
    const Cx<S> Z0 = {cx[0], cy[0]}, Z1 = {cx[1], cy[1]}, Z2 = {cx[2], cy[2]};
    const double a11 = lx[0] * lx[0], a12 = lx[0], a21 = lx[1] * lx[1], a22 = lx[1],
                 a31 = lx[2] * lx[2], a32 = lx[2];
    const double det = a11 * (a22 - a32) - a12 * (a21 - a31) + (a21 * a32 - a22 * a31);
    const Cx<S> A = cdivr(cadd(csub(cscale(Z0, a22 - a32), cscale(csub(Z1, Z2), a12)),
                               csub(cscale(Z1, a32), cscale(Z2, a22))), det);
    const Cx<S> B = cdivr(cadd(csub(cscale(csub(Z1, Z2), a11), cscale(Z0, a21 - a31)),
                               csub(cscale(Z2, a21), cscale(Z1, a31))), det);
    const Cx<S> C = cdivr(cadd(csub(cscale(csub(cscale(Z2, a22), cscale(Z1, a32)), a11),
                                    cscale(csub(cscale(Z2, a21), cscale(Z1, a31)), a12)),
                               cscale(Z0, a21 * a32 - a22 * a31)), det);
    if (cabs(A) < 1e-14) {
      if (cabs(B) < 1e-30) { zeta = 0.0; eta = 1.0; return; }
      const Cx<S> r = cdiv(csub(y, C), B);
      zeta = r.re;
      eta = fabs(r.im);
      return;
    }
    const Cx<S> disc = csqrt_(csub(cmul(B, B), cmul(cscale(A, 4.0), csub(C, y))));
    const Cx<S> twoA = cscale(A, 2.0);
    const Cx<S> mB = {-B.re, -B.im};
    const Cx<S> r1 = cdiv(cadd(mB, disc), twoA);
    const Cx<S> r2 = cdiv(csub(mB, disc), twoA);
    const Cx<S> r = (fabs(r1.im) <= fabs(r2.im)) ? r1 : r2;
    zeta = r.re;
    eta = fabs(r.im);
  }
}

// Derivative-only Newton step at a converged root: with f(x) ~ 0 the update
// x - f/f'(x) changes the derivatives to those of the root, -f_theta / f_x
// (implicit function theorem), whatever the iteration count; the value is
// kept so that the rule's branch decisions stay those of the plain evaluation.
__host__ __device__ inline void newton_tangent(double&, const double&, double) {}
template <int D>
__host__ __device__ inline void newton_tangent(ad::Dual<D>& x, const ad::Dual<D>& f, double fp) {
  for (int k = 0; k < D; ++k) x.d[k] -= f.d[k] / fp;
}
__host__ __device__ inline void newton_tangent2(double&, double&, const double&, const double&,
                                                double, double, double, double, double) {}
template <int D>
__host__ __device__ inline void newton_tangent2(ad::Dual<D>& x, ad::Dual<D>& y,
                                                const ad::Dual<D>& f1, const ad::Dual<D>& f2,
                                                double j11, double j12, double j21, double j22,
                                                double det) {
  for (int k = 0; k < D; ++k) {
    const double a = f1.d[k], b = f2.d[k];
    x.d[k] -= (j22 * a - j12 * b) / det;
    y.d[k] -= (-j21 * a + j11 * b) / det;
  }
}

// p3c inverse pole.
template <typename S>
__host__ __device__ void p3c_inv_pole(const S& zeta, const S& eta, S& zt_out, S& et_out) {
  if (fabs(eta) < 1e-14) {
    S x = zeta;
    for (int it = 0; it < 25; ++it) {
      const S g = (((x * x * x) - (3.0 * zeta) * (x * x)) + 3.0 * x) - zeta;
      const S gp = ((3.0 * (x * x)) - (6.0 * zeta) * x) + 3.0;
      if (fabs(gp) < 1e-30) break;
      x -= g / gp;
      if (fabs(g) < 1e-14) break;
    }
    {  // derivative of the root: one Newton step on the derivatives only
      const S g = (((x * x * x) - (3.0 * zeta) * (x * x)) + 3.0 * x) - zeta;
      const double gp = value(((3.0 * (x * x)) - (6.0 * zeta) * x) + 3.0);
      if (::fabs(gp) >= 1e-30) newton_tangent(x, g, gp);
    }
    zt_out = x;
    et_out = 0.0;
    return;
  }
  const double sgn = eta > 0.0 ? 1.0 : (eta < 0.0 ? -1.0 : 0.0);
  S et = cbrt(fabs(eta)) * sgn;
  for (int it = 0; it < 20; ++it) {
    const S g = ((2.0 * (et * et * et)) - (3.0 * eta) * (et * et)) - eta;
    const S gp = (6.0 * (et * et)) - (6.0 * eta) * et;
    if (fabs(gp) < 1e-30) break;
    et -= g / gp;
  }
  et += cbrt(eta) * sqrt(zeta * zeta + eta * eta);
  S zt = zeta;
  for (int it = 0; it < 30; ++it) {
    const S den = (1.0 + 3.0 * (zt * zt)) + 3.0 * (et * et);
    const S nz = ((3.0 + zt * zt) + 3.0 * (et * et)) * zt;
    const S ne = 2.0 * (et * et * et);
    const S f1 = nz / den - zeta;
    const S f2 = ne / den - eta;
    if (fabs(f1) + fabs(f2) < 1e-14) break;
    const S denz = 6.0 * zt, dene = 6.0 * et;
    const S nz_z = (3.0 + 3.0 * (zt * zt)) + 3.0 * (et * et);
    const S nz_e = (6.0 * et) * zt;
    const S ne_e = 6.0 * (et * et);
    const S den2 = den * den;
    const S j11 = (nz_z * den - nz * denz) / den2;
    const S j12 = (nz_e * den - nz * dene) / den2;
    const S j21 = (0.0 * den - ne * denz) / den2;
    const S j22 = (ne_e * den - ne * dene) / den2;
    const S detJ = j11 * j22 - j12 * j21;
    if (fabs(detJ) < 1e-30) break;
    zt -= (j22 * f1 - j12 * f2) / detJ;
    et -= (-j21 * f1 + j11 * f2) / detJ;
  }
  {  // derivatives of the root (implicit function): one Newton step on them only
    const S den = (1.0 + 3.0 * (zt * zt)) + 3.0 * (et * et);
    const S nz = ((3.0 + zt * zt) + 3.0 * (et * et)) * zt;
    const S ne = 2.0 * (et * et * et);
    const S f1 = nz / den - zeta;
    const S f2 = ne / den - eta;
    const double z = value(zt), t = value(et);
    const double dv = (1.0 + 3.0 * (z * z)) + 3.0 * (t * t);
    const double nzv = ((3.0 + z * z) + 3.0 * (t * t)) * z, nev = 2.0 * (t * t * t);
    const double d2 = dv * dv;
    const double j11 = (((3.0 + 3.0 * (z * z)) + 3.0 * (t * t)) * dv - nzv * (6.0 * z)) / d2;
    const double j12 = (((6.0 * t) * z) * dv - nzv * (6.0 * t)) / d2;
    const double j21 = (0.0 - nev * (6.0 * z)) / d2;
    const double j22 = ((6.0 * (t * t)) * dv - nev * (6.0 * t)) / d2;
    const double detJ = j11 * j22 - j12 * j21;
    if (::fabs(detJ) >= 1e-30) newton_tangent2(zt, et, f1, f2, j11, j12, j21, j22, detJ);
  }
  zt_out = zt;
  et_out = et;
}

template <typename S>
struct P3C {
  S zt, et, A, B;
};
template <typename S>
__host__ __device__ inline P3C<S> p3c_make(const S& zeta, const S& eta) {
  P3C<S> c;
  p3c_inv_pole(zeta, eta, c.zt, c.et);
  const S den = (1.0 + 3.0 * (c.zt * c.zt)) + 3.0 * (c.et * c.et);
  c.A = 3.0 / den;
  c.B = ((3.0 + c.zt * c.zt) * c.zt) / den;
  return c;
}
template <typename S, typename U>
__host__ __device__ inline void p3c_node(const P3C<S>& c, const U& s, const U& w, S& x, S& wo) {
  const S ds = s - c.zt;
  x = c.A * ((ds * ds * ds) / 3.0 + (c.et * c.et) * s) + c.B;
  wo = w * (c.A * (ds * ds + c.et * c.et));
}

template <typename S>
__host__ __device__ inline bool eq41_collinear(const S& zeta, const S& eta, bool has_den,
                                               const S& den) {
  if (fabs(zeta) <= 1.0) return false;
  if (fabs(eta) < 1e-8) return true;
  if (has_den && fabs(den) < 1e-10) return true;
  return false;
}

template <typename S>
__host__ __device__ inline void mobius_node(const S& zeta, double u, double w, S& x, S& wo) {
  if (fabs(zeta) <= 1.0) { x = u; wo = w; return; }
  const S A = zeta * zeta - 1.0, B = -zeta;
  const S d = u - B;
  x = zeta - A / d;
  wo = w * (A / (d * d));
}


template <typename S>
__host__ __device__ void tangent_node(const S& zeta0, const S& eta0, double u, double w, S& x,
                                      S& wo) {
  if (eq41_collinear(zeta0, eta0, false, S(0.0))) { mobius_node(zeta0, u, w, x, wo); return; }
  const S zeta = zeta0;
  const S eta = fmax(eta0, 1e-14);
  const S tp = atan((1.0 - zeta) / eta);
  const S tm = atan((-1.0 - zeta) / eta);
  const S den = tp - tm;
  if (eq41_collinear(zeta, eta, true, den)) { mobius_node(zeta, u, w, x, wo); return; }
  const S A = (2.0 * eta) / den;
  const S B = -(tp + tm) / den;
  const double halfpi = kPi / 2.0 - 1e-8;
  const S phi = clampS((eta * (u - B)) / A, -halfpi, halfpi);
  const S tn = tan(phi);
  x = zeta + eta * tn;
  wo = w * (((eta * eta) / A) * (1.0 + tn * tn));
}

template <typename S>
struct TanPlan {
  int kind;  // 0 Mobius, 1 p3c only, 2 compose, 3 split
  S zeta, eta, At, Bt;
  S up;      // compose: the tangent map's pole u (uL or uR), where p3c clusters
  S r;       // pole form: den / eta, with den = 2 / At (see tanp3c_plan)
  P3C<S> c1, c2;
  S mid1, h1, mid2, h2;
  bool tan1, tan2;
  bool pole_form;  // compose with |zeta| > 1 and eta^2 < (zeta^2 - 1) / 2
};

// atan(z) / z and tan(z) / z, without the cancellation of the quotient at
// small z (series to the z^6 term below |z| = 1e-3, where it is exact in FP64).
template <typename S>
__host__ __device__ inline S atanc(const S& z) {
  if (fabs(z) < 1e-3) {
    const S z2 = z * z;
    return 1.0 + z2 * (-1.0 / 3.0 + z2 * (1.0 / 5.0 + z2 * (-1.0 / 7.0)));
  }
  return atan(z) / z;
}
template <typename S>
__host__ __device__ inline S tanc(const S& z) {
  if (fabs(z) < 1e-3) {
    const S z2 = z * z;
    return 1.0 + z2 * (1.0 / 3.0 + z2 * (2.0 / 15.0 + z2 * (17.0 / 315.0)));
  }
  return tan(z) / z;
}

// Tangent map back from u to the element: x = zeta + eta tan((u - Bt) / At),
// weight (eta / At) (1 + tan^2).
//
// Beyond the element (|zeta| > 1) with a small eta, At and Bt grow like
// 1/eta and theta = (u - Bt) / At stays within ~eta of -+pi/2: the value of
// tan(theta) carries the rounding of u - Bt magnified by 1/eta, and x is the
// difference of terms ~1/eta. That is the reference's arithmetic and stays
// the value; a dual evaluation takes its derivatives from the same map with
// the 1/eta factors cancelled analytically. With up the pole (uL or uR),
// phi = (u - up) / At = (u - up) eta r / 2 (r = den / eta) and
// theta = phi -+ pi/2, so tan(theta) = -1 / tan(phi) and
//   x = zeta - 2 / ((u - up) r tanc(phi)),
//   w = w_u 2 (1 + tan^2 phi) / ((u - up)^2 r tanc(phi)^2).
template <typename S>
__host__ __device__ inline void tan_back(const TanPlan<S>& P, const S& ut, const S& wt, S& x,
                                         S& wo) {
  const double halfpi = kPi / 2.0 - 1e-8;
  const S raw = (ut - P.Bt) / P.At;
  const S th = clampS(raw, -halfpi, halfpi);
  const S tn = tan(th);
  x = P.zeta + P.eta * tn;
  wo = wt * ((P.eta / P.At) * (1.0 + tn * tn));
  if constexpr (is_dual<S>::value) {
    if (P.pole_form && raw >= -halfpi && raw <= halfpi) {
      const S du = ut - P.up;
      const S phi = (du * P.eta) * (P.r * 0.5);
      const S tc = tanc(phi);
      const S tp = tan(phi);
      const S xs = P.zeta - 2.0 / ((du * P.r) * tc);
      const S ws = wt * ((2.0 * (1.0 + tp * tp)) / (((du * du) * P.r) * (tc * tc)));
      x = with_value(xs, value(x));
      wo = with_value(ws, value(wo));
    }
  }
}

template <typename S>
__host__ __device__ void tanp3c_plan(const S& zeta, const S& eta0, TanPlan<S>& P) {
  P.zeta = zeta;
  P.eta = eta0;
  P.tan1 = P.tan2 = false;
  P.pole_form = false;
  if (eq41_collinear(zeta, eta0, false, S(0.0))) { P.kind = 0; return; }
  const S eta = fmax(eta0, 1e-14);
  P.eta = eta;
  const S ap = atan((1.0 - zeta) / eta);
  const S am = atan((1.0 + zeta) / eta);
  const S den = ap + am;
  if (eq41_collinear(zeta, eta, true, den)) { P.kind = 0; return; }
  if (fabs(den) < 1e-30) { P.kind = 1; P.c1 = p3c_make(zeta, eta); return; }
  const S At = 2.0 / den;
  const S Bt = ((-At) * (ap - am)) / 2.0;
  const S uL = Bt - (At * kPi) / 2.0;
  const S uR = Bt + (At * kPi) / 2.0;
  P.At = At;
  P.Bt = Bt;
  if (fabs(zeta) < 1.0 && -1.0 < Bt && Bt < 1.0) {
    P.kind = 3;
    P.mid1 = (Bt + -1.0) / 2.0; P.h1 = (Bt - -1.0) / 2.0;
    P.mid2 = (1.0 + Bt) / 2.0;  P.h2 = (1.0 - Bt) / 2.0;
    P.tan1 = P.h1 <= 1e-14;
    P.tan2 = P.h2 <= 1e-14;
    if (!P.tan1) P.c1 = p3c_make((uL - P.mid1) / P.h1, S(1e-14));
    if (!P.tan2) P.c2 = p3c_make((uR - P.mid2) / P.h2, S(1e-14));
    return;
  }
  P.kind = 2;
  P.up = zeta >= 1.0 ? uL : uR;
  if constexpr (is_dual<S>::value) {
    // uL = Bt - At pi/2 (uR = Bt + At pi/2) is O(1) while At, Bt ~ 1/eta when
    // |zeta| > 1: with a = atan(eta/(|zeta|-1)), b = atan(eta/(|zeta|+1)),
    // den = a - b = atan(z1), z1 = 2 eta / (zeta^2 - 1 + eta^2), and
    // a + b = atan(z2), z2 = 2 eta |zeta| / (zeta^2 - 1 - eta^2), so the pole
    // -sign(zeta) (a + b) / (a - b) and r = den / eta are, with no 1/eta,
    //   up = -zeta (zeta^2 - 1 + eta^2) / (zeta^2 - 1 - eta^2) atanc(z2) / atanc(z1),
    //   r  = 2 atanc(z1) / (zeta^2 - 1 + eta^2).
    // Their derivatives come from these forms (the values stay the reference's).
    const S z2m1 = zeta * zeta - 1.0, e2 = eta * eta;
    if (fabs(zeta) > 1.0 && e2 < 0.5 * z2m1) {
      const S qp = z2m1 + e2, qm = z2m1 - e2;
      const S a1 = atanc((2.0 * eta) / qp);
      const S a2 = atanc(((2.0 * eta) * fabs(zeta)) / qm);
      P.up = with_value(((-zeta) * (qp / qm)) * (a2 / a1), value(P.up));
      P.r = (2.0 * a1) / qp;
      P.pole_form = true;
    }
  }
  P.c1 = p3c_make(P.up, S(1e-14));
}

template <typename S>
__host__ __device__ inline int tanp3c_count(const TanPlan<S>& P, int npg) {
  return P.kind == 3 ? 2 * npg : npg;
}

template <typename S>
__host__ __device__ void tanp3c_node(const TanPlan<S>& P, int q, int npg, const double* u,
                                     const double* w, S& x, S& wo) {
  switch (P.kind) {
    case 0: mobius_node(P.zeta, u[q], w[q], x, wo); return;
    case 1: p3c_node(P.c1, S(u[q]), S(w[q]), x, wo); return;
    case 2: {
      S ut, wt;
      p3c_node(P.c1, S(u[q]), S(w[q]), ut, wt);
      tan_back(P, ut, wt, x, wo);
      return;
    }
    default: {
      const bool left = q < npg;
      const int qq = left ? q : q - npg;
      if (left ? P.tan1 : P.tan2) { tangent_node(P.zeta, P.eta, u[qq], w[qq], x, wo); return; }
      const P3C<S>& c = left ? P.c1 : P.c2;
      const S mid = left ? P.mid1 : P.mid2, h = left ? P.h1 : P.h2;
      S s, ws;
      p3c_node(c, S(u[qq]), S(w[qq]), s, ws);
      tan_back(P, mid + h * s, ws * h, x, wo);
      return;
    }
  }
}

// Off-element near pair: h_k, g_k (k < NN) by the tan-p3c rule.
//   cx, cy    collocation nodes of the element (pole fit)
//   gx, gy    geometric nodes (equispaced xi = -1..1)
//   nref      outward normal at the element's first node (orientation only)
template <typename S, int NN>
__host__ __device__ void near_pair(const QuadConst& qc, const S* cx, const S* cy, const S* gx,
                                   const S* gy, const S& px, const S& py, double nrefx,
                                   double nrefy, S* hs, S* gs) {
  const double cU = 0.5 * (1.0 / (2.0 * kPi)) / qc.k;
  const double cT = 1.0 / (2.0 * kPi);
  S zeta, eta;
  complex_pole<S, NN>(qc.lp_x, cx, cy, px, py, zeta, eta);
  TanPlan<S> P;
  tanp3c_plan(zeta, eta, P);
  const int npg = qc.npg;
  const int nq = tanp3c_count(P, npg);
  double flip = 1.0;
  {
    S x0, w0;
    tanp3c_node(P, 0, npg, qc.gl_x, qc.gl_w, x0, w0);
    S Ng[NN], dNg[NN];
    shapefun<S, NN>(qc.eq_x, qc.eq_bw, qc.eq_D, x0, Ng, dNg);
    S dxx = 0.0, dxy = 0.0;
    for (int k = 0; k < NN; ++k) { dxx += dNg[k] * gx[k]; dxy += dNg[k] * gy[k]; }
    const S J = sqrt(dxx * dxx + dxy * dxy);
    const S nx = dxy / J, ny = -(dxx / J);
    if (nx * nrefx + ny * nrefy < 0.0) flip = -1.0;
  }
  for (int k = 0; k < NN; ++k) { hs[k] = 0.0; gs[k] = 0.0; }
  for (int q = 0; q < nq; ++q) {
    S xq, wq;
    tanp3c_node(P, q, npg, qc.gl_x, qc.gl_w, xq, wq);
    S N[NN], dN[NN], Ng[NN], dNg[NN];
    shapefun<S, NN>(qc.lp_x, qc.lp_bw, qc.lp_D, xq, N, dN);
    shapefun<S, NN>(qc.eq_x, qc.eq_bw, qc.eq_D, xq, Ng, dNg);
    S pgx = 0.0, pgy = 0.0, dxx = 0.0, dxy = 0.0;
    for (int k = 0; k < NN; ++k) {
      pgx += Ng[k] * gx[k]; pgy += Ng[k] * gy[k];
      dxx += dNg[k] * gx[k]; dxy += dNg[k] * gy[k];
    }
    const S rx = pgx - px, ry = pgy - py;
    const S J = sqrt(dxx * dxx + dxy * dxy);
    S nx = dxy / J, ny = -(dxx / J);
    if (flip < 0.0) { nx = -nx; ny = -ny; }
    const S R2 = rx * rx + ry * ry;
    const S U = -log(R2) * cU;
    const S T = (rx * nx + ry * ny) * (cT / R2);
    const S wi = J * wq;
    for (int k = 0; k < NN; ++k) {
      const S Nj = N[k] * wi;
      hs[k] += T * Nj;
      gs[k] += U * Nj;
    }
  }
}

// Parent coordinate of the closest point of the collocation interpolant to
// (px, py), clamped to the open interval (-1, 1).
template <int NN>
__host__ __device__ double closest_param(const QuadConst& qc, const double* cx, const double* cy,
                                         double px, double py) {
  const double* lx = qc.lp_x;
  const double dxs = cx[NN - 1] - cx[0], dys = cy[NN - 1] - cy[0];
  const double nd = ::sqrt(dxs * dxs + dys * dys);
  double xi = (lx[NN - 1] - lx[0]) * (dxs * (px - cx[0]) + dys * (py - cy[0])) /
                  (nd * nd + kEps) + lx[0];
  xi = clampS(xi, -1.0, 1.0);
  for (int it = 0; it < 25; ++it) {
    double N[NN], dN[NN], d2N[NN];
    shapefun<double, NN>(lx, qc.lp_bw, qc.lp_D, xi, N, dN);
    for (int j = 0; j < NN; ++j) {
      double s = 0.0;
      for (int q = 0; q < NN; ++q) s += dN[q] * qc.lp_D[q * NN + j];
      d2N[j] = s;
    }
    double X = 0, Y = 0, DX = 0, DY = 0, D2X = 0, D2Y = 0;
    for (int k = 0; k < NN; ++k) {
      X += N[k] * cx[k]; Y += N[k] * cy[k];
      DX += dN[k] * cx[k]; DY += dN[k] * cy[k];
      D2X += d2N[k] * cx[k]; D2Y += d2N[k] * cy[k];
    }
    const double rx = X - px, ry = Y - py;
    const double res = rx * DX + ry * DY;
    const double Jn = (DX * DX + DY * DY) + (rx * D2X + ry * D2Y);
    if (::fabs(Jn) < 1e-30) break;
    const double d = res / Jn;
    xi -= d;
    if (::fabs(d) < 1e-14) break;
  }
  if (xi < -1.0 || xi > 1.0) {
    double N[NN], dN[NN];
    double X0 = 0, Y0 = 0, X1 = 0, Y1 = 0;
    shapefun<double, NN>(lx, qc.lp_bw, qc.lp_D, -1.0, N, dN);
    for (int k = 0; k < NN; ++k) { X0 += N[k] * cx[k]; Y0 += N[k] * cy[k]; }
    shapefun<double, NN>(lx, qc.lp_bw, qc.lp_D, 1.0, N, dN);
    for (int k = 0; k < NN; ++k) { X1 += N[k] * cx[k]; Y1 += N[k] * cy[k]; }
    const double d0 = ::sqrt((X0 - px) * (X0 - px) + (Y0 - py) * (Y0 - py));
    const double d1 = ::sqrt((X1 - px) * (X1 - px) + (Y1 - py) * (Y1 - py));
    xi = (d0 <= d1) ? -1.0 : 1.0;
  }
  return clampS(xi, -0x1.fffffffffffffp-1, 0x1.fffffffffffffp-1);
}

// Kernel samples F = (U, T) N_k J at parent coordinate xm of the element.
template <typename S, int NN>
__host__ __device__ inline void singular_sample(const QuadConst& qc, double xm, const S* gx,
                                                const S* gy, const S& px, const S& py,
                                                double nrefx, double nrefy, S* Fg, S* Fh) {
  const double cU = 0.5 * (1.0 / (2.0 * kPi)) / qc.k;
  const double cT = 1.0 / (2.0 * kPi);
  double N[NN], dN[NN], Ng[NN], dNg[NN];
  shapefun<double, NN>(qc.lp_x, qc.lp_bw, qc.lp_D, xm, N, dN);
  shapefun<double, NN>(qc.eq_x, qc.eq_bw, qc.eq_D, xm, Ng, dNg);
  S pgx = 0.0, pgy = 0.0, dxx = 0.0, dxy = 0.0;
  for (int k = 0; k < NN; ++k) {
    pgx += Ng[k] * gx[k]; pgy += Ng[k] * gy[k];
    dxx += dNg[k] * gx[k]; dxy += dNg[k] * gy[k];
  }
  for (int k = 0; k < NN; ++k) { Fg[k] = 0.0; Fh[k] = 0.0; }
  const S J = sqrt(dxx * dxx + dxy * dxy);
  const S rx = pgx - px, ry = pgy - py;
  if (J < 1e-30 || sqrt(rx * rx + ry * ry) < 1e-30) return;
  S nx = dxy / J, ny = -(dxx / J);
  if (nx * nrefx + ny * nrefy < 0.0) { nx = -nx; ny = -ny; }
  const S R2 = rx * rx + ry * ry;
  const S U = -log(R2) * cU;
  const S T = (rx * nx + ry * ny) * (cT / R2);
  for (int k = 0; k < NN; ++k) {
    const S NjJ = N[k] * J;
    Fg[k] = U * NjJ;
    Fh[k] = T * NjJ;
  }
}

// On-element pair: h_k, g_k by the interpolated-Guiggiani rule (ninterp = 20)
// at the source's parent coordinate a: G with a log (order 0) Laurent term,
// H with a CPV (order -1) term.
template <typename S, int NN>
__host__ __device__ void singular_pair(const QuadConst& qc, double a, const S* gx, const S* gy,
                                       const S& px, const S& py, double nrefx, double nrefy,
                                       S* hs, S* gs) {
  // rule constants (independent of the geometry)
  bool keep[kNinterp];
  int m = 0;
  for (int q = 0; q < kNinterp; ++q) {
    keep[q] = !(::fabs(qc.g20_x[q] - a) <= 1e-12);
    m += keep[q] ? 1 : 0;
  }
  double Ni[kNinterp];
  double msum = 0.0;
  for (int q = 0; q < kNinterp; ++q) {
    double bw = 0.0;
    if (keep[q]) {
      if (m == kNinterp) {
        bw = qc.g20_bw[q];
      } else {
        double prod = 1.0;
        for (int j = 0; j < kNinterp; ++j) {
          if (j == q || !keep[j]) continue;
          prod *= qc.g20_x[q] - qc.g20_x[j];
        }
        bw = 1.0 / prod;
      }
    }
    Ni[q] = keep[q] ? bw / (a - qc.g20_x[q]) : 0.0;
    msum += Ni[q];
  }
  double sL = 0.0, sL2 = 0.0, SwL = 0.0;
  int nused = 0;
  for (int q = 0; q < kNinterp; ++q) {
    Ni[q] = keep[q] ? Ni[q] * (1.0 / msum) : 0.0;
    const double delta = qc.g20_x[q] - a;
    const bool used = keep[q] && !(::fabs(delta) < 1e-14);
    if (used) {
      const double lg = ::log(::fabs(delta));
      sL += lg;
      sL2 += lg * lg;
      ++nused;
    }
    if (keep[q]) SwL += qc.g20_w[q] * ::log(::fabs(delta));
  }
  const double detA = sL2 * nused - sL * sL;
  const double hfp1 = ::log(::fabs((1.0 - a) / (1.0 + a)));
  const double logp =
      ((1.0 - a) * ::log(::fabs(1.0 - a)) + (1.0 + a) * ::log(::fabs(1.0 + a))) - 2.0;
  // pass 1: Laurent coefficients F_-1 = sum N_m(a) F_m delta_m
  S Fm1g[NN], F1h[NN];
  for (int k = 0; k < NN; ++k) { Fm1g[k] = 0.0; F1h[k] = 0.0; }
  for (int q = 0; q < kNinterp; ++q) {
    if (!keep[q]) continue;
    const double delta = qc.g20_x[q] - a;
    S Fg[NN], Fh[NN];
    singular_sample<S, NN>(qc, qc.g20_x[q], gx, gy, px, py, nrefx, nrefy, Fg, Fh);
    for (int k = 0; k < NN; ++k) {
      Fm1g[k] += Ni[q] * (Fg[k] * delta);
      F1h[k] += Ni[q] * (Fh[k] * delta);
    }
  }
  // pass 2: regular parts and the least-squares log coefficient of G
  S sW[NN], sWL[NN], SwW[NN], rh[NN];
  for (int k = 0; k < NN; ++k) { sW[k] = 0.0; sWL[k] = 0.0; SwW[k] = 0.0; rh[k] = 0.0; }
  for (int q = 0; q < kNinterp; ++q) {
    if (!keep[q]) continue;
    const double delta = qc.g20_x[q] - a;
    const bool used = !(::fabs(delta) < 1e-14);
    const double lg = used ? ::log(::fabs(delta)) : 0.0;
    const double wq = qc.g20_w[q];
    S Fg[NN], Fh[NN];
    singular_sample<S, NN>(qc, qc.g20_x[q], gx, gy, px, py, nrefx, nrefy, Fg, Fh);
    for (int k = 0; k < NN; ++k) {
      const S W = Fg[k] - Fm1g[k] / delta;
      if (used) { sW[k] += W; sWL[k] += W * lg; }
      SwW[k] += wq * W;
      rh[k] += wq * (Fh[k] - F1h[k] / delta);
    }
  }
  for (int k = 0; k < NN; ++k) {
    S F0 = 0.0;
    if (nused >= 2 && !(::fabs(detA) < 1e-30)) F0 = (double(nused) * sWL[k] - sL * sW[k]) / detA;
    gs[k] = (SwW[k] - F0 * SwL) + (Fm1g[k] * hfp1 + F0 * logp);
    hs[k] = rh[k] + F1h[k] * hfp1;
  }
}

}  // namespace rule
}  // namespace bemhip
