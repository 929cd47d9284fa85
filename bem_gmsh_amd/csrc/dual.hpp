// Forward-mode dual numbers with D derivative directions, for the adjoint.
//
// Dual<D> carries a value and its derivatives with respect to D seeded
// inputs. The elementary functions below propagate exact derivatives; the
// comparisons act on the value only, so a branch in generic code takes the
// same side as the plain-double evaluation of the same expression (the
// derivative of a piecewise rule is the derivative of the branch taken).
//
// The type and its functions live in namespace bemhip::ad so that generic
// code calling sqrt/log/... unqualified resolves plain doubles to the
// ordinary math functions and Dual arguments to these overloads (by
// argument-dependent lookup).
#pragma once

#include <cmath>

namespace bemhip {
namespace ad {

template <int D>
struct Dual {
  double v;
  double d[D];
  __host__ __device__ Dual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < D; ++k) d[k] = 0.0;
  }
  __host__ __device__ Dual(double x) : v(x) {  // NOLINT: constants convert implicitly
#pragma unroll
    for (int k = 0; k < D; ++k) d[k] = 0.0;
  }
  __host__ __device__ static Dual seed(double x, int dir) {
    Dual r(x);
    r.d[dir] = 1.0;
    return r;
  }
};

// r = f(a) with f'(a) = s
template <int D>
__host__ __device__ inline Dual<D> chain1(double fv, double s, const Dual<D>& a) {
  Dual<D> r(fv);
#pragma unroll
  for (int k = 0; k < D; ++k) r.d[k] = s * a.d[k];
  return r;
}

template <int D>
__host__ __device__ inline Dual<D> operator-(const Dual<D>& a) {
  return chain1(-a.v, -1.0, a);
}
template <int D>
__host__ __device__ inline Dual<D> operator+(const Dual<D>& a, const Dual<D>& b) {
  Dual<D> r(a.v + b.v);
#pragma unroll
  for (int k = 0; k < D; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int D>
__host__ __device__ inline Dual<D> operator-(const Dual<D>& a, const Dual<D>& b) {
  Dual<D> r(a.v - b.v);
#pragma unroll
  for (int k = 0; k < D; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int D>
__host__ __device__ inline Dual<D> operator*(const Dual<D>& a, const Dual<D>& b) {
  Dual<D> r(a.v * b.v);
#pragma unroll
  for (int k = 0; k < D; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int D>
__host__ __device__ inline Dual<D> operator/(const Dual<D>& a, const Dual<D>& b) {
  const double q = a.v / b.v;
  Dual<D> r(q);
#pragma unroll
  for (int k = 0; k < D; ++k) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
  return r;
}
template <int D>
__host__ __device__ inline Dual<D> operator+(const Dual<D>& a, double b) {
  Dual<D> r = a;
  r.v += b;
  return r;
}
template <int D>
__host__ __device__ inline Dual<D> operator+(double a, const Dual<D>& b) {
  return b + a;
}
template <int D>
__host__ __device__ inline Dual<D> operator-(const Dual<D>& a, double b) {
  Dual<D> r = a;
  r.v -= b;
  return r;
}
template <int D>
__host__ __device__ inline Dual<D> operator-(double a, const Dual<D>& b) {
  return chain1(a - b.v, -1.0, b);
}
template <int D>
__host__ __device__ inline Dual<D> operator*(const Dual<D>& a, double b) {
  return chain1(a.v * b, b, a);
}
template <int D>
__host__ __device__ inline Dual<D> operator*(double a, const Dual<D>& b) {
  return chain1(a * b.v, a, b);
}
template <int D>
__host__ __device__ inline Dual<D> operator/(const Dual<D>& a, double b) {
  return chain1(a.v / b, 1.0 / b, a);
}
template <int D>
__host__ __device__ inline Dual<D> operator/(double a, const Dual<D>& b) {
  const double q = a / b.v;
  return chain1(q, -q / b.v, b);
}
template <int D>
__host__ __device__ inline Dual<D>& operator+=(Dual<D>& a, const Dual<D>& b) {
  a = a + b;
  return a;
}
template <int D>
__host__ __device__ inline Dual<D>& operator-=(Dual<D>& a, const Dual<D>& b) {
  a = a - b;
  return a;
}
template <int D>
__host__ __device__ inline Dual<D>& operator*=(Dual<D>& a, const Dual<D>& b) {
  a = a * b;
  return a;
}

// comparisons on the value
#define BEMHIP_DUAL_CMP(OP)                                                               \
  template <int D>                                                                        \
  __host__ __device__ inline bool operator OP(const Dual<D>& a, const Dual<D>& b) {       \
    return a.v OP b.v;                                                                    \
  }                                                                                       \
  template <int D>                                                                        \
  __host__ __device__ inline bool operator OP(const Dual<D>& a, double b) {               \
    return a.v OP b;                                                                      \
  }                                                                                       \
  template <int D>                                                                        \
  __host__ __device__ inline bool operator OP(double a, const Dual<D>& b) {               \
    return a OP b.v;                                                                      \
  }
BEMHIP_DUAL_CMP(<)
BEMHIP_DUAL_CMP(>)
BEMHIP_DUAL_CMP(<=)
BEMHIP_DUAL_CMP(>=)
BEMHIP_DUAL_CMP(==)
#undef BEMHIP_DUAL_CMP

template <int D>
__host__ __device__ inline Dual<D> sqrt(const Dual<D>& a) {
  const double s = ::sqrt(a.v);
  return chain1(s, 0.5 / s, a);
}
template <int D>
__host__ __device__ inline Dual<D> log(const Dual<D>& a) {
  return chain1(::log(a.v), 1.0 / a.v, a);
}
template <int D>
__host__ __device__ inline Dual<D> atan(const Dual<D>& a) {
  return chain1(::atan(a.v), 1.0 / (1.0 + a.v * a.v), a);
}
template <int D>
__host__ __device__ inline Dual<D> tan(const Dual<D>& a) {
  const double t = ::tan(a.v);
  return chain1(t, 1.0 + t * t, a);
}
template <int D>
__host__ __device__ inline Dual<D> cbrt(const Dual<D>& a) {
  const double c = ::cbrt(a.v);
  return chain1(c, c == 0.0 ? 0.0 : 1.0 / (3.0 * c * c), a);
}
template <int D>
__host__ __device__ inline Dual<D> fabs(const Dual<D>& a) {
  return a.v < 0.0 ? -a : a;
}
template <int D>
__host__ __device__ inline Dual<D> hypot(const Dual<D>& a, const Dual<D>& b) {
  const double h = ::hypot(a.v, b.v);
  Dual<D> r(h);
  if (h > 0.0) {
#pragma unroll
    for (int k = 0; k < D; ++k) r.d[k] = (a.v * a.d[k] + b.v * b.d[k]) / h;
  }
  return r;
}
template <int D>
__host__ __device__ inline Dual<D> fmax(const Dual<D>& a, double b) {
  return a.v >= b ? a : Dual<D>(b);
}
template <int D>
__host__ __device__ inline Dual<D> fmin(const Dual<D>& a, const Dual<D>& b) {
  return a.v <= b.v ? a : b;
}
// |a| with the sign of b's value
template <int D>
__host__ __device__ inline Dual<D> copysign(const Dual<D>& a, const Dual<D>& b) {
  const bool neg = ::copysign(1.0, b.v) < 0.0;
  const Dual<D> m = fabs(a);
  return neg ? -m : m;
}
// arg(x + i y): d = (x dy - y dx) / (x^2 + y^2)
template <int D>
__host__ __device__ inline Dual<D> atan2(const Dual<D>& y, const Dual<D>& x) {
  Dual<D> r(::atan2(y.v, x.v));
  const double r2 = x.v * x.v + y.v * y.v;
  if (r2 > 0.0) {
#pragma unroll
    for (int k = 0; k < D; ++k) r.d[k] = (x.v * y.d[k] - y.v * x.d[k]) / r2;
  }
  return r;
}

}  // namespace ad

// value of a scalar (double or dual)
__host__ __device__ inline double value(double x) { return x; }
template <int D>
__host__ __device__ inline double value(const ad::Dual<D>& x) {
  return x.v;
}

// true for the dual types (derivative-carrying scalars)
template <typename S>
struct is_dual {
  static constexpr bool value = false;
};
template <int D>
struct is_dual<ad::Dual<D>> {
  static constexpr bool value = true;
};

// The derivatives of `deriv` with the value v: a quantity whose value comes
// from one (reference) formula and whose derivatives come from another,
// mathematically identical one that does not cancel.
__host__ __device__ inline double with_value(double, double v) { return v; }
template <int D>
__host__ __device__ inline ad::Dual<D> with_value(const ad::Dual<D>& deriv, double v) {
  ad::Dual<D> r = deriv;
  r.v = v;
  return r;
}

}  // namespace bemhip
