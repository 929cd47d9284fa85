// Test-only host build of the near-singular rule templates (near_rule.hpp),
// the single source of the tan-p3c / Guiggiani steps used by the assembly
// kernels (plain FP64) and the adjoint (dual numbers). Compiled with g++ by
// tests/test_near_rule_host.py; never part of libbemhip.so.
#define __host__
#define __device__
#include <cmath>
#include <cstring>

#include "near_rule.hpp"

using namespace bemhip;

namespace {

QuadConst g_qc;

void legendre(int n, double x, double* pn, double* dpn) {
  double p0 = 1.0, p1 = x;
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

void gauss(int n, double* x, double* w) {
  if (n == 1) { x[0] = 0.0; w[0] = 2.0; return; }
  const int m = (n + 1) / 2;
  for (int i = 0; i < m; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pn, dpn;
    for (int it = 0; it < 100; ++it) {
      legendre(n, z, &pn, &dpn);
      const double dz = pn / dpn;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    legendre(n, z, &pn, &dpn);
    x[n - 1 - i] = z;
    x[i] = -z;
    w[n - 1 - i] = w[i] = 2.0 / ((1.0 - z * z) * dpn * dpn);
  }
  if (n % 2 == 1) {
    double pn, dpn;
    x[m - 1] = 0.0;
    legendre(n, 0.0, &pn, &dpn);
    w[m - 1] = 2.0 / (dpn * dpn);
  }
}

void diff_matrix(int n, const double* x, const double* bw, double* D) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      D[i * n + j] = (bw[j] / bw[i]) / (x[i] - x[j]);
      s += D[i * n + j];
    }
    D[i * n + i] = -s;
  }
}

constexpr int NN = 3;             // quadratic elements
constexpr int D = 2 + 2 * NN;     // source (x, y) + geometric nodes (x, y)

}  // namespace

extern "C" {

// Quadrature constants for p = 2, npg points, conductivity k.
void rule_init(int npg, double k, const double* elem_w) {
  QuadConst* q = &g_qc;
  std::memset(q, 0, sizeof(*q));
  q->npg = npg;
  q->p = 2;
  gauss(npg, q->gl_x, q->gl_w);
  gauss(kNinterp, q->g20_x, q->g20_w);
  for (int i = 0; i < kNinterp; ++i) {
    double prod = 1.0;
    for (int j = 0; j < kNinterp; ++j)
      if (j != i) prod *= q->g20_x[i] - q->g20_x[j];
    q->g20_bw[i] = 1.0 / prod;
  }
  double w[NN];
  gauss(NN, q->lp_x, w);
  for (int i = 0; i < NN; ++i)
    q->lp_bw[i] = ((i % 2 == 0) ? 1.0 : -1.0) * std::sqrt((1.0 - q->lp_x[i] * q->lp_x[i]) * w[i]);
  diff_matrix(NN, q->lp_x, q->lp_bw, q->lp_D);
  q->eq_x[0] = -1.0; q->eq_x[1] = 0.0; q->eq_x[2] = 1.0;
  q->eq_bw[0] = 1.0; q->eq_bw[1] = -2.0; q->eq_bw[2] = 1.0;
  diff_matrix(NN, q->eq_x, q->eq_bw, q->eq_D);
  for (int j = 0; j < NN; ++j) q->ew[j] = elem_w[j];
  q->k = k;
  q->near_factor = 1.5;
}

// h, g of one pair and their derivatives d/d(source x, y, geometric nodes
// x0, y0, x1, y1, x2, y2), as the adjoint evaluates them: collocation values
// as stored, their derivatives through x_k = sum_m Ngeo_m(xi_k) X_m. kloc >= 0
// marks an on-element pair (the source is collocation node kloc).
void rule_pair_dual(const double* gxv, const double* gyv, const double* cxv, const double* cyv,
                    double px, double py, double nrx, double nry, int kloc, double* h, double* g,
                    double* dh, double* dg) {
  using Dl = ad::Dual<D>;
  double Ng[NN][NN], dNg[NN][NN];
  for (int k = 0; k < NN; ++k)
    rule::shapefun<double, NN>(g_qc.eq_x, g_qc.eq_bw, g_qc.eq_D, g_qc.lp_x[k], Ng[k], dNg[k]);
  Dl gx[NN], gy[NN], cx[NN], cy[NN];
  for (int m = 0; m < NN; ++m) {
    gx[m] = Dl::seed(gxv[m], 2 + 2 * m);
    gy[m] = Dl::seed(gyv[m], 3 + 2 * m);
  }
  for (int k = 0; k < NN; ++k) {
    Dl sx = 0.0, sy = 0.0;
    for (int m = 0; m < NN; ++m) { sx += Ng[k][m] * gx[m]; sy += Ng[k][m] * gy[m]; }
    sx.v = cxv[k];
    sy.v = cyv[k];
    cx[k] = sx;
    cy[k] = sy;
  }
  Dl hs[NN], gs[NN];
  if (kloc >= 0) {
    const double a = rule::closest_param<NN>(g_qc, cxv, cyv, cxv[kloc], cyv[kloc]);
    rule::singular_pair<Dl, NN>(g_qc, a, gx, gy, cx[kloc], cy[kloc], nrx, nry, hs, gs);
  } else {
    rule::near_pair<Dl, NN>(g_qc, cx, cy, gx, gy, Dl::seed(px, 0), Dl::seed(py, 1), nrx, nry,
                            hs, gs);
  }
  for (int k = 0; k < NN; ++k) {
    h[k] = hs[k].v;
    g[k] = gs[k].v;
    for (int t = 0; t < D; ++t) {
      dh[k * D + t] = hs[k].d[t];
      dg[k * D + t] = gs[k].d[t];
    }
  }
}

// The same pair in plain FP64 (the assembly's arithmetic).
void rule_pair(const double* gxv, const double* gyv, const double* cxv, const double* cyv,
               double px, double py, double nrx, double nry, int kloc, double* h, double* g) {
  if (kloc >= 0) {
    const double a = rule::closest_param<NN>(g_qc, cxv, cyv, cxv[kloc], cyv[kloc]);
    rule::singular_pair<double, NN>(g_qc, a, gxv, gyv, cxv[kloc], cyv[kloc], nrx, nry, h, g);
  } else {
    rule::near_pair<double, NN>(g_qc, cxv, cyv, gxv, gyv, px, py, nrx, nry, h, g);
  }
}

}  // extern "C"
