// Dense 2-D isotropic elasticity (Kelvin) for the BEM path (gfx950):
// H_G_full_direct for a Vectorial problem (src/Core/Assembly_full.jl:120-175)
// and the per-dof mixed-BC system (src/Core/Boundary_conditions.jl:153-175).
//
// Layout: node j's component b is dof 2j + b. H (2nt x 2nt) and G (2nt x 2n)
// are FP64, column-major with the padded leading dimension ld64 of 2nt rows,
// so the rows 2i, 2i + 1 of a column are adjacent (one 16-byte store).
//
// Kernels, in launch order on the library stream:
//   near_classify / near_scan / near_fill   the Laplace assembly's near-pair
//                   CSR (the classification depends on the geometry only).
//   kelvin_far_kernel<TS, NN, STORE>   rows x 32-element column blocks: the
//                   pair test of near_classify and the far nodal lumping of
//                   _far_nodal_vec! (:812-822) as 2x2 blocks T * w, U * w; near
//                   blocks are written 0; FP64 row sums of H per (row, column)
//                   component. FP64 storage keeps the reference's operation
//                   order; FP32 storage and the matrix-free row sums use one
//                   division per pair.
//   kelvin_near_reg_kernel<NN, G> / kelvin_near_sing_kernel<NN>   one lane
//                   group per near pair: the tan-p3c rule (vectorial
//                   _accumulate_kernels!, :697-713) and the interpolated
//                   Guiggiani rule (vectorial _integrate_singular_guiggiani!,
//                   :426-457; orders (0, -1) per component) with the Kelvin
//                   kernel; [pair][node][component] values in the pair list.
//   kelvin_finalize  rigid-body free term (_vectorial_free_term!, :777-786):
//                   H[ii, ii] block = -(sum over the other nodes' blocks) per
//                   component, and the near blocks into H and G.
//   kelvin_rhs_part, kelvin_swap_kernel   applyBC: b (chunks of dofs, each in
//                   the reference's loop order, then a fixed-order sum of the
//                   chunks) and A = H with Dirichlet columns -G.
//   kelvin_apply_part / kelvin_apply_sum   y = A x from the stored H, G with
//                   the BC column selection (chunks of columns, fixed-order
//                   sum): the FP64 residual of the mixed-precision refinement.
//
// Compiled with -ffp-contract=off (the un-fused FP64 operation order).

#include "common.hpp"
#include "assembly_shared.hpp"

#include <cmath>

namespace bemhip {

namespace {

typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

// The two rows (2i, 2i + 1) of one column: one 16-byte (FP64) or 8-byte (FP32)
// store; the FP32 storage rounds the FP64 values.
template <typename TS>
__device__ __forceinline__ void store_pair_nt(TS* p, double a, double b) {
  if constexpr (sizeof(TS) == 8)
    __builtin_nontemporal_store(d2v{a, b}, (d2v*)p);
  else
    __builtin_nontemporal_store(f2v{(float)a, (float)b}, (f2v*)p);
}
template <typename TS>
__device__ __forceinline__ void store_pair(TS* p, double a, double b) {
  if constexpr (sizeof(TS) == 8)
    *(d2v*)p = d2v{a, b};
  else
    *(f2v*)p = f2v{(float)a, (float)b};
}

// Kelvin constants of fundamental(props::Elasticity, r, n)
// (src/Elasticity/Fundamental.jl:28-57): prod1 = 4 pi (1 - nu), base =
// 2 prod1 mu, c34 = 3 - 4 nu, c12 = 1 - 2 nu.
struct Kelvin {
  double prod1, base, c34, c12;
  double inv_base, inv_prod1;  // 1 / base, 1 / prod1 (operator passes)
};

Kelvin make_kelvin(const bemhip_ctx* c) {
  Kelvin K;
  K.prod1 = (4.0 * kPi) * (1.0 - c->nu);
  K.base = (2.0 * K.prod1) * c->mu;
  K.c34 = 3.0 - 4.0 * c->nu;
  K.c12 = 1.0 - 2.0 * c->nu;
  K.inv_base = 1.0 / K.base;
  K.inv_prod1 = 1.0 / K.prod1;
  return K;
}

// U (c = 2a + b; symmetric) and T at r = x_field - x_source, n the field
// normal, with the reference's operation order.
__device__ __forceinline__ void kelvin_UT(const Kelvin& K, double rx, double ry, double nx,
                                          double ny, double* U, double* T) {
  const double R2 = rx * rx + ry * ry;
  const double R = sqrt(R2);
  const double dr1 = rx / R, dr2 = ry / R;
  const double drdn = dr1 * nx + dr2 * ny;
  const double prod2 = K.c34 * log(1.0 / R);
  U[0] = (prod2 + dr1 * dr1) / K.base;
  U[1] = (dr1 * dr2) / K.base;
  U[2] = U[1];
  U[3] = (prod2 + dr2 * dr2) / K.base;
  const double fat = 1.0 / (K.prod1 * R);
  T[0] = ((-drdn) * (K.c12 + 2.0 * (dr1 * dr1))) * fat;
  T[3] = ((-drdn) * (K.c12 + 2.0 * (dr2 * dr2))) * fat;
  T[1] = (-(((drdn * 2.0) * dr1) * dr2 - K.c12 * (dr1 * ny - dr2 * nx))) * fat;
  T[2] = (-(((drdn * 2.0) * dr1) * dr2 - K.c12 * (dr2 * nx - dr1 * ny))) * fat;
}

// The same U, T with one division (1 / R) and the constant reciprocals: for the
// FP64 operator passes, whose sums are checked to a relative 1e-13, not the
// stored entries (a few ulps from kelvin_UT's values).
__device__ __forceinline__ void kelvin_UT_op(const Kelvin& K, double rx, double ry, double nx,
                                             double ny, double* U, double* T) {
  const double R = sqrt(rx * rx + ry * ry);
  const double rinv = 1.0 / R;
  const double dr1 = rx * rinv, dr2 = ry * rinv;
  const double drdn = dr1 * nx + dr2 * ny;
  const double prod2 = -K.c34 * log(R);
  U[0] = (prod2 + dr1 * dr1) * K.inv_base;
  U[1] = (dr1 * dr2) * K.inv_base;
  U[2] = U[1];
  U[3] = (prod2 + dr2 * dr2) * K.inv_base;
  const double fat = K.inv_prod1 * rinv;
  T[0] = ((-drdn) * (K.c12 + 2.0 * (dr1 * dr1))) * fat;
  T[3] = ((-drdn) * (K.c12 + 2.0 * (dr2 * dr2))) * fat;
  T[1] = (-(((drdn * 2.0) * dr1) * dr2 - K.c12 * (dr1 * ny - dr2 * nx))) * fat;
  T[2] = (-(((drdn * 2.0) * dr1) * dr2 - K.c12 * (dr2 * nx - dr1 * ny))) * fat;
}

// ---------------------------------------------------------------------------
// Far field. Block = 256 row nodes x 32 elements, element data staged in LDS.
// Lane i owns rows 2i, 2i + 1: the store of a column's two rows is 16 bytes,
// a wave's stores are 1 KB contiguous. hpart: [4][ncb][nt] row sums of H.
// ---------------------------------------------------------------------------
// STORE = false: the row sums only (the matrix-free assembly's diagonal).
template <typename TS, int NN, bool STORE = true>
__global__ __launch_bounds__(kFarThreads) void kelvin_far_kernel(Geo g, Consts cs, Kelvin K,
                                                                 const double* __restrict__ cbox,
                                                                 const double* __restrict__ rbox,
                                                                 TS* __restrict__ H,
                                                                 TS* __restrict__ G,
                                                                 double* __restrict__ hpart) {
  __shared__ double s_cx[kEB * NN], s_cy[kEB * NN], s_nx[kEB * NN], s_ny[kEB * NN],
      s_jw[kEB * NN], s_L[kEB];
  __shared__ float s_lim2f[kEB];
  const int cb = blockIdx.y;
  const int ncb = gridDim.y;
  const int e0 = cb * kEB;
  const int ecount = min(kEB, g.ne - e0);
  stage_class_block<NN>(g, cs, e0, ecount, s_cx, s_cy, s_L, s_lim2f);
  for (int t = threadIdx.x; t < ecount * NN; t += blockDim.x) {
    const int j = e0 * NN + t;
    s_nx[t] = g.nx[j]; s_ny[t] = g.ny[j];
    s_jw[t] = g.jw[j];
  }
  // all-far tiles (the box test of near_classify) skip the pair tests
  const bool all_far = tile_all_far(cbox, rbox, cb, blockIdx.x, cs.nf);
  __syncthreads();
  const int i = blockIdx.x * kFarThreads + threadIdx.x;
  if (i >= g.nt) return;
  const double px = g.rx[i], py = g.ry[i];
  const size_t ld = g.ld;
  double rs[4] = {0.0, 0.0, 0.0, 0.0};
  for (int el = 0; el < ecount; ++el) {
    double rx[NN], ry[NN];
    const int col0 = (e0 + el) * NN;
    bool nr = false;
    if (all_far) {
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        rx[k] = s_cx[el * NN + k] - px;
        ry[k] = s_cy[el * NN + k] - py;
      }
    } else {
      float rxf[NN], ryf[NN], d2f[NN];
      const bool onel = (i + g.row0 >= col0) && (i + g.row0 < col0 + NN);
      nr = classify_pair<NN>(px, py, s_cx + el * NN, s_cy + el * NN, s_L[el], s_lim2f[el], cs.nf,
                             onel, rx, ry, rxf, ryf, d2f);
    }
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      const size_t c0 = (size_t)(2 * (col0 + k)) * ld + 2 * (size_t)i;  // column 2j, row 2i
      double h[4] = {0.0, 0.0, 0.0, 0.0}, gv[4] = {0.0, 0.0, 0.0, 0.0};
      if (!nr) {
        const int t = el * NN + k;
        double U[4], T[4];
        // FP64 storage: the reference's operation order; FP32 storage and the
        // row sums alone: one division per pair (entries within FP64 ulps)
        if constexpr (sizeof(TS) == 8 && STORE)
          kelvin_UT(K, rx[k], ry[k], s_nx[t], s_ny[t], U, T);
        else
          kelvin_UT_op(K, rx[k], ry[k], s_nx[t], s_ny[t], U, T);
        const double w = s_jw[t];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          h[c] = T[c] * w;
          gv[c] = U[c] * w;
          rs[c] += h[c];
        }
      }
      if (STORE) {  // column 2j: rows (2i, 2i + 1) = (c 0, c 2); column 2j + 1: (c 1, c 3)
        store_pair_nt(H + c0, h[0], h[2]);
        store_pair_nt(H + c0 + ld, h[1], h[3]);
        store_pair_nt(G + c0, gv[0], gv[2]);
        store_pair_nt(G + c0 + ld, gv[1], gv[3]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) hpart[((size_t)c * ncb + cb) * g.nt + i] = rs[c];
}

// ---------------------------------------------------------------------------
// Near pairs (off-element): tan-p3c rule, G lanes per pair. Values per pair:
// ph/pg[pair][k][c], k the element node, c = 2a + b.
// ---------------------------------------------------------------------------
template <int NN, int G>
__global__ __launch_bounds__(256) void kelvin_near_reg_kernel(
    Geo g, Kelvin K, const QuadConst* __restrict__ qc, int64_t count,
    const int* __restrict__ list, const int* __restrict__ prow, const int* __restrict__ pelem,
    double* __restrict__ ph, double* __restrict__ pg) {
  constexpr int NV = 4 * NN;
  const int64_t k_raw =
      ((int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * (kWave / G) +
      (threadIdx.x & (kWave - 1)) / G;
  const bool active = k_raw < count;
  const int64_t kk = active ? k_raw : count - 1;  // idle groups redo the last pair
  const int64_t pair = list ? (int64_t)list[kk] : kk;
  const int lane = threadIdx.x & (G - 1);
  const int i = prow[pair];
  const int e = pelem[pair];
  const double px = g.rx[i], py = g.ry[i];
  double cx[NN], cy[NN], gxx[NN], gyy[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    cx[k] = g.x[e * NN + k]; cy[k] = g.y[e * NN + k];
    gxx[k] = g.gx[e * NN + k]; gyy[k] = g.gy[e * NN + k];
  }
  const double nrefx = g.nx[e * NN], nrefy = g.ny[e * NN];
  const double* lx = qc->lp_x;
  const double* lbw = qc->lp_bw;
  const double* lD = qc->lp_D;
  const double* ex = qc->eq_x;
  const double* ebw = qc->eq_bw;
  const double* eD = qc->eq_D;
  double hs[NV], gs[NV];
  {
    double zeta, eta;
    complex_pole<NN>(lx, cx, cy, px, py, zeta, eta);
    TanPlan P;
    tanp3c_plan(zeta, eta, P);
    const int npg = qc->npg;
    const int nq = tanp3c_count(P, npg);
    // normal orientation from the first rule node (src/Core/Assembly_full.jl:664-667)
    double flip = 1.0;
    {
      double x0, w0;
      tanp3c_node(P, 0, npg, qc->gl_x, qc->gl_w, x0, w0);
      double Ng[NN], dNg[NN];
      shapefun<NN>(ex, ebw, eD, x0, Ng, dNg);
      double dxx = 0.0, dxy = 0.0;
#pragma unroll
      for (int k = 0; k < NN; ++k) { dxx += dNg[k] * gxx[k]; dxy += dNg[k] * gyy[k]; }
      const double J = sqrt(dxx * dxx + dxy * dxy);
      const double nx = dxy / J, ny = -(dxx / J);
      if (nx * nrefx + ny * nrefy < 0.0) flip = -1.0;
    }
    double ha[NV], ga[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { ha[v] = 0.0; ga[v] = 0.0; }
    for (int q0 = 0; q0 < nq; q0 += G) {
      const int q = q0 + lane;
      if (q < nq) {
        double xq, wq;
        tanp3c_node(P, q, npg, qc->gl_x, qc->gl_w, xq, wq);
        double N[NN], dN[NN], Ng[NN], dNg[NN];
        shapefun<NN>(lx, lbw, lD, xq, N, dN);
        shapefun<NN>(ex, ebw, eD, xq, Ng, dNg);
        double pgx = 0.0, pgy = 0.0, dxx = 0.0, dxy = 0.0;
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          pgx += Ng[k] * gxx[k]; pgy += Ng[k] * gyy[k];
          dxx += dNg[k] * gxx[k]; dxy += dNg[k] * gyy[k];
        }
        const double rx = pgx - px, ry = pgy - py;
        const double J = sqrt(dxx * dxx + dxy * dxy);
        double nx = dxy / J, ny = -(dxx / J);
        if (flip < 0.0) { nx = -nx; ny = -ny; }
        double U[4], T[4];
        kelvin_UT(K, rx, ry, nx, ny, U, T);
        const double wi = J * wq;
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          const double Nj = N[k] * wi;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            ha[k * 4 + c] += T[c] * Nj;
            ga[k * 4 + c] += U[c] * Nj;
          }
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) { hs[v] = group_sum<G>(ha[v]); gs[v] = group_sum<G>(ga[v]); }
  }
  if (active && lane < NV) {
    double hv = hs[0], gv = gs[0];
#pragma unroll
    for (int v = 1; v < NV; ++v)
      if (lane == v) { hv = hs[v]; gv = gs[v]; }
    ph[(size_t)pair * NV + lane] = hv;
    pg[(size_t)pair * NV + lane] = gv;
  }
}

// ---------------------------------------------------------------------------
// On-element pairs: interpolated Guiggiani over 20 nodes, 32 lanes per pair
// (two pairs per wavefront), every component through the scalar rule.
// ---------------------------------------------------------------------------
template <int NN>
__global__ __launch_bounds__(256) void kelvin_near_sing_kernel(
    Geo g, Kelvin K, const QuadConst* __restrict__ qc, int64_t count,
    const int* __restrict__ list, const int* __restrict__ prow, const int* __restrict__ pelem,
    double* __restrict__ ph, double* __restrict__ pg) {
  constexpr int G = 32;
  constexpr int NV = 4 * NN;
  const int64_t k_raw =
      ((int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * 2 +
      ((threadIdx.x & (kWave - 1)) >> 5);
  const bool active = k_raw < count;
  const int64_t kk = active ? k_raw : count - 1;  // idle groups redo the last pair
  const int64_t pair = list ? (int64_t)list[kk] : kk;
  const int lane = threadIdx.x & (G - 1);
  const int gshift = threadIdx.x & G;  // 0 or 32: bit offset of this group in a ballot
  const int i = prow[pair];
  const int e = pelem[pair];
  const double px = g.rx[i], py = g.ry[i];
  double cx[NN], cy[NN], gxx[NN], gyy[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    cx[k] = g.x[e * NN + k]; cy[k] = g.y[e * NN + k];
    gxx[k] = g.gx[e * NN + k]; gyy[k] = g.gy[e * NN + k];
  }
  const double nrefx = g.nx[e * NN], nrefy = g.ny[e * NN];
  const double* lx = qc->lp_x;
  const double* lbw = qc->lp_bw;
  const double* lD = qc->lp_D;
  const double* ex = qc->eq_x;
  const double* ebw = qc->eq_bw;
  const double* eD = qc->eq_D;
  double hs[NV], gs[NV];
  {
    // closest point on the collocation interpolant (src/Core/Integration.jl:323-349)
    double a;
    {
      const double dxs = cx[NN - 1] - cx[0], dys = cy[NN - 1] - cy[0];
      const double nd = sqrt(dxs * dxs + dys * dys);
      double xi = (lx[NN - 1] - lx[0]) * (dxs * (px - cx[0]) + dys * (py - cy[0])) /
                      (nd * nd + kEps) + lx[0];
      xi = clampd(xi, -1.0, 1.0);
      for (int it = 0; it < 25; ++it) {
        double N[NN], dN[NN], d2N[NN];
        shapefun<NN>(lx, lbw, lD, xi, N, dN);
#pragma unroll
        for (int j = 0; j < NN; ++j) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < NN; ++q) s += dN[q] * lD[q * NN + j];
          d2N[j] = s;
        }
        double X = 0, Y = 0, DX = 0, DY = 0, D2X = 0, D2Y = 0;
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          X += N[k] * cx[k]; Y += N[k] * cy[k];
          DX += dN[k] * cx[k]; DY += dN[k] * cy[k];
          D2X += d2N[k] * cx[k]; D2Y += d2N[k] * cy[k];
        }
        const double rx = X - px, ry = Y - py;
        const double res = rx * DX + ry * DY;
        const double Jn = (DX * DX + DY * DY) + (rx * D2X + ry * D2Y);
        if (fabs(Jn) < 1e-30) break;
        const double d = res / Jn;
        xi -= d;
        if (fabs(d) < 1e-14) break;
      }
      if (xi < -1.0 || xi > 1.0) {
        double N[NN], dN[NN];
        double X0 = 0, Y0 = 0, X1 = 0, Y1 = 0;
        shapefun<NN>(lx, lbw, lD, -1.0, N, dN);
#pragma unroll
        for (int k = 0; k < NN; ++k) { X0 += N[k] * cx[k]; Y0 += N[k] * cy[k]; }
        shapefun<NN>(lx, lbw, lD, 1.0, N, dN);
#pragma unroll
        for (int k = 0; k < NN; ++k) { X1 += N[k] * cx[k]; Y1 += N[k] * cy[k]; }
        const double d0 = sqrt((X0 - px) * (X0 - px) + (Y0 - py) * (Y0 - py));
        const double d1 = sqrt((X1 - px) * (X1 - px) + (Y1 - py) * (Y1 - py));
        xi = (d0 <= d1) ? -1.0 : 1.0;
      }
      a = xi;
    }
    a = clampd(a, __longlong_as_double(0xBFEFFFFFFFFFFFFFll),
               __longlong_as_double(0x3FEFFFFFFFFFFFFFll));
    // interpolated-Guiggiani rule (src/Core/Integration.jl:1654-1670)
    const bool lane_on = lane < kNinterp;
    const double xm = lane_on ? qc->g20_x[lane] : 0.0;
    const bool keep = lane_on && !(fabs(xm - a) <= 1e-12);
    const uint32_t keep_mask = (uint32_t)(__ballot(keep) >> gshift);
    const int m = __popc(keep_mask);
    double bw = 0.0;
    if (keep) {
      if (m == kNinterp) {
        bw = qc->g20_bw[lane];
      } else {
        double prod = 1.0;
        for (int j = 0; j < kNinterp; ++j) {
          const double xj = qc->g20_x[j];
          if (j == lane || fabs(xj - a) <= 1e-12) continue;
          prod *= xm - xj;
        }
        bw = 1.0 / prod;
      }
    }
    const double dm = a - xm;
    const double bnum = keep ? bw / dm : 0.0;
    const double msum = group_sum<G>(bnum);
    const double Ni = keep ? bnum * (1.0 / msum) : 0.0;
    const double delta = xm - a;
    // kernel samples F = (U, T) * N_k * J (src/Core/Assembly_full.jl:437-447)
    double Fg[NV], Fh[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { Fg[v] = 0.0; Fh[v] = 0.0; }
    if (keep) {
      double N[NN], dN[NN], Ng[NN], dNg[NN];
      shapefun<NN>(lx, lbw, lD, xm, N, dN);
      shapefun<NN>(ex, ebw, eD, xm, Ng, dNg);
      double pgx = 0.0, pgy = 0.0, dxx = 0.0, dxy = 0.0;
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        pgx += Ng[k] * gxx[k]; pgy += Ng[k] * gyy[k];
        dxx += dNg[k] * gxx[k]; dxy += dNg[k] * gyy[k];
      }
      const double J = sqrt(dxx * dxx + dxy * dxy);
      const double rx = pgx - px, ry = pgy - py;
      if (!(J < 1e-30) && !(sqrt(rx * rx + ry * ry) < 1e-30)) {
        double nx = dxy / J, ny = -(dxx / J);
        if (nx * nrefx + ny * nrefy < 0.0) { nx = -nx; ny = -ny; }
        double U[4], T[4];
        kelvin_UT(K, rx, ry, nx, ny, U, T);
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          const double NjJ = N[k] * J;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            Fg[k * 4 + c] = U[c] * NjJ;
            Fh[k * 4 + c] = T[c] * NjJ;
          }
        }
      }
    }
    // log-coefficient fit for G (order 0) and CPV coefficient for H (order -1)
    const bool used = keep && !(fabs(delta) < 1e-14);
    const double lg = used ? log(fabs(delta)) : 0.0;
    const double sL = group_sum<G>(lg);
    const double sL2 = group_sum<G>(lg * lg);
    const int nused = __popc((uint32_t)(__ballot(used) >> gshift));
    const double detA = sL2 * nused - sL * sL;
    const double wq = keep ? qc->g20_w[lane] : 0.0;
    const double hfp1 = log(fabs((1.0 - a) / (1.0 + a)));
    const double logp = ((1.0 - a) * log(fabs(1.0 - a)) + (1.0 + a) * log(fabs(1.0 + a))) - 2.0;
    const double lgd = keep ? log(fabs(delta)) : 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double Fm1g = group_sum<G>(keep ? Ni * (Fg[v] * delta) : 0.0);
      const double F1h = group_sum<G>(keep ? Ni * (Fh[v] * delta) : 0.0);
      const double wi = keep ? Fg[v] - Fm1g / delta : 0.0;
      const double sW = group_sum<G>(used ? wi : 0.0);
      const double sWL = group_sum<G>(used ? wi * lg : 0.0);
      double F0 = 0.0;
      if (nused >= 2 && !(fabs(detA) < 1e-30)) F0 = (nused * sWL - sL * sW) / detA;
      const double regg = keep ? wq * ((Fg[v] - Fm1g / delta) - F0 * lgd) : 0.0;
      const double regh = keep ? wq * (Fh[v] - F1h / delta) : 0.0;
      gs[v] = group_sum<G>(regg) + (Fm1g * hfp1 + F0 * logp);
      hs[v] = group_sum<G>(regh) + F1h * hfp1;
    }
  }
  if (active && lane < NV) {
    double hv = hs[0], gv = gs[0];
#pragma unroll
    for (int v = 1; v < NV; ++v)
      if (lane == v) { hv = hs[v]; gv = gs[v]; }
    ph[(size_t)pair * NV + lane] = hv;
    pg[(size_t)pair * NV + lane] = gv;
  }
}

// ---------------------------------------------------------------------------
// Rigid-body free term and near blocks, 64 row nodes per 256-thread block:
// D_c = -(far row sums + near values of the other nodes), c = 2a + b; the
// near blocks of row i into H and G, node i's own H block = D.
// ---------------------------------------------------------------------------
template <typename TS>
__global__ __launch_bounds__(256) void kelvin_finalize(int nt, int nn, int ncb, size_t ld,
                                                       const double* __restrict__ part,
                                                       const int* __restrict__ row_ptr,
                                                       const int* __restrict__ pair_elem,
                                                       const double* __restrict__ ph,
                                                       const double* __restrict__ pg,
                                                       double* __restrict__ diag,
                                                       TS* __restrict__ H,
                                                       TS* __restrict__ G) {
  __shared__ double s_red[4][4][kFinRows];
  const int i = blockIdx.x * kFinRows + (threadIdx.x & (kWave - 1));
  double s[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) s[c] = sum_partials(part + (size_t)c * ncb * nt, ncb, nt, i, s_red[c]);
  if (threadIdx.x >= kWave || i >= nt) return;  // wave 0 finishes its 64 rows
  const int p0 = row_ptr[i], p1 = row_ptr[i + 1];
  for (int pr = p0; pr < p1; ++pr) {
    const int c0 = pair_elem[pr] * nn;
    for (int k = 0; k < nn; ++k)
      if (c0 + k != i)
        for (int c = 0; c < 4; ++c) s[c] += ph[((size_t)pr * nn + k) * 4 + c];
  }
  const double D[4] = {-s[0], -s[1], -s[2], -s[3]};
#pragma unroll
  for (int c = 0; c < 4; ++c) diag[4 * (size_t)i + c] = D[c];
  if (!H) return;  // operator-only assembly: the FP64 diagonal blocks
  for (int pr = p0; pr < p1; ++pr) {
    const int c0 = pair_elem[pr] * nn;
    for (int k = 0; k < nn; ++k) {
      const int j = c0 + k;
      const double* hv = ph + ((size_t)pr * nn + k) * 4;
      const double* gv = pg + ((size_t)pr * nn + k) * 4;
      const size_t at = (size_t)(2 * j) * ld + 2 * (size_t)i;
      if (j == i) {
        store_pair(H + at, D[0], D[2]);
        store_pair(H + at + ld, D[1], D[3]);
      } else {
        store_pair(H + at, hv[0], hv[2]);
        store_pair(H + at + ld, hv[1], hv[3]);
      }
      store_pair(G + at, gv[0], gv[2]);
      store_pair(G + at + ld, gv[1], gv[3]);
    }
  }
  const size_t ad = (size_t)(2 * i) * ld + 2 * (size_t)i;  // internal rows: no own pair
  store_pair(H + ad, D[0], D[2]);
  store_pair(H + ad + ld, D[1], D[3]);
}

// ---------------------------------------------------------------------------
// FP64 operator pass without stored FP64 matrices (mixed storage):
// y = H wH + G wG with the far blocks recomputed in FP64 (kelvin_UT_op), the near blocks from the pair list and the FP64 diagonal
// blocks. kelvin_op64_kernel: partial sums of rows 2i, 2i + 1 per 32-element
// column block, part[a][cb][i]; kelvin_op_finalize adds them in order with the
// near and diagonal terms.
// ---------------------------------------------------------------------------
template <int NN>
__global__ __launch_bounds__(kFarThreads) void kelvin_op64_kernel(Geo g, Consts cs, Kelvin K,
                                                                  const double* __restrict__ wH,
                                                                  const double* __restrict__ wG,
                                                                  double* __restrict__ part) {
  __shared__ double s_cx[kEB * NN], s_cy[kEB * NN], s_nx[kEB * NN], s_ny[kEB * NN],
      s_jw[kEB * NN], s_L[kEB];
  __shared__ double s_w[4][kEB * NN];  // wH, wG of columns 2j and 2j + 1
  __shared__ float s_lim2f[kEB];
  const int cb = blockIdx.y;
  const int ncb = gridDim.y;
  const int e0 = cb * kEB;
  const int ecount = min(kEB, g.ne - e0);
  stage_class_block<NN>(g, cs, e0, ecount, s_cx, s_cy, s_L, s_lim2f);
  for (int t = threadIdx.x; t < ecount * NN; t += blockDim.x) {
    const int j = e0 * NN + t;
    s_nx[t] = g.nx[j]; s_ny[t] = g.ny[j];
    s_jw[t] = g.jw[j];
    s_w[0][t] = wH[2 * j];
    s_w[1][t] = wH[2 * j + 1];
    s_w[2][t] = wG[2 * j];
    s_w[3][t] = wG[2 * j + 1];
  }
  __syncthreads();
  const int i = blockIdx.x * kFarThreads + threadIdx.x;
  if (i >= g.nt) return;
  const double px = g.rx[i], py = g.ry[i];
  double acc0 = 0.0, acc1 = 0.0;
  for (int el = 0; el < ecount; ++el) {
    double rx[NN], ry[NN];
    float rxf[NN], ryf[NN], d2f[NN];
    const int col0 = (e0 + el) * NN;
    const bool onel = (i + g.row0 >= col0) && (i + g.row0 < col0 + NN);
    if (classify_pair<NN>(px, py, s_cx + el * NN, s_cy + el * NN, s_L[el], s_lim2f[el], cs.nf,
                          onel, rx, ry, rxf, ryf, d2f))
      continue;
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      const int t = el * NN + k;
      double U[4], T[4];
      kelvin_UT_op(K, rx[k], ry[k], s_nx[t], s_ny[t], U, T);
      const double w = s_jw[t];
      const double h0 = T[0] * w, h1 = T[1] * w, h2 = T[2] * w, h3 = T[3] * w;
      const double g0 = U[0] * w, g1 = U[1] * w, g2 = U[2] * w, g3 = U[3] * w;
      acc0 += ((h0 * s_w[0][t] + h1 * s_w[1][t]) + g0 * s_w[2][t]) + g1 * s_w[3][t];
      acc1 += ((h2 * s_w[0][t] + h3 * s_w[1][t]) + g2 * s_w[2][t]) + g3 * s_w[3][t];
    }
  }
  part[(size_t)cb * g.nt + i] = acc0;
  part[((size_t)ncb + cb) * g.nt + i] = acc1;
}

__global__ __launch_bounds__(256) void kelvin_op_finalize(
    int nt, int n, int nn, int ncb, const double* __restrict__ part,
    const int* __restrict__ row_ptr, const int* __restrict__ pair_elem,
    const double* __restrict__ ph, const double* __restrict__ pg, const double* __restrict__ diag,
    const double* __restrict__ wH, const double* __restrict__ wG, double* __restrict__ y) {
  __shared__ double s_red[2][4][kFinRows];
  const int i = blockIdx.x * kFinRows + (threadIdx.x & (kWave - 1));
  double s0 = sum_partials(part, ncb, nt, i, s_red[0]);
  double s1 = sum_partials(part + (size_t)ncb * nt, ncb, nt, i, s_red[1]);
  if (threadIdx.x >= kWave || i >= nt) return;  // wave 0 finishes its 64 rows
  const double* D = diag + 4 * (size_t)i;
  for (int pr = row_ptr[i]; pr < row_ptr[i + 1]; ++pr) {
    const int c0 = pair_elem[pr] * nn;
    for (int k = 0; k < nn; ++k) {
      const int j = c0 + k;
      const double* h = (j == i) ? D : ph + ((size_t)pr * nn + k) * 4;
      const double* gv = pg + ((size_t)pr * nn + k) * 4;
      const double a = wH[2 * j], b = wH[2 * j + 1], ga = wG[2 * j], gb = wG[2 * j + 1];
      s0 += ((h[0] * a + h[1] * b) + gv[0] * ga) + gv[1] * gb;
      s1 += ((h[2] * a + h[3] * b) + gv[2] * ga) + gv[3] * gb;
    }
  }
  if (i >= n) {  // internal row: its own block is in no pair
    s0 += D[0] * wH[2 * i] + D[1] * wH[2 * i + 1];
    s1 += D[2] * wH[2 * i] + D[3] * wH[2 * i + 1];
  }
  y[2 * (size_t)i] = s0;
  y[2 * (size_t)i + 1] = s1;
}

// Operator weights per dof (boundary dofs j < nb): apply (x): wH = BC 0 ? 0 : x,
// wG = BC 0 ? -x : 0; right-hand side (bv): wH = BC 0 ? -bv : 0, wG = BC 0 ?
// 0 : bv. Internal dofs: wH = x (apply) or 0 (right-hand side).
__global__ void kelvin_weights_kernel(int ns, int nb, int rhs, const int* __restrict__ bc,
                                      const double* __restrict__ v, double* __restrict__ wH,
                                      double* __restrict__ wG) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  if (j >= nb) { wH[j] = rhs ? 0.0 : v[j]; return; }
  const bool dir = bc[j] == 0;
  if (rhs) {
    wH[j] = dir ? -v[j] : 0.0;
    wG[j] = dir ? 0.0 : v[j];
  } else {
    wH[j] = dir ? 0.0 : v[j];
    wG[j] = dir ? -v[j] : 0.0;
  }
}

// y = A x with A's columns read from H or -G: part[chunk][r] over columns
// [chunk * kApplyCols, (chunk + 1) * kApplyCols), then the chunks summed in
// order (kelvin_apply_sum).
constexpr int kApplyCols = 256;

// b_r = sum over the boundary dofs j of -H_rj BV_j (BC 0) or + G_rj BV_j
// (src/Core/Boundary_conditions.jl:163-172): each chunk of columns adds its
// terms in the reference's dof order, the chunks are summed in order.
__global__ __launch_bounds__(256) void kelvin_rhs_part(int ns, int nb, size_t ld,
                                                       const int* __restrict__ bc,
                                                       const double* __restrict__ bv,
                                                       const double* __restrict__ H,
                                                       const double* __restrict__ G,
                                                       double* __restrict__ part) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * kApplyCols;
  const int j1 = min(nb, j0 + kApplyCols);
  if (r >= ns) return;
  double s = 0.0;
  for (int j = j0; j < j1; ++j) {
    const size_t at = (size_t)j * ld + r;
    if (bc[j] == 0)
      s -= H[at] * bv[j];
    else
      s += G[at] * bv[j];
  }
  part[(size_t)blockIdx.y * ns + r] = s;
}

// A = H with the Dirichlet columns (BC 0) replaced by -G; workgroup row y
// takes columns y, y + gridDim.y, ...
template <typename TS>
__global__ void kelvin_swap_kernel(int ns, int nb, size_t ld, const int* __restrict__ bc,
                                   const TS* __restrict__ H, const TS* __restrict__ G,
                                   TS* __restrict__ A) {
  for (int j = blockIdx.y; j < ns; j += gridDim.y) {
    const bool dir = (j < nb) && bc[j] == 0;
    const TS* src = dir ? G + (size_t)j * ld : H + (size_t)j * ld;
    TS* dst = A + (size_t)j * ld;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ns; r += gridDim.x * blockDim.x)
      dst[r] = dir ? -src[r] : src[r];
  }
}

__global__ __launch_bounds__(256) void kelvin_apply_part(int ns, int nb, size_t ld,
                                                         const int* __restrict__ bc,
                                                         const double* __restrict__ H,
                                                         const double* __restrict__ G,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ part) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * kApplyCols;
  const int j1 = min(ns, j0 + kApplyCols);
  if (r >= ns) return;
  double s = 0.0;
  for (int j = j0; j < j1; ++j) {
    const size_t at = (size_t)j * ld + r;
    s += ((j < nb) && bc[j] == 0) ? -G[at] * x[j] : H[at] * x[j];
  }
  part[(size_t)blockIdx.y * ns + r] = s;
}

__global__ __launch_bounds__(256) void kelvin_apply_sum(int ns, int nchunk,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ns) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * ns + r];
  y[r] = s;
}

// Diagonal of A per dof r = 2i + a: H's own-block entry D_aa for traction-known
// and internal dofs, -G_rr (node i's own entry of its on-element pair) for
// displacement-known (BC 0) dofs; 1 where that is zero or not finite.
__global__ void kelvin_jacobi_kernel(int nt, int n, int nn, const int* __restrict__ bc,
                                     const double* __restrict__ diag,
                                     const int* __restrict__ row_ptr,
                                     const int* __restrict__ pair_elem,
                                     const double* __restrict__ pg, double* __restrict__ d) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 2 * nt) return;
  const int i = r >> 1, a = r & 1;
  double v = diag[4 * (size_t)i + 3 * a];
  if (i < n && bc[r] == 0) {
    const int own = i / nn;
    v = 0.0;
    for (int pr = row_ptr[i]; pr < row_ptr[i + 1]; ++pr)
      if (pair_elem[pr] == own) v = -pg[((size_t)pr * nn + (i - own * nn)) * 4 + 3 * a];
  }
  d[r] = (v != 0.0 && isfinite(v)) ? v : 1.0;
}

template <int NN>
int launch_kelvin_pairs(bemhip_ctx* c, hipStream_t stream, const Geo& g, const Kelvin& K,
                        int64_t nsing, const int* sing, int64_t nreg, const int* reg) {
  const QuadConst* qc = c->qc.ptr;
  if (nsing > 0) {
    const unsigned blocks = (unsigned)((nsing + 7) / 8);  // 2 pairs per wave, 4 waves
    hipLaunchKernelGGL((kelvin_near_sing_kernel<NN>), dim3(blocks), dim3(256), 0, stream, g, K,
                       qc, nsing, sing, c->pair_row.ptr, c->pair_elem.ptr, c->pair_h.ptr,
                       c->pair_g.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  if (nreg > 0) {
    // G >= npg lanes per pair (a split rule's 2 npg nodes take two rounds);
    // G >= 4 NN output lanes
    const int need = c->npg;
    if (need <= 16) {
      const unsigned blocks = (unsigned)((nreg + 15) / 16);
      hipLaunchKernelGGL((kelvin_near_reg_kernel<NN, 16>), dim3(blocks), dim3(256), 0, stream, g,
                         K, qc, nreg, reg, c->pair_row.ptr, c->pair_elem.ptr, c->pair_h.ptr,
                         c->pair_g.ptr);
    } else if (need <= 32) {
      const unsigned blocks = (unsigned)((nreg + 7) / 8);
      hipLaunchKernelGGL((kelvin_near_reg_kernel<NN, 32>), dim3(blocks), dim3(256), 0, stream, g,
                         K, qc, nreg, reg, c->pair_row.ptr, c->pair_elem.ptr, c->pair_h.ptr,
                         c->pair_g.ptr);
    } else {
      const unsigned blocks = (unsigned)((nreg + 3) / 4);
      hipLaunchKernelGGL((kelvin_near_reg_kernel<NN, 64>), dim3(blocks), dim3(256), 0, stream, g,
                         K, qc, nreg, reg, c->pair_row.ptr, c->pair_elem.ptr, c->pair_h.ptr,
                         c->pair_g.ptr);
    }
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  return BEMHIP_OK;
}

Consts kelvin_consts(const bemhip_ctx* c) {
  Consts cs;
  cs.cU = 0.0;
  cs.cT = 0.0;
  cs.nf = c->near_factor;
  return cs;
}

}  // namespace

// Kelvin assembly (FP64 storage), one stream: classification -> pair count ->
// pair CSR -> near rules -> far field -> kelvin_finalize.
int kelvin_assemble_impl(bemhip_ctx* c) {
  const Geo g = make_geo(c);
  const Consts cs = kelvin_consts(c);
  const Kelvin K = make_kelvin(c);
  const size_t nt = (size_t)c->nt;
  const size_t ns = (size_t)c->nsys(), nb = (size_t)c->nbnd();
  const int ncb = (c->ne + kEB - 1) / kEB;
  const int nrb = (int)((nt + kFarThreads - 1) / kFarThreads);
  const bool fp64 = (c->precision == BEMHIP_FP64);
  const bool mfree = (c->precision == BEMHIP_MATRIX_FREE);
  // a matrix-free assembly stores no matrix: every stored one goes; the stored
  // modes keep the other storage's matrices until memory runs short
  // (ensure_matrix)
  if (mfree) {
    c->H64.release(); c->G64.release(); c->A64.release();
    c->H32.release(); c->G32.release(); c->A32.release();
    // near bits (4 B), the four FP64 far row-sum components (32 B) and the
    // products' two partial sums (16 B) per (row node, 32-element column block)
    const int st0 = check_matrix_free_scratch(c, "assemble (elasticity)", 4 + 32 + 16, (size_t)ncb,
                                              nt);
    if (st0) return st0;
  }
  BEMHIP_HIP_TRY(c, c->nearbits.ensure((size_t)ncb * nt));
  BEMHIP_HIP_TRY(c, c->diag64.ensure(4 * nt));
  BEMHIP_HIP_TRY(c, c->hpart.ensure(4 * (size_t)ncb * nt));
  BEMHIP_HIP_TRY(c, c->row_cnt.ensure(nt));
  BEMHIP_HIP_TRY(c, c->row_ptr.ensure(nt + 1));
  BEMHIP_HIP_TRY(c, c->blkbox.ensure((size_t)ncb * 5));
  BEMHIP_HIP_TRY(c, c->rowbox.ensure((size_t)nrb * 4));
  BEMHIP_HIP_TRY(c, c->tiles.ensure((size_t)nrb * ncb));
  if (mfree) {
    // no stored H, G: the far pass below computes the row sums only
  } else if (fp64) {
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->H64, c->ld64 * ns));
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->G64, c->ld64 * nb, &c->H64));
    if (c->ni > 0)  // columns of the internal nodes: zero but the diagonal blocks
      BEMHIP_HIP_TRY(c, hipMemsetAsync(c->H64.ptr + nb * c->ld64, 0,
                                       sizeof(double) * c->ld64 * (ns - nb), c->stream));
  } else {
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->H32, c->ld32 * ns));
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->G32, c->ld32 * nb, &c->H32));
    if (c->ni > 0)
      BEMHIP_HIP_TRY(c, hipMemsetAsync(c->H32.ptr + nb * c->ld32, 0,
                                       sizeof(float) * c->ld32 * (ns - nb), c->stream));
  }
  // events: 0 start, 1 classified, 2 far done, 3 far start, 4 pairs counted,
  // 5 near done, 6 finalized
  hipEvent_t ev[7];
  for (auto& e : ev) BEMHIP_HIP_TRY(c, hipEventCreate(&e));
  struct Destroy {
    hipEvent_t* e;
    ~Destroy() { for (int k = 0; k < 7; ++k) hipEventDestroy(e[k]); }
  } destroy{ev};
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[0], c->stream));
  {
    const int cpb = 32;  // column blocks per classify workgroup
    dim3 grid((unsigned)nrb, (unsigned)((ncb + cpb - 1) / cpb));
    // column- and row-block boxes (the all-far tile test of near_classify and
    // kelvin_far_kernel); row_cnt cleared on the way
    const unsigned nbx = (unsigned)(ncb + nrb);
    if (c->nn == 2) {
      hipLaunchKernelGGL((block_boxes<2>), dim3(nbx), dim3(64), 0, c->stream, g, cs, ncb,
                         c->blkbox.ptr, c->rowbox.ptr, c->row_cnt.ptr);
      hipLaunchKernelGGL((near_classify<2>), grid, dim3(kFarThreads), 0, c->stream, g, cs, ncb,
                         cpb, c->blkbox.ptr, c->nearbits.ptr, c->row_cnt.ptr, c->tiles.ptr);
    } else {
      hipLaunchKernelGGL((block_boxes<3>), dim3(nbx), dim3(64), 0, c->stream, g, cs, ncb,
                         c->blkbox.ptr, c->rowbox.ptr, c->row_cnt.ptr);
      hipLaunchKernelGGL((near_classify<3>), grid, dim3(kFarThreads), 0, c->stream, g, cs, ncb,
                         cpb, c->blkbox.ptr, c->nearbits.ptr, c->row_cnt.ptr, c->tiles.ptr);
    }
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[1], c->stream));
  // the pair count (one workgroup) before the far field fills the GPU
  BEMHIP_HIP_TRY(c, ensure_count_word(c));
  hipLaunchKernelGGL(near_scan, dim3(1), dim3(kScanWaves * kWave), 0, c->stream, (int)nt,
                     c->row_cnt.ptr, c->row_ptr.ptr, c->d_count);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[4], c->stream));
  // pair CSR and the near rules
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  const int total = *(volatile int*)c->h_count;
  c->npairs = total;
  const size_t nv = 4 * (size_t)c->nn;
  BEMHIP_HIP_TRY(c, c->pair_row.ensure((size_t)std::max(total, 1)));
  BEMHIP_HIP_TRY(c, c->pair_elem.ensure((size_t)std::max(total, 1)));
  BEMHIP_HIP_TRY(c, c->pair_h.ensure((size_t)std::max(total, 1) * nv));
  BEMHIP_HIP_TRY(c, c->pair_g.ensure((size_t)std::max(total, 1) * nv));
  const int64_t nsing = std::min<int64_t>(nt, c->n);  // one on-element pair per boundary row
  const int64_t nreg = (int64_t)total - nsing;
  BEMHIP_HIP_TRY(c, c->sing_idx.ensure((size_t)std::max<int64_t>(nsing, 1)));
  BEMHIP_HIP_TRY(c, c->reg_idx.ensure((size_t)std::max<int64_t>(nreg, 1)));
  hipLaunchKernelGGL(near_fill, dim3(nrb), dim3(kFarThreads), 0, c->stream, (int)nt, ncb, c->nn,
                     0, c->n, c->nearbits.ptr, c->tiles.ptr, c->row_ptr.ptr, c->pair_row.ptr,
                     c->pair_elem.ptr, c->sing_idx.ptr, c->reg_idx.ptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  int st = (c->nn == 2) ? launch_kelvin_pairs<2>(c, c->stream, g, K, nsing, c->sing_idx.ptr,
                                                 nreg, c->reg_idx.ptr)
                        : launch_kelvin_pairs<3>(c, c->stream, g, K, nsing, c->sing_idx.ptr,
                                                 nreg, c->reg_idx.ptr);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[5], c->stream));
  // the far field after the pair kernels: run beside it, the occupancy-2 FP64
  // pair kernels wait for the far kernel's workgroups (measured: 0.46 ms alone,
  // 3.8 ms beside it, against 3.5 ms for the far kernel)
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[3], c->stream));
  {
    dim3 grid((unsigned)nrb, (unsigned)ncb);
    if (mfree && c->nn == 2)
      hipLaunchKernelGGL((kelvin_far_kernel<double, 2, false>), grid, dim3(kFarThreads), 0,
                         c->stream, g, cs, K, c->blkbox.ptr, c->rowbox.ptr,
                         nullptr, nullptr, c->hpart.ptr);
    else if (mfree)
      hipLaunchKernelGGL((kelvin_far_kernel<double, 3, false>), grid, dim3(kFarThreads), 0,
                         c->stream, g, cs, K, c->blkbox.ptr, c->rowbox.ptr,
                         nullptr, nullptr, c->hpart.ptr);
    else if (fp64 && c->nn == 2)
      hipLaunchKernelGGL((kelvin_far_kernel<double, 2>), grid, dim3(kFarThreads), 0, c->stream,
                         g, cs, K, c->blkbox.ptr, c->rowbox.ptr,
                         c->H64.ptr, c->G64.ptr, c->hpart.ptr);
    else if (fp64)
      hipLaunchKernelGGL((kelvin_far_kernel<double, 3>), grid, dim3(kFarThreads), 0, c->stream,
                         g, cs, K, c->blkbox.ptr, c->rowbox.ptr,
                         c->H64.ptr, c->G64.ptr, c->hpart.ptr);
    else if (c->nn == 2)
      hipLaunchKernelGGL((kelvin_far_kernel<float, 2>), grid, dim3(kFarThreads), 0, c->stream,
                         g, cs, K, c->blkbox.ptr, c->rowbox.ptr,
                         c->H32.ptr, c->G32.ptr, c->hpart.ptr);
    else
      hipLaunchKernelGGL((kelvin_far_kernel<float, 3>), grid, dim3(kFarThreads), 0, c->stream,
                         g, cs, K, c->blkbox.ptr, c->rowbox.ptr,
                         c->H32.ptr, c->G32.ptr, c->hpart.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[2], c->stream));
  const dim3 fgrid((unsigned)((nt + kFinRows - 1) / kFinRows));
  if (mfree)  // the FP64 diagonal blocks only
    hipLaunchKernelGGL((kelvin_finalize<double>), fgrid, dim3(256), 0, c->stream, (int)nt, c->nn,
                       ncb, (size_t)0, c->hpart.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                       c->pair_h.ptr, c->pair_g.ptr, c->diag64.ptr, (double*)nullptr,
                       (double*)nullptr);
  else if (fp64)
    hipLaunchKernelGGL((kelvin_finalize<double>), fgrid, dim3(256), 0, c->stream, (int)nt, c->nn,
                       ncb, c->ld64, c->hpart.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                       c->pair_h.ptr, c->pair_g.ptr, c->diag64.ptr, c->H64.ptr, c->G64.ptr);
  else
    hipLaunchKernelGGL((kelvin_finalize<float>), fgrid, dim3(256), 0, c->stream, (int)nt, c->nn,
                       ncb, c->ld32, c->hpart.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                       c->pair_h.ptr, c->pair_g.ptr, c->diag64.ptr, c->H32.ptr, c->G32.ptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[6], c->stream));
  BEMHIP_HIP_TRY(c, hipEventSynchronize(ev[6]));
  auto el = [&](int a, int b) {
    float t = 0.f;
    hipEventElapsedTime(&t, ev[a], ev[b]);
    return (double)t;
  };
  c->ms["assemble_classify"] = el(0, 1);
  c->ms["assemble_far"] = el(3, 2);
  c->ms["assemble_scan"] = el(1, 4);
  c->ms["assemble_near"] = el(4, 5);
  c->ms["assemble_finalize"] = el(2, 6);
  c->stats["near_pairs"] = (double)total;
  return BEMHIP_OK;
}

// One FP64 operator pass y = H wH + G wG (wH over 2nt dofs, wG over 2n) from
// the geometry, the pair list and the FP64 diagonal blocks.
static int kelvin_operator_pass(bemhip_ctx* c, const double* wH, const double* wG, double* y) {
  const Geo g = make_geo(c);
  const Consts cs = kelvin_consts(c);
  const Kelvin K = make_kelvin(c);
  const int nt = c->nt;
  const int ncb = (c->ne + kEB - 1) / kEB;
  // part1: bemhip_solve keeps its residual and product vectors in part2
  BEMHIP_HIP_TRY(c, c->part1.ensure(2 * (size_t)ncb * nt));
  const dim3 grid((unsigned)((nt + kFarThreads - 1) / kFarThreads), (unsigned)ncb);
  if (c->nn == 2)
    hipLaunchKernelGGL((kelvin_op64_kernel<2>), grid, dim3(kFarThreads), 0, c->stream, g, cs, K,
                       wH, wG, c->part1.ptr);
  else
    hipLaunchKernelGGL((kelvin_op64_kernel<3>), grid, dim3(kFarThreads), 0, c->stream, g, cs, K,
                       wH, wG, c->part1.ptr);
  hipLaunchKernelGGL(kelvin_op_finalize, dim3((unsigned)((nt + kFinRows - 1) / kFinRows)),
                     dim3(256), 0, c->stream, nt, c->n, c->nn, ncb, c->part1.ptr, c->row_ptr.ptr,
                     c->pair_elem.ptr, c->pair_h.ptr, c->pair_g.ptr, c->diag64.ptr, wH, wG, y);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

int kelvin_apply_bc_impl(bemhip_ctx* c) {
  const int ns = c->nsys(), nb = c->nbnd();
  const bool fp64 = (c->precision == BEMHIP_FP64);
  BEMHIP_HIP_TRY(c, c->b64.ensure((size_t)ns));
  if (fp64) {
    ScopedTimer t(c, "apply_bc_rhs");
    const int nchunk = (nb + kApplyCols - 1) / kApplyCols;
    BEMHIP_HIP_TRY(c, c->part1.ensure((size_t)nchunk * ns));
    const unsigned rb = (unsigned)((ns + 255) / 256);
    hipLaunchKernelGGL(kelvin_rhs_part, dim3(rb, (unsigned)nchunk), dim3(256), 0, c->stream, ns,
                       nb, c->ld64, c->d_bc.ptr, c->d_bv.ptr, c->H64.ptr, c->G64.ptr,
                       c->part1.ptr);
    hipLaunchKernelGGL(kelvin_apply_sum, dim3(rb), dim3(256), 0, c->stream, ns, nchunk,
                       c->part1.ptr, c->b64.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  } else {
    // FP64 right-hand side from the operator pass (no FP64 H, G stored)
    ScopedTimer t(c, "apply_bc_rhs");
    BEMHIP_HIP_TRY(c, c->wH2.ensure((size_t)ns));
    BEMHIP_HIP_TRY(c, c->wG2.ensure((size_t)nb));
    hipLaunchKernelGGL(kelvin_weights_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                       c->stream, ns, nb, 1, c->d_bc.ptr, c->d_bv.ptr, c->wH2.ptr, c->wG2.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    const int st = kelvin_operator_pass(c, c->wH2.ptr, c->wG2.ptr, c->b64.ptr);
    if (st) return st;
  }
  if (c->precision == BEMHIP_MATRIX_FREE) return BEMHIP_OK;  // A is applied, never stored
  {
    ScopedTimer t(c, "apply_bc_swap");
    dim3 grid((unsigned)std::min((ns + 255) / 256, 64), (unsigned)std::min(ns, 32768));
    if (fp64) {
      BEMHIP_HIP_TRY(c, ensure_matrix(c, c->A64, c->ld64 * (size_t)ns, &c->H64, &c->G64));
      hipLaunchKernelGGL((kelvin_swap_kernel<double>), grid, dim3(256), 0, c->stream, ns, nb,
                         c->ld64, c->d_bc.ptr, c->H64.ptr, c->G64.ptr, c->A64.ptr);
    } else {
      BEMHIP_HIP_TRY(c, ensure_matrix(c, c->A32, c->ld32 * (size_t)ns, &c->H32, &c->G32));
      hipLaunchKernelGGL((kelvin_swap_kernel<float>), grid, dim3(256), 0, c->stream, ns, nb,
                         c->ld32, c->d_bc.ptr, c->H32.ptr, c->G32.ptr, c->A32.ptr);
    }
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  return BEMHIP_OK;
}

// Jacobi diagonal of the current system (2nt values) for GMRES.
int kelvin_jacobi_diag(bemhip_ctx* c, double* d) {
  const int ns = c->nsys();
  hipLaunchKernelGGL(kelvin_jacobi_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                     c->stream, c->nt, c->n, c->nn, c->d_bc.ptr, c->diag64.ptr, c->row_ptr.ptr,
                     c->pair_elem.ptr, c->pair_g.ptr, d);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

// y = A x in FP64: from the stored FP64 H, G (FP64 assembly) or the operator
// pass (mixed and matrix-free assemblies).
int kelvin_operator_apply(bemhip_ctx* c, const double* d_x, double* d_y) {
  const int ns = c->nsys(), nb = c->nbnd();
  if (c->precision != BEMHIP_FP64) {
    BEMHIP_HIP_TRY(c, c->wH.ensure((size_t)ns));
    BEMHIP_HIP_TRY(c, c->wG.ensure((size_t)nb));
    hipLaunchKernelGGL(kelvin_weights_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                       c->stream, ns, nb, 0, c->d_bc.ptr, d_x, c->wH.ptr, c->wG.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    return kelvin_operator_pass(c, c->wH.ptr, c->wG.ptr, d_y);
  }
  const int nchunk = (ns + kApplyCols - 1) / kApplyCols;
  BEMHIP_HIP_TRY(c, c->part1.ensure((size_t)nchunk * ns));
  const unsigned rb = (unsigned)((ns + 255) / 256);
  hipLaunchKernelGGL(kelvin_apply_part, dim3(rb, (unsigned)nchunk), dim3(256), 0, c->stream, ns,
                     nb, c->ld64, c->d_bc.ptr, c->H64.ptr, c->G64.ptr, d_x, c->part1.ptr);
  hipLaunchKernelGGL(kelvin_apply_sum, dim3(rb), dim3(256), 0, c->stream, ns, nchunk,
                     c->part1.ptr, d_y);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

}  // namespace bemhip
