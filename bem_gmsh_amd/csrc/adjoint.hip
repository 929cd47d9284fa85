// Adjoint shape gradient of a boundary-flux misfit (no reference counterpart;
// it differentiates the path assemble -> applyBC -> solve of
// src/Core/Assembly_full.jl:61-85, src/Core/Boundary_conditions.jl:20-75,
// src/Core/Solver.jl:24-64 with respect to the elements' geometric nodes).
//
//   J(X) = 1/2 sum_{j<n} w_j (q_j - qhat_j)^2,   w_j = J_j w_k (or given)
//
// The solved system is r(x, X) = H(X) T - G(X) q = 0 (A x = b after the
// column swap), with the unknowns x inside T and q. With A^T lam = dJ/dx,
//
//   dJ/dX = dJ/dX|explicit - d/dX [ sum_i lam_i r_i ]   (lam, T, q frozen).
//
// The rigid-body diagonal H_ii = -sum_{j!=i} H_ij turns sum_j H_ij T_j into
// sum_{j!=i} H_ij (T_j - T_i), so each off-diagonal entry carries the
// coefficient a_ij = -lam_i (T_j - T_i) and each G entry c_ij = lam_i q_j:
//
//   psi = sum_{i, j!=i} a_ij H_ij + sum_{i,j} c_ij G_ij,   dJ/dX = explicit + dpsi/dX.
//
// Kernels:
//   adj_pair_kernel   near / on-element pairs: the tan-p3c or interpolated-
//                     Guiggiani rule evaluated in forward-mode dual numbers
//                     (source point + the element's geometric nodes), so the
//                     derivative is that of the rule the assembly used,
//                     pole fit and map included.
//   col_pass_kernel   far pairs, closed-form derivatives of the lumped
//                     entries, one thread per element over a tile of rows;
//                     the source-side derivative is reduced across the wave
//                     per row. The same pass with MODE_TRANS computes the FP64
//                     transposed operator A^T v for the refined transposed solve.
//   reductions + adj_chain_kernel   node sensitivities -> geometric nodes
//                     through the collocation map x_k = sum_m Ngeo_m(xi_k) X_m.

#include "common.hpp"
#include "near_rule.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace bemhip {

namespace {

constexpr int kColTile = 256;  // most rows per tile of the column pass (LDS arrays)
constexpr int kColEl = 64;     // elements per workgroup of the column pass (one wave)
constexpr int MODE_GRAD = 0;
constexpr int MODE_TRANS = 1;
constexpr double kTwoPi = 2.0 * 3.141592653589793;

Geo make_geo_adj(const bemhip_ctx* c) {
  Geo g;
  g.p = c->p; g.nn = c->nn; g.ne = c->ne; g.n = c->n; g.ni = c->ni; g.nt = c->nt;
  g.x = c->x.ptr; g.y = c->y.ptr; g.nx = c->nx.ptr; g.ny = c->ny.ptr; g.jw = c->jw.ptr;
  g.elen = c->elen.ptr; g.gx = c->gx.ptr; g.gy = c->gy.ptr;
  g.rx = c->x.ptr; g.ry = c->y.ptr; g.row0 = 0;
  g.ld = (c->precision == BEMHIP_FP64) ? c->ld64 : c->ld32;
  return g;
}

// Equispaced geometric basis and its derivative at the collocation parent
// coordinates: Ng[k][m] = Ngeo_m(xi_k), dNg[k][m] = dNgeo_m(xi_k).
template <int NN>
__device__ void geo_tables(const QuadConst& qc, double (*Ng)[NN], double (*dNg)[NN]) {
  for (int k = 0; k < NN; ++k)
    rule::shapefun<double, NN>(qc.eq_x, qc.eq_bw, qc.eq_D, qc.lp_x[k], Ng[k], dNg[k]);
}

// One thread per near pair: d psi_pair / d(source x, y, geometric nodes of e).
template <int NN>
__global__ __launch_bounds__(64) void adj_pair_kernel(Geo g, const QuadConst* __restrict__ qcp,
                                                      int64_t npairs,
                                                      const int* __restrict__ prow,
                                                      const int* __restrict__ pelem,
                                                      const double* __restrict__ Tf,
                                                      const double* __restrict__ qf,
                                                      const double* __restrict__ lam,
                                                      double* __restrict__ grad) {
  constexpr int D = 2 + 2 * NN;
  using Dl = ad::Dual<D>;
  const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pair >= npairs) return;
  const QuadConst& qc = *qcp;
  const int i = prow[pair], e = pelem[pair];
  double Ng[NN][NN], dNg[NN][NN];
  geo_tables<NN>(qc, Ng, dNg);
  Dl gx[NN], gy[NN], cx[NN], cy[NN];
  for (int m = 0; m < NN; ++m) {
    gx[m] = Dl::seed(g.gx[e * NN + m], 2 + 2 * m);
    gy[m] = Dl::seed(g.gy[e * NN + m], 3 + 2 * m);
  }
  for (int k = 0; k < NN; ++k) {  // stored collocation values, derivatives of the map
    Dl sx = 0.0, sy = 0.0;
    for (int m = 0; m < NN; ++m) { sx += Ng[k][m] * gx[m]; sy += Ng[k][m] * gy[m]; }
    sx.v = g.x[e * NN + k];
    sy.v = g.y[e * NN + k];
    cx[k] = sx;
    cy[k] = sy;
  }
  const double nrx = g.nx[e * NN], nry = g.ny[e * NN];
  Dl hs[NN], gs[NN];
  const bool singular = (i >= e * NN) && (i < (e + 1) * NN);
  if (singular) {
    double cxd[NN], cyd[NN];
    for (int k = 0; k < NN; ++k) { cxd[k] = g.x[e * NN + k]; cyd[k] = g.y[e * NN + k]; }
    const double a = rule::closest_param<NN>(qc, cxd, cyd, g.x[i], g.y[i]);
    rule::singular_pair<Dl, NN>(qc, a, gx, gy, cx[i - e * NN], cy[i - e * NN], nrx, nry, hs, gs);
  } else {
    rule::near_pair<Dl, NN>(qc, cx, cy, gx, gy, Dl::seed(g.x[i], 0), Dl::seed(g.y[i], 1), nrx,
                            nry, hs, gs);
  }
  const double li = lam[i], Ti = Tf[i];
  Dl psi = 0.0;
  for (int k = 0; k < NN; ++k) {
    const int j = e * NN + k;
    const double ak = (j == i) ? 0.0 : -li * (Tf[j] - Ti);
    const double ck = li * qf[j];
    psi += ak * hs[k] + ck * gs[k];
  }
  for (int t = 0; t < D; ++t) grad[(size_t)pair * D + t] = psi.d[t];
}

// Column pass over far pairs, one thread per element e, rows [r0, r0 + tile).
// MODE_GRAD: accumulates per element node j the derivatives of psi with
//   respect to x_j (2), the scaled normal n_j J_j (2) and J_j (1), plus the
//   near pairs' element-side gradients (2 NN); the source-side derivative of
//   each row is reduced across the wave into rowpart[group][2][nt].
// MODE_TRANS: accumulates sum_i v_i H_ij and sum_i v_i G_ij for the element's
//   nodes (far entries recomputed as the forward FP64 operator does, near
//   entries from the stored FP64 pair values, the diagonal from diag64).
template <int NN, int MODE>
__global__ __launch_bounds__(kColEl) void col_pass_kernel(
    Geo g, const QuadConst* __restrict__ qcp, const uint32_t* __restrict__ nearbits,
    const int* __restrict__ row_ptr, const int* __restrict__ pelem,
    const double* __restrict__ pgrad, const double* __restrict__ pair_h,
    const double* __restrict__ pair_g, const double* __restrict__ diag,
    const double* __restrict__ Jn, const double* __restrict__ Tf, const double* __restrict__ qf,
    const double* __restrict__ lam, double* __restrict__ colpart, double* __restrict__ rowpart,
    int tile) {
  constexpr int D = 2 + 2 * NN;
  constexpr int NC = (MODE == MODE_GRAD) ? 7 * NN : 2 * NN;
  __shared__ double s_x[kColTile], s_y[kColTile], s_l[kColTile], s_T[kColTile];
  __shared__ uint32_t s_bits[2][kColTile];
  __shared__ double s_row[2][kColTile];
  const QuadConst& qc = *qcp;
  const int lane = threadIdx.x;
  const int e = blockIdx.x * kColEl + lane;
  const int r0 = blockIdx.y * tile;
  const int nt = g.nt, ne = g.ne;
  const int ncb = (ne + 31) / 32;
  const int cb0 = 2 * blockIdx.x;
  for (int t = lane; t < tile; t += kColEl) {
    const int i = r0 + t;
    const bool ok = i < nt;
    s_x[t] = ok ? g.x[i] : 0.0;
    s_y[t] = ok ? g.y[i] : 0.0;
    s_l[t] = ok ? lam[i] : 0.0;
    s_T[t] = (ok && MODE == MODE_GRAD) ? Tf[i] : 0.0;
    s_bits[0][t] = (ok && cb0 < ncb) ? nearbits[(size_t)cb0 * nt + i] : 0u;
    s_bits[1][t] = (ok && cb0 + 1 < ncb) ? nearbits[(size_t)(cb0 + 1) * nt + i] : 0u;
    s_row[0][t] = 0.0;
    s_row[1][t] = 0.0;
  }
  __syncthreads();
  const bool ev = e < ne;
  const double cU = 0.5 * (1.0 / kTwoPi) / qc.k;
  const double cT = 1.0 / kTwoPi;
  double xj[NN], yj[NN], nxj[NN], nyj[NN], Jj[NN], jwj[NN], wk[NN], Tj[NN], qj[NN];
  double acc[NC];
#pragma unroll
  for (int t = 0; t < NC; ++t) acc[t] = 0.0;
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    const int j = ev ? e * NN + k : 0;
    xj[k] = g.x[j]; yj[k] = g.y[j];
    nxj[k] = g.nx[j]; nyj[k] = g.ny[j];
    Jj[k] = (MODE == MODE_GRAD) ? Jn[j] : 0.0;
    jwj[k] = g.jw[j];
    wk[k] = qc.ew[k];
    Tj[k] = (MODE == MODE_GRAD) ? Tf[j] : 0.0;
    qj[k] = (MODE == MODE_GRAD) ? qf[j] : 0.0;
  }
  const int tmax = min(tile, nt - r0);
  for (int t = 0; t < tmax; ++t) {
    const int i = r0 + t;
    const double px = s_x[t], py = s_y[t], li = s_l[t], Ti = s_T[t];
    double sgx = 0.0, sgy = 0.0;
    if (ev) {
      const bool nr = (s_bits[lane >> 5][t] >> (lane & 31)) & 1u;
      if (nr) {
        int q = row_ptr[i];
        const int q1 = row_ptr[i + 1];
        while (q < q1 && pelem[q] != e) ++q;
        if (q < q1) {
          if (MODE == MODE_GRAD) {
#pragma unroll
            for (int m = 0; m < 2 * NN; ++m) acc[5 * NN + m] += pgrad[(size_t)q * D + 2 + m];
          } else {
#pragma unroll
            for (int k = 0; k < NN; ++k) {
              const int j = e * NN + k;
              const double h = (j == i) ? diag[i] : pair_h[(size_t)q * NN + k];
              acc[k] += li * h;
              acc[NN + k] += li * pair_g[(size_t)q * NN + k];
            }
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          const double rx = xj[k] - px, ry = yj[k] - py;
          const double R2 = rx * rx + ry * ry;
          if (MODE == MODE_GRAD) {
            const double inv = 1.0 / R2;
            const double mtx = nxj[k] * Jj[k], mty = nyj[k] * Jj[k];
            const double rm = rx * mtx + ry * mty;
            const double a = -li * (Tj[k] - Ti);
            const double cc = li * qj[k];
            const double A = a * cT * wk[k];
            const double C2 = 2.0 * (cc * cU * jwj[k]) * inv;
            const double t2 = 2.0 * rm * inv * inv;
            const double grx = A * (mtx * inv - t2 * rx) - C2 * rx;
            const double gry = A * (mty * inv - t2 * ry) - C2 * ry;
            acc[2 * k] += grx;
            acc[2 * k + 1] += gry;
            acc[2 * NN + 2 * k] += A * inv * rx;
            acc[2 * NN + 2 * k + 1] += A * inv * ry;
            acc[4 * NN + k] += cc * cU * wk[k] * (-log(R2));
            sgx -= grx;
            sgy -= gry;
          } else {
            const double U = -log(R2) * cU;
            const double T = (rx * nxj[k] + ry * nyj[k]) * (cT / R2);
            acc[k] += li * (T * jwj[k]);
            acc[NN + k] += li * (U * jwj[k]);
          }
        }
      }
    }
    if (MODE == MODE_GRAD) {
      sgx = wave_sum(sgx);
      sgy = wave_sum(sgy);
      if (lane == 0) { s_row[0][t] = sgx; s_row[1][t] = sgy; }
    }
  }
  if (ev) {
    const size_t base = (size_t)blockIdx.y * NC * ne + e;
#pragma unroll
    for (int t = 0; t < NC; ++t) colpart[base + (size_t)t * ne] = acc[t];
  }
  if (MODE == MODE_GRAD) {
    __syncthreads();
    for (int t = lane; t < tmax; t += kColEl) {
      rowpart[((size_t)blockIdx.x * 2 + 0) * nt + r0 + t] = s_row[0][t];
      rowpart[((size_t)blockIdx.x * 2 + 1) * nt + r0 + t] = s_row[1][t];
    }
  }
}

// colsum[c][e] = sum over row tiles of colpart[tile][c][e] (fixed order).
__global__ void col_reduce_kernel(int64_t count, int ntiles, const double* __restrict__ colpart,
                                  double* __restrict__ colsum) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= count) return;
  double s = 0.0;
  for (int t = 0; t < ntiles; ++t) s += colpart[(size_t)t * count + idx];
  colsum[idx] = s;
}

// gsrc[c][i] = sum over element groups of rowpart + near pairs' source parts.
__global__ void row_reduce_kernel(int nt, int ngroups, int D, const double* __restrict__ rowpart,
                                  const int* __restrict__ row_ptr,
                                  const double* __restrict__ pgrad, double* __restrict__ gsrc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  double sx = 0.0, sy = 0.0;
  for (int gi = 0; gi < ngroups; ++gi) {
    sx += rowpart[((size_t)gi * 2 + 0) * nt + i];
    sy += rowpart[((size_t)gi * 2 + 1) * nt + i];
  }
  for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
    sx += pgrad[(size_t)q * D + 0];
    sy += pgrad[(size_t)q * D + 1];
  }
  gsrc[i] = sx;
  gsrc[nt + i] = sy;
}

// Node sensitivities -> dJ/dX for the element's geometric nodes, out[(e*NN+m)*2 + c].
template <int NN>
__global__ void adj_chain_kernel(Geo g, const QuadConst* __restrict__ qcp,
                                 const double* __restrict__ colsum,
                                 const double* __restrict__ gsrc,
                                 const double* __restrict__ gJexp, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int ne = g.ne, nt = g.nt;
  if (e >= ne) return;
  const QuadConst& qc = *qcp;
  double Ng[NN][NN], dNg[NN][NN];
  geo_tables<NN>(qc, Ng, dNg);
  double X[NN], Y[NN], ox[NN], oy[NN];
  for (int m = 0; m < NN; ++m) {
    X[m] = g.gx[e * NN + m];
    Y[m] = g.gy[e * NN + m];
    ox[m] = colsum[(size_t)(5 * NN + 2 * m) * ne + e];
    oy[m] = colsum[(size_t)(5 * NN + 2 * m + 1) * ne + e];
  }
  for (int k = 0; k < NN; ++k) {
    const int j = e * NN + k;
    const double Px = colsum[(size_t)(2 * k) * ne + e] + gsrc[j];
    const double Py = colsum[(size_t)(2 * k + 1) * ne + e] + gsrc[nt + j];
    const double gmx = colsum[(size_t)(2 * NN + 2 * k) * ne + e];
    const double gmy = colsum[(size_t)(2 * NN + 2 * k + 1) * ne + e];
    const double gJ = colsum[(size_t)(4 * NN + k) * ne + e] + (gJexp ? gJexp[j] : 0.0);
    double dx = 0.0, dy = 0.0;
    for (int m = 0; m < NN; ++m) { dx += dNg[k][m] * X[m]; dy += dNg[k][m] * Y[m]; }
    const double J = sqrt(dx * dx + dy * dy);
    // stored normal n = s (dy, -dx) / J
    const double s = (g.nx[j] * dy - g.ny[j] * dx) < 0.0 ? -1.0 : 1.0;
    const double gdx = -s * gmy + gJ * dx / J;
    const double gdy = s * gmx + gJ * dy / J;
    for (int m = 0; m < NN; ++m) {
      ox[m] += Ng[k][m] * Px + dNg[k][m] * gdx;
      oy[m] += Ng[k][m] * Py + dNg[k][m] * gdy;
    }
  }
  for (int m = 0; m < NN; ++m) {
    out[((size_t)e * NN + m) * 2] = ox[m];
    out[((size_t)e * NN + m) * 2 + 1] = oy[m];
  }
}

// y_j = A^T v: Neumann and internal columns sum_i v_i H_ij, Dirichlet -sum_i v_i G_ij.
__global__ void trans_finalize_kernel(int nt, int n, int nn, int ne, const int* __restrict__ bc,
                                      const double* __restrict__ colsum,
                                      const double* __restrict__ diag,
                                      const double* __restrict__ v, double* __restrict__ y) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  if (j >= n) { y[j] = diag[j] * v[j]; return; }
  const int e = j / nn, k = j - e * nn;
  y[j] = (bc[j] == 0) ? -colsum[(size_t)(nn + k) * ne + e] : colsum[(size_t)k * ne + e];
}

__global__ void resid_kernel(int n, const double* __restrict__ b, const double* __restrict__ y,
                             double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

__global__ void add_kernel(int n, const double* __restrict__ d, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}

struct ColWork {
  double* colpart;
  double* rowpart;
  double* colsum;
  int ntiles, ngroups;
  int tile;  // rows per workgroup
};

template <int NN, int MODE>
void launch_col_pass(bemhip_ctx* c, const double* pgrad, const double* Tf, const double* qf,
                     const double* lam, const ColWork& w) {
  const Geo g = make_geo_adj(c);
  dim3 grid((unsigned)w.ngroups, (unsigned)w.ntiles);
  hipLaunchKernelGGL((col_pass_kernel<NN, MODE>), grid, dim3(kColEl), 0, c->stream, g,
                     c->qc.ptr, c->nearbits.ptr, c->row_ptr.ptr, c->pair_elem.ptr, pgrad,
                     c->pair_h.ptr, c->pair_g.ptr, c->diag64.ptr, c->Jn.ptr, Tf, qf, lam,
                     w.colpart, w.rowpart, w.tile);
  const int NC = (MODE == MODE_GRAD) ? 7 * NN : 2 * NN;
  const int64_t count = (int64_t)NC * c->ne;
  hipLaunchKernelGGL(col_reduce_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                     c->stream, count, w.ntiles, w.colpart, w.colsum);
}

// Device scratch for the column passes, carved from ctx->adj_work.
// Rows per workgroup: 256, halved (down to 32) while the grid has fewer than
// 16 workgroups (one wave each) per CU, so that small problems still fill the
// chip and hide the FP64 log / division latency.
int col_work(bemhip_ctx* c, int nc, double* base, ColWork& w, size_t& used) {
  w.ngroups = (c->ne + kColEl - 1) / kColEl;
  w.tile = kColTile;
  const int64_t want = 16 * (int64_t)std::max(c->num_cu, 1);
  while (w.tile > 32 && (int64_t)w.ngroups * ((c->nt + w.tile - 1) / w.tile) < want) w.tile /= 2;
  w.ntiles = (c->nt + w.tile - 1) / w.tile;
  w.colpart = base;
  w.colsum = w.colpart + (size_t)w.ntiles * nc * c->ne;
  w.rowpart = w.colsum + (size_t)nc * c->ne;
  used = (size_t)w.ntiles * nc * c->ne + (size_t)nc * c->ne + (size_t)w.ngroups * 2 * c->nt;
  return BEMHIP_OK;
}

}  // namespace

// y = A^T v in FP64 (same entries as operator_apply, transposed).
int operator_apply_trans(bemhip_ctx* c, const double* d_v, double* d_y) {
  ColWork w;
  size_t used = 0;
  col_work(c, 2 * c->nn, nullptr, w, used);
  BEMHIP_HIP_TRY(c, c->part1.ensure(used));
  col_work(c, 2 * c->nn, c->part1.ptr, w, used);
  if (c->nn == 2)
    launch_col_pass<2, MODE_TRANS>(c, nullptr, nullptr, nullptr, d_v, w);
  else
    launch_col_pass<3, MODE_TRANS>(c, nullptr, nullptr, nullptr, d_v, w);
  const int tb = 256;
  hipLaunchKernelGGL(trans_finalize_kernel, dim3((c->nt + tb - 1) / tb), dim3(tb), 0, c->stream,
                     c->nt, c->n, c->nn, c->ne, c->d_bc.ptr, w.colsum, c->diag64.ptr, d_v, d_y);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

// A^T lam = g (g overwritten by lam): FP64 factors directly; FP32 factors with
// FP64 iterative refinement on the transposed FP64 operator, switching to
// GMRES-IR (the FP32 factors, transposed, as the preconditioner) when the
// corrections stop halving above 1e-10 ||lam||, as the forward solve does.
static int solve_transposed(bemhip_ctx* c, double* d_g, double tol, int maxit, int* iters) {
  const int nt = c->nt;
  if (c->sol_mode == BEMHIP_SOLVE_FP64) {
    if (iters) *iters = 0;
    return lu_solve_trans_device(c, nt, 0, c->A64.ptr, (int)c->ld64, c->perm.ptr, d_g);
  }
  BEMHIP_HIP_TRY(c, c->vec64b.ensure(3 * (size_t)nt));
  double* rhs = c->vec64b.ptr;
  double* y = rhs + nt;
  double* r = y + nt;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(rhs, d_g, sizeof(double) * nt, hipMemcpyDeviceToDevice,
                                   c->stream));
  int st = lu_solve_trans_device(c, nt, 1, c->A32.ptr, (int)c->ld32, c->perm32.ptr, d_g);
  if (st) return st;
  const int tb = 256, nb = (nt + tb - 1) / tb;
  int it = 0;
  bool converged = false;
  double dn_prev = INFINITY;
  for (it = 0; it < maxit; ++it) {
    if ((st = operator_apply_trans(c, d_g, y))) return st;
    hipLaunchKernelGGL(resid_kernel, dim3(nb), dim3(tb), 0, c->stream, nt, rhs, y, r);
    if ((st = lu_solve_trans_device(c, nt, 1, c->A32.ptr, (int)c->ld32, c->perm32.ptr, r)))
      return st;
    double dn = 0.0, xn = 0.0;
    if ((st = absmax2(c, nt, r, d_g, &dn, &xn))) return st;
    if (!std::isfinite(dn)) break;
    hipLaunchKernelGGL(add_kernel, dim3(nb), dim3(tb), 0, c->stream, nt, r, d_g);
    if (dn <= tol * xn) { converged = true; ++it; break; }
    if (it >= 1 && dn > 0.5 * dn_prev && dn <= 1e-10 * xn) { converged = true; ++it; break; }
    if (it >= 1 && dn > 0.5 * dn_prev) { ++it; break; }
    dn_prev = dn;
  }
  c->stats.erase("adjoint_gmres_ir");
  if (!converged && maxit > 0) {
    int outer = 0, inner = 0;
    if ((st = gmres_ir(c, nt, true, rhs, d_g, r, y, tol, maxit, &outer, &inner, &converged)))
      return st;
    c->stats["adjoint_gmres_ir"] = 1.0;
    it += outer;
  }
  BEMHIP_HIP_TRY(c, hipGetLastError());
  if (iters) *iters = it;
  if (!converged && maxit > 0)
    return fail(c, BEMHIP_ERR_NOCONV,
                "adjoint: transposed refinement did not converge, nor did GMRES-IR");
  return BEMHIP_OK;
}

}  // namespace bemhip

using namespace bemhip;

extern "C" int bemhip_adjoint(bemhip_ctx* c, const double* q_target, const double* w,
                              double* dJ_dgeo, double* J_out) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "adjoint: NULL context");
  if (!q_target || !dJ_dgeo) return fail(c, BEMHIP_ERR_ARG, "adjoint: NULL q_target or dJ_dgeo");
  if (!c->has_sol) return fail(c, BEMHIP_ERR_STATE, "adjoint: solve first");
  if (c->dof != 1) return fail(c, BEMHIP_ERR_STATE, "adjoint: Laplace problems only");
  if (c->sol_mode == BEMHIP_SOLVE_FP64 ? !c->factored64 : !c->factored32)
    return fail(c, BEMHIP_ERR_STATE, "adjoint: the factors of the last solve are gone");
  hipSetDevice(c->device);
  const int nt = c->nt, n = c->n, nn = c->nn, ne = c->ne;
  const int D = 2 + 2 * nn;
  // host: T, q from x, the objective, dJ/dx and the explicit dJ/dJ_j
  std::vector<double> x(nt), Jh(n);
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(x.data(), c->xsol.ptr, sizeof(double) * nt,
                                   hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(Jh.data(), c->Jn.ptr, sizeof(double) * n,
                                   hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  std::vector<double> T(nt), q(n), gx(nt, 0.0), gJ(n, 0.0);
  double Jv = 0.0;
  for (int j = 0; j < nt; ++j) {
    if (j >= n) { T[j] = x[j]; continue; }
    const bool dir = c->h_bc[j] == 0;
    T[j] = dir ? c->h_bv[j] : x[j];
    q[j] = dir ? x[j] : c->h_bv[j];
    const double wk = c->h_elem_w[j % nn];
    const double wj = w ? w[j] : Jh[j] * wk;
    const double d = q[j] - q_target[j];
    Jv += 0.5 * wj * d * d;
    if (dir) gx[j] = wj * d;
    if (!w) gJ[j] = 0.5 * wk * d * d;
  }
  // device scratch: T, q, lam, explicit dJ/dJ, pair gradients, column pass, output
  ColWork cw;
  size_t colw = 0;
  col_work(c, 7 * nn, nullptr, cw, colw);
  const size_t npg = (size_t)c->npairs * D;
  const size_t need = (size_t)nt + n + nt + n + npg + colw + 2 * (size_t)nt + (size_t)ne * nn * 2;
  BEMHIP_HIP_TRY(c, c->adj_work.ensure(need));
  double* dT = c->adj_work.ptr;
  double* dq = dT + nt;
  double* dlam = dq + n;
  double* dgJ = dlam + nt;
  double* dpg = dgJ + n;
  double* dcol = dpg + npg;
  col_work(c, 7 * nn, dcol, cw, colw);
  double* dgsrc = dcol + colw;
  double* dout = dgsrc + 2 * (size_t)nt;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(dT, T.data(), sizeof(double) * nt, hipMemcpyHostToDevice,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(dq, q.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(dlam, gx.data(), sizeof(double) * nt, hipMemcpyHostToDevice,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(dgJ, gJ.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                   c->stream));
  struct Ev {
    hipEvent_t e = nullptr;
    Ev() { hipEventCreate(&e); }
    ~Ev() { hipEventDestroy(e); }
  } e0, e1, e2;
  hipEvent_t t0 = e0.e, t1 = e1.e, t2 = e2.e;
  hipEventRecord(t0, c->stream);
  int iters = 0;
  int st = solve_transposed(c, dlam, 1e-10, 30, &iters);
  if (st) return st;
  hipEventRecord(t1, c->stream);
  const Geo g = make_geo_adj(c);
  if (c->npairs > 0) {
    const unsigned nb = (unsigned)((c->npairs + 63) / 64);
    if (nn == 2)
      hipLaunchKernelGGL(adj_pair_kernel<2>, dim3(nb), dim3(64), 0, c->stream, g, c->qc.ptr,
                         c->npairs, c->pair_row.ptr, c->pair_elem.ptr, dT, dq, dlam, dpg);
    else
      hipLaunchKernelGGL(adj_pair_kernel<3>, dim3(nb), dim3(64), 0, c->stream, g, c->qc.ptr,
                         c->npairs, c->pair_row.ptr, c->pair_elem.ptr, dT, dq, dlam, dpg);
  }
  if (nn == 2)
    launch_col_pass<2, MODE_GRAD>(c, dpg, dT, dq, dlam, cw);
  else
    launch_col_pass<3, MODE_GRAD>(c, dpg, dT, dq, dlam, cw);
  const int tb = 256;
  hipLaunchKernelGGL(row_reduce_kernel, dim3((nt + tb - 1) / tb), dim3(tb), 0, c->stream, nt,
                     cw.ngroups, D, cw.rowpart, c->row_ptr.ptr, dpg, dgsrc);
  if (nn == 2)
    hipLaunchKernelGGL(adj_chain_kernel<2>, dim3((ne + tb - 1) / tb), dim3(tb), 0, c->stream, g,
                       c->qc.ptr, cw.colsum, dgsrc, w ? nullptr : dgJ, dout);
  else
    hipLaunchKernelGGL(adj_chain_kernel<3>, dim3((ne + tb - 1) / tb), dim3(tb), 0, c->stream, g,
                       c->qc.ptr, cw.colsum, dgsrc, w ? nullptr : dgJ, dout);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  hipEventRecord(t2, c->stream);
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(dJ_dgeo, dout, sizeof(double) * ne * nn * 2,
                                   hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  float ms_solve = 0, ms_grad = 0;
  hipEventElapsedTime(&ms_solve, t0, t1);
  hipEventElapsedTime(&ms_grad, t1, t2);
  c->ms["adjoint_solve"] = ms_solve;
  c->ms["adjoint_grad"] = ms_grad;
  c->ms["adjoint"] = ms_solve + ms_grad;
  c->stats["adjoint_refine_iters"] = iters;
  if (J_out) *J_out = Jv;
  return BEMHIP_OK;
}
