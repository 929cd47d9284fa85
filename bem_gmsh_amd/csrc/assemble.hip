// Influence-matrix assembly for the dense 2-D Laplace BEM path (gfx950).
//
// Kernels (three streams: the far field, the near-pair chain and the
// on-element rule run concurrently):
//   block_boxes     node bounding boxes of the 32-element column blocks and of
//                   the 256-row blocks (the all-far tile test)
//   near_classify   near/far classification of every (row, element) pair
//                   (src/Core/Assembly_full.jl:736-754) into a bit per pair and
//                   per-row pair counts; column blocks far from the rows'
//                   bounding box are settled without pair tests.
//   far_kernel      rows x 32-element column blocks: the same classification
//                   (skipped on all-far tiles) and the far nodal lumping
//                   (:801-810) written column-major with lanes over rows
//                   (coalesced), FP64 row sums.
//   near_scan / near_fill   counts and near bits -> row-sorted CSR of pairs.
//   near_reg_kernel / near_sing_kernel   one lane group per near pair:
//                   tan-p3c rule for off-element pairs (:649-695,
//                   src/Core/Integration.jl:68-93, :1117-1365), interpolated
//                   Guiggiani for on-element pairs (:397-424,
//                   src/Core/Integration.jl:1597-1845); lanes over quadrature
//                   points, group reductions; the two kinds in separate
//                   launches (the on-element rule by row, before the pair
//                   list exists; sing_scatter moves its values into the list).
//   op64_finalize   rigid-body diagonal H_ii = -sum_{j!=i} H_ij (:384-390) and
//                   the near / on-element entries into H and G.
//   op64_kernel / op64_finalize   FP64 operator passes that recompute the far
//                   entries on the fly (row sums, right-hand side, A*x).
//
// This file is compiled with -ffp-contract=off: the FP64 parity path keeps
// the un-fused operation order of the reference's FP64 arithmetic.

#include "common.hpp"
#include "assembly_shared.hpp"

#include <cmath>
#include <cstring>

namespace bemhip {

namespace {

// ---------------------------------------------------------------------------
// Far-field tile kernel.
// Block = 256 rows x 32 elements. Element data staged in LDS; each lane owns
// one row i, so every store H[i + j*nt] of a wave is 64 consecutive entries.
// Tiles whose row box lies farther than (1 + 1e-6) x the largest cutoff from
// the column block's node box hold only far pairs (the box test of
// near_classify, the same decision): they skip the pair tests. In the other
// tiles the FP32 pre-test settles the clearly-far pairs and the rest (near
// pairs and those within 0.1% of the cutoff) take the exact FP64 test.
// FP64 storage: the entries with the FP64 operation order of the reference.
// Mixed storage: the FP64 far H term (Newton reciprocal, about 1 ulp) goes
// into the FP64 row sum for the diagonal and, rounded, into H; G from the FP32
// logarithm of the rounded squared distance. The row sum is the diagonal of
// the FP64 operator that the refinement iterates with: FP64 accuracy matters
// there (an ill-conditioned system amplifies a relative 1e-13 into the
// solution's digits).
// H and G are written with non-temporal stores: nothing reads them back
// before the kernel ends.
// ---------------------------------------------------------------------------

template <typename TS, int NN>
__global__ __launch_bounds__(kFarThreads) void far_kernel(Geo g, Consts cs,
                                                          const double* __restrict__ cbox,
                                                          const double* __restrict__ rbox,
                                                          TS* __restrict__ H, TS* __restrict__ G,
                                                          double* __restrict__ hpart) {
  constexpr bool kMixed = sizeof(TS) == 4;
  // wave priority 2: the HBM-bound far field issues ahead of the FP64 near
  // rules running beside it (priority 0), which take the cycles it leaves;
  // the short latency-bound kernels of the pair-list chain run at 3
  __builtin_amdgcn_s_setprio(2);
  __shared__ double s_cx[kEB * NN], s_cy[kEB * NN], s_nx[kEB * NN], s_ny[kEB * NN],
      s_jw[kEB * NN], s_L[kEB];
  __shared__ float s_gwf[kEB * NN], s_lim2f[kEB];
  const int cb = blockIdx.y;
  const int e0 = cb * kEB;
  const int ecount = min(kEB, g.ne - e0);
  stage_class_block<NN>(g, cs, e0, ecount, s_cx, s_cy, s_L, s_lim2f);
  for (int t = threadIdx.x; t < ecount * NN; t += blockDim.x) {
    const int j = e0 * NN + t;
    s_nx[t] = g.nx[j]; s_ny[t] = g.ny[j];
    if constexpr (kMixed) {
      s_jw[t] = cs.cT * g.jw[j];                                   // T weight
      s_gwf[t] = (float)(-0.69314718055994531 * cs.cU * g.jw[j]);  // log2 -> G weight
    } else {
      s_jw[t] = g.jw[j];
    }
  }
  const bool all_far = tile_all_far(cbox, rbox, cb, blockIdx.x, cs.nf);
  __syncthreads();
  const int i = blockIdx.x * kFarThreads + threadIdx.x;
  if (i >= g.nt) return;
  const double px = g.rx[i], py = g.ry[i];
  const size_t ld = g.ld;
  double rowsum = 0.0;
  // one far entry pair (H, G) of node t from the FP64 offsets
  auto far_entry = [&](double rx, double ry, int t, TS* Hp, TS* Gp) {
    if constexpr (kMixed) {
      const double d2 = fma(rx, rx, ry * ry);
      const double T = fma(rx, s_nx[t], ry * s_ny[t]) * rcp_f64(d2);
      rowsum = fma(T, s_jw[t], rowsum);
      __builtin_nontemporal_store((float)(T * s_jw[t]), Hp);
      __builtin_nontemporal_store(__builtin_amdgcn_logf((float)d2) * s_gwf[t], Gp);
    } else {
      const double d2 = rx * rx + ry * ry;
      const double U = -log(d2) * cs.cU;
      const double T = (rx * s_nx[t] + ry * s_ny[t]) * (cs.cT / d2);
      const double w = s_jw[t];
      const double h = T * w;
      __builtin_nontemporal_store(h, Hp);
      __builtin_nontemporal_store(U * w, Gp);
      rowsum += h;
    }
  };
  if (all_far) {
    for (int el = 0; el < ecount; ++el) {
      const int col0 = (e0 + el) * NN;
      TS* Hc = H + (size_t)col0 * ld + i;
      TS* Gc = G + (size_t)col0 * ld + i;
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        const int t = el * NN + k;
        far_entry(s_cx[t] - px, s_cy[t] - py, t, Hc + k * ld, Gc + k * ld);
      }
    }
  } else {
    for (int el = 0; el < ecount; ++el) {
      double rx[NN], ry[NN];
      float rxf[NN], ryf[NN], d2f[NN];
      const int col0 = (e0 + el) * NN;
      const bool onel = (i + g.row0 >= col0) && (i + g.row0 < col0 + NN);
      const bool nr = classify_pair<NN>(px, py, s_cx + el * NN, s_cy + el * NN, s_L[el],
                                        s_lim2f[el], cs.nf, onel, rx, ry, rxf, ryf, d2f);
      TS* Hc = H + (size_t)col0 * ld + i;
      TS* Gc = G + (size_t)col0 * ld + i;
      if (nr) {
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          __builtin_nontemporal_store(TS(0), Hc + k * ld);
          __builtin_nontemporal_store(TS(0), Gc + k * ld);
        }
      } else {
#pragma unroll
        for (int k = 0; k < NN; ++k) far_entry(rx[k], ry[k], el * NN + k, Hc + k * ld, Gc + k * ld);
      }
    }
  }
  if (hpart) hpart[(size_t)cb * g.nt + i] = rowsum;
}

// ---------------------------------------------------------------------------
// FP64 operator pass: for every far pair recompute H_ij, G_ij exactly as
// far_kernel<double> and accumulate acc1 += H wH1_j + G wG1_j (and acc2 with
// wH2, wG2 when nvec == 2). Partial sums per 32-element column block. With
// the block boxes of the assembly (cbox, rbox; null: none), all-far tiles skip
// the pair tests as in far_kernel (the same decision, the same arithmetic).
// ---------------------------------------------------------------------------
template <int NN>
__global__ __launch_bounds__(kFarThreads) void op64_kernel(
    Geo g, Consts cs, int nvec, const double* __restrict__ wH1, const double* __restrict__ wG1,
    const double* __restrict__ wH2, const double* __restrict__ wG2, double* __restrict__ part1,
    double* __restrict__ part2, const double* __restrict__ cbox,
    const double* __restrict__ rbox) {
  __shared__ double s_cx[kEB * NN], s_cy[kEB * NN], s_nx[kEB * NN], s_ny[kEB * NN],
      s_jw[kEB * NN], s_L[kEB];
  __shared__ double s_w[4][kEB * NN];
  __shared__ float s_lim2f[kEB];
  const int cb = blockIdx.y;
  const int e0 = cb * kEB;
  const int ecount = min(kEB, g.ne - e0);
  for (int t = threadIdx.x; t < ecount * NN; t += blockDim.x) {
    const int j = e0 * NN + t;
    s_cx[t] = g.x[j]; s_cy[t] = g.y[j];
    s_nx[t] = g.nx[j]; s_ny[t] = g.ny[j];
    s_jw[t] = g.jw[j];
    s_w[0][t] = wH1[j];
    s_w[1][t] = wG1[j];
    if (nvec == 2) { s_w[2][t] = wH2[j]; s_w[3][t] = wG2[j]; }
  }
  for (int t = threadIdx.x; t < ecount; t += blockDim.x) {
    const double L = g.elen[e0 + t];
    s_L[t] = L;
    const double lim = cs.nf * L;
    s_lim2f[t] = (cs.nf > 1e20) ? INFINITY : (float)((lim * lim) * (1.0 + 1e-3));
  }
  const bool all_far = tile_all_far(cbox, rbox, cb, blockIdx.x, cs.nf);
  __syncthreads();
  const int i = blockIdx.x * kFarThreads + threadIdx.x;
  if (i >= g.nt) return;
  const double px = g.rx[i], py = g.ry[i];
  double acc1 = 0.0, acc2 = 0.0;
  for (int el = 0; el < ecount; ++el) {
    const double* cx = s_cx + el * NN;
    const double* cy = s_cy + el * NN;
    double rx[NN], ry[NN], d2[NN];
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      rx[k] = cx[k] - px;
      ry[k] = cy[k] - py;
      d2[k] = rx[k] * rx[k] + ry[k] * ry[k];
    }
    if (!all_far) {
      float rxf[NN], ryf[NN], d2f[NN];
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        rxf[k] = (float)rx[k];
        ryf[k] = (float)ry[k];
        d2f[k] = rxf[k] * rxf[k] + ryf[k] * ryf[k];
      }
      const int col0 = (e0 + el) * NN;
      const bool onel = (i + g.row0 >= col0) && (i + g.row0 < col0 + NN);
      if (onel) continue;
      if (!clearly_far<NN>(rxf, ryf, d2f, s_lim2f[el]) &&
          is_near<NN>(px, py, cx, cy, d2, s_L[el], cs.nf))
        continue;
    }
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      const int t = el * NN + k;
      // A column with zero H (G) weights needs no T (U): the term would be an
      // exact zero. The test is uniform over the workgroup (weights per column).
      const bool needH = s_w[0][t] != 0.0 || (nvec == 2 && s_w[2][t] != 0.0);
      const bool needG = s_w[1][t] != 0.0 || (nvec == 2 && s_w[3][t] != 0.0);
      const double w = s_jw[t];
      double h = 0.0, gg = 0.0;
      if (needG) {
        const double U = -log(d2[k]) * cs.cU;
        gg = U * w;
      }
      if (needH) {
        const double T = (rx[k] * s_nx[t] + ry[k] * s_ny[t]) * (cs.cT / d2[k]);
        h = T * w;
      }
      acc1 += h * s_w[0][t] + gg * s_w[1][t];
      if (nvec == 2) acc2 += h * s_w[2][t] + gg * s_w[3][t];
    }
  }
  const size_t ld = (size_t)g.nt;
  part1[(size_t)cb * ld + i] = acc1;
  if (nvec == 2) part2[(size_t)cb * ld + i] = acc2;
}

// Finalize an operator pass for row i (64 rows per 256-thread block):
//   mode 0 (row sums):  diag_i = -(far + near H entries except column i)
//                       written to diag[] (and to H's diagonal when given);
//   mode 1 (apply):     y_i = far + near (self column: G_ii wG_i + diag_i wH_i)
//                       + diag_i wH_i for internal rows.

template <typename TS>
__global__ __launch_bounds__(256) void op64_finalize(
    int nt, int n, int nn, int ncb, int mode, size_t ldh, const double* __restrict__ part,
    const int* __restrict__ row_ptr, const int* __restrict__ pair_elem,
    const double* __restrict__ pair_h, const double* __restrict__ pair_g,
    const double* __restrict__ wH, const double* __restrict__ wG, double* __restrict__ diag,
    double* __restrict__ y, TS* __restrict__ Hm, TS* __restrict__ Gm) {
  __shared__ double s_red[4][kFinRows];
  const int i = blockIdx.x * kFinRows + (threadIdx.x & (kWave - 1));
  double s = sum_partials(part, ncb, nt, i, s_red);
  if (threadIdx.x >= kWave || i >= nt) return;  // wave 0 finishes its 64 rows
  const int p0 = row_ptr[i], p1 = row_ptr[i + 1];
  if (mode == 0) {
    for (int pr = p0; pr < p1; ++pr) {
      const int c0 = pair_elem[pr] * nn;
      for (int k = 0; k < nn; ++k)
        if (c0 + k != i) s += pair_h[(size_t)pr * nn + k];
    }
    const double d = -s;
    diag[i] = d;
    if (Hm) {  // the near / on-element entries of row i, and the diagonal
      for (int pr = p0; pr < p1; ++pr) {
        const int c0 = pair_elem[pr] * nn;
        for (int k = 0; k < nn; ++k) {
          const size_t at = (size_t)(c0 + k) * ldh + i;
          Hm[at] = TS((c0 + k == i) ? d : pair_h[(size_t)pr * nn + k]);
          Gm[at] = TS(pair_g[(size_t)pr * nn + k]);
        }
      }
      Hm[(size_t)i * ldh + i] = TS(d);
    }
  } else {
    for (int pr = p0; pr < p1; ++pr) {
      const int c0 = pair_elem[pr] * nn;
      for (int k = 0; k < nn; ++k) {
        const int col = c0 + k;
        const double h = (col == i) ? diag[i] : pair_h[(size_t)pr * nn + k];
        s += h * wH[col] + pair_g[(size_t)pr * nn + k] * wG[col];
      }
    }
    if (i >= n) s += diag[i] * wH[i];
    y[i] = s;
  }
}


// Near classification from candidate lists (FMM assembly): the targets of
// leaf lf (tgt_order[tgt_lo[lf] .. tgt_hi[lf])) against the candidate
// elements of the leaf (cand_idx[cand_ptr[lf] .. cand_ptr[lf + 1]), ascending),
// with the pair test of near_classify (same arithmetic, same decisions).
// FILL = false: row_cnt[i] = near pairs of row i. FILL = true: the CSR
// pair_row / pair_elem at ptr[i] and the on-element / other pair lists, as
// near_fill writes them.
template <int NN, bool FILL>
__global__ __launch_bounds__(64) void cand_near_kernel(
    Geo g, Consts cs, const int* __restrict__ tgt_lo, const int* __restrict__ tgt_hi,
    const int* __restrict__ tgt_order, const int* __restrict__ cand_ptr,
    const int* __restrict__ cand_idx, int* __restrict__ row_cnt, const int* __restrict__ ptr,
    int* __restrict__ prow, int* __restrict__ pelem, int* __restrict__ sing,
    int* __restrict__ reg) {
  const int lf = blockIdx.x;
  const int nsing = max(0, min(g.nt, g.n - g.row0));
  for (int q = tgt_lo[lf] + (int)threadIdx.x; q < tgt_hi[lf]; q += 64) {
    const int i = tgt_order[q];
    const double px = g.rx[i], py = g.ry[i];
    const int own = (i < nsing) ? (i + g.row0) / NN : -1;
    int cnt = 0;
    int off = FILL ? ptr[i] : 0;
    int before = min(i, nsing);
    for (int t = cand_ptr[lf]; t < cand_ptr[lf + 1]; ++t) {
      const int e = cand_idx[t];
      double cx[NN], cy[NN], rx[NN], ry[NN];
      float rxf[NN], ryf[NN], d2f[NN];
#pragma unroll
      for (int k = 0; k < NN; ++k) { cx[k] = g.x[e * NN + k]; cy[k] = g.y[e * NN + k]; }
      const double L = g.elen[e];
      const double lim = cs.nf * L;
      const float lim2f = (cs.nf > 1e20) ? INFINITY : (float)((lim * lim) * (1.0 + 1e-3));
      const int col0 = e * NN;
      const bool onel = (i + g.row0 >= col0) && (i + g.row0 < col0 + NN);
      if (!classify_pair<NN>(px, py, cx, cy, L, lim2f, cs.nf, onel, rx, ry, rxf, ryf, d2f))
        continue;
      if (!FILL) { ++cnt; continue; }
      prow[off] = i;
      pelem[off] = e;
      if (e == own) {
        sing[i] = off;
        ++before;
      } else {
        reg[off - before] = off;
      }
      ++off;
    }
    if (!FILL) row_cnt[i] = cnt;
  }
}


// ---------------------------------------------------------------------------
// Near-pair integrals, one group of lanes per pair. near_fill lists the two
// kinds of pairs apart, so that no wavefront runs both rules:
//   near_reg_kernel<NN, G>  off-element pairs: tan-p3c rule over npg points
//                           (2 npg when split), G lanes per pair (G = 16, 32
//                           or 64, at least npg), 64 / G pairs per wavefront;
//   near_sing_kernel<NN>    on-element pairs: interpolated Guiggiani over 20
//                           nodes, 32 lanes per pair, two pairs per wavefront.
// Lanes run over quadrature points; the values go to the pair list at the
// pair's CSR index. list == nullptr: the k-th pair is CSR index k.
// near_sing_kernel with prow == pelem == nullptr takes the k-th pair as row k
// and the element holding its node, (k + row0) / NN, and writes slot k: the
// on-element pairs need no classification, so the assembly starts them first.
// ---------------------------------------------------------------------------

template <int NN, int G>
__global__ __launch_bounds__(256) void near_reg_kernel(Geo g, const QuadConst* __restrict__ qc,
                                                       int64_t count,
                                                       const int* __restrict__ list,
                                                       const int* __restrict__ prow,
                                                       const int* __restrict__ pelem,
                                                       double* __restrict__ ph,
                                                       double* __restrict__ pg) {
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
  const double cU = 0.5 * (1.0 / (2.0 * kPi)) / qc->k;
  const double cT = 1.0 / (2.0 * kPi);
  double hs[NN], gs[NN];
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
    // at most two rule nodes per lane (nq <= 2 npg <= 2 G): placed first, so
    // that the plan's registers are free during the integration
    double xqs[2], wqs[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      xqs[r] = 0.0; wqs[r] = 0.0;
      if (r * G + lane < nq) tanp3c_node(P, r * G + lane, npg, qc->gl_x, qc->gl_w, xqs[r], wqs[r]);
    }
    double ha[NN], ga[NN];
#pragma unroll
    for (int k = 0; k < NN; ++k) { ha[k] = 0.0; ga[k] = 0.0; }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (r * G + lane < nq) {
        const double xq = xqs[r], wq = wqs[r];
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
        const double R2 = rx * rx + ry * ry;
        const double U = -log(R2) * cU;
        const double T = (rx * nx + ry * ny) * (cT / R2);
        const double wi = J * wq;
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          const double Nj = N[k] * wi;
          ha[k] += T * Nj;
          ga[k] += U * Nj;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NN; ++k) { hs[k] = group_sum<G>(ha[k]); gs[k] = group_sum<G>(ga[k]); }
  }
  if (active && lane < NN) {
    double hv = hs[0], gv = gs[0];
#pragma unroll
    for (int k = 1; k < NN; ++k)
      if (lane == k) { hv = hs[k]; gv = gs[k]; }
    ph[(size_t)pair * NN + lane] = hv;
    pg[(size_t)pair * NN + lane] = gv;
  }
}

template <int NN>
__global__ __launch_bounds__(256) void near_sing_kernel(Geo g, const QuadConst* __restrict__ qc,
                                                        int64_t count,
                                                        const int* __restrict__ list,
                                                        const int* __restrict__ prow,
                                                        const int* __restrict__ pelem,
                                                        double* __restrict__ ph,
                                                        double* __restrict__ pg) {
  constexpr int G = 32;
  const int64_t k_raw =
      ((int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * 2 +
      ((threadIdx.x & (kWave - 1)) >> 5);
  const bool active = k_raw < count;
  const int64_t kk = active ? k_raw : count - 1;  // idle groups redo the last pair
  const int64_t pair = list ? (int64_t)list[kk] : kk;
  const int lane = threadIdx.x & (G - 1);
  const int gshift = threadIdx.x & G;  // 0 or 32: bit offset of this group in a ballot
  const int i = prow ? prow[pair] : (int)pair;
  const int e = pelem ? pelem[pair] : (int)((pair + g.row0) / NN);
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
  const double cU = 0.5 * (1.0 / (2.0 * kPi)) / qc->k;
  const double cT = 1.0 / (2.0 * kPi);
  double hs[NN], gs[NN];
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
    // kernel samples F = (U, T) * N_k * J  (src/Core/Assembly_full.jl:629-643)
    double Fg[NN], Fh[NN];
#pragma unroll
    for (int k = 0; k < NN; ++k) { Fg[k] = 0.0; Fh[k] = 0.0; }
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
        const double R2 = rx * rx + ry * ry;
        const double U = -log(R2) * cU;
        const double T = (rx * nx + ry * ny) * (cT / R2);
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          const double NjJ = N[k] * J;
          Fg[k] = U * NjJ;
          Fh[k] = T * NjJ;
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
    for (int k = 0; k < NN; ++k) {
      const double Fm1g = group_sum<G>(keep ? Ni * (Fg[k] * delta) : 0.0);
      const double F1h = group_sum<G>(keep ? Ni * (Fh[k] * delta) : 0.0);
      const double wi = keep ? Fg[k] - Fm1g / delta : 0.0;
      const double sW = group_sum<G>(used ? wi : 0.0);
      const double sWL = group_sum<G>(used ? wi * lg : 0.0);
      double F0 = 0.0;
      if (nused >= 2 && !(fabs(detA) < 1e-30)) F0 = (nused * sWL - sL * sW) / detA;
      const double regg = keep ? wq * ((Fg[k] - Fm1g / delta) - F0 * lgd) : 0.0;
      const double regh = keep ? wq * (Fh[k] - F1h / delta) : 0.0;
      gs[k] = group_sum<G>(regg) + (Fm1g * hfp1 + F0 * logp);
      hs[k] = group_sum<G>(regh) + F1h * hfp1;
    }
  }
  if (active && lane < NN) {
    double hv = hs[0], gv = gs[0];
#pragma unroll
    for (int k = 1; k < NN; ++k)
      if (lane == k) { hv = hs[k]; gv = gs[k]; }
    ph[(size_t)pair * NN + lane] = hv;
    pg[(size_t)pair * NN + lane] = gv;
  }
}

// On-element values by row (near_sing_kernel's row mode) into the pair list:
// slot r goes to CSR index sing[r].
__global__ void sing_scatter(int nsing, int nn, const int* __restrict__ sing,
                             const double* __restrict__ sh, const double* __restrict__ sg,
                             double* __restrict__ ph, double* __restrict__ pg) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsing * nn) return;
  const int r = t / nn, k = t - r * nn;
  const size_t dst = (size_t)sing[r] * nn + k;
  ph[dst] = sh[t];
  pg[dst] = sg[t];
}

// Column swap A = H with Dirichlet columns -G (src/Core/Boundary_conditions.jl:54-64).
template <typename TS>
__global__ void bc_swap_kernel(int nt, int n, size_t ld, const int* __restrict__ bc,
                               const TS* __restrict__ H, const TS* __restrict__ G,
                               TS* __restrict__ A) {
  const int j = blockIdx.y;
  const bool dir = (j < n) && bc[j] == 0;
  const TS* src = dir ? G + (size_t)j * ld : H + (size_t)j * ld;
  TS* dst = A + (size_t)j * ld;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += gridDim.x * blockDim.x)
    dst[i] = dir ? -src[i] : src[i];
}

// Dynamic LDS that the FP32 far_kernel reserves and does not use: with its
// 5.4 KB of element data a workgroup then holds 37.4 KB, so at most 4 of them
// share a CU's 160 KB. The HBM-write-bound FP32 far field runs as fast at 4
// workgroups per CU as at 8 (measured: 1.62-1.66 ms either way at N = 32769),
// and the wave slots and registers left free let the near-pair chain, which
// runs beside it, make progress (mixed assembly 1.96 -> 1.82 ms). The FP64
// far field (a logarithm and a division per entry) needs its full occupancy.
constexpr unsigned kFarLdsPad = 32768;

template <typename TS, int NN>
int launch_far(bemhip_ctx* c, hipStream_t stream, const Geo& g, const Consts& cs,
               const double* cbox, const double* rbox, TS* H, TS* G, double* hpart) {
  dim3 grid((g.nt + kFarThreads - 1) / kFarThreads, (g.ne + kEB - 1) / kEB);
  const unsigned pad = sizeof(TS) == 4 ? kFarLdsPad : 0u;
  hipLaunchKernelGGL((far_kernel<TS, NN>), grid, dim3(kFarThreads), pad, stream, g, cs, cbox, rbox,
                     H, G, hpart);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

// Near-pair integrals of a CSR pair list: the on-element pairs (sing, nsing)
// and the others (reg, nreg; null = CSR order) in launches of their own.
template <int NN>
int launch_near_pairs(bemhip_ctx* c, hipStream_t stream, const Geo& g, const QuadConst* qc,
                      int npg, int64_t nsing, const int* sing, int64_t nreg, const int* reg,
                      const int* prow, const int* pelem, double* ph, double* pg) {
  constexpr int waves = 256 / kWave;
  if (nsing > 0) {
    const int64_t blocks = (nsing + 2 * waves - 1) / (2 * waves);
    hipLaunchKernelGGL((near_sing_kernel<NN>), dim3((unsigned)blocks), dim3(256), 0, stream, g,
                       qc, nsing, sing, prow, pelem, ph, pg);
  }
  if (nreg > 0) {
    if (npg <= 16) {
      const int64_t blocks = (nreg + 4 * waves - 1) / (4 * waves);
      hipLaunchKernelGGL((near_reg_kernel<NN, 16>), dim3((unsigned)blocks), dim3(256), 0, stream,
                         g, qc, nreg, reg, prow, pelem, ph, pg);
    } else if (npg <= 32) {
      const int64_t blocks = (nreg + 2 * waves - 1) / (2 * waves);
      hipLaunchKernelGGL((near_reg_kernel<NN, 32>), dim3((unsigned)blocks), dim3(256), 0, stream,
                         g, qc, nreg, reg, prow, pelem, ph, pg);
    } else {
      const int64_t blocks = (nreg + waves - 1) / waves;
      hipLaunchKernelGGL((near_reg_kernel<NN, 64>), dim3((unsigned)blocks), dim3(256), 0, stream,
                         g, qc, nreg, reg, prow, pelem, ph, pg);
    }
  }
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

}  // namespace

// x = a, y = b over n entries.
__global__ void fill2_kernel(int n, double a, double* __restrict__ x, double b,
                             double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { x[i] = a; y[i] = b; }
}

// Pinned host word for the near-pair count, written by near_scan through its
// device address.
hipError_t ensure_count_word(bemhip_ctx* c) {
  if (c->h_count && c->d_count) return hipSuccess;
  if (!c->h_count) {
    hipError_t e = hipHostMalloc((void**)&c->h_count, sizeof(int));
    if (e != hipSuccess) return e;
  }
  return hipHostGetDevicePointer((void**)&c->d_count, c->h_count, 0);
}

Geo make_geo(bemhip_ctx* c) {
  Geo g;
  g.p = c->p; g.nn = c->nn; g.ne = c->ne; g.n = c->n; g.ni = c->ni; g.nt = c->nt;
  g.x = c->x.ptr; g.y = c->y.ptr; g.nx = c->nx.ptr; g.ny = c->ny.ptr; g.jw = c->jw.ptr;
  g.elen = c->elen.ptr; g.gx = c->gx.ptr; g.gy = c->gy.ptr;
  g.rx = c->x.ptr; g.ry = c->y.ptr; g.row0 = 0;
  g.ld = (c->precision == BEMHIP_FP64) ? c->ld64 : c->ld32;
  return g;
}

static Consts make_consts(const bemhip_ctx* c) {
  Consts cs;
  const double inv2pi = 1.0 / (2.0 * kPi);
  cs.cU = (0.5 * inv2pi) / c->k;
  cs.cT = inv2pi;
  cs.nf = c->near_factor;
  return cs;
}

// Full assembly on three streams:
//   stream 3: near_sing (row mode) ----------------------------------------+
//   stream 1: near_classify -> near-pair CSR -> near_reg -> sing_scatter -+-> finalize
//   stream 2: far_kernel -------------------------------------------------/
// The far field (HBM-write bound) and the near/singular integrals (FP64 VALU
// bound) run concurrently; the on-element rule needs no classification (row
// i's own element is (i + row0) / nn), so it starts first. far_kernel writes
// 0 into the near entries, so the near kernels keep their values in the pair
// list and the finalize pass, after all, writes them into H and G together
// with the rigid-body diagonal.
int assemble_impl(bemhip_ctx* c) {
  const Geo g = make_geo(c);
  const Consts cs = make_consts(c);
  const size_t nt = (size_t)c->nt;
  const int ncb = (c->ne + kEB - 1) / kEB;
  const bool fp64 = (c->precision == BEMHIP_FP64);
  const bool fmm = (c->precision == BEMHIP_FMM);
  const bool mfree = (c->precision == BEMHIP_MATRIX_FREE) || fmm;
  if (fmm) {
    const int st0 = fmm_build(c);
    if (st0) return st0;
  }
  if (mfree && !fmm) {
    // near bits (4 B), FP64 far row sums (8 B) and the products' partial sums
    // (8 B) per (row, 32-element column block): refuse what does not fit
    c->H64.release(); c->G64.release(); c->H32.release(); c->G32.release();
    c->A64.release(); c->A32.release();
    const int st0 = check_matrix_free_scratch(c, "assemble", 4 + 8 + 8, (size_t)ncb, nt);
    if (st0) return st0;
  }
  if (fmm)  // near pairs from the tree's candidates: no near bits
    c->nearbits.release();
  else
    BEMHIP_HIP_TRY(c, c->nearbits.ensure((size_t)ncb * nt));
  BEMHIP_HIP_TRY(c, c->hpart.ensure(fmm ? nt : (size_t)ncb * nt));  // FMM: one partial per row
  BEMHIP_HIP_TRY(c, c->diag64.ensure(nt));
  BEMHIP_HIP_TRY(c, c->row_cnt.ensure(nt));
  BEMHIP_HIP_TRY(c, c->row_ptr.ensure(nt + 1));
  if (mfree) {
    // no stored H, G: the far field of the diagonal's row sums comes from an
    // FP64 operator pass with unit H weights (the same sums as the FP64 far kernel)
    c->H64.release(); c->G64.release(); c->H32.release(); c->G32.release();
    BEMHIP_HIP_TRY(c, c->wH.ensure(nt));
    BEMHIP_HIP_TRY(c, c->wG.ensure(nt));
    hipLaunchKernelGGL(fill2_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream,
                       (int)nt, 1.0, c->wH.ptr, 0.0, c->wG.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  } else if (fp64) {
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->H64, c->ld64 * nt));
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->G64, c->ld64 * (size_t)c->n, &c->H64));
    if (c->ni > 0)  // columns of the internal nodes: zero but the diagonal
      BEMHIP_HIP_TRY(c, hipMemsetAsync(c->H64.ptr + (size_t)c->n * c->ld64, 0,
                                       sizeof(double) * c->ld64 * c->ni, c->stream));
  } else {
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->H32, c->ld32 * nt));
    BEMHIP_HIP_TRY(c, ensure_matrix(c, c->G32, c->ld32 * (size_t)c->n, &c->H32));
    if (c->ni > 0)
      BEMHIP_HIP_TRY(c, hipMemsetAsync(c->H32.ptr + (size_t)c->n * c->ld32, 0,
                                       sizeof(float) * c->ld32 * c->ni, c->stream));
  }
  // events: 0 start, 1 classified, 2 far done, 3 pairs start, 4 pairs done,
  // 5 near done, 6 diagonal done, 7 far start, 8 finalize start, 9 on-element
  // start, 10 on-element done, 11 block boxes done
  constexpr int kEvents = 12;
  hipEvent_t ev[kEvents];
  for (auto& e : ev) BEMHIP_HIP_TRY(c, hipEventCreate(&e));
  struct Destroy {
    hipEvent_t* e;
    ~Destroy() { for (int k = 0; k < kEvents; ++k) hipEventDestroy(e[k]); }
  } destroy{ev};
  const int nfb = (int)((nt + kFinRows - 1) / kFinRows);
  hipStream_t s2 = c->stream2;

  // stream 1: classification with the per-row counts; then the far field on
  // stream 2 (the longer chain: HBM-write bound) while stream 1 builds the
  // near-pair CSR and runs the near integrals.
  const int nrb = (int)((nt + kFarThreads - 1) / kFarThreads);  // row blocks
  BEMHIP_HIP_TRY(c, ensure_count_word(c));
  const int64_t nsing = std::min<int64_t>(nt, c->n);  // one on-element pair per boundary row
  BEMHIP_HIP_TRY(c, c->sing_h.ensure((size_t)std::max<int64_t>(nsing, 1) * c->nn));
  BEMHIP_HIP_TRY(c, c->sing_g.ensure((size_t)std::max<int64_t>(nsing, 1) * c->nn));
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[0], c->stream));
  const FmmPlan& FP = c->fmm;
  if (fmm) {
    if (c->nn == 2)
      hipLaunchKernelGGL((cand_near_kernel<2, false>), dim3(FP.nleaf), dim3(64), 0, c->stream, g,
                         cs, FP.tgt_lo.ptr, FP.tgt_hi.ptr, FP.tgt_order.ptr, FP.cand_ptr.ptr,
                         FP.cand_idx.ptr, c->row_cnt.ptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr);
    else
      hipLaunchKernelGGL((cand_near_kernel<3, false>), dim3(FP.nleaf), dim3(64), 0, c->stream, g,
                         cs, FP.tgt_lo.ptr, FP.tgt_hi.ptr, FP.tgt_order.ptr, FP.cand_ptr.ptr,
                         FP.cand_idx.ptr, c->row_cnt.ptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  } else {
    BEMHIP_HIP_TRY(c, c->blkbox.ensure((size_t)ncb * 5));
    BEMHIP_HIP_TRY(c, c->rowbox.ensure((size_t)nrb * 4));
    BEMHIP_HIP_TRY(c, c->tiles.ensure((size_t)nrb * ncb));
  }
  if (!fmm) {
    // node boxes of the column blocks and row blocks: the all-far tile test
    // of near_classify and far_kernel (also clears row_cnt)
    const unsigned nb = (unsigned)(ncb + nrb);
    if (c->nn == 2)
      hipLaunchKernelGGL((block_boxes<2>), dim3(nb), dim3(64), 0, c->stream, g, cs, ncb,
                         c->blkbox.ptr, c->rowbox.ptr, c->row_cnt.ptr);
    else
      hipLaunchKernelGGL((block_boxes<3>), dim3(nb), dim3(64), 0, c->stream, g, cs, ncb,
                         c->blkbox.ptr, c->rowbox.ptr, c->row_cnt.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[11], c->stream));
  // stream 3: the on-element rule by row, in line for the CUs right after the
  // block boxes (the far field waits for those)
  hipStream_t s3 = c->stream3;
  BEMHIP_HIP_TRY(c, hipStreamWaitEvent(s3, ev[0], 0));
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[9], s3));
  if (nsing > 0) {
    const int64_t blocks = (nsing + 7) / 8;  // two rows per wavefront, four wavefronts
    if (c->nn == 2)
      hipLaunchKernelGGL((near_sing_kernel<2>), dim3((unsigned)blocks), dim3(256), 0, s3, g,
                         c->qc.ptr, nsing, nullptr, nullptr, nullptr, c->sing_h.ptr,
                         c->sing_g.ptr);
    else
      hipLaunchKernelGGL((near_sing_kernel<3>), dim3((unsigned)blocks), dim3(256), 0, s3, g,
                         c->qc.ptr, nsing, nullptr, nullptr, nullptr, c->sing_h.ptr,
                         c->sing_g.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[10], s3));
  if (!fmm) {
    const int cpb = 32;  // column blocks per classify workgroup
    dim3 grid((unsigned)((nt + kFarThreads - 1) / kFarThreads), (unsigned)((ncb + cpb - 1) / cpb));
    if (c->nn == 2)
      hipLaunchKernelGGL((near_classify<2>), grid, dim3(kFarThreads), 0, c->stream, g, cs, ncb,
                         cpb, c->blkbox.ptr, c->nearbits.ptr, c->row_cnt.ptr, c->tiles.ptr);
    else
      hipLaunchKernelGGL((near_classify<3>), grid, dim3(kFarThreads), 0, c->stream, g, cs, ncb,
                         cpb, c->blkbox.ptr, c->nearbits.ptr, c->row_cnt.ptr, c->tiles.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[1], c->stream));
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[3], c->stream));
  // stream 2: the far field (needs the geometry, the block boxes and the unit
  // weights): it starts with the classification, not after it
  int st = BEMHIP_OK;
  if (!fmm) {
  BEMHIP_HIP_TRY(c, hipStreamWaitEvent(s2, ev[11], 0));
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[7], s2));
  if (mfree) {
    dim3 grid((unsigned)nrb, (unsigned)ncb);
    if (c->nn == 2)
      hipLaunchKernelGGL(op64_kernel<2>, grid, dim3(kFarThreads), 0, s2, g, cs, 1, c->wH.ptr,
                         c->wG.ptr, nullptr, nullptr, c->hpart.ptr, nullptr, c->blkbox.ptr,
                         c->rowbox.ptr);
    else
      hipLaunchKernelGGL(op64_kernel<3>, grid, dim3(kFarThreads), 0, s2, g, cs, 1, c->wH.ptr,
                         c->wG.ptr, nullptr, nullptr, c->hpart.ptr, nullptr, c->blkbox.ptr,
                         c->rowbox.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    st = BEMHIP_OK;
  } else if (fp64)
    st = (c->nn == 2) ? launch_far<double, 2>(c, s2, g, cs, c->blkbox.ptr, c->rowbox.ptr,
                                              c->H64.ptr, c->G64.ptr, c->hpart.ptr)
                      : launch_far<double, 3>(c, s2, g, cs, c->blkbox.ptr, c->rowbox.ptr,
                                              c->H64.ptr, c->G64.ptr, c->hpart.ptr);
  else
    st = (c->nn == 2) ? launch_far<float, 2>(c, s2, g, cs, c->blkbox.ptr, c->rowbox.ptr,
                                             c->H32.ptr, c->G32.ptr, c->hpart.ptr)
                      : launch_far<float, 3>(c, s2, g, cs, c->blkbox.ptr, c->rowbox.ptr,
                                             c->H32.ptr, c->G32.ptr, c->hpart.ptr);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[2], s2));
  }
  hipLaunchKernelGGL(near_scan, dim3(1), dim3(kScanWaves * kWave), 0, c->stream, (int)nt,
                     c->row_cnt.ptr, c->row_ptr.ptr, c->d_count);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  const int total = *(volatile int*)c->h_count;
  c->npairs = total;
  BEMHIP_HIP_TRY(c, c->pair_row.ensure((size_t)std::max(total, 1)));
  BEMHIP_HIP_TRY(c, c->pair_elem.ensure((size_t)std::max(total, 1)));
  BEMHIP_HIP_TRY(c, c->pair_h.ensure((size_t)std::max(total, 1) * c->nn));
  BEMHIP_HIP_TRY(c, c->pair_g.ensure((size_t)std::max(total, 1) * c->nn));
  const int64_t nreg = (int64_t)total - nsing;
  BEMHIP_HIP_TRY(c, c->sing_idx.ensure((size_t)std::max<int64_t>(nsing, 1)));
  BEMHIP_HIP_TRY(c, c->reg_idx.ensure((size_t)std::max<int64_t>(nreg, 1)));
  if (fmm) {
    if (c->nn == 2)
      hipLaunchKernelGGL((cand_near_kernel<2, true>), dim3(FP.nleaf), dim3(64), 0, c->stream, g,
                         cs, FP.tgt_lo.ptr, FP.tgt_hi.ptr, FP.tgt_order.ptr, FP.cand_ptr.ptr,
                         FP.cand_idx.ptr, nullptr, c->row_ptr.ptr, c->pair_row.ptr,
                         c->pair_elem.ptr, c->sing_idx.ptr, c->reg_idx.ptr);
    else
      hipLaunchKernelGGL((cand_near_kernel<3, true>), dim3(FP.nleaf), dim3(64), 0, c->stream, g,
                         cs, FP.tgt_lo.ptr, FP.tgt_hi.ptr, FP.tgt_order.ptr, FP.cand_ptr.ptr,
                         FP.cand_idx.ptr, nullptr, c->row_ptr.ptr, c->pair_row.ptr,
                         c->pair_elem.ptr, c->sing_idx.ptr, c->reg_idx.ptr);
  } else {
    hipLaunchKernelGGL(near_fill, dim3(nrb), dim3(kFarThreads), 0, c->stream, (int)nt, ncb, c->nn,
                       0, c->n, c->nearbits.ptr, c->tiles.ptr, c->row_ptr.ptr, c->pair_row.ptr,
                       c->pair_elem.ptr, c->sing_idx.ptr, c->reg_idx.ptr);
  }
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[4], c->stream));
  if (fmm) {  // stream 2: FMM far row sums of H (P2P needs the near-pair CSR)
    BEMHIP_HIP_TRY(c, hipStreamWaitEvent(s2, ev[4], 0));
    BEMHIP_HIP_TRY(c, hipEventRecord(ev[7], s2));
    st = fmm_far(c, s2, c->wH.ptr, c->wG.ptr, c->hpart.ptr);
    if (st) return st;
    BEMHIP_HIP_TRY(c, hipEventRecord(ev[2], s2));
  }
  // Near entries: far_kernel writes 0 into the near entries of H and G, so the
  // near values stay in the pair list and the finalize pass writes them into
  // H and G after all. The on-element values (stream 3) join the list here.
  st = (c->nn == 2)
               ? launch_near_pairs<2>(c, c->stream, g, c->qc.ptr, c->npg, 0, nullptr, nreg,
                                      c->reg_idx.ptr, c->pair_row.ptr, c->pair_elem.ptr,
                                      c->pair_h.ptr, c->pair_g.ptr)
               : launch_near_pairs<3>(c, c->stream, g, c->qc.ptr, c->npg, 0, nullptr, nreg,
                                      c->reg_idx.ptr, c->pair_row.ptr, c->pair_elem.ptr,
                                      c->pair_h.ptr, c->pair_g.ptr);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipStreamWaitEvent(c->stream, ev[10], 0));
  if (nsing > 0) {
    const int cnt = (int)nsing * c->nn;
    hipLaunchKernelGGL(sing_scatter, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0,
                       c->stream, (int)nsing, c->nn, c->sing_idx.ptr, c->sing_h.ptr,
                       c->sing_g.ptr, c->pair_h.ptr, c->pair_g.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[5], c->stream));
  BEMHIP_HIP_TRY(c, hipStreamWaitEvent(c->stream, ev[2], 0));
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[8], c->stream));
  // join: near values into H, G and the rigid-body diagonal
  if (mfree)
    hipLaunchKernelGGL((op64_finalize<double>), dim3(nfb), dim3(256), 0, c->stream, (int)nt,
                       c->n, c->nn, fmm ? 1 : ncb, 0, (size_t)0, c->hpart.ptr, c->row_ptr.ptr,
                       c->pair_elem.ptr, c->pair_h.ptr, c->pair_g.ptr, nullptr, nullptr,
                       c->diag64.ptr, nullptr, (double*)nullptr, (double*)nullptr);
  else if (fp64)
    hipLaunchKernelGGL((op64_finalize<double>), dim3(nfb), dim3(256), 0, c->stream, (int)nt,
                       c->n, c->nn, ncb, 0, c->ld64, c->hpart.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                       c->pair_h.ptr, c->pair_g.ptr, nullptr, nullptr, c->diag64.ptr, nullptr,
                       c->H64.ptr, c->G64.ptr);
  else
    hipLaunchKernelGGL((op64_finalize<float>), dim3(nfb), dim3(256), 0, c->stream, (int)nt,
                       c->n, c->nn, ncb, 0, c->ld32, c->hpart.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                       c->pair_h.ptr, c->pair_g.ptr, nullptr, nullptr, c->diag64.ptr, nullptr,
                       c->H32.ptr, c->G32.ptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipEventRecord(ev[6], c->stream));
  BEMHIP_HIP_TRY(c, hipEventSynchronize(ev[6]));
  auto el = [&](int a, int b) {
    float t = 0.f;
    hipEventElapsedTime(&t, ev[a], ev[b]);
    return (double)t;
  };
  c->ms["assemble_classify"] = el(0, 1);
  c->ms["assemble_far"] = el(7, 2);
  c->ms["assemble_pairs"] = el(3, 4);
  c->ms["assemble_near"] = el(4, 5);
  c->ms["assemble_diag"] = el(8, 6);
  c->ms["assemble_sing"] = el(9, 10);
  return BEMHIP_OK;
}

__global__ void apply_weights_kernel(int nt, int n, const int* __restrict__ bc,
                                     const double* __restrict__ x, double* __restrict__ wH,
                                     double* __restrict__ wG) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  const bool dir = (j < n) && bc[j] == 0;
  wH[j] = dir ? 0.0 : x[j];
  wG[j] = dir ? -x[j] : 0.0;
}

__global__ void rhs_weights_kernel(int nt, int n, const int* __restrict__ bc,
                                   const double* __restrict__ bv, double* __restrict__ wH,
                                   double* __restrict__ wG) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  if (j >= n) { wH[j] = 0.0; wG[j] = 0.0; return; }
  const bool dir = bc[j] == 0;
  wH[j] = dir ? -bv[j] : 0.0;
  wG[j] = dir ? 0.0 : bv[j];
}

// One FP64 operator pass y = sum_j H_ij wH_j + G_ij wG_j over all columns, with
// the far entries recomputed, near entries from the pair list and the
// rigid-body diagonal in column i (mode 1 of op64_finalize).
static int operator_pass(bemhip_ctx* c, const double* wH, const double* wG, double* d_y) {
  const Geo g = make_geo(c);
  const Consts cs = make_consts(c);
  const size_t nt = (size_t)c->nt;
  const int nfb = (int)((nt + kFinRows - 1) / kFinRows);
  if (c->precision == BEMHIP_FMM) {  // far field by the FMM, one partial per row
    BEMHIP_HIP_TRY(c, c->part1.ensure(nt));
    const int st = fmm_far(c, c->stream, wH, wG, c->part1.ptr);
    if (st) return st;
    hipLaunchKernelGGL((op64_finalize<double>), dim3(nfb), dim3(256), 0, c->stream, (int)nt, c->n,
                       c->nn, 1, 1, (size_t)0, c->part1.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                       c->pair_h.ptr, c->pair_g.ptr, wH, wG, c->diag64.ptr, d_y,
                       (double*)nullptr, (double*)nullptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    return BEMHIP_OK;
  }
  const int ncb = (c->ne + kEB - 1) / kEB;
  BEMHIP_HIP_TRY(c, c->part1.ensure((size_t)ncb * nt));
  dim3 grid((unsigned)((nt + kFarThreads - 1) / kFarThreads), (unsigned)ncb);
  // the block boxes of the (non-FMM) assembly this operator belongs to
  if (c->nn == 2)
    hipLaunchKernelGGL(op64_kernel<2>, grid, dim3(kFarThreads), 0, c->stream, g, cs, 1, wH, wG,
                       nullptr, nullptr, c->part1.ptr, nullptr, c->blkbox.ptr, c->rowbox.ptr);
  else
    hipLaunchKernelGGL(op64_kernel<3>, grid, dim3(kFarThreads), 0, c->stream, g, cs, 1, wH, wG,
                       nullptr, nullptr, c->part1.ptr, nullptr, c->blkbox.ptr, c->rowbox.ptr);
  hipLaunchKernelGGL((op64_finalize<double>), dim3(nfb), dim3(256), 0, c->stream, (int)nt, c->n,
                     c->nn, ncb, 1, (size_t)0, c->part1.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                     c->pair_h.ptr, c->pair_g.ptr, wH, wG, c->diag64.ptr, d_y,
                     (double*)nullptr, (double*)nullptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

// y = A64 x for the mixed-BC system (Dirichlet columns of A are -G).
int operator_apply(bemhip_ctx* c, const double* d_x, double* d_y) {
  if (c->dof == 2) return kelvin_operator_apply(c, d_x, d_y);
  const size_t nt = (size_t)c->nt;
  BEMHIP_HIP_TRY(c, c->wH2.ensure(nt));
  BEMHIP_HIP_TRY(c, c->wG2.ensure(nt));
  const int tb = 256, nb = (int)((nt + tb - 1) / tb);
  hipLaunchKernelGGL(apply_weights_kernel, dim3(nb), dim3(tb), 0, c->stream, c->nt, c->n,
                     c->d_bc.ptr, d_x, c->wH2.ptr, c->wG2.ptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return operator_pass(c, c->wH2.ptr, c->wG2.ptr, d_y);
}

// applyBC on the device: b from an FP64 operator pass (identical in both
// precisions), A from the column swap of the stored H, G.
int apply_bc_impl(bemhip_ctx* c) {
  if (c->dof == 2) return kelvin_apply_bc_impl(c);
  const size_t nt = (size_t)c->nt;
  BEMHIP_HIP_TRY(c, c->wH2.ensure(nt));
  BEMHIP_HIP_TRY(c, c->wG2.ensure(nt));
  BEMHIP_HIP_TRY(c, c->b64.ensure(nt));
  const int tb = 256, nb = (int)((nt + tb - 1) / tb);
  {
    ScopedTimer t(c, "apply_bc_rhs");
    hipLaunchKernelGGL(rhs_weights_kernel, dim3(nb), dim3(tb), 0, c->stream, c->nt, c->n,
                       c->d_bc.ptr, c->d_bv.ptr, c->wH2.ptr, c->wG2.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    int st = operator_pass(c, c->wH2.ptr, c->wG2.ptr, c->b64.ptr);
    if (st) return st;
  }
  if (c->precision == BEMHIP_MATRIX_FREE || c->precision == BEMHIP_FMM)
    return BEMHIP_OK;  // A is applied, never stored
  {
    ScopedTimer t(c, "apply_bc_swap");
    dim3 grid((unsigned)std::min<size_t>((nt + 255) / 256, 64), (unsigned)nt);
    if (c->precision == BEMHIP_FP64) {
      BEMHIP_HIP_TRY(c, ensure_matrix(c, c->A64, c->ld64 * nt, &c->H64, &c->G64));
      hipLaunchKernelGGL((bc_swap_kernel<double>), grid, dim3(256), 0, c->stream, c->nt, c->n,
                         c->ld64, c->d_bc.ptr, c->H64.ptr, c->G64.ptr, c->A64.ptr);
    } else {
      BEMHIP_HIP_TRY(c, ensure_matrix(c, c->A32, c->ld32 * nt, &c->H32, &c->G32));
      hipLaunchKernelGGL((bc_swap_kernel<float>), grid, dim3(256), 0, c->stream, c->nt, c->n,
                         c->ld32, c->d_bc.ptr, c->H32.ptr, c->G32.ptr, c->A32.ptr);
    }
    BEMHIP_HIP_TRY(c, hipGetLastError());
  }
  return BEMHIP_OK;
}

// ---------------------------------------------------------------------------
// Potential at points inside the domain (src/Laplace/Lubrication.jl:399-417;
// the same formula in the reference's tests, test/laplace.jl:36-53):
//   u(p) = (sum_j g_j q_j - sum_j h_j T_j) / c,   c = -sum_j h_j,
// h_j, g_j the integrals of T*, U* times the shape function of boundary node
// j over its element, seen from p. Row sums per point: s1 = sum_j h_j w[0][j]
// + g_j w[1][j] (w[0] = -T, w[1] = q) and s2 = sum_j h_j; u = s1 / (-s2).
// part: [2][ncb][npts] far partial sums (null when every pair is near).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void interior_finalize(
    int npts, int n, int nn, int ncb, const double* __restrict__ part,
    const int* __restrict__ row_ptr, const int* __restrict__ pair_elem,
    const double* __restrict__ ph, const double* __restrict__ pg, const double* __restrict__ w,
    double* __restrict__ out) {
  __shared__ double s_red1[4][kFinRows], s_red2[4][kFinRows];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int i0 = blockIdx.x * kFinRows;
  if (part) {  // uniform over the launch
    const int i = i0 + lane;
    const double f1 = sum_partials(part, ncb, npts, i, s_red1);
    const double f2 = sum_partials(part + (size_t)ncb * npts, ncb, npts, i, s_red2);
    __syncthreads();  // every wave has read the sums of its lane's point
    if (wv == 0) { s_red1[0][lane] = f1; s_red2[0][lane] = f2; }
  } else if (wv == 0) {
    s_red1[0][lane] = 0.0;
    s_red2[0][lane] = 0.0;
  }
  __syncthreads();
  const double* wT = w;
  const double* wq = w + n;
  // near pairs: wave wv takes points i0 + wv, i0 + wv + 4, ...; its lanes take
  // the point's pairs in strides of 64 and the wave sums them (fixed order)
  for (int t = wv; t < kFinRows; t += 4) {
    const int i = i0 + t;
    if (i >= npts) break;  // uniform over the wave
    double s1 = 0.0, s2 = 0.0;
    for (int pr = row_ptr[i] + lane; pr < row_ptr[i + 1]; pr += kWave) {
      const int c0 = pair_elem[pr] * nn;
      for (int k = 0; k < nn; ++k) {
        const double h = ph[(size_t)pr * nn + k];
        s1 += h * wT[c0 + k] + pg[(size_t)pr * nn + k] * wq[c0 + k];
        s2 += h;
      }
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) out[i] = (s_red1[0][t] + s1) / (-(s_red2[0][t] + s2));
  }
}

template <int NN>
static int interior_chunk(bemhip_ctx* c, const Geo& g, const Consts& cs, const QuadConst* d_qc,
                          int npg, const double* d_w, bool all_near, double* d_out,
                          hipEvent_t start,
                          hipEvent_t far_done, int64_t* npairs_out) {
  const int npts = g.nt;
  const int ncb = (g.ne + kEB - 1) / kEB;
  const int nrb = (npts + kFarThreads - 1) / kFarThreads;
  hipStream_t s1 = c->stream, s2 = c->stream2;
  // node boxes of the point blocks (the all-far tile test of the far field)
  // and the points' near-pair counts cleared
  hipLaunchKernelGGL((block_boxes<NN>), dim3((unsigned)nrb), dim3(64), 0, s1, g, cs, 0,
                     (double*)nullptr, c->ev_rowbox.ptr, c->ev_cnt.ptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipEventRecord(start, s1));
  // far field: FP64 contraction of the far pairs with (-T, q) and (1, 0), stream 2
  if (!all_near) {
    BEMHIP_HIP_TRY(c, hipStreamWaitEvent(s2, start, 0));
    double* part = c->ev_part.ptr;
    dim3 grid((unsigned)nrb, (unsigned)ncb);
    const int n = g.n;
    hipLaunchKernelGGL(op64_kernel<NN>, grid, dim3(kFarThreads), 0, s2, g, cs, 2, d_w, d_w + n,
                       d_w + 2 * (size_t)n, d_w + 3 * (size_t)n, part,
                       part + (size_t)ncb * npts, c->ev_box.ptr, c->ev_rowbox.ptr);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    BEMHIP_HIP_TRY(c, hipEventRecord(far_done, s2));
  }
  // near pairs: classification, CSR, near rule (stream 1)
  const int cpb = 32;
  dim3 cgrid((unsigned)nrb, (unsigned)((ncb + cpb - 1) / cpb));
  hipLaunchKernelGGL((near_classify<NN>), cgrid, dim3(kFarThreads), 0, s1, g, cs, ncb, cpb,
                     c->ev_box.ptr, c->ev_bits.ptr, c->ev_cnt.ptr, c->ev_tiles.ptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  hipLaunchKernelGGL(near_scan, dim3(1), dim3(kScanWaves * kWave), 0, s1, npts, c->ev_cnt.ptr,
                     c->ev_ptr.ptr, c->d_count);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(s1));
  const int total = *(volatile int*)c->h_count;
  const size_t cap = (size_t)std::max(total, 1);
  BEMHIP_HIP_TRY(c, c->ev_prow.ensure(cap));
  BEMHIP_HIP_TRY(c, c->ev_pelem.ensure(cap));
  BEMHIP_HIP_TRY(c, c->ev_ph.ensure(cap * NN));
  BEMHIP_HIP_TRY(c, c->ev_pg.ensure(cap * NN));
  hipLaunchKernelGGL(near_fill, dim3(nrb), dim3(kFarThreads), 0, s1, npts, ncb, NN, g.row0, g.n,
                     c->ev_bits.ptr, c->ev_tiles.ptr, c->ev_ptr.ptr, c->ev_prow.ptr,
                     c->ev_pelem.ptr, (int*)nullptr, (int*)nullptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  // no point is an element's node: every pair takes the tan-p3c rule, in CSR order
  const int st = launch_near_pairs<NN>(c, s1, g, d_qc, npg, 0, nullptr, total, nullptr,
                                       c->ev_prow.ptr, c->ev_pelem.ptr, c->ev_ph.ptr,
                                       c->ev_pg.ptr);
  if (st) return st;
  if (!all_near) BEMHIP_HIP_TRY(c, hipStreamWaitEvent(s1, far_done, 0));
  hipLaunchKernelGGL(interior_finalize, dim3((npts + kFinRows - 1) / kFinRows), dim3(256), 0, s1,
                     npts, g.n,
                     NN, ncb, all_near ? (const double*)nullptr : c->ev_part.ptr, c->ev_ptr.ptr,
                     c->ev_pelem.ptr, c->ev_ph.ptr, c->ev_pg.ptr, d_w, d_out);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  *npairs_out += total;
  return BEMHIP_OK;
}

// Points in chunks: the pair counts are 32-bit (at most npts x ne pairs per
// chunk) and an all-near chunk keeps its pair list below 2^26 pairs.
int interior_impl(bemhip_ctx* c, int npts, const double* d_px, const double* d_py,
                  const double* d_w, const QuadConst* d_qc, int npg, double near_factor,
                  double* d_out) {
  if (npts == 0) return BEMHIP_OK;
  Geo g = make_geo(c);
  g.row0 = c->nt;  // no point row is a node of an element
  Consts cs = make_consts(c);
  cs.nf = near_factor;
  const bool all_near = near_factor > 1e20;
  const int ne = c->ne;
  const int ncb = (ne + kEB - 1) / kEB;
  int64_t chunk = std::min<int64_t>(npts, 65536);
  chunk = std::min<int64_t>(chunk, std::max<int64_t>(1, (int64_t)INT32_MAX / ne - 1));
  if (all_near) chunk = std::min<int64_t>(chunk, std::max<int64_t>(1, ((int64_t)1 << 26) / ne));
  const int nrb = (int)((chunk + kFarThreads - 1) / kFarThreads);
  BEMHIP_HIP_TRY(c, c->ev_box.ensure((size_t)ncb * 5));
  BEMHIP_HIP_TRY(c, c->ev_bits.ensure((size_t)ncb * chunk));
  BEMHIP_HIP_TRY(c, c->ev_cnt.ensure((size_t)chunk));
  BEMHIP_HIP_TRY(c, c->ev_ptr.ensure((size_t)chunk + 1));
  BEMHIP_HIP_TRY(c, c->ev_tiles.ensure((size_t)nrb * ncb));
  BEMHIP_HIP_TRY(c, c->ev_rowbox.ensure((size_t)nrb * 4));
  if (!all_near) BEMHIP_HIP_TRY(c, c->ev_part.ensure(2 * (size_t)ncb * chunk));
  BEMHIP_HIP_TRY(c, ensure_count_word(c));
  if (c->nn == 2)
    hipLaunchKernelGGL((elem_block_boxes<2>), dim3(ncb), dim3(64), 0, c->stream, g,
                       cs, ncb, c->ev_box.ptr);
  else
    hipLaunchKernelGGL((elem_block_boxes<3>), dim3(ncb), dim3(64), 0, c->stream, g,
                       cs, ncb, c->ev_box.ptr);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  hipEvent_t ev[2];
  for (auto& e : ev) BEMHIP_HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct Destroy {
    hipEvent_t* e;
    ~Destroy() { hipEventDestroy(e[0]); hipEventDestroy(e[1]); }
  } destroy{ev};
  int64_t npairs = 0;
  for (int64_t p0 = 0; p0 < npts; p0 += chunk) {
    const int m = (int)std::min<int64_t>(chunk, npts - p0);
    Geo gc = g;
    gc.nt = m;
    gc.rx = d_px + p0;
    gc.ry = d_py + p0;
    const int st = (c->nn == 2)
                       ? interior_chunk<2>(c, gc, cs, d_qc, npg, d_w, all_near, d_out + p0,
                                           ev[0], ev[1], &npairs)
                       : interior_chunk<3>(c, gc, cs, d_qc, npg, d_w, all_near, d_out + p0,
                                           ev[0], ev[1], &npairs);
    if (st) return st;
  }
  c->stats["interior_pairs"] = (double)npairs;
  c->stats["interior_chunk"] = (double)chunk;
  return BEMHIP_OK;
}

}  // namespace bemhip
