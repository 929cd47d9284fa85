// Far field of the FP64 operator by a fast multipole method (BEMHIP_FMM).
//
// The far entries are the nodal lumping of src/Core/Assembly_full.jl:801-810:
// point monopoles (G) and dipoles (H) at the boundary nodes. With weights wH,
// wG per node j, the far part of row i is
//   phi(z_i) = sum_j wG_j jw_j U(z_i, zeta_j) + wH_j jw_j T(z_i, zeta_j)
//            = Re sum_j [ Q_j log(z_i - zeta_j) + D_j / (z_i - zeta_j) ],
//   Q_j = -wG_j jw_j / (2 pi k),  D_j = -(wH_j jw_j / (2 pi)) nu_j,
// nu_j the unit normal as a complex number (U = -log r / (2 pi k),
// T = (r . n) / (2 pi r^2), r = zeta_j - z_i). The complex-variable FMM of
// Greengard and Rokhlin (J. Comput. Phys. 73, 1987) evaluates it:
// multipole expansions a_0 log(z - c) + sum_k a_k / (z - c)^k of each box's
// sources, shifted up the quadtree (M2M), converted to local Taylor expansions
// of well-separated boxes of the same level (M2L, the interaction list: the
// children of the parent's neighbours that are not neighbours), shifted down
// (L2L), evaluated at the targets (L2P), plus the direct sum over the 3 x 3
// neighbour leaves (P2P). Coefficients are scaled by the box side (a_k / r^k,
// b_l r^l), so every power that appears is of a number of modulus <= 1/2
// (M2L: |r / z0| <= 1/2 for centres two sides apart) or <= sqrt(2)/4 (M2M,
// L2L); with p = 48 terms the truncation error is below (0.55)^49 ~ 2e-13 of
// the interacting mass.
//
// Near pairs: the leaf side is at least (near_factor + 1) x the longest
// element, so every node of an element that is near a target (distance to the
// element <= near_factor L_e) lies in the target's 3 x 3 leaf neighbourhood;
// P2P skips those sources (the elements of the row's near-pair list) and the
// node of the row itself; the near / on-element values and the rigid-body
// diagonal are added afterwards by the operator's finalize pass, as for the
// recomputed far field. The same neighbourhoods give each leaf its candidate
// elements for the near classification, so the FMM path needs no all-pairs
// classification. The result equals the far part of the dense FP64 operator to
// the FMM's truncation error.

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace bemhip {

namespace {

constexpr int kFmmP = 48;          // expansion terms (coefficients 0..kFmmP)
constexpr int kFmmLeafOcc = 32;    // target mean of boundary nodes per non-empty leaf
constexpr double kPi2 = 6.283185307179586;

struct cpx {
  double re, im;
};
__device__ __forceinline__ cpx cmk(double r, double i) { return {r, i}; }
__device__ __forceinline__ cpx cadd(cpx a, cpx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cpx cmul(cpx a, cpx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cpx cscal(cpx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cpx cinv(cpx a) {
  const double d = a.re * a.re + a.im * a.im;
  return {a.re / d, -a.im / d};
}
__device__ __forceinline__ cpx cpow(cpx a, int k) {  // a^k, k >= 0
  cpx r = {1.0, 0.0};
  while (k) {
    if (k & 1) r = cmul(r, a);
    a = cmul(a, a);
    k >>= 1;
  }
  return r;
}
__device__ __forceinline__ cpx ld2(const double* p) { return {p[0], p[1]}; }
__device__ __forceinline__ void st2(double* p, cpx v) { p[0] = v.re; p[1] = v.im; }

// Source strengths of node j (see the file comment).
__device__ __forceinline__ void strengths(const Geo& g, double inv2pik, const double* wH,
                                          const double* wG, int j, double& Q, cpx& D) {
  const double w = g.jw[j];
  Q = -wG[j] * w * inv2pik;
  const double s = -wH[j] * w * (1.0 / kPi2);
  D = {s * g.nx[j], s * g.ny[j]};
}

// P2M: one wave per leaf, lane k = coefficient k of the scaled multipole
// a_k / r^k = sum_j [-Q_j t_j^k / k + D_j t_j^(k-1) / r], t_j = (zeta_j - c) / r.
__global__ __launch_bounds__(64) void p2m_kernel(Geo g, double inv2pik, const double* __restrict__ wH,
                                                  const double* __restrict__ wG, int leaf0, double r,
                                                  const double* __restrict__ center,
                                                  const int* __restrict__ src_lo,
                                                  const int* __restrict__ src_hi,
                                                  const int* __restrict__ src_order,
                                                  double* __restrict__ mpole) {
  const int leaf = blockIdx.x, b = leaf0 + leaf, k = threadIdx.x;
  if (k > kFmmP) return;
  const cpx c = ld2(center + 2 * (size_t)b);
  cpx acc = {0.0, 0.0};
  for (int q = src_lo[leaf]; q < src_hi[leaf]; ++q) {
    const int j = src_order[q];
    double Q;
    cpx D;
    strengths(g, inv2pik, wH, wG, j, Q, D);
    if (k == 0) {
      acc.re += Q;
    } else {
      const cpx t = cmk((g.x[j] - c.re) / r, (g.y[j] - c.im) / r);
      const cpx tk1 = cpow(t, k - 1);
      acc = cadd(acc, cadd(cscal(cmul(tk1, t), -Q / k), cscal(cmul(D, tk1), 1.0 / r)));
    }
  }
  st2(mpole + 2 * ((size_t)b * (kFmmP + 1) + k), acc);
}

// M2M: parent (side rp) from its children (side rp / 2), u = (c_child - c_parent) / rp:
// b_l = sum_children [-a_0 u^l / l + sum_{k=1..l} a_k 2^-k u^(l-k) C(l-1, k-1)], b_0 = sum a_0.
__global__ __launch_bounds__(64) void m2m_kernel(int box0, double rp, const double* __restrict__ center,
                                                  const int* __restrict__ child_lo,
                                                  const int* __restrict__ child_hi,
                                                  const double* __restrict__ binom,
                                                  double* __restrict__ mpole) {
  constexpr int NB = 2 * kFmmP + 1;
  const int b = box0 + blockIdx.x, l = threadIdx.x;
  if (l > kFmmP) return;
  const cpx cp = ld2(center + 2 * (size_t)b);
  cpx acc = {0.0, 0.0};
  for (int ch = child_lo[b]; ch < child_hi[b]; ++ch) {
    const double* a = mpole + 2 * (size_t)ch * (kFmmP + 1);
    const cpx a0 = ld2(a);
    if (l == 0) { acc = cadd(acc, a0); continue; }
    const cpx cc = ld2(center + 2 * (size_t)ch);
    const cpx u = cmk((cc.re - cp.re) / rp, (cc.im - cp.im) / rp);
    cpx s = cscal(cmul(a0, cpow(u, l)), -1.0 / l);
    // sum_{k=1..l} a_k 2^-k u^(l-k) C(l-1, k-1): u^(l-k) from u^(l-1) down
    cpx upow = {1.0, 0.0};  // u^(l-k) for k = l
    double half = 1.0;
    for (int k = 1; k <= l; ++k) half *= 0.5;  // 2^-l
    for (int k = l; k >= 1; --k) {
      const cpx ak = ld2(a + 2 * k);
      s = cadd(s, cscal(cmul(ak, upow), half * binom[(l - 1) * NB + (k - 1)]));
      upow = cmul(upow, u);
      half *= 2.0;
    }
    acc = cadd(acc, s);
  }
  st2(mpole + 2 * ((size_t)b * (kFmmP + 1) + l), acc);
}

// M2L at one level (side r): for each source box s of the interaction list,
// rho = r / (c_s - c_b):
//   b_0 += a_0 log(-(c_s - c_b)) + sum_k a_k rho^k (-1)^k
//   b_l += -a_0 rho^l / l + sum_k a_k rho^(k+l) C(l+k-1, k-1) (-1)^k   (l >= 1)
__global__ __launch_bounds__(64) void m2l_kernel(int box0, double r, const double* __restrict__ center,
                                                  const int* __restrict__ il_ptr,
                                                  const int* __restrict__ il_idx,
                                                  const double* __restrict__ binom,
                                                  const double* __restrict__ mpole,
                                                  double* __restrict__ local) {
  constexpr int NB = 2 * kFmmP + 1;
  __shared__ double s_a[2 * (kFmmP + 1)];
  const int b = box0 + blockIdx.x, l = threadIdx.x;
  const cpx cb = ld2(center + 2 * (size_t)b);
  cpx acc = {0.0, 0.0};
  for (int q = il_ptr[b]; q < il_ptr[b + 1]; ++q) {
    const int sb = il_idx[q];
    __syncthreads();
    if (l <= kFmmP) {
      s_a[2 * l] = mpole[2 * ((size_t)sb * (kFmmP + 1) + l)];
      s_a[2 * l + 1] = mpole[2 * ((size_t)sb * (kFmmP + 1) + l) + 1];
    }
    __syncthreads();
    if (l > kFmmP) continue;
    const cpx cs = ld2(center + 2 * (size_t)sb);
    const cpx z0 = cmk(cs.re - cb.re, cs.im - cb.im);
    const cpx rho = cscal(cinv(z0), r);
    const cpx a0 = cmk(s_a[0], s_a[1]);
    cpx s;
    cpx rp = (l == 0) ? cmk(1.0, 0.0) : cpow(rho, l);  // rho^l
    if (l == 0) {
      const double mod = sqrt(z0.re * z0.re + z0.im * z0.im);
      const cpx lg = cmk(log(mod), atan2(-z0.im, -z0.re));  // log(-z0)
      s = cmul(a0, lg);
    } else {
      s = cscal(cmul(a0, rp), -1.0 / l);
    }
    cpx pw = cmul(rp, rho);  // rho^(k+l), k = 1
    double sg = -1.0;
    for (int k = 1; k <= kFmmP; ++k) {
      const cpx ak = cmk(s_a[2 * k], s_a[2 * k + 1]);
      const double cf = (l == 0) ? sg : sg * binom[(l + k - 1) * NB + (k - 1)];
      s = cadd(s, cscal(cmul(ak, pw), cf));
      pw = cmul(pw, rho);
      sg = -sg;
    }
    acc = cadd(acc, s);
  }
  if (l <= kFmmP) {
    double* o = local + 2 * ((size_t)b * (kFmmP + 1) + l);
    o[0] += acc.re;
    o[1] += acc.im;
  }
}

// L2L: child (side rc = rp / 2) from its parent, u = (c_child - c_parent) / rp:
// b'_m += 2^-m sum_{l >= m} b_l C(l, m) u^(l - m).
__global__ __launch_bounds__(64) void l2l_kernel(int box0, double rp, const double* __restrict__ center,
                                                  const int* __restrict__ parent,
                                                  const double* __restrict__ binom,
                                                  double* __restrict__ local) {
  constexpr int NB = 2 * kFmmP + 1;
  const int b = box0 + blockIdx.x, m = threadIdx.x;
  if (m > kFmmP) return;
  const int pb = parent[b];
  const cpx cc = ld2(center + 2 * (size_t)b), cp = ld2(center + 2 * (size_t)pb);
  const cpx u = cmk((cc.re - cp.re) / rp, (cc.im - cp.im) / rp);
  const double* bl = local + 2 * (size_t)pb * (kFmmP + 1);
  cpx s = {0.0, 0.0};
  cpx upow = {1.0, 0.0};
  for (int l = m; l <= kFmmP; ++l) {
    s = cadd(s, cscal(cmul(ld2(bl + 2 * l), upow), binom[l * NB + m]));
    upow = cmul(upow, u);
  }
  double sc = 1.0;
  for (int q = 0; q < m; ++q) sc *= 0.5;
  double* o = local + 2 * ((size_t)b * (kFmmP + 1) + m);
  o[0] += s.re * sc;
  o[1] += s.im * sc;
}

// L2P + P2P: one wave per leaf, lanes over its targets. out[i] = Re of the
// local expansion at z_i plus the direct sum over the sources of the 3 x 3
// neighbour leaves, without the row's own node and the sources of the
// elements in the row's near-pair list.
__global__ __launch_bounds__(64) void l2p_p2p_kernel(
    Geo g, double inv2pik, const double* __restrict__ wH, const double* __restrict__ wG, int leaf0,
    double r, const double* __restrict__ center, const double* __restrict__ local,
    const int* __restrict__ tgt_lo, const int* __restrict__ tgt_hi,
    const int* __restrict__ tgt_order, const int* __restrict__ src_lo,
    const int* __restrict__ src_hi, const int* __restrict__ src_order,
    const int* __restrict__ nb_ptr, const int* __restrict__ nb_idx,
    const int* __restrict__ row_ptr, const int* __restrict__ pair_elem, int nn,
    double* __restrict__ out) {
  __shared__ double s_b[2 * (kFmmP + 1)];
  const int leaf = blockIdx.x, b = leaf0 + leaf, lane = threadIdx.x;
  if (lane <= kFmmP) {
    s_b[2 * lane] = local[2 * ((size_t)b * (kFmmP + 1) + lane)];
    s_b[2 * lane + 1] = local[2 * ((size_t)b * (kFmmP + 1) + lane) + 1];
  }
  __syncthreads();
  const cpx c = ld2(center + 2 * (size_t)b);
  for (int q = tgt_lo[leaf] + lane; q < tgt_hi[leaf]; q += 64) {
    const int i = tgt_order[q];
    const double zx = g.x[i], zy = g.y[i];
    // local expansion (Horner in t = (z - c) / r)
    const cpx t = cmk((zx - c.re) / r, (zy - c.im) / r);
    cpx f = cmk(s_b[2 * kFmmP], s_b[2 * kFmmP + 1]);
    for (int l = kFmmP - 1; l >= 0; --l) f = cadd(cmul(f, t), cmk(s_b[2 * l], s_b[2 * l + 1]));
    double phi = f.re;
    const int p0 = row_ptr[i], p1 = row_ptr[i + 1];  // the row's near elements
    for (int nbq = nb_ptr[leaf]; nbq < nb_ptr[leaf + 1]; ++nbq) {
      const int sl = nb_idx[nbq];
      for (int sq = src_lo[sl]; sq < src_hi[sl]; ++sq) {
        const int j = src_order[sq];
        if (j == i) continue;
        const int e = j / nn;
        bool nr = false;
        for (int pr = p0; pr < p1; ++pr) nr |= pair_elem[pr] == e;
        if (nr) continue;
        double Q;
        cpx D;
        strengths(g, inv2pik, wH, wG, j, Q, D);
        const double rx = zx - g.x[j], ry = zy - g.y[j];
        const double d2 = rx * rx + ry * ry;
        // Q log|z - zeta| + Re(D / (z - zeta))
        phi += 0.5 * Q * log(d2) + (D.re * rx + D.im * ry) / d2;
      }
    }
    out[i] = phi;
  }
}

inline uint64_t spread2(uint32_t v) {  // bits of v at the even positions
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
inline uint32_t compact2(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return (uint32_t)x;
}
inline uint64_t morton(uint32_t ix, uint32_t iy) { return spread2(ix) | (spread2(iy) << 1); }

template <typename T>
int upload(bemhip_ctx* c, DevBuf<T>& d, const std::vector<T>& h) {
  BEMHIP_HIP_TRY(c, d.ensure(std::max<size_t>(h.size(), 1)));
  if (!h.empty())
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(d.ptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice,
                                     c->stream));
  return BEMHIP_OK;
}

}  // namespace

// Quadtree plan from the collocation points (targets) and boundary nodes
// (sources). Fails with BEMHIP_ERR_ARG when the boundary is too coarse for a
// tree of at least two levels below the root.
int fmm_build(bemhip_ctx* c) {
  FmmPlan& P = c->fmm;
  P.built = false;
  const int nt = c->nt, n = c->n, ne = c->ne;
  std::vector<double> x(nt), y(nt), elen(ne);
  BEMHIP_HIP_TRY(c, hipMemcpy(x.data(), c->x.ptr, sizeof(double) * nt, hipMemcpyDeviceToHost));
  BEMHIP_HIP_TRY(c, hipMemcpy(y.data(), c->y.ptr, sizeof(double) * nt, hipMemcpyDeviceToHost));
  BEMHIP_HIP_TRY(c, hipMemcpy(elen.data(), c->elen.ptr, sizeof(double) * ne, hipMemcpyDeviceToHost));
  if (!(c->near_factor <= 1e20))
    return fail(c, BEMHIP_ERR_ARG, "assemble: the FMM needs a finite near_factor");
  double xmin = x[0], xmax = x[0], ymin = y[0], ymax = y[0];
  for (int i = 1; i < nt; ++i) {
    xmin = std::min(xmin, x[i]); xmax = std::max(xmax, x[i]);
    ymin = std::min(ymin, y[i]); ymax = std::max(ymax, y[i]);
  }
  const double ext = std::max({xmax - xmin, ymax - ymin, 1e-300});
  P.S = ext * (1.0 + 1e-9);
  P.x0 = xmin - 0.5e-9 * ext;
  P.y0 = ymin - 0.5e-9 * ext;
  const double maxL = *std::max_element(elen.begin(), elen.end());
  const double hmin = (c->near_factor + 1.0) * maxL * 1.05;
  if (nt >= (1 << 24)) return fail(c, BEMHIP_ERR_ARG, "assemble: the FMM takes fewer than 2^24 rows");
  int Lmax = 0;  // at most 20 levels: 40-bit codes beside 24-bit indices in one sort key
  while (Lmax < 20 && P.S / std::ldexp(1.0, Lmax + 1) >= hmin) ++Lmax;
  if (Lmax < 2)
    return fail(c, BEMHIP_ERR_ARG,
                "assemble: boundary too coarse for the FMM (use BEMHIP_MATRIX_FREE)");
  // Morton codes at the finest admissible level, one sort of (code, index)
  // keys; coarser levels are the codes shifted right by two bits per level
  const int Lm = Lmax;
  std::vector<uint64_t> cmax(nt);
  {
    const double h = P.S / std::ldexp(1.0, Lm);
    const double mx = std::ldexp(1.0, Lm) - 1.0;
    for (int i = 0; i < nt; ++i) {
      const double fx = std::min(mx, std::max(0.0, std::floor((x[i] - P.x0) / h)));
      const double fy = std::min(mx, std::max(0.0, std::floor((y[i] - P.y0) / h)));
      cmax[i] = morton((uint32_t)fx, (uint32_t)fy);
    }
  }
  std::vector<uint64_t> key(nt);  // code at Lm in the high bits, index in the low 24
  for (int i = 0; i < nt; ++i) key[i] = (cmax[i] << 24) | (uint64_t)i;
  std::sort(key.begin(), key.end());
  std::vector<int> tgt(nt), src;
  src.reserve(n);
  for (int q = 0; q < nt; ++q) {
    tgt[q] = (int)(key[q] & 0xffffffull);
    if (tgt[q] < n) src.push_back(tgt[q]);
  }
  // depth: the finest level that keeps the near zone inside the neighbours,
  // stopping once the mean occupancy of the source leaves drops to kFmmLeafOcc
  int L = 2;
  for (; L < Lm; ++L) {
    const int sh = 2 * (Lm - L);
    size_t nu = 0;
    uint64_t prev = ~0ull;
    for (int j : src) {
      const uint64_t cd = cmax[j] >> sh;
      if (cd != prev) { ++nu; prev = cd; }
    }
    if ((double)n / (double)nu <= kFmmLeafOcc) break;
  }
  P.L = L;
  P.p = kFmmP;
  std::vector<uint64_t> tcode(nt);
  for (int i = 0; i < nt; ++i) tcode[i] = cmax[i] >> (2 * (Lm - L));
  // boxes per level (codes sorted), from the leaves up
  std::vector<std::vector<uint64_t>> codes(L + 1);
  codes[L] = tcode;  // targets include the boundary nodes (sources)
  std::sort(codes[L].begin(), codes[L].end());
  codes[L].erase(std::unique(codes[L].begin(), codes[L].end()), codes[L].end());
  for (int l = L - 1; l >= 2; --l) {
    codes[l].resize(codes[l + 1].size());
    for (size_t q = 0; q < codes[l + 1].size(); ++q) codes[l][q] = codes[l + 1][q] >> 2;
    codes[l].erase(std::unique(codes[l].begin(), codes[l].end()), codes[l].end());
  }
  P.lvl_off.assign(L + 2, 0);
  P.lvl_n.assign(L + 1, 0);
  int tot = 0;
  for (int l = 2; l <= L; ++l) {
    P.lvl_off[l] = tot;
    P.lvl_n[l] = (int)codes[l].size();
    tot += P.lvl_n[l];
  }
  P.lvl_off[L + 1] = tot;
  P.nbox = tot;
  P.nleaf = P.lvl_n[L];
  auto find = [&](int l, uint64_t code) -> int {  // global index or -1
    const auto& v = codes[l];
    auto it = std::lower_bound(v.begin(), v.end(), code);
    return (it != v.end() && *it == code) ? P.lvl_off[l] + (int)(it - v.begin()) : -1;
  };
  // leaves: source / target ranges
  const int nl = P.nleaf;
  std::vector<int> slo(nl), shi(nl), tlo(nl), thi(nl);
  {
    size_t qs = 0, qt = 0;
    for (int lf = 0; lf < nl; ++lf) {
      const uint64_t cd = codes[L][lf];
      slo[lf] = (int)qs;
      while (qs < src.size() && tcode[src[qs]] == cd) ++qs;
      shi[lf] = (int)qs;
      tlo[lf] = (int)qt;
      while (qt < tgt.size() && tcode[tgt[qt]] == cd) ++qt;
      thi[lf] = (int)qt;
    }
  }
  // has sources, per box
  std::vector<char> has_src(tot, 0);
  for (int lf = 0; lf < nl; ++lf) has_src[P.lvl_off[L] + lf] = shi[lf] > slo[lf];
  std::vector<int> clo(tot, 0), chi(tot, 0), par(tot, -1);
  std::vector<double> cen(2 * (size_t)tot);
  for (int l = L; l >= 2; --l) {
    const double h = P.S / std::ldexp(1.0, l);
    for (int q = 0; q < P.lvl_n[l]; ++q) {
      const int b = P.lvl_off[l] + q;
      const uint64_t cd = codes[l][q];
      cen[2 * b] = P.x0 + (compact2(cd) + 0.5) * h;
      cen[2 * b + 1] = P.y0 + (compact2(cd >> 1) + 0.5) * h;
      if (l > 2) par[b] = find(l - 1, cd >> 2);
      if (l < L) {
        const auto& v = codes[l + 1];
        const auto lo = std::lower_bound(v.begin(), v.end(), cd << 2);
        const auto hi = std::lower_bound(v.begin(), v.end(), (cd + 1) << 2);
        clo[b] = P.lvl_off[l + 1] + (int)(lo - v.begin());
        chi[b] = P.lvl_off[l + 1] + (int)(hi - v.begin());
        for (int ch = clo[b]; ch < chi[b]; ++ch) has_src[b] |= has_src[ch];
      }
    }
  }
  // interaction lists
  std::vector<int> ilp(tot + 1, 0), ili;
  for (int l = 2; l <= L; ++l) {
    for (int q = 0; q < P.lvl_n[l]; ++q) {
      const int b = P.lvl_off[l] + q;
      const uint64_t cd = codes[l][q];
      const int ix = (int)compact2(cd), iy = (int)compact2(cd >> 1);
      const int px = ix >> 1, py = iy >> 1, np = 1 << (l - 1);
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int qx = px + dx, qy = py + dy;
          if (qx < 0 || qy < 0 || qx >= np || qy >= np) continue;
          for (int cy = 0; cy < 2; ++cy)
            for (int cx = 0; cx < 2; ++cx) {
              const int sx = 2 * qx + cx, sy = 2 * qy + cy;
              if (std::abs(sx - ix) <= 1 && std::abs(sy - iy) <= 1) continue;
              const int sb = find(l, morton((uint32_t)sx, (uint32_t)sy));
              if (sb >= 0 && has_src[sb]) ili.push_back(sb);
            }
        }
      ilp[b + 1] = (int)ili.size();
    }
  }
  // neighbour leaves with sources
  std::vector<int> nbp(nl + 1, 0), nbi;
  const int nside = 1 << L;
  for (int lf = 0; lf < nl; ++lf) {
    const uint64_t cd = codes[L][lf];
    const int ix = (int)compact2(cd), iy = (int)compact2(cd >> 1);
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int sx = ix + dx, sy = iy + dy;
        if (sx < 0 || sy < 0 || sx >= nside || sy >= nside) continue;
        const int sb = find(L, morton((uint32_t)sx, (uint32_t)sy));
        if (sb >= 0 && has_src[sb]) nbi.push_back(sb - P.lvl_off[L]);
      }
    nbp[lf + 1] = (int)nbi.size();
  }
  // candidate elements per leaf: elements with a node in the 3 x 3 neighbourhood
  // (every element near a target of the leaf is one of them)
  std::vector<int> cdp(nl + 1, 0), cdi;
  {
    std::vector<int> tmp;
    for (int lf = 0; lf < nl; ++lf) {
      tmp.clear();
      for (int q = nbp[lf]; q < nbp[lf + 1]; ++q) {
        const int sl = nbi[q];
        for (int k = slo[sl]; k < shi[sl]; ++k) tmp.push_back(src[k] / c->nn);
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      cdi.insert(cdi.end(), tmp.begin(), tmp.end());
      cdp[lf + 1] = (int)cdi.size();
    }
  }
  // binomials C(a, b), a, b <= 2p
  constexpr int NB = 2 * kFmmP + 1;
  std::vector<double> bin((size_t)NB * NB, 0.0);
  for (int a = 0; a < NB; ++a) {
    bin[(size_t)a * NB] = 1.0;
    for (int b = 1; b <= a; ++b)
      bin[(size_t)a * NB + b] = bin[(size_t)(a - 1) * NB + b - 1] +
                                (b <= a - 1 ? bin[(size_t)(a - 1) * NB + b] : 0.0);
  }
  int st;
  if ((st = upload(c, P.center, cen)) || (st = upload(c, P.child_lo, clo)) ||
      (st = upload(c, P.child_hi, chi)) || (st = upload(c, P.parent, par)) ||
      (st = upload(c, P.il_ptr, ilp)) || (st = upload(c, P.il_idx, ili)) ||
      (st = upload(c, P.src_lo, slo)) || (st = upload(c, P.src_hi, shi)) ||
      (st = upload(c, P.tgt_lo, tlo)) || (st = upload(c, P.tgt_hi, thi)) ||
      (st = upload(c, P.src_order, src)) || (st = upload(c, P.tgt_order, tgt)) ||
      (st = upload(c, P.nb_ptr, nbp)) || (st = upload(c, P.nb_idx, nbi)) ||
      (st = upload(c, P.cand_ptr, cdp)) || (st = upload(c, P.cand_idx, cdi)) ||
      (st = upload(c, P.binom, bin)))
    return st;
  BEMHIP_HIP_TRY(c, P.mpole.ensure(2 * (size_t)tot * (kFmmP + 1)));
  BEMHIP_HIP_TRY(c, P.local.ensure(2 * (size_t)tot * (kFmmP + 1)));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->stats["fmm_levels"] = L;
  c->stats["fmm_boxes"] = tot;
  c->stats["fmm_leaves"] = nl;
  c->stats["fmm_m2l_pairs"] = (double)ili.size();
  c->stats["fmm_candidates"] = (double)cdi.size();
  P.built = true;
  return BEMHIP_OK;
}

Geo make_geo(bemhip_ctx* c);

// out[i] = far part of row i (all rows) for node weights wH, wG, on stream s.
int fmm_far(bemhip_ctx* c, hipStream_t s, const double* wH, const double* wG, double* out) {
  FmmPlan& P = c->fmm;
  if (!P.built) return fail(c, BEMHIP_ERR_STATE, "fmm: no plan (assemble with BEMHIP_FMM)");
  const Geo g = make_geo(c);
  const double inv2pik = 1.0 / (kPi2 * c->k);
  const int L = P.L;
  const size_t ncoef = 2 * (size_t)P.nbox * (kFmmP + 1);
  BEMHIP_HIP_TRY(c, hipMemsetAsync(P.mpole.ptr, 0, sizeof(double) * ncoef, s));
  BEMHIP_HIP_TRY(c, hipMemsetAsync(P.local.ptr, 0, sizeof(double) * ncoef, s));
  auto side = [&](int l) { return P.S / std::ldexp(1.0, l); };
  hipLaunchKernelGGL(p2m_kernel, dim3(P.nleaf), dim3(64), 0, s, g, inv2pik, wH, wG, P.lvl_off[L],
                     side(L), P.center.ptr, P.src_lo.ptr, P.src_hi.ptr, P.src_order.ptr,
                     P.mpole.ptr);
  for (int l = L - 1; l >= 2; --l)
    hipLaunchKernelGGL(m2m_kernel, dim3(P.lvl_n[l]), dim3(64), 0, s, P.lvl_off[l], side(l),
                       P.center.ptr, P.child_lo.ptr, P.child_hi.ptr, P.binom.ptr, P.mpole.ptr);
  for (int l = 2; l <= L; ++l) {
    if (l > 2)
      hipLaunchKernelGGL(l2l_kernel, dim3(P.lvl_n[l]), dim3(64), 0, s, P.lvl_off[l], side(l - 1),
                         P.center.ptr, P.parent.ptr, P.binom.ptr, P.local.ptr);
    hipLaunchKernelGGL(m2l_kernel, dim3(P.lvl_n[l]), dim3(64), 0, s, P.lvl_off[l], side(l),
                       P.center.ptr, P.il_ptr.ptr, P.il_idx.ptr, P.binom.ptr, P.mpole.ptr,
                       P.local.ptr);
  }
  hipLaunchKernelGGL(l2p_p2p_kernel, dim3(P.nleaf), dim3(64), 0, s, g, inv2pik, wH, wG,
                     P.lvl_off[L], side(L), P.center.ptr, P.local.ptr, P.tgt_lo.ptr, P.tgt_hi.ptr,
                     P.tgt_order.ptr, P.src_lo.ptr, P.src_hi.ptr, P.src_order.ptr, P.nb_ptr.ptr,
                     P.nb_idx.ptr, c->row_ptr.ptr, c->pair_elem.ptr, c->nn, out);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

}  // namespace bemhip
