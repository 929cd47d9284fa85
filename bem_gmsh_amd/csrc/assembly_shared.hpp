// Device code shared by the Laplace and Kelvin (2-D elasticity) assemblies:
// the near/far pair classification (src/Core/Assembly_full.jl:736-754), the
// near-pair CSR build, the row sums of column-block partials and the entry
// points of the near-singular rule machinery (complex pole, tan-p3c and Mobius
// maps, barycentric bases; src/Core/Integration.jl:68-93, :1117-1365,
// :1597-1845), which near_rule.hpp holds as templates shared with the adjoint.
// Everything is in an anonymous namespace: each translation unit that includes
// this header gets its own copy.
//
// Translation units including this header are compiled with -ffp-contract=off.

#pragma once

#include "common.hpp"
#include "near_rule.hpp"

#include <cmath>

namespace bemhip {
namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kEps = 2.220446049250313e-16;

// ---------------------------------------------------------------------------
// Classification of one (row, element) pair. rx/ry: node - source; d2 per node.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dist_chord(double px, double py, double ax, double ay,
                                             double bx, double by) {
  double abx = bx - ax, aby = by - ay;
  double t = ((px - ax) * abx + (py - ay) * aby) / ((abx * abx + aby * aby) + kEps);
  t = t > 1.0 ? 1.0 : (t < 0.0 ? 0.0 : t);
  double qx = ax + t * abx, qy = ay + t * aby;
  double dx = px - qx, dy = py - qy;
  return sqrt(dx * dx + dy * dy);
}

template <int NN>
__device__ __forceinline__ bool is_near(double px, double py, const double* cx,
                                        const double* cy, const double* d2, double L,
                                        double nf) {
  if (nf > 1e20) return true;  // includes Inf
  double lim = nf * L;
  double lim2 = lim * lim;
  double dn2 = INFINITY;
  double lox = cx[0], loy = cy[0], hix = cx[0], hiy = cy[0];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    lox = fmin(lox, cx[k]); loy = fmin(loy, cy[k]);
    hix = fmax(hix, cx[k]); hiy = fmax(hiy, cy[k]);
    if (d2[k] < dn2) dn2 = d2[k];
  }
  if (dn2 < lim2) return true;
  double s = 0.0;
  if (px < lox) { double d = lox - px; s += d * d; }
  else if (px > hix) { double d = px - hix; s += d * d; }
  if (py < loy) { double d = loy - py; s += d * d; }
  else if (py > hiy) { double d = py - hiy; s += d * d; }
  if (sqrt(s) >= lim) return false;
  double d = INFINITY;
#pragma unroll
  for (int k = 0; k + 1 < NN; ++k)
    d = fmin(d, dist_chord(px, py, cx[k], cy[k], cx[k + 1], cy[k + 1]));
  return d < lim;
}

struct Consts {
  double cU;  // 0.5/(2 pi k)
  double cT;  // 1/(2 pi)
  double nf;  // near factor
};

// Conservative FP32 pre-test: true only when the pair is far by a relative
// margin (lim2m = lim^2 (1 + 1e-3)) on both the node-distance and the
// bounding-box tests, so that the exact FP64 test (is_near) would also call
// it far. r*f are the FP32 roundings of the FP64 node - source offsets.
template <int NN>
__device__ __forceinline__ bool clearly_far(const float* rxf, const float* ryf, const float* d2f,
                                            float lim2m) {
  float dn2 = d2f[0], lox = rxf[0], hix = rxf[0], loy = ryf[0], hiy = ryf[0];
#pragma unroll
  for (int k = 1; k < NN; ++k) {
    dn2 = fminf(dn2, d2f[k]);
    lox = fminf(lox, rxf[k]); hix = fmaxf(hix, rxf[k]);
    loy = fminf(loy, ryf[k]); hiy = fmaxf(hiy, ryf[k]);
  }
  const float dx = lox > 0.f ? lox : (hix < 0.f ? -hix : 0.f);
  const float dy = loy > 0.f ? loy : (hiy < 0.f ? -hiy : 0.f);
  return dn2 > lim2m && (dx * dx + dy * dy) > lim2m;
}

// 1/x in FP64 from the FP32 reciprocal and two Newton steps (about 1 ulp).
__device__ __forceinline__ double rcp_f64(double x) {
  double r = (double)__builtin_amdgcn_rcpf((float)x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}

// Classification of pair (row at (px, py), element with collocation nodes
// cx/cy): the node - source offsets (FP64 and their FP32 roundings) and the
// near decision: on-element pairs are near; otherwise the FP32 pre-test settles
// the clearly-far pairs and the rest take the exact FP64 test. Shared by
// near_classify and far_kernel so that both make the same decision.
template <int NN>
__device__ __forceinline__ bool classify_pair(double px, double py, const double* cx,
                                              const double* cy, double L, float lim2f,
                                              double nf, bool onel, double* rx, double* ry,
                                              float* rxf, float* ryf, float* d2f) {
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    rx[k] = cx[k] - px;
    ry[k] = cy[k] - py;
    rxf[k] = (float)rx[k];
    ryf[k] = (float)ry[k];
    d2f[k] = rxf[k] * rxf[k] + ryf[k] * ryf[k];
  }
  if (onel) return true;
  if (clearly_far<NN>(rxf, ryf, d2f, lim2f)) return false;
  double d2[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) d2[k] = rx[k] * rx[k] + ry[k] * ry[k];
  return is_near<NN>(px, py, cx, cy, d2, L, nf);
}

// Element data of one 32-element column block in LDS for the classification.
template <int NN>
__device__ __forceinline__ void stage_class_block(const Geo& g, const Consts& cs, int e0,
                                                  int ecount, double* s_cx, double* s_cy,
                                                  double* s_L, float* s_lim2f) {
  for (int t = threadIdx.x; t < ecount * NN; t += blockDim.x) {
    const int j = e0 * NN + t;
    s_cx[t] = g.x[j];
    s_cy[t] = g.y[j];
  }
  for (int t = threadIdx.x; t < ecount; t += blockDim.x) {
    const double L = g.elen[e0 + t];
    s_L[t] = L;
    const double lim = cs.nf * L;
    // +inf (no far lumping) when the cutoff is disabled or overflows
    s_lim2f[t] = (cs.nf > 1e20) ? INFINITY : (float)((lim * lim) * (1.0 + 1e-3));
  }
}

// ---------------------------------------------------------------------------
// Bounding box of the collocation nodes of each 32-element column block and
// its largest cutoff nf * L: box[cb] = {lox, loy, hix, hiy, maxlim}.
// ---------------------------------------------------------------------------
// One wavefront per column block: lanes over its nodes and elements, then lane
// reductions (min / max are exact, so the order does not matter).
template <int NN>
__device__ __forceinline__ void elem_box_wave(const Geo& g, const Consts& cs, int cb,
                                              double* __restrict__ box) {
  const int e0 = cb * kEB, e1 = min(g.ne, e0 + kEB);
  double lox = INFINITY, loy = INFINITY, hix = -INFINITY, hiy = -INFINITY, ml = 0.0;
  for (int t = e0 * NN + threadIdx.x; t < e1 * NN; t += 64) {
    const double x = g.x[t], y = g.y[t];
    lox = fmin(lox, x); hix = fmax(hix, x);
    loy = fmin(loy, y); hiy = fmax(hiy, y);
  }
  for (int e = e0 + threadIdx.x; e < e1; e += 64) ml = fmax(ml, cs.nf * g.elen[e]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lox = fmin(lox, __shfl_xor(lox, off, 64));
    hix = fmax(hix, __shfl_xor(hix, off, 64));
    loy = fmin(loy, __shfl_xor(loy, off, 64));
    hiy = fmax(hiy, __shfl_xor(hiy, off, 64));
    ml = fmax(ml, __shfl_xor(ml, off, 64));
  }
  if (threadIdx.x == 0) {
    double* b = box + (size_t)cb * 5;
    b[0] = lox; b[1] = loy; b[2] = hix; b[3] = hiy; b[4] = ml;
  }
}

template <int NN>
__global__ __launch_bounds__(64) void elem_block_boxes(Geo g, Consts cs, int ncb,
                                                       double* __restrict__ box) {
  if ((int)blockIdx.x < ncb) elem_box_wave<NN>(g, cs, blockIdx.x, box);
}

// Node bounding boxes of the column blocks (cbox, as elem_block_boxes) and of
// the 256-row blocks (rbox[rb] = {lox, hix, loy, hiy}), one wavefront per
// block in one launch: workgroups [0, ncb) take the column blocks, the rest
// the row blocks, clearing the rows' near-pair counts on the way.
template <int NN>
__global__ __launch_bounds__(64) void block_boxes(Geo g, Consts cs, int ncb,
                                                  double* __restrict__ cbox,
                                                  double* __restrict__ rbox,
                                                  int* __restrict__ cnt) {
  if ((int)blockIdx.x < ncb) {
    elem_box_wave<NN>(g, cs, blockIdx.x, cbox);
    return;
  }
  const int rb = blockIdx.x - ncb;
  const int r0 = rb * kFarThreads;
  double lox = INFINITY, hix = -INFINITY, loy = INFINITY, hiy = -INFINITY;
  for (int r = r0 + threadIdx.x; r < min(g.nt, r0 + kFarThreads); r += 64) {
    cnt[r] = 0;
    const double x = g.rx[r], y = g.ry[r];
    lox = fmin(lox, x); hix = fmax(hix, x);
    loy = fmin(loy, y); hiy = fmax(hiy, y);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lox = fmin(lox, __shfl_xor(lox, off, 64));
    hix = fmax(hix, __shfl_xor(hix, off, 64));
    loy = fmin(loy, __shfl_xor(loy, off, 64));
    hiy = fmax(hiy, __shfl_xor(hiy, off, 64));
  }
  if (threadIdx.x == 0) {
    double* b = rbox + (size_t)rb * 4;
    b[0] = lox; b[1] = hix; b[2] = loy; b[3] = hiy;
  }
}

// True when every pair of the tile (row block rb, column block cb) is far:
// the row block's node box lies farther than (1 + 1e-6) x the column block's
// largest cutoff from its node box (uniform over a workgroup).
__device__ __forceinline__ bool tile_all_far(const double* __restrict__ cbox,
                                             const double* __restrict__ rbox, int cb, int rb,
                                             double nf) {
  if (!cbox || !rbox || nf > 1e20) return false;
  const double* bx = cbox + (size_t)cb * 5;
  const double* r = rbox + (size_t)rb * 4;
  const double dx = fmax(0.0, fmax(bx[0] - r[1], r[0] - bx[2]));
  const double dy = fmax(0.0, fmax(bx[1] - r[3], r[2] - bx[3]));
  const double ml = bx[4];
  return (dx * dx + dy * dy) > (ml * ml) * (1.0 + 1e-6);
}

// ---------------------------------------------------------------------------
// Near bits of every (row, element) pair (src/Core/Assembly_full.jl:736-754)
// and the per-row near-pair counts: the classification of far_kernel without
// the far entries, so that the near-pair list and the near/singular integrals
// do not wait for the far field.
// Block = 256 rows x a range of 32-element column blocks. A column block whose
// node bounding box lies farther than (1 + 1e-6) x its largest cutoff from the
// rows' bounding box holds only far pairs (every node distance and every
// element bounding-box distance is at least the box distance, so the exact
// test answers far); its bit words are written as zero without the pair tests.
// Other column blocks take the pair tests of far_kernel (same arithmetic, same
// decisions). row_cnt must be zero on entry; tiles[row block][column block]
// records which column blocks took the pair tests (the others hold no near bit).
// ---------------------------------------------------------------------------
template <int NN>
__global__ __launch_bounds__(kFarThreads) void near_classify(Geo g, Consts cs, int ncb,
                                                             int cb_per_block,
                                                             const double* __restrict__ box,
                                                             uint32_t* __restrict__ nearbits,
                                                             int* __restrict__ row_cnt,
                                                             uint8_t* __restrict__ tiles) {
  __builtin_amdgcn_s_setprio(3);  // pair-list chain: ahead of the far field
  __shared__ double s_cx[kEB * NN], s_cy[kEB * NN], s_L[kEB];
  __shared__ float s_lim2f[kEB];
  __shared__ double s_red[4][kFarThreads / 64];
  const int i = blockIdx.x * kFarThreads + threadIdx.x;
  const bool ok = i < g.nt;
  const double px = ok ? g.rx[i] : 0.0, py = ok ? g.ry[i] : 0.0;
  // bounding box of this block's rows
  double rlox = ok ? px : INFINITY, rhix = ok ? px : -INFINITY;
  double rloy = ok ? py : INFINITY, rhiy = ok ? py : -INFINITY;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    rlox = fmin(rlox, __shfl_xor(rlox, off, 64));
    rhix = fmax(rhix, __shfl_xor(rhix, off, 64));
    rloy = fmin(rloy, __shfl_xor(rloy, off, 64));
    rhiy = fmax(rhiy, __shfl_xor(rhiy, off, 64));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[0][wv] = rlox; s_red[1][wv] = rhix; s_red[2][wv] = rloy; s_red[3][wv] = rhiy;
  }
  __syncthreads();
  rlox = s_red[0][0]; rhix = s_red[1][0]; rloy = s_red[2][0]; rhiy = s_red[3][0];
#pragma unroll
  for (int q = 1; q < kFarThreads / 64; ++q) {
    rlox = fmin(rlox, s_red[0][q]); rhix = fmax(rhix, s_red[1][q]);
    rloy = fmin(rloy, s_red[2][q]); rhiy = fmax(rhiy, s_red[3][q]);
  }
  const int cb0 = blockIdx.y * cb_per_block, cb1 = min(ncb, cb0 + cb_per_block);
  int cnt = 0;
  for (int cb = cb0; cb < cb1; ++cb) {
    const double* bx = box + (size_t)cb * 5;
    const double dx = fmax(0.0, fmax(bx[0] - rhix, rlox - bx[2]));
    const double dy = fmax(0.0, fmax(bx[1] - rhiy, rloy - bx[3]));
    const double ml = bx[4];
    const bool all_far = !(cs.nf > 1e20) && (dx * dx + dy * dy) > (ml * ml) * (1.0 + 1e-6);
    if (threadIdx.x == 0) tiles[(size_t)blockIdx.x * ncb + cb] = all_far ? 0 : 1;
    uint32_t bits = 0u;
    if (!all_far) {  // uniform over the block
      const int e0 = cb * kEB;
      const int ecount = min(kEB, g.ne - e0);
      __syncthreads();
      stage_class_block<NN>(g, cs, e0, ecount, s_cx, s_cy, s_L, s_lim2f);
      __syncthreads();
      if (ok) {
        for (int el = 0; el < ecount; ++el) {
          double rx[NN], ry[NN];
          float rxf[NN], ryf[NN], d2f[NN];
          const int col0 = (e0 + el) * NN;
          const bool onel = (i + g.row0 >= col0) && (i + g.row0 < col0 + NN);
          if (classify_pair<NN>(px, py, s_cx + el * NN, s_cy + el * NN, s_L[el], s_lim2f[el],
                                cs.nf, onel, rx, ry, rxf, ryf, d2f))
            bits |= 1u << el;
        }
      }
    }
    if (ok) nearbits[(size_t)cb * g.nt + i] = bits;
    cnt += __popc(bits);
  }
  if (ok && cnt) atomicAdd(&row_cnt[i], cnt);
}

// Row sums of column-block partials part[cb * nt + i] for kFinRows rows per
// 256-thread block: wave w adds blocks w, w + 4, w + 8, ... in order, then the
// four wave sums are added in order (a fixed order for every launch).
constexpr int kFinRows = 64;
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int ncb, int nt,
                                               int i, double (*s_red)[kFinRows]) {
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  double s = 0.0;
  if (i < nt) {
#pragma unroll 8
    for (int cb = wv; cb < ncb; cb += 4) s += part[(size_t)cb * nt + i];
  }
  s_red[wv][lane] = s;
  __syncthreads();
  return ((s_red[0][lane] + s_red[1][lane]) + s_red[2][lane]) + s_red[3][lane];
}

// ---------------------------------------------------------------------------
// Near-pair CSR from the bit matrix [word][row].
// ---------------------------------------------------------------------------
// Exclusive scan of cnt[0..nt) into ptr[0..nt] with one workgroup of
// kScanWaves wavefronts (one per SIMD: it fits beside a kernel that holds most
// of a CU): each wave owns a contiguous segment, reads it 64 rows at a time
// (coalesced, four loads in flight) for its total, then scans it with lane
// shuffles from the sum of the segments before it. The total also goes to
// host_total (pinned host memory, may be null), so the host reads it without
// a copy on the stream.
constexpr int kScanWaves = 4;
__global__ __launch_bounds__(kScanWaves * kWave) void near_scan(int nt,
                                                                const int* __restrict__ cnt,
                                                                int* __restrict__ ptr,
                                                                int* __restrict__ host_total) {
  __builtin_amdgcn_s_setprio(3);  // pair-list chain: ahead of the far field
  __shared__ int s_tot[kScanWaves];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int seg = (nt + kScanWaves - 1) / kScanWaves;
  const int lo = min(nt, wv * seg), hi = min(nt, lo + seg);
  int tot = 0;
#pragma unroll 4
  for (int r = lo + lane; r < hi; r += kWave) tot += cnt[r];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) tot += __shfl_xor(tot, off, kWave);
  if (lane == 0) s_tot[wv] = tot;
  __syncthreads();
  int run = 0;
  for (int w = 0; w < wv; ++w) run += s_tot[w];
  for (int r0 = lo; r0 < hi; r0 += 4 * kWave) {
    int v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = r0 + u * kWave + lane;
      v[u] = (r < hi) ? cnt[r] : 0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = r0 + u * kWave + lane;
      int x = v[u];  // inclusive scan over the wave
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
      }
      if (r < hi) ptr[r] = run + x - v[u];
      run += __shfl(x, kWave - 1, kWave);
    }
  }
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kScanWaves; ++w) t += s_tot[w];
    ptr[nt] = t;
    if (host_total) __hip_atomic_store(host_total, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// One thread per row, in the row blocks of near_classify: only the column
// blocks that took the pair tests can hold near bits. The pairs are also
// listed by kind: sing[i] = CSR index of the on-element pair of row i (the
// rows whose global index row0 + i is a boundary node, one such pair each,
// i < nsing) and reg[] = CSR indices of the other pairs in order (null: not
// written).
__global__ __launch_bounds__(kFarThreads) void near_fill(int nt, int nwords, int nn, int row0,
                                                         int n, const uint32_t* __restrict__ bits,
                                                         const uint8_t* __restrict__ tiles,
                                                         const int* __restrict__ ptr,
                                                         int* __restrict__ prow,
                                                         int* __restrict__ pelem,
                                                         int* __restrict__ sing,
                                                         int* __restrict__ reg) {
  __builtin_amdgcn_s_setprio(3);  // pair-list chain: ahead of the far field
  const int i = blockIdx.x * kFarThreads + threadIdx.x;
  const uint8_t* tl = tiles + (size_t)blockIdx.x * nwords;
  const int nsing = max(0, min(nt, n - row0));
  const int own = (i < nsing) ? (i + row0) / nn : -1;  // element holding row i's node
  int off = (i < nt) ? ptr[i] : 0;
  int before = min(i, nsing);  // on-element pairs ahead of this row's next pair
  for (int w = 0; w < nwords; ++w) {
    if (!tl[w]) continue;  // uniform over the block
    uint32_t b = (i < nt) ? bits[(size_t)w * nt + i] : 0u;
    while (b) {
      const int t = __ffs(b) - 1;
      b &= b - 1u;
      const int e = w * 32 + t;
      prow[off] = i;
      pelem[off] = e;
      if (e == own) {
        sing[i] = off;
        ++before;
      } else if (reg) {
        reg[off - before] = off;
      }
      ++off;
    }
  }
}

// ---------------------------------------------------------------------------
// Near-singular and singular element integrals: the pole, tan-p3c and basis
// steps are the templated rule of near_rule.hpp (one source for the assembly
// and the adjoint), instantiated here in plain FP64.
// ---------------------------------------------------------------------------
using TanPlan = rule::TanPlan<double>;
using rule::tanp3c_count;
using rule::tanp3c_node;
using rule::tanp3c_plan;

__device__ __forceinline__ double clampd(double x, double lo, double hi) {
  return rule::clampS(x, lo, hi);
}

// Barycentric basis and its derivative (src/Core/Interpolation.jl:148-178, :215-218).
template <int NN>
__device__ __forceinline__ void shapefun(const double* x0, const double* bw, const double* D,
                                         double xx, double* N, double* dN) {
  rule::shapefun<double, NN>(x0, bw, D, xx, N, dN);
}

// Complex pole zeta + i eta of z(xi) through the collocation nodes
// (src/Core/Integration.jl:1117-1160) for p = 1, 2.
template <int NN>
__device__ __forceinline__ void complex_pole(const double* lx, const double* cx,
                                             const double* cy, double px, double py,
                                             double& zeta, double& eta) {
  rule::complex_pole<double, NN>(lx, cx, cy, px, py, zeta, eta);
}

// Sum over the G lanes of a group (xor offsets < G stay in the group).
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

}  // namespace
}  // namespace bemhip
