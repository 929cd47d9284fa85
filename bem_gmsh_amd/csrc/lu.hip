// Dense LU with partial pivoting (P A = L U, unit L) on gfx950, FP64 or FP32,
// in place on a column-major matrix, plus the triangular solves.
//
// This is the LU seam of bem_linsolve (src/Core/LinearSolveUtils.jl:44-49),
// whose default for a dense square system is an LAPACK getrf/getrs.
//
// Blocked right-looking algorithm, panel width kb:
//   panel_kernel   cooperative launch over all CUs: the m x kb panel is held in
//                  LDS, split by rows across workgroups; one grid barrier per
//                  column for the partial-pivot search (GEPP, ties -> lowest row).
//   perm_build / perm_apply   the panel's kb interchanges turned into one gather
//                  list, applied to every column outside the panel (a wave per
//                  column, coalesced over the contiguous destination rows).
//   trsm_kernel    U12 = L11^-1 A12, column strips in LDS.
//   gemm_kernel    A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 /
//                  v_mfma_f32_16x16x4f32), 128x128 block tiles, LDS double buffer.
// Solves: row permutation as a gather, then blocked forward / backward
// substitution with FP64 accumulation.

#include "common.hpp"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <climits>
#include <cmath>

namespace cg = cooperative_groups;

namespace bemhip {

namespace {

// ---------------------------------------------------------------------------
// Panel factorization.
// ---------------------------------------------------------------------------
constexpr int kPanelThreads = 512;

template <typename T>
struct PanelArgs {
  T* A;
  int lda;
  int m;          // panel rows (global rows k0 .. k0+m-1)
  int kb;         // panel width
  int k0;         // global index of the panel's first row/column
  int rows_per;   // rows per workgroup
  int G;          // workgroups
  int* ipiv;      // ipiv[k0 + c] = global pivot row of column k0 + c
  double* st_val; // [2][G] |candidate|
  int* st_row;    // [2][G] candidate row (panel-relative), INT_MAX if none
  T* st_data;     // [2][G][kb] candidate row values
  T* st_rowc;     // [2][kb] values of row c before the interchange
  int* info;      // first zero pivot: global column + 1 (0 if none)
};

__device__ __forceinline__ void better(double& v, int& r, double ov, int orr) {
  if (ov > v || (ov == v && orr < r)) { v = ov; r = orr; }
}

template <typename T>
__global__ __launch_bounds__(kPanelThreads) void panel_kernel(PanelArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);  // [kb][rows_per], column-major
  __shared__ double red_v[kPanelThreads / 64];
  __shared__ int red_r[kPanelThreads / 64];
  __shared__ int red_w[kPanelThreads / 64];
  __shared__ int s_lwin;
  __shared__ int s_gwin_wg, s_gwin_row;
  __shared__ double s_gwin_val;
  __shared__ T s_prow[256];
  cg::grid_group grid = cg::this_grid();

  const int wg = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwaves = kPanelThreads / 64;
  const int rp = a.rows_per;
  const int r0 = wg * rp;
  const int nr = max(0, min(rp, a.m - r0));
  const int kb = a.kb;
  const size_t lda = (size_t)a.lda;
  T* Ap = a.A + (size_t)a.k0 + (size_t)a.k0 * lda;  // panel origin

  for (int idx = tid; idx < nr * kb; idx += kPanelThreads) {
    const int r = idx % nr, c = idx / nr;
    tile[c * rp + r] = Ap[(size_t)(r0 + r) + (size_t)c * lda];
  }
  __syncthreads();

  for (int c = 0; c < kb; ++c) {
    const int par = c & 1;
    // 1. local candidate: max |A(r, c)| over rows r >= c
    double best = -1.0;
    int brow = INT_MAX;
    const int lr0 = max(0, c - r0);
    for (int r = lr0 + tid; r < nr; r += kPanelThreads) {
      const double v = fabs((double)tile[c * rp + r]);
      better(best, brow, v, r0 + r);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ov = __shfl_xor(best, off, 64);
      const int orr = __shfl_xor(brow, off, 64);
      better(best, brow, ov, orr);
    }
    if (lane == 0) { red_v[wave] = best; red_r[wave] = brow; }
    __syncthreads();
    if (tid == 0) {
      double b = red_v[0];
      int br = red_r[0];
      for (int w = 1; w < nwaves; ++w) better(b, br, red_v[w], red_r[w]);
      a.st_val[par * a.G + wg] = b;
      a.st_row[par * a.G + wg] = br;
      s_lwin = (br == INT_MAX) ? -1 : br - r0;
    }
    __syncthreads();
    // 2. stage the candidate row and (owner of row c) the current row c
    const int lw = s_lwin;
    if (lw >= 0)
      for (int cc = tid; cc < kb; cc += kPanelThreads)
        a.st_data[((size_t)par * a.G + wg) * kb + cc] = tile[cc * rp + lw];
    if (c >= r0 && c < r0 + nr)
      for (int cc = tid; cc < kb; cc += kPanelThreads)
        a.st_rowc[(size_t)par * kb + cc] = tile[cc * rp + (c - r0)];
    grid.sync();
    // 3. global winner (every workgroup computes the same answer)
    {
      double b = -1.0;
      int br = INT_MAX, bw = -1;
      for (int w = tid; w < a.G; w += kPanelThreads) {
        const double v = a.st_val[par * a.G + w];
        const int r = a.st_row[par * a.G + w];
        if (r != INT_MAX && (v > b || (v == b && r < br))) { b = v; br = r; bw = w; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(b, off, 64);
        const int orr = __shfl_xor(br, off, 64);
        const int ow = __shfl_xor(bw, off, 64);
        if (ov > b || (ov == b && orr < br)) { b = ov; br = orr; bw = ow; }
      }
      __syncthreads();
      if (lane == 0) { red_v[wave] = b; red_r[wave] = br; red_w[wave] = bw; }
      __syncthreads();
      if (tid == 0) {
        double bb = red_v[0];
        int rr = red_r[0], ww = red_w[0];
        for (int w = 1; w < nwaves; ++w) {
          if (red_v[w] > bb || (red_v[w] == bb && red_r[w] < rr)) {
            bb = red_v[w]; rr = red_r[w]; ww = red_w[w];
          }
        }
        s_gwin_val = bb;
        s_gwin_row = rr;
        s_gwin_wg = ww;
      }
      __syncthreads();
    }
    const int p = s_gwin_row;     // panel-relative pivot row
    const int pw = s_gwin_wg;
    const bool zero = !(s_gwin_val > 0.0);
    for (int cc = tid; cc < kb; cc += kPanelThreads)
      s_prow[cc] = a.st_data[((size_t)par * a.G + pw) * kb + cc];
    __syncthreads();
    if (zero) {
      if (wg == 0 && tid == 0) {
        a.ipiv[a.k0 + c] = a.k0 + c;
        atomicCAS(a.info, 0, a.k0 + c + 1);
      }
      continue;  // LAPACK: no interchange, no scaling for a zero column
    }
    // 4. interchange rows c and p inside the panel
    if (p != c) {
      if (p >= r0 && p < r0 + nr)
        for (int cc = tid; cc < kb; cc += kPanelThreads)
          tile[cc * rp + (p - r0)] = a.st_rowc[(size_t)par * kb + cc];
      if (c >= r0 && c < r0 + nr)
        for (int cc = tid; cc < kb; cc += kPanelThreads) tile[cc * rp + (c - r0)] = s_prow[cc];
    }
    if (wg == 0 && tid == 0) a.ipiv[a.k0 + c] = a.k0 + p;
    __syncthreads();
    // 5. multipliers and rank-1 update of rows below c
    const T piv = s_prow[c];
    const int lr1 = max(0, c + 1 - r0);
    const int na = nr - lr1;
    if (na > 0) {
      for (int r = lr1 + tid; r < nr; r += kPanelThreads) tile[c * rp + r] = tile[c * rp + r] / piv;
      __syncthreads();
      const int ncol = kb - c - 1;
      for (int idx = tid; idx < na * ncol; idx += kPanelThreads) {
        const int r = lr1 + idx % na;
        const int cc = c + 1 + idx / na;
        tile[cc * rp + r] -= tile[c * rp + r] * s_prow[cc];
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < nr * kb; idx += kPanelThreads) {
    const int r = idx % nr, c = idx / nr;
    Ap[(size_t)(r0 + r) + (size_t)c * lda] = tile[c * rp + r];
  }
}

// ---------------------------------------------------------------------------
// Row interchanges of one panel as a gather list (dst[t] <- src[t]).
// One wavefront; the kb transpositions are replayed on a position->content map.
// ---------------------------------------------------------------------------
constexpr int kMaxPanel = 256;

__global__ __launch_bounds__(64) void perm_build(const int* __restrict__ ipiv, int k0, int kb,
                                                 int* __restrict__ list_dst,
                                                 int* __restrict__ list_src,
                                                 int* __restrict__ list_len) {
  __shared__ int pos[2 * kMaxPanel];
  __shared__ int content[2 * kMaxPanel];
  __shared__ int s_nout;
  const int lane = threadIdx.x;
  for (int t = lane; t < kb; t += 64) { pos[t] = k0 + t; content[t] = k0 + t; }
  if (lane == 0) s_nout = 0;
  __syncthreads();
  for (int s = 0; s < kb; ++s) {
    const int r = ipiv[k0 + s];
    int idx;
    if (r < k0 + kb) {
      idx = r - k0;
    } else {
      const int nout = s_nout;
      int found = -1;
      for (int t0 = 0; t0 < nout; t0 += 64) {
        const int t = t0 + lane;
        const bool hit = (t < nout) && pos[kb + t] == r;
        const uint64_t m = __ballot(hit);
        if (m) { found = t0 + __ffsll((unsigned long long)m) - 1; break; }
      }
      if (found < 0) {
        if (lane == 0) { pos[kb + nout] = r; content[kb + nout] = r; s_nout = nout + 1; }
        found = nout;
      }
      idx = kb + found;
    }
    __syncthreads();
    if (lane == 0 && idx != s) {
      const int tmp = content[s];
      content[s] = content[idx];
      content[idx] = tmp;
    }
    __syncthreads();
  }
  const int total = kb + s_nout;
  // compact the entries that moved
  int base = 0;
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    const bool mv = (t < total) && pos[t] != content[t];
    const uint64_t m = __ballot(mv);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (mv) { list_dst[base + before] = pos[t]; list_src[base + before] = content[t]; }
    base += __popcll(m);
  }
  if (lane == 0) *list_len = base;
}

// Apply the gather list to columns [c0, c1) and [c2, c3): one wave per column.
template <typename T>
__global__ __launch_bounds__(256) void perm_apply(T* __restrict__ A, int lda, int c0, int c1,
                                                  int c2, int c3, const int* __restrict__ ldst,
                                                  const int* __restrict__ lsrc,
                                                  const int* __restrict__ llen) {
  const int len = *llen;
  const int n1 = c1 - c0, ntot = n1 + (c3 - c2);
  const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const bool active = wid < ntot;
  const int col = !active ? 0 : ((wid < n1) ? c0 + wid : c2 + (wid - n1));
  T* Ac = A + (size_t)col * lda;
  T v[8];  // up to 512 entries per column
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int t = q * 64 + lane;
    v[q] = (active && t < len) ? Ac[lsrc[t]] : T(0);
  }
  // every load of this column completes before any store to it
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int t = q * 64 + lane;
    if (active && t < len) Ac[ldst[t]] = v[q];
  }
}

// ---------------------------------------------------------------------------
// U12 = L11^-1 A12 (L11 unit lower, kb x kb). One workgroup per NC columns.
// ---------------------------------------------------------------------------
template <typename T, int NC>
__global__ __launch_bounds__(256) void trsm_kernel(int w, int ncols, const T* __restrict__ L,
                                                   int lda, T* __restrict__ B, int ldb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);  // [NC][w]
  const int cbase = blockIdx.x * NC;
  const int nc = min(NC, ncols - cbase);
  const int tid = threadIdx.x;
  for (int idx = tid; idx < w * nc; idx += 256) {
    const int r = idx % w, c = idx / w;
    tile[c * w + r] = B[(size_t)r + (size_t)(cbase + c) * ldb];
  }
  __syncthreads();
  constexpr int RG = 256 / NC;
  const int c = tid % NC, rg = tid / NC;
  for (int s = 0; s < w - 1; ++s) {
    if (c < nc) {
      const T xs = tile[c * w + s];
      const T* Ls = L + (size_t)s * lda;
      for (int r = s + 1 + rg; r < w; r += RG) tile[c * w + r] -= Ls[r] * xs;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < w * nc; idx += 256) {
    const int r = idx % w, cc = idx / w;
    B[(size_t)r + (size_t)(cbase + cc) * ldb] = tile[cc * w + r];
  }
}

// ---------------------------------------------------------------------------
// C -= A B (column-major) on MFMA 16x16x4. Block 128x128, 4 waves (2x2), each
// wave 64x64 = 4x4 MFMA tiles. The MFMA is issued as Out = B^T-fragment x
// A-fragment so that each output register holds 16 consecutive rows of C.
// ---------------------------------------------------------------------------
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

constexpr int kBM = 128, kBN = 128, kBK = 16;

template <typename T>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const T* __restrict__ A,
                                                   int lda, const T* __restrict__ B, int ldb,
                                                   T* __restrict__ C, int ldc) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int PAD = (sizeof(T) == 4) ? 4 : 2;
  __shared__ T As[2][kBK][kBM + PAD];
  __shared__ T Bs[2][kBK][kBN + PAD];
  const int bm = blockIdx.x * kBM, bn = blockIdx.y * kBN;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // global -> register staging: A tile (128 rows x 16 cols), B tile (16 x 128)
  const int a_col = tid >> 4, a_row = (tid & 15) * 8;
  const int b_col = tid >> 1, b_row = (tid & 1) * 8;
  T ra[8], rb[8];
  auto gload = [&](int kt) {
    const int ka = kt + a_col;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int r = bm + a_row + q;
      ra[q] = (r < M && ka < K) ? A[(size_t)r + (size_t)ka * lda] : T(0);
    }
    const int cb = bn + b_col;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kr = kt + b_row + q;
      rb[q] = (cb < N && kr < K) ? B[(size_t)kr + (size_t)cb * ldb] : T(0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; ++q) As[buf][a_col][a_row + q] = ra[q];
#pragma unroll
    for (int q = 0; q < 8; ++q) Bs[buf][b_row + q][b_col] = rb[q];
  };

  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int kt = 0; kt < K; kt += kBK) {
    const bool more = kt + kBK < K;
    if (more) gload(kt + kBK);
#pragma unroll
    for (int kk = 0; kk < kBK / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      T fa[4], fb[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        fa[t] = As[buf][k][wm + t * 16 + (lane & 15)];
        fb[t] = Bs[buf][k][wn + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = Mfma<T>::op(fb[tj], fa[ti], acc[ti][tj]);
    }
    if (more) {
      sstore(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
#pragma unroll
  for (int ti = 0; ti < 4; ++ti) {
    const int i = bm + wm + ti * 16 + (lane & 15);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // Output row m of the 16x16 MFMA block held in register r of this lane:
        // f32: m = 4*(lane/16) + r; f64: m = 4*r + lane/16 (rows interleaved).
        const int m = (sizeof(T) == 4) ? (lane >> 4) * 4 + r : r * 4 + (lane >> 4);
        const int j = bn + wn + tj * 16 + m;
        if (i < M && j < N) C[(size_t)i + (size_t)j * ldc] -= acc[ti][tj][r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Triangular solves with FP64 accumulation (factors in T).
// ---------------------------------------------------------------------------
constexpr int kSolveNB = 128;

__global__ void gather_kernel(int n, const int* __restrict__ perm, const double* __restrict__ b,
                              double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = b[perm[i]];
}

// Solve the diagonal block [s0, s0+nb) of L (unit, lower) or U (upper).
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsv_diag(int nb, int s0, const T* __restrict__ F,
                                                 int lda, double* __restrict__ y) {
  __shared__ double ys[kSolveNB];
  const int tid = threadIdx.x;
  for (int r = tid; r < nb; r += 256) ys[r] = y[s0 + r];
  __syncthreads();
  const T* Fb = F + (size_t)s0 + (size_t)s0 * lda;
  if (LOWER) {
    for (int s = 0; s < nb - 1; ++s) {
      const double xs = ys[s];
      for (int r = s + 1 + tid; r < nb; r += 256) ys[r] -= (double)Fb[r + (size_t)s * lda] * xs;
      __syncthreads();
    }
  } else {
    for (int s = nb - 1; s >= 0; --s) {
      if (tid == 0) ys[s] = ys[s] / (double)Fb[s + (size_t)s * lda];
      __syncthreads();
      const double xs = ys[s];
      for (int r = tid; r < s; r += 256) ys[r] -= (double)Fb[r + (size_t)s * lda] * xs;
      __syncthreads();
    }
  }
  for (int r = tid; r < nb; r += 256) y[s0 + r] = ys[r];
}

// y[r] -= sum_{s in [s0, s0+nb)} F[r, s] y[s] for rows r in [r0, r1).
template <typename T>
__global__ __launch_bounds__(256) void trsv_update(int r0, int r1, int s0, int nb,
                                                   const T* __restrict__ F, int lda,
                                                   double* __restrict__ y) {
  __shared__ double ys[kSolveNB];
  for (int t = threadIdx.x; t < nb; t += 256) ys[t] = y[s0 + t];
  __syncthreads();
  const int r = r0 + blockIdx.x * 256 + threadIdx.x;
  if (r >= r1) return;
  double acc = 0.0;
  const T* Fr = F + (size_t)r + (size_t)s0 * lda;
  for (int s = 0; s < nb; ++s) acc += (double)Fr[(size_t)s * lda] * ys[s];
  y[r] -= acc;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct LuWork {
  double* st_val;
  int* st_row;
  void* st_data;
  void* st_rowc;
  int* ldst;
  int* lsrc;
  int* llen;
  int* info;
};

template <typename T>
int choose_panel(bemhip_ctx* c, int n, int* kb_out, int* G_out, int* rows_out) {
  const int ncu = std::max(1, c->num_cu);
  const size_t lds_cap = 150 * 1024;
  int kb = (sizeof(T) == 8) ? 128 : 256;
  for (; kb >= 16; kb /= 2) {
    int G = std::min(ncu, std::max(1, (n + 127) / 128));
    int rows = (n + G - 1) / G;
    if ((size_t)rows * kb * sizeof(T) <= lds_cap) {
      *kb_out = kb;
      *G_out = G;
      *rows_out = rows;
      return BEMHIP_OK;
    }
  }
  return fail(c, BEMHIP_ERR_ARG, "lu: matrix too large for the LDS-resident panel");
}

template <typename T>
int lu_factor_t(bemhip_ctx* c, int n, T* dA, int lda, int* dipiv, int* info_host) {
  int kb0, G0, rows0;
  int st = choose_panel<T>(c, n, &kb0, &G0, &rows0);
  if (st) return st;
  // work buffers
  const size_t need = sizeof(double) * 2 * G0 + sizeof(int) * 2 * G0 +
                      sizeof(T) * 2 * (size_t)G0 * kb0 + sizeof(T) * 2 * kb0 +
                      sizeof(int) * (4 * kMaxPanel + 8) + 1024;
  BEMHIP_HIP_TRY(c, c->lu_work.ensure(need));
  char* wp = c->lu_work.ptr;
  auto carve = [&](size_t bytes) {
    char* p = wp;
    wp += (bytes + 255) & ~size_t(255);
    return p;
  };
  LuWork w;
  w.st_val = (double*)carve(sizeof(double) * 2 * G0);
  w.st_row = (int*)carve(sizeof(int) * 2 * G0);
  w.st_data = carve(sizeof(T) * 2 * (size_t)G0 * kb0);
  w.st_rowc = carve(sizeof(T) * 2 * kb0);
  w.ldst = (int*)carve(sizeof(int) * 2 * kMaxPanel);
  w.lsrc = (int*)carve(sizeof(int) * 2 * kMaxPanel);
  w.llen = (int*)carve(sizeof(int) * 8);
  w.info = w.llen + 4;
  if ((size_t)(wp - c->lu_work.ptr) > c->lu_work.count)
    return fail(c, BEMHIP_ERR_STATE, "lu: work buffer carve overflow");
  BEMHIP_HIP_TRY(c, hipMemsetAsync(w.info, 0, sizeof(int), c->stream));

  size_t max_lds = 0;  // largest panel over the whole factorization
  for (int k0 = 0; k0 < n; k0 += kb0) {
    const int m = n - k0, kb = std::min(kb0, n - k0);
    const int G = std::min(G0, std::max(1, (m + 127) / 128));
    const int rows = (m + G - 1) / G;
    max_lds = std::max(max_lds, (size_t)rows * kb * sizeof(T));
  }
  if (max_lds > 160 * 1024) return fail(c, BEMHIP_ERR_STATE, "lu: panel exceeds LDS");
  BEMHIP_HIP_TRY(c, hipFuncSetAttribute((const void*)panel_kernel<T>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)max_lds));
  int occ = 0;
  BEMHIP_HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ, (const void*)panel_kernel<T>, kPanelThreads, max_lds));
  if (occ < 1) return fail(c, BEMHIP_ERR_STATE, "lu: panel kernel cannot be resident");
  constexpr int NC = (sizeof(T) == 8) ? 16 : 32;
  const size_t trsm_lds = (size_t)kb0 * NC * sizeof(T);
  BEMHIP_HIP_TRY(c, hipFuncSetAttribute((const void*)trsm_kernel<T, NC>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)trsm_lds));
  double t_panel = 0, t_swap = 0, t_trsm = 0, t_gemm = 0;
  const int npanels = (n + kb0 - 1) / kb0;
  std::vector<hipEvent_t> ev((size_t)npanels * 5);
  for (auto& x : ev) hipEventCreate(&x);
  const bool timing = true;
  for (int k0 = 0; k0 < n; k0 += kb0) {
    hipEvent_t* e = &ev[(size_t)(k0 / kb0) * 5];
    const int kb = std::min(kb0, n - k0);
    const int m = n - k0;
    const int G = std::min(G0, std::max(1, (m + 127) / 128));
    const int rows = (m + G - 1) / G;
    if (timing) hipEventRecord(e[0], c->stream);
    PanelArgs<T> pa;
    pa.A = dA; pa.lda = lda; pa.m = m; pa.kb = kb; pa.k0 = k0; pa.rows_per = rows; pa.G = G;
    pa.ipiv = dipiv; pa.st_val = w.st_val; pa.st_row = w.st_row; pa.st_data = (T*)w.st_data;
    pa.st_rowc = (T*)w.st_rowc; pa.info = w.info;
    void* args[] = {&pa};
    const size_t lds = (size_t)rows * kb * sizeof(T);
    BEMHIP_HIP_TRY(c, hipLaunchCooperativeKernel((const void*)panel_kernel<T>, dim3(G),
                                                 dim3(kPanelThreads), args, lds, c->stream));
    if (timing) hipEventRecord(e[1], c->stream);
    // interchanges outside the panel
    hipLaunchKernelGGL(perm_build, dim3(1), dim3(64), 0, c->stream, dipiv, k0, kb, w.ldst,
                       w.lsrc, w.llen);
    const int ncols_out = n - kb;
    if (ncols_out > 0) {
      hipLaunchKernelGGL((perm_apply<T>), dim3((ncols_out + 3) / 4), dim3(256), 0, c->stream,
                         dA, lda, 0, k0, k0 + kb, n, w.ldst, w.lsrc, w.llen);
    }
    BEMHIP_HIP_TRY(c, hipGetLastError());
    if (timing) hipEventRecord(e[2], c->stream);
    const int nrest = n - k0 - kb;
    if (nrest > 0) {
      T* L11 = dA + (size_t)k0 + (size_t)k0 * lda;
      T* A12 = dA + (size_t)k0 + (size_t)(k0 + kb) * lda;
      hipLaunchKernelGGL((trsm_kernel<T, NC>), dim3((nrest + NC - 1) / NC), dim3(256),
                         (size_t)kb * NC * sizeof(T), c->stream, kb, nrest, L11, lda, A12, lda);
      if (timing) hipEventRecord(e[3], c->stream);
      T* L21 = dA + (size_t)(k0 + kb) + (size_t)k0 * lda;
      T* A22 = dA + (size_t)(k0 + kb) + (size_t)(k0 + kb) * lda;
      dim3 grid((nrest + kBM - 1) / kBM, (nrest + kBN - 1) / kBN);
      hipLaunchKernelGGL((gemm_kernel<T>), grid, dim3(256), 0, c->stream, nrest, nrest, kb, L21,
                         lda, A12, lda, A22, lda);
      BEMHIP_HIP_TRY(c, hipGetLastError());
      if (timing) hipEventRecord(e[4], c->stream);
    }
  }
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int k0 = 0; k0 < n; k0 += kb0) {
    hipEvent_t* e = &ev[(size_t)(k0 / kb0) * 5];
    const int nrest = n - k0 - std::min(kb0, n - k0);
    float t01 = 0, t12 = 0, t23 = 0, t34 = 0;
    hipEventElapsedTime(&t01, e[0], e[1]);
    hipEventElapsedTime(&t12, e[1], e[2]);
    if (nrest > 0) {
      hipEventElapsedTime(&t23, e[2], e[3]);
      hipEventElapsedTime(&t34, e[3], e[4]);
    }
    t_panel += t01; t_swap += t12; t_trsm += t23; t_gemm += t34;
  }
  for (auto& x : ev) hipEventDestroy(x);
  c->ms[sizeof(T) == 8 ? "lu64_panel" : "lu32_panel"] = t_panel;
  c->ms[sizeof(T) == 8 ? "lu64_swap" : "lu32_swap"] = t_swap;
  c->ms[sizeof(T) == 8 ? "lu64_trsm" : "lu32_trsm"] = t_trsm;
  c->ms[sizeof(T) == 8 ? "lu64_gemm" : "lu32_gemm"] = t_gemm;
  c->stats[sizeof(T) == 8 ? "lu64_kb" : "lu32_kb"] = kb0;
  c->stats[sizeof(T) == 8 ? "lu64_panel_wgs" : "lu32_panel_wgs"] = G0;
  int info = 0;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(&info, w.info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (info_host) *info_host = info;
  return BEMHIP_OK;
}

template <typename T>
int lu_solve_t(bemhip_ctx* c, int n, const T* dLU, int lda, const int* dperm, double* d_b) {
  BEMHIP_HIP_TRY(c, c->vec64c.ensure((size_t)n));
  double* y = c->vec64c.ptr;
  const int tb = 256;
  hipLaunchKernelGGL(gather_kernel, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n, dperm,
                     d_b, y);
  for (int s0 = 0; s0 < n; s0 += kSolveNB) {
    const int nb = std::min(kSolveNB, n - s0);
    hipLaunchKernelGGL((trsv_diag<T, true>), dim3(1), dim3(256), 0, c->stream, nb, s0, dLU, lda, y);
    const int r0 = s0 + nb;
    if (r0 < n)
      hipLaunchKernelGGL((trsv_update<T>), dim3((n - r0 + 255) / 256), dim3(256), 0, c->stream,
                         r0, n, s0, nb, dLU, lda, y);
  }
  const int nblk = (n + kSolveNB - 1) / kSolveNB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int s0 = bi * kSolveNB;
    const int nb = std::min(kSolveNB, n - s0);
    hipLaunchKernelGGL((trsv_diag<T, false>), dim3(1), dim3(256), 0, c->stream, nb, s0, dLU, lda, y);
    if (s0 > 0)
      hipLaunchKernelGGL((trsv_update<T>), dim3((s0 + 255) / 256), dim3(256), 0, c->stream, 0,
                         s0, s0, nb, dLU, lda, y);
  }
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(d_b, y, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                   c->stream));
  return BEMHIP_OK;
}

}  // namespace

int lu_factor_device(bemhip_ctx* c, int n, int dtype, void* dA, int lda, int* dipiv,
                     int* info_host) {
  if (dtype == 0) return lu_factor_t<double>(c, n, (double*)dA, lda, dipiv, info_host);
  return lu_factor_t<float>(c, n, (float*)dA, lda, dipiv, info_host);
}

int lu_solve_device(bemhip_ctx* c, int n, int dtype, const void* dLU, int lda, const int* dperm,
                    double* d_b) {
  if (dtype == 0) return lu_solve_t<double>(c, n, (const double*)dLU, lda, dperm, d_b);
  return lu_solve_t<float>(c, n, (const float*)dLU, lda, dperm, d_b);
}

// Sequential LAPACK-style interchanges -> gather permutation Pb[i] = b[perm[i]].
int build_perm(bemhip_ctx* c, int n, const int* dipiv, int* dperm) {
  std::vector<int> ip(n), pm(n);
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(ip.data(), dipiv, sizeof(int) * n, hipMemcpyDeviceToHost,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) pm[i] = i;
  for (int i = 0; i < n; ++i) {
    const int r = ip[i];
    if (r < 0 || r >= n) return fail(c, BEMHIP_ERR_STATE, "lu: invalid pivot index");
    std::swap(pm[i], pm[r]);
  }
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(dperm, pm.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return BEMHIP_OK;
}

}  // namespace bemhip
