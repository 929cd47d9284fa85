// Dense LU with partial pivoting (P A = L U, unit L) on gfx950, FP64 or FP32,
// in place on a column-major matrix, plus the triangular solves.
//
// This is the LU seam of bem_linsolve (src/Core/LinearSolveUtils.jl:44-49),
// whose default for a dense square system is an LAPACK getrf/getrs.
//
// Blocked right-looking algorithm, panel width 128:
//   panel_kernel   cooperative launch, one workgroup per 64 panel rows, the
//                  panel held in registers; one grid barrier per column for the
//                  partial-pivot search (GEPP, ties -> lowest row).
//   perm_build / perm_apply   the panel's kb interchanges turned into one gather
//                  list, applied to every column outside the panel (a wave per
//                  column, coalesced over the contiguous destination rows).
//   trsm_rec       U12 = L11^-1 A12, recursive halving: 32-row diagonal blocks
//                  in LDS, off-diagonal blocks on the MFMA GEMM.
//   gemm_kernel    A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 /
//                  v_mfma_f32_16x16x4f32), 128x128 block tiles, LDS double buffer.
// Solves: row permutation as a gather, then blocked forward / backward
// substitution with precomputed inverses of the 128x128 diagonal blocks and
// FP64 accumulation (one launch per block). Transposed solves (the adjoint)
// use the dot-product form over the factor's contiguous columns.

#include "common.hpp"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <climits>
#include <cmath>

namespace cg = cooperative_groups;

namespace bemhip {

namespace {

// ---------------------------------------------------------------------------
// Panel factorization (GEPP) with the panel held in registers.
//
// Workgroup = NW waves x 64 lanes. Lane l of every wave owns panel row
// blockIdx.x*64 + l; wave w owns the CG panel columns [w*CG, (w+1)*CG), so a
// thread holds CG values of one row. Panel width KB = NW*CG (4 x 32 = 128).
// Per column c: the owning wave finds the local candidate (max |a|, lowest
// row on ties), stages it, one grid barrier, every wave picks the global
// winner, the pivot row's values (uniform across lanes) drive the interchange
// and the rank-1 update of the rows below.
// ---------------------------------------------------------------------------
constexpr int kPB = 256;   // largest panel width handled by the interchange lists
constexpr int kPR = 64;    // rows per workgroup

template <typename T>
struct PanelArgs {
  T* A;
  int lda;
  int m;          // panel rows (global rows k0 .. k0+m-1)
  int kb;         // panel width (<= KB of the kernel)
  int k0;         // global index of the panel's first row/column
  int G;          // workgroups
  int* ipiv;      // ipiv[k0 + c] = global pivot row of column k0 + c
  double* st_val; // [2][G] |candidate|
  int* st_row;    // [2][G] candidate row (panel-relative), INT_MAX if none
  T* st_data;     // [2][G][KB] candidate row values
  T* st_rowc;     // [2][KB] values of row c before the interchange
  int* info;      // first zero pivot: global column + 1 (0 if none)
  unsigned* bar;  // grid barrier counter (zero at launch)
};

__device__ __forceinline__ bool better(double ov, int orr, double v, int r) {
  return ov > v || (ov == v && orr < r);
}

// Grid-wide barrier on a monotonic counter: every workgroup adds 1 per use
// and waits until the counter reaches `target` (= uses so far x workgroups).
// Every wave first waits for its own outstanding global stores (they are then
// in L2); after the workgroup barrier, thread 0's agent-scope release writes L2
// back and publishes them. Thread 0's acquire (which also invalidates the CU's
// caches) precedes every wave's later loads through the second barrier.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned target) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target)
      __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
}

template <typename T, int NW, int CG>
__global__ __launch_bounds__(NW * 64, 3) void panel_kernel(PanelArgs<T> a) {
  constexpr int KB = NW * CG;  // panel width of this instantiation
  __shared__ T s_l[kPR];
  __shared__ T s_stage[2][KB];
  __shared__ T s_pr[KB];   // pivot row (each wave fills and reads its own columns)
  __shared__ T s_rc[KB];   // row c before the interchange
  __shared__ int s_cand;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int blk = blockIdx.x;
  const int rbase = blk * kPR;
  const int r = rbase + lane;  // panel-relative row
  const bool valid = r < a.m;
  const size_t lda = (size_t)a.lda;
  T* Ap = a.A + (size_t)a.k0 + (size_t)a.k0 * lda;
  const int G = a.G;
  T v[CG];
#pragma unroll
  for (int j = 0; j < CG; ++j) {
    const int col = w * CG + j;
    v[j] = (valid && col < a.kb) ? Ap[(size_t)r + (size_t)col * lda] : T(0);
  }
  unsigned target = 0;
  for (int g = 0; g < NW; ++g) {
    if (g * CG >= a.kb) break;
#pragma unroll
    for (int jc = 0; jc < CG; ++jc) {
      const int c = g * CG + jc;
      if (c < a.kb) {
      const int par = c & 1;
      // 1. local candidate in wave g
      if (w == g) {
        double bv = (valid && r >= c) ? fabs((double)v[jc]) : -1.0;
        int br = (valid && r >= c) ? r : INT_MAX;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          const double ov = __shfl_xor(bv, off, 64);
          const int orr = __shfl_xor(br, off, 64);
          if (better(ov, orr, bv, br)) { bv = ov; br = orr; }
        }
        if (lane == 0) {
          a.st_val[par * G + blk] = bv;
          a.st_row[par * G + blk] = br;
          s_cand = (br == INT_MAX) ? -1 : br - rbase;
        }
      }
      __syncthreads();
      const int cl = s_cand;
      const bool own_c = (c >= rbase) && (c < rbase + kPR);
      if (cl >= 0 && lane == cl) {
#pragma unroll
        for (int j = 0; j < CG; ++j) s_stage[0][w * CG + j] = v[j];
      }
      if (own_c && lane == c - rbase) {
#pragma unroll
        for (int j = 0; j < CG; ++j) s_stage[1][w * CG + j] = v[j];
      }
      __syncthreads();
      if (cl >= 0)
        for (int t = tid; t < KB; t += NW * 64)
          a.st_data[((size_t)par * G + blk) * KB + t] = s_stage[0][t];
      if (own_c)
        for (int t = tid; t < KB; t += NW * 64) a.st_rowc[(size_t)par * KB + t] = s_stage[1][t];
      // 2. grid barrier
      target += (unsigned)G;
      grid_barrier(a.bar, target);
      // 3. global winner (each wave independently)
      double bv = -1.0;
      int br = INT_MAX, bb = 0;
      for (int q = lane; q < G; q += 64) {
        const double vq = a.st_val[par * G + q];
        const int rq = a.st_row[par * G + q];
        if (rq != INT_MAX && better(vq, rq, bv, br)) { bv = vq; br = rq; bb = q; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int orr = __shfl_xor(br, off, 64);
        const int ob = __shfl_xor(bb, off, 64);
        if (better(ov, orr, bv, br)) { bv = ov; br = orr; bb = ob; }
      }
      const int p = __builtin_amdgcn_readfirstlane(br);
      const int pb = __builtin_amdgcn_readfirstlane(bb);
      if (!(bv > 0.0)) {  // zero column: no interchange, no scaling (LAPACK)
        if (blk == 0 && tid == 0) {
          a.ipiv[a.k0 + c] = a.k0 + c;
          atomicCAS(a.info, 0, a.k0 + c + 1);
        }
      } else {
      // 4. pivot row of this wave's columns into LDS (lane j loads column j;
      //    LDS operations of one wave complete in order), interchange
      T* pr = s_pr + w * CG;
      if (lane < CG) pr[lane] = a.st_data[((size_t)par * G + pb) * KB + w * CG + lane];
      if (p != c) {
        T* rc = s_rc + w * CG;
        if (lane < CG) rc[lane] = a.st_rowc[(size_t)par * KB + w * CG + lane];
        __builtin_amdgcn_wave_barrier();
        if (valid && r == p) {
#pragma unroll
          for (int j = 0; j < CG; ++j) v[j] = rc[j];
        }
        if (valid && r == c) {
#pragma unroll
          for (int j = 0; j < CG; ++j) v[j] = pr[j];
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (blk == 0 && tid == 0) a.ipiv[a.k0 + c] = a.k0 + p;
      // 5. multipliers (wave g) and rank-1 update of rows below c
      if (w == g) {
        const bool below = valid && r > c;
        const T l = below ? v[jc] / pr[jc] : T(0);
        if (below) v[jc] = l;
        s_l[lane] = l;
      }
      __syncthreads();
      const T l = s_l[lane];
      // (pivot-row reads in chunks of 8 columns: bounded register live ranges)
      if (w > g) {
#pragma unroll
        for (int j0 = 0; j0 < CG; j0 += 8) {
#pragma unroll
          for (int j = j0; j < j0 + 8; ++j) v[j] -= l * pr[j];
          __builtin_amdgcn_sched_barrier(0);
        }
      } else if (w == g) {
#pragma unroll
        for (int j0 = 0; j0 < CG; j0 += 8) {
          if (j0 + 8 > jc + 1) {
#pragma unroll
            for (int j = (j0 > jc + 1 ? j0 : jc + 1); j < j0 + 8; ++j) v[j] -= l * pr[j];
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      }  // nonzero pivot
      }  // c < kb
    }
  }
#pragma unroll
  for (int j = 0; j < CG; ++j) {
    const int col = w * CG + j;
    if (valid && col < a.kb) Ap[(size_t)r + (size_t)col * lda] = v[j];
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
// Base case of the recursive TRSM: X = L^-1 B for a unit-lower w x w block
// (w <= 32) staged in LDS; one workgroup per NC columns of B.
// ---------------------------------------------------------------------------
constexpr int kTrsmBase = 32;

template <typename T, int NC>
__global__ __launch_bounds__(256) void trsm_base_kernel(int w, int ncols, const T* __restrict__ L,
                                                        int lda, T* __restrict__ B, int ldb) {
  __shared__ T Ls[kTrsmBase][kTrsmBase + 1];
  __shared__ T tile[NC][kTrsmBase + 1];
  const int cbase = blockIdx.x * NC;
  const int nc = min(NC, ncols - cbase);
  const int tid = threadIdx.x;
  for (int idx = tid; idx < w * w; idx += 256) {
    const int r = idx % w, k = idx / w;
    Ls[r][k] = L[(size_t)r + (size_t)k * lda];
  }
  for (int idx = tid; idx < w * nc; idx += 256) {
    const int r = idx % w, c = idx / w;
    tile[c][r] = B[(size_t)r + (size_t)(cbase + c) * ldb];
  }
  __syncthreads();
  // one thread per column: sequential forward substitution in LDS
  for (int c = tid; c < nc; c += 256) {
    for (int s = 0; s < w - 1; ++s) {
      const T xs = tile[c][s];
      for (int r = s + 1; r < w; ++r) tile[c][r] -= Ls[r][s] * xs;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < w * nc; idx += 256) {
    const int r = idx % w, c = idx / w;
    B[(size_t)r + (size_t)(cbase + c) * ldb] = tile[c][r];
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
__global__ __launch_bounds__(256, 2) void gemm_kernel(int M, int N, int K, const T* __restrict__ A,
                                                   int lda, const T* __restrict__ B, int ldb,
                                                   T* __restrict__ C, int ldc) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int PAD = (sizeof(T) == 4) ? 16 : 2;
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
// Triangular solves with FP64 accumulation (factors in T), blocked by kSolveNB
// with the inverses of the diagonal blocks precomputed once per factorization.
// ---------------------------------------------------------------------------
constexpr int kSolveNB = 128;

__global__ void gather_kernel(int n, const int* __restrict__ perm, const double* __restrict__ b,
                              double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = b[perm[i]];
}

// Inverse of diagonal block b of L (unit lower) or U (upper): one workgroup of
// kSolveNB threads per block; thread c builds column c in LDS. The factor's
// entries are read at lane-uniform addresses.
template <typename T, bool LOWER>
__global__ __launch_bounds__(kSolveNB) void tri_inv_kernel(int n, const T* __restrict__ F, int lda,
                                                          T* __restrict__ Dinv) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);  // [kSolveNB][kSolveNB], X[i*NB + c]
  const int b = blockIdx.x;
  const int s0 = b * kSolveNB;
  const int nb = min(kSolveNB, n - s0);
  const int c = threadIdx.x;
  const T* Fb = F + (size_t)s0 + (size_t)s0 * lda;
  if (LOWER) {
    for (int i = 0; i < nb; ++i) {
      double acc = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) acc -= (double)Fb[i + (size_t)k * lda] * (double)X[k * kSolveNB + c];
      X[i * kSolveNB + c] = (T)acc;
    }
  } else {
    for (int i = nb - 1; i >= 0; --i) {
      double acc = (i == c) ? 1.0 : 0.0;
      for (int k = i + 1; k < nb; ++k) acc -= (double)Fb[i + (size_t)k * lda] * (double)X[k * kSolveNB + c];
      X[i * kSolveNB + c] = (T)(acc / (double)Fb[i + (size_t)i * lda]);
    }
  }
  T* D = Dinv + (size_t)b * kSolveNB * kSolveNB;
  for (int i = 0; i < nb; ++i) D[i + (size_t)c * kSolveNB] = (c < nb) ? X[i * kSolveNB + c] : T(0);
}

// One block step of a triangular solve: z = Dinv_b y_b (written to x_b), then
// y_r -= sum_k F[r, s0+k] z_k for rows after (LOWER) or before (upper) the block.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsv_block_kernel(int n, int s0, int nb,
                                                         const T* __restrict__ F, int lda,
                                                         const T* __restrict__ D,
                                                         double* __restrict__ y,
                                                         double* __restrict__ x) {
  __shared__ double ys[kSolveNB];
  __shared__ double z[kSolveNB];
  const int tid = threadIdx.x;
  if (tid < nb) ys[tid] = y[s0 + tid];
  __syncthreads();
  if (tid < nb) {
    double acc = 0.0;
    for (int k = 0; k < nb; ++k) acc += (double)D[tid + (size_t)k * kSolveNB] * ys[k];
    z[tid] = acc;
    if (blockIdx.x == 0) x[s0 + tid] = acc;
  }
  __syncthreads();
  const int r = (LOWER ? s0 + nb : 0) + blockIdx.x * 256 + tid;
  const int rend = LOWER ? n : s0;
  if (r >= rend) return;
  double acc = 0.0;
  const T* Fr = F + (size_t)r + (size_t)s0 * lda;
  for (int k = 0; k < nb; ++k) acc += (double)Fr[(size_t)k * lda] * z[k];
  y[r] -= acc;
}

// Transposed solves A^T x = b with A = P^T L U: U^T w = b (forward), L^T z = w
// (backward), x = P^T z. Left-looking dot-product form: block bi needs, for
// each of its columns j, s_j = sum_r F[r, j] v[r] over the already solved
// rows r, which reads column j of the factor contiguously (one workgroup per
// column), then v_b = Dinv_b^T (b_b - s).
template <typename T>
__global__ __launch_bounds__(256) void dot_cols_kernel(const T* __restrict__ F, int lda, int c0,
                                                       int r0, int r1,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ s) {
  __shared__ double red[4];
  const int j = c0 + blockIdx.x;
  const T* Fc = F + (size_t)j * lda;
  double acc = 0.0;
  for (int r = r0 + threadIdx.x; r < r1; r += 256) acc += (double)Fc[r] * v[r];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) s[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// v[s0 + i] = sum_k D[k + i*NB] (b[s0 + k] - s[k]), i < nb (D^T applied).
template <typename T>
__global__ __launch_bounds__(kSolveNB) void diag_t_kernel(int s0, int nb, const T* __restrict__ D,
                                                          const double* __restrict__ b,
                                                          const double* __restrict__ s,
                                                          double* __restrict__ v) {
  __shared__ double t[kSolveNB];
  const int i = threadIdx.x;
  if (i < nb) t[i] = b[s0 + i] - s[i];
  __syncthreads();
  if (i >= nb) return;
  const T* Dc = D + (size_t)i * kSolveNB;
  double acc = 0.0;
  for (int k = 0; k < nb; ++k) acc += (double)Dc[k] * t[k];
  v[s0 + i] = acc;
}

__global__ void scatter_kernel(int n, const int* __restrict__ perm, const double* __restrict__ z,
                               double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[perm[i]] = z[i];
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
template <typename T>
int gemm_sub(bemhip_ctx* c, int M, int N, int K, const T* A, int lda, const T* B, int ldb, T* C,
             int ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return BEMHIP_OK;
  dim3 grid((M + kBM - 1) / kBM, (N + kBN - 1) / kBN);
  hipLaunchKernelGGL((gemm_kernel<T>), grid, dim3(256), 0, c->stream, M, N, K, A, lda, B, ldb,
                     C, ldc);
  return BEMHIP_OK;
}

// B = L^-1 B, L unit lower w x w: split in halves, the off-diagonal block on MFMA.
template <typename T>
int trsm_rec(bemhip_ctx* c, int w, int ncols, const T* L, int lda, T* B, int ldb) {
  if (w <= kTrsmBase) {
    constexpr int NC = 64;
    hipLaunchKernelGGL((trsm_base_kernel<T, NC>), dim3((ncols + NC - 1) / NC), dim3(256), 0,
                       c->stream, w, ncols, L, lda, B, ldb);
    return BEMHIP_OK;
  }
  const int w1 = ((w / 2 + kTrsmBase - 1) / kTrsmBase) * kTrsmBase;
  int st = trsm_rec(c, w1, ncols, L, lda, B, ldb);
  if (st) return st;
  st = gemm_sub(c, w - w1, ncols, w1, L + w1, lda, B, ldb, B + w1, ldb);
  if (st) return st;
  return trsm_rec(c, w - w1, ncols, L + (size_t)w1 + (size_t)w1 * lda, lda, B + w1, ldb);
}

template <typename T>
int lu_factor_t(bemhip_ctx* c, int n, T* dA, int lda, int* dipiv, int* info_host) {
  // Panel width 128 = 4 waves x 32 columns for both precisions (the panel of
  // 32769 rows stays register-resident with 3 workgroups per CU).
  constexpr int NW = 4;
  constexpr int CG = 32;
  constexpr int KB = NW * CG;
  const int Gmax = (n + kPR - 1) / kPR;
  // work buffers
  const size_t need = sizeof(double) * 2 * Gmax + sizeof(int) * 2 * Gmax +
                      sizeof(T) * 2 * (size_t)Gmax * KB + sizeof(T) * 2 * KB +
                      sizeof(int) * (4 * kMaxPanel + 16) + 8 * 256;
  BEMHIP_HIP_TRY(c, c->lu_work.ensure(need));
  char* wp = c->lu_work.ptr;
  auto carve = [&](size_t bytes) {
    char* p = wp;
    wp += (bytes + 255) & ~size_t(255);
    return p;
  };
  double* st_val = (double*)carve(sizeof(double) * 2 * Gmax);
  int* st_row = (int*)carve(sizeof(int) * 2 * Gmax);
  T* st_data = (T*)carve(sizeof(T) * 2 * (size_t)Gmax * KB);
  T* st_rowc = (T*)carve(sizeof(T) * 2 * KB);
  int* ldst = (int*)carve(sizeof(int) * 2 * kMaxPanel);
  int* lsrc = (int*)carve(sizeof(int) * 2 * kMaxPanel);
  int* llen = (int*)carve(sizeof(int) * 16);
  int* info = llen + 4;
  unsigned* bar = (unsigned*)(llen + 8);
  if ((size_t)(wp - c->lu_work.ptr) > c->lu_work.count)
    return fail(c, BEMHIP_ERR_STATE, "lu: work buffer carve overflow");
  BEMHIP_HIP_TRY(c, hipMemsetAsync(info, 0, sizeof(int), c->stream));
  int occ = 0;
  BEMHIP_HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ, (const void*)panel_kernel<T, NW, CG>, NW * 64, 0));
  if ((int64_t)occ * std::max(1, c->num_cu) < Gmax)
    return fail(c, BEMHIP_ERR_STATE, "lu: panel workgroups cannot all be resident");
  c->dinv_of = nullptr;

  const int npanels = (n + KB - 1) / KB;
  std::vector<hipEvent_t> ev((size_t)npanels * 5);
  for (auto& x : ev) hipEventCreate(&x);
  for (int k0 = 0; k0 < n; k0 += KB) {
    hipEvent_t* e = &ev[(size_t)(k0 / KB) * 5];
    const int kb = std::min(KB, n - k0);
    const int m = n - k0;
    const int G = (m + kPR - 1) / kPR;
    hipEventRecord(e[0], c->stream);
    BEMHIP_HIP_TRY(c, hipMemsetAsync(bar, 0, sizeof(unsigned), c->stream));
    PanelArgs<T> pa;
    pa.A = dA; pa.lda = lda; pa.m = m; pa.kb = kb; pa.k0 = k0; pa.G = G;
    pa.ipiv = dipiv; pa.st_val = st_val; pa.st_row = st_row; pa.st_data = st_data;
    pa.st_rowc = st_rowc; pa.info = info; pa.bar = bar;
    void* args[] = {&pa};
    BEMHIP_HIP_TRY(c, hipLaunchCooperativeKernel((const void*)panel_kernel<T, NW, CG>, dim3(G),
                                                 dim3(NW * 64), args, 0, c->stream));
    hipEventRecord(e[1], c->stream);
    // interchanges outside the panel
    hipLaunchKernelGGL(perm_build, dim3(1), dim3(64), 0, c->stream, dipiv, k0, kb, ldst, lsrc,
                       llen);
    const int ncols_out = n - kb;
    if (ncols_out > 0)
      hipLaunchKernelGGL((perm_apply<T>), dim3((ncols_out + 3) / 4), dim3(256), 0, c->stream,
                         dA, lda, 0, k0, k0 + kb, n, ldst, lsrc, llen);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    hipEventRecord(e[2], c->stream);
    const int nrest = n - k0 - kb;
    if (nrest > 0) {
      T* L11 = dA + (size_t)k0 + (size_t)k0 * lda;
      T* A12 = dA + (size_t)k0 + (size_t)(k0 + kb) * lda;
      int st = trsm_rec(c, kb, nrest, (const T*)L11, lda, A12, lda);
      if (st) return st;
      hipEventRecord(e[3], c->stream);
      T* L21 = dA + (size_t)(k0 + kb) + (size_t)k0 * lda;
      T* A22 = dA + (size_t)(k0 + kb) + (size_t)(k0 + kb) * lda;
      gemm_sub(c, nrest, nrest, kb, (const T*)L21, lda, (const T*)A12, lda, A22, lda);
      BEMHIP_HIP_TRY(c, hipGetLastError());
      hipEventRecord(e[4], c->stream);
    }
  }
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  double t_panel = 0, t_swap = 0, t_trsm = 0, t_gemm = 0;
  for (int k0 = 0; k0 < n; k0 += KB) {
    hipEvent_t* e = &ev[(size_t)(k0 / KB) * 5];
    const int nrest = n - k0 - std::min(KB, n - k0);
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
  const char* pfx = (sizeof(T) == 8) ? "lu64_" : "lu32_";
  c->ms[std::string(pfx) + "panel"] = t_panel;
  c->ms[std::string(pfx) + "swap"] = t_swap;
  c->ms[std::string(pfx) + "trsm"] = t_trsm;
  c->ms[std::string(pfx) + "gemm"] = t_gemm;
  c->stats[std::string(pfx) + "kb"] = KB;
  c->stats[std::string(pfx) + "panel_wgs"] = Gmax;
  int infoh = 0;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(&infoh, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (info_host) *info_host = infoh;
  return BEMHIP_OK;
}

template <typename T>
int ensure_dinv(bemhip_ctx* c, int n, const T* dLU, int lda) {
  const int nblk = (n + kSolveNB - 1) / kSolveNB;
  const size_t blk_elems = (size_t)kSolveNB * kSolveNB;
  BEMHIP_HIP_TRY(c, c->dinv.ensure(2 * nblk * blk_elems * sizeof(T)));
  T* DL = (T*)c->dinv.ptr;
  T* DU = DL + nblk * blk_elems;
  if (c->dinv_of != (const void*)dLU || c->dinv_n != n) {
    const size_t lds = blk_elems * sizeof(T);
    BEMHIP_HIP_TRY(c, hipFuncSetAttribute((const void*)tri_inv_kernel<T, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    BEMHIP_HIP_TRY(c, hipFuncSetAttribute((const void*)tri_inv_kernel<T, false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((tri_inv_kernel<T, true>), dim3(nblk), dim3(kSolveNB), lds, c->stream, n,
                       dLU, lda, DL);
    hipLaunchKernelGGL((tri_inv_kernel<T, false>), dim3(nblk), dim3(kSolveNB), lds, c->stream, n,
                       dLU, lda, DU);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    c->dinv_of = dLU;
    c->dinv_n = n;
  }
  return BEMHIP_OK;
}

// A^T x = b in place of d_b, with P A = L U (perm: (P v)[i] = v[perm[i]]).
template <typename T>
int lu_solve_trans_t(bemhip_ctx* c, int n, const T* dLU, int lda, const int* dperm,
                     double* d_b) {
  int st = ensure_dinv<T>(c, n, dLU, lda);
  if (st) return st;
  const int nblk = (n + kSolveNB - 1) / kSolveNB;
  const size_t blk_elems = (size_t)kSolveNB * kSolveNB;
  const T* DL = (const T*)c->dinv.ptr;
  const T* DU = DL + nblk * blk_elems;
  BEMHIP_HIP_TRY(c, c->vec64c.ensure(2 * (size_t)n + kSolveNB));
  double* w = c->vec64c.ptr;
  double* z = c->vec64c.ptr + n;
  double* s = c->vec64c.ptr + 2 * (size_t)n;
  for (int bi = 0; bi < nblk; ++bi) {  // U^T w = b: rows [0, s0) of U's columns
    const int s0 = bi * kSolveNB, nb = std::min(kSolveNB, n - s0);
    hipLaunchKernelGGL((dot_cols_kernel<T>), dim3(nb), dim3(256), 0, c->stream, dLU, lda, s0, 0,
                       s0, w, s);
    hipLaunchKernelGGL((diag_t_kernel<T>), dim3(1), dim3(kSolveNB), 0, c->stream, s0, nb,
                       DU + bi * blk_elems, d_b, s, w);
  }
  for (int bi = nblk - 1; bi >= 0; --bi) {  // L^T z = w: rows [s0+nb, n) of L's columns
    const int s0 = bi * kSolveNB, nb = std::min(kSolveNB, n - s0);
    hipLaunchKernelGGL((dot_cols_kernel<T>), dim3(nb), dim3(256), 0, c->stream, dLU, lda, s0,
                       s0 + nb, n, z, s);
    hipLaunchKernelGGL((diag_t_kernel<T>), dim3(1), dim3(kSolveNB), 0, c->stream, s0, nb,
                       DL + bi * blk_elems, w, s, z);
  }
  const int tb = 256;
  hipLaunchKernelGGL(scatter_kernel, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n, dperm,
                     z, d_b);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  return BEMHIP_OK;
}

template <typename T>
int lu_solve_t(bemhip_ctx* c, int n, const T* dLU, int lda, const int* dperm, double* d_b) {
  int st = ensure_dinv<T>(c, n, dLU, lda);
  if (st) return st;
  const int nblk = (n + kSolveNB - 1) / kSolveNB;
  const size_t blk_elems = (size_t)kSolveNB * kSolveNB;
  const T* DL = (const T*)c->dinv.ptr;
  const T* DU = DL + nblk * blk_elems;
  BEMHIP_HIP_TRY(c, c->vec64c.ensure(2 * (size_t)n));
  double* y = c->vec64c.ptr;
  double* x = c->vec64c.ptr + n;
  const int tb = 256;
  hipLaunchKernelGGL(gather_kernel, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n, dperm,
                     d_b, y);
  for (int bi = 0; bi < nblk; ++bi) {  // forward: L x = y
    const int s0 = bi * kSolveNB, nb = std::min(kSolveNB, n - s0);
    const int rows = n - s0 - nb;
    hipLaunchKernelGGL((trsv_block_kernel<T, true>), dim3(std::max(1, (rows + 255) / 256)),
                       dim3(256), 0, c->stream, n, s0, nb, dLU, lda, DL + bi * blk_elems, y, x);
  }
  for (int bi = nblk - 1; bi >= 0; --bi) {  // backward: U y = x
    const int s0 = bi * kSolveNB, nb = std::min(kSolveNB, n - s0);
    hipLaunchKernelGGL((trsv_block_kernel<T, false>), dim3(std::max(1, (s0 + 255) / 256)),
                       dim3(256), 0, c->stream, n, s0, nb, dLU, lda, DU + bi * blk_elems, x, y);
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

int lu_solve_trans_device(bemhip_ctx* c, int n, int dtype, const void* dLU, int lda,
                          const int* dperm, double* d_b) {
  if (dtype == 0) return lu_solve_trans_t<double>(c, n, (const double*)dLU, lda, dperm, d_b);
  return lu_solve_trans_t<float>(c, n, (const float*)dLU, lda, dperm, d_b);
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
