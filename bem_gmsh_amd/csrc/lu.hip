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
//   build_gather_list / perm_apply   the panel's kb interchanges turned into one gather
//                  list, applied to every column outside the panel (a wave per
//                  column, coalesced over the contiguous destination rows).
//   trsm_rec       U12 = L11^-1 A12, recursive halving: 32-row diagonal blocks
//                  in LDS, off-diagonal blocks on the MFMA GEMM.
//   gemm_kernel    A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 /
//                  v_mfma_f32_16x16x4f32), 128x128 block tiles, LDS double buffer.
// Solves: row permutation as a gather, then blocked forward / backward
// substitution with precomputed inverses of the 128x128 diagonal blocks and
// FP64 accumulation, one launch per triangle (row blocks chained by flags). Transposed solves (the adjoint)
// use the dot-product form over the factor's contiguous columns.

#include "common.hpp"

#include <hip/hip_cooperative_groups.h>
#include <mutex>
#include <string>

#include <algorithm>
#include <type_traits>
#include <climits>
#include <limits>
#include <cmath>

namespace cg = cooperative_groups;

namespace bemhip {

namespace {

// ---------------------------------------------------------------------------
// Panel factorization (GEPP) with the panel held in registers.
//
// Workgroup = 8 waves x 64 lanes = kPR = 512 panel rows; thread (wave w, lane
// l) holds the whole panel row blockIdx.x*512 + 64w + l (W values: W = 128 for
// FP32, 64 for FP64, 128 VGPRs either way), rotated so that register k holds
// column (c + k) mod W at column step c: the current column is always
// register 0 and no register is ever indexed at run time (a run-time index
// would send the row to scratch memory). LDS copies of rows use the same
// rotated order; the exchanged records use natural column order. Per column c:
//   1. each wave finds its candidate (max |a| over its rows >= c, lowest row
//      on ties) and stages the candidate row in LDS; the row at position c is
//      staged by its owner;
//   2. wave 0 picks the workgroup candidate and publishes it: a 64-bit key
//      word (|a| key, tag, row) and the candidate row as tagged data words,
//      plus (workgroup 0) row c;
//   3. wave 0 polls the key words of all workgroups until every one carries
//      this column's tag, picks the winner (max |a|, lowest row), and reads the
//      winner's row (and, in the workgroup that owns the pivot row, row c);
//   4. every wave interchanges, scales and applies the rank-1 update.
// Exchange protocol: every cross-workgroup word is written and read with
// relaxed agent-scope atomics (coherent across XCDs without cache-wide
// write-back or invalidation) and validates itself by its tag, so no ordering
// between words is needed. Key words and records are double-buffered by column
// parity: a workgroup can only overwrite a buffer after every other workgroup
// has published the next column, i.e. has finished reading this one. Key tags
// are c + 1 (the key buffer is cleared before each launch); data-word tags add
// a per-launch epoch. Every wait is bounded: past kSpinLimit polls the waiting
// workgroup raises the abort word, every workgroup leaves at its next wait,
// and the host reports an error.
//
// FP32 keys are the exact |a| bits. FP64 keys are |a| rounded toward zero to
// FP32 (monotone); workgroups tied on the maximum key are resolved on the exact
// FP64 values read from their records.
// ---------------------------------------------------------------------------
constexpr int kPW = 8;          // waves per panel workgroup
constexpr int kPR = kPW * 64;   // rows per panel workgroup
constexpr int kPU = 8;          // panel columns per unrolled chunk
constexpr unsigned kSpinLimit = 1u << 20;

template <typename T>
struct PanelArgs {
  T* A;
  int lda;
  int m;          // panel rows (global rows k0 .. k0+m-1)
  int kb;         // panel width (<= W of the kernel)
  int k0;         // global index of the panel's first row/column
  int G;          // workgroups
  int gstride;    // key / record buffer stride in workgroups (>= G)
  unsigned epoch; // per-launch tag base of the data words
  int* ipiv;      // ipiv[k0 + c] = global pivot row of column k0 + c
  int* info;      // first zero pivot: global column + 1 (0 if none)
  unsigned long long* keys;   // [2][gstride] key words (zero at launch)
  unsigned long long* keys_next;  // the next launch's key buffer, cleared by this one
  unsigned long long* recs;   // [2][gstride][RW] candidate rows (tagged words)
  unsigned long long* rowc;   // [2][RW] row c before the interchange (tagged words)
  unsigned* abort_flag;       // raised by a workgroup whose wait ran out
  int* list_dst;              // the panel's interchanges as a gather list, built at the
  int* list_src;              // end by workgroup 0 (kb <= 128; list_len == nullptr: not built)
  int* list_len;
};

__device__ __forceinline__ void st_ag(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_ag(const unsigned long long* p) {
  return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_ag32(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_ag32(const unsigned* p) {
  return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tagged data words: FP32 one word per value, FP64 two (low, high halves).
template <typename T>
struct TagWords;
template <>
struct TagWords<float> {
  static constexpr int kWords = 1;
  static __device__ __forceinline__ void put(unsigned long long* w, float v, unsigned tag) {
    st_ag(w, ((unsigned long long)tag << 32) | __float_as_uint(v));
  }
  // returns false if the word does not carry `tag` yet
  static __device__ __forceinline__ bool get(const unsigned long long* w, unsigned tag, float& v) {
    const unsigned long long x = ld_ag(w);
    v = __uint_as_float((unsigned)x);
    return (unsigned)(x >> 32) == tag;
  }
  static __device__ __forceinline__ unsigned key(float a) { return __float_as_uint(a); }
};
template <>
struct TagWords<double> {
  static constexpr int kWords = 2;
  static __device__ __forceinline__ void put(unsigned long long* w, double v, unsigned tag) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    st_ag(w, ((unsigned long long)tag << 32) | (b & 0xffffffffull));
    st_ag(w + 1, ((unsigned long long)tag << 32) | (b >> 32));
  }
  static __device__ __forceinline__ bool get(const unsigned long long* w, unsigned tag, double& v) {
    const unsigned long long lo = ld_ag(w), hi = ld_ag(w + 1);
    v = __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
    return (unsigned)(lo >> 32) == tag && (unsigned)(hi >> 32) == tag;
  }
  static __device__ __forceinline__ unsigned key(double a) {
    return __float_as_uint(__double2float_rz(a));
  }
};

// ranking of (|a|, row): larger |a| first, then the lower row
template <typename T>
__device__ __forceinline__ bool better(T ov, int orr, T v, int r) {
  return ov > v || (ov == v && orr < r);
}

// Reads the kb-value tagged record at `w` into s in the rotated order of column
// step c (s[k] = column (c + k) mod W, zeros past kb); one whole wave. Returns
// false if the wait ran out (abort raised).
template <typename T, int W>
__device__ bool read_record(const unsigned long long* w, unsigned tag, int kb, int c, T* s,
                            unsigned* abort_flag) {
  constexpr int NW = TagWords<T>::kWords;
  const int lane = threadIdx.x & 63;
  for (unsigned spin = 0;; ++spin) {
    bool ok = true;
    T val[W / 64 > 0 ? W / 64 : 1];
#pragma unroll
    for (int q = 0; q < (W + 63) / 64; ++q) {
      const int j = lane + 64 * q;
      val[q] = T(0);
      if (j < kb && j < W) ok = TagWords<T>::get(w + (size_t)j * NW, tag, val[q]) && ok;
    }
    if (__all(ok)) {
#pragma unroll
      for (int q = 0; q < (W + 63) / 64; ++q) {
        const int j = lane + 64 * q;
        if (j < W) s[(j - c + W) % W] = val[q];
      }
      return true;
    }
    if (spin > kSpinLimit || ld_ag32(abort_flag)) {
      if (lane == 0) atomicExch(abort_flag, 1u);
      return false;
    }
  }
}

// The gather list of perm_build for panels of at most 128 columns, with the
// position -> content map in the wave's registers (lane l holds entries l, l + 64, l + 128,
// l + 192: the kb panel positions, then the rows from outside the panel in the
// order met) and the kb transpositions replayed with lane reads: no LDS round
// trip per interchange.
__device__ __forceinline__ int sel4(const int (&a)[4], int q) {
  return q == 0 ? a[0] : (q == 1 ? a[1] : (q == 2 ? a[2] : a[3]));
}
// One full wave; lane l passes the pivots of columns l and l + 64 (piv0, piv1).
__device__ void build_gather_list(int piv0, int piv1, int k0, int kb, int* __restrict__ list_dst,
                                  int* __restrict__ list_src, int* __restrict__ list_len) {
  const int lane = threadIdx.x & 63;
  int pos[4], con[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int t = lane + 64 * q;
    pos[q] = con[q] = (t < kb) ? k0 + t : -1;
  }
  int nout = 0;
  for (int s = 0; s < kb; ++s) {
    const int r = __builtin_amdgcn_readlane(s < 64 ? piv0 : piv1, s & 63);
    int idx;
    if (r < k0 + kb) {
      idx = r - k0;
    } else {
      // the outside row's entry: the one already met, else a new one
      int found = INT_MAX;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int t = lane + 64 * q;
        if (t >= kb && t < kb + nout && pos[q] == r) found = min(found, t);
      }
      const unsigned long long hit = __ballot(found != INT_MAX);
      if (hit) {
        idx = __builtin_amdgcn_readlane(found, (int)__ffsll((long long)hit) - 1);
      } else {
        idx = kb + nout;
        ++nout;
        if (lane == (idx & 63)) {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q == (idx >> 6)) { pos[q] = r; con[q] = r; }
        }
      }
    }
    if (idx != s) {
      const int cs = __builtin_amdgcn_readlane(sel4(con, s >> 6), s & 63);
      const int ci = __builtin_amdgcn_readlane(sel4(con, idx >> 6), idx & 63);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (lane == (s & 63) && q == (s >> 6)) con[q] = ci;
        if (lane == (idx & 63) && q == (idx >> 6)) con[q] = cs;
      }
    }
  }
  // compact the entries that moved
  const int total = kb + nout;
  int base = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int t = lane + 64 * q;
    const bool mv = (t < total) && pos[q] != con[q];
    const unsigned long long m = __ballot(mv);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (mv) { list_dst[base + before] = pos[q]; list_src[base + before] = con[q]; }
    base += __popcll(m);
  }
  if (lane == 0) *list_len = base;
}

__global__ __launch_bounds__(64) void perm_build_fast(const int* __restrict__ ipiv, int k0, int kb,
                                                      int* __restrict__ list_dst,
                                                      int* __restrict__ list_src,
                                                      int* __restrict__ list_len) {
  const int lane = threadIdx.x;
  build_gather_list(lane < kb ? ipiv[k0 + lane] : 0, lane + 64 < kb ? ipiv[k0 + 64 + lane] : 0, k0,
                    kb, list_dst, list_src, list_len);
}

template <typename T, int W>
__global__ __launch_bounds__(kPR, 1) void panel_kernel(PanelArgs<T> a) {
  using TW = TagWords<T>;
  constexpr int NWD = TW::kWords;
  constexpr int RW = W * NWD;  // data words per record
  __shared__ __align__(16) T s_wrow[kPW][W];  // each wave's candidate row
  __shared__ __align__(16) T s_rowc[W];       // row at position c (workgroup 0)
  // s_rows[0]: the pivot row read from the winner's record; s_rows[1]: row c
  // (for the owner of the pivot row, and the pivot row itself when p == c)
  __shared__ __align__(16) T s_rows[2][W];
  T* const s_pr = s_rows[0];
  T* const s_rc = s_rows[1];
  __shared__ T s_wval[kPW];
  __shared__ int s_wrowi[kPW];
  __shared__ int s_p, s_ok, s_abort, s_use_rc;
  __shared__ int s_piv[W];  // workgroup 0: ipiv of the panel's columns
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int blk = blockIdx.x;
  const int rbase = blk * kPR;
  const int r = rbase + w * 64 + lane;  // panel-relative row
  const bool valid = r < a.m;
  const size_t lda = (size_t)a.lda;
  T* Ap = a.A + (size_t)a.k0 + (size_t)a.k0 * lda;
  const int G = a.G, kb = a.kb;
  T v[W];
#pragma unroll
  for (int j = 0; j < W; ++j) v[j] = (valid && j < kb) ? Ap[(size_t)r + (size_t)j * lda] : T(0);
  if (tid == 0) s_abort = 0;
  // Clear this workgroup's key slots of the next launch's buffer: the previous
  // launch (the buffer's last user) has finished, and the next launch of the
  // factorization has no more workgroups than this one.
  if (tid < 2) a.keys_next[(size_t)tid * a.gstride + blk] = 0ull;
  bool stop = false;
  int rot = 0;  // register k holds column (rot + k) mod W; rot = c0 within a chunk

  // Columns in chunks of kPU: within a chunk starting at c0, column c0 + u is
  // register u (u known at compile time), and the registers rotate by kPU per
  // chunk. LDS copies of rows use the rotation of the chunk (index k = column
  // (c0 + k) mod W). Registers in [W - c0, W) hold finished columns; with c0 a
  // multiple of kPU they form whole kPU-register groups, skipped as a unit.
  for (int c0 = 0; c0 < kb && !stop; c0 += kPU) {
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int c = c0 + u;
      if (c >= kb || stop) break;
      const int par = c & 1;
      const unsigned tag8 = (unsigned)(c + 1);
      const unsigned tag32 = (a.epoch << 8) | tag8;
      // 1. wave candidate over rows >= c (column c is register u)
      T bv = (valid && r >= c) ? fabs(v[u]) : T(-1);
      int br = (valid && r >= c) ? r : INT_MAX;
      if constexpr (sizeof(T) == 4) {
        // |a| bits order like |a| (+1: an eligible zero beats no row); one
        // DPP wave reduction, then the lowest lane (= lowest row) holding it
        const unsigned key = br != INT_MAX ? __float_as_uint(bv) + 1u : 0u;
        const unsigned kmax = __ockl_wfred_max_u32(key);
        const unsigned long long hit = __ballot(key == kmax && kmax != 0u);
        bv = hit ? __uint_as_float(kmax - 1u) : T(-1);
        br = hit ? rbase + w * 64 + (int)__ffsll((long long)hit) - 1 : INT_MAX;
      } else {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          const T ov = __shfl_xor(bv, off, 64);
          const int orr = __shfl_xor(br, off, 64);
          if (better(ov, orr, bv, br)) { bv = ov; br = orr; }
        }
      }
      if (lane == 0) { s_wval[w] = bv; s_wrowi[w] = br; }
      if (br != INT_MAX && r == br) {
#pragma unroll
        for (int j = 0; j < W; ++j) s_wrow[w][j] = v[j];
      }
      if (r == c) {
#pragma unroll
        for (int j = 0; j < W; ++j) s_rowc[j] = v[j];
      }
      __syncthreads();
      // 2.-3. wave 0: publish, poll, fetch the pivot row
      if (w == 0) {
        T cv = T(-1);
        int cr = INT_MAX, cw = 0;
        if (lane < kPW) { cv = s_wval[lane]; cr = s_wrowi[lane]; cw = lane; }
#pragma unroll
        for (int off = 4; off >= 1; off >>= 1) {
          const T ov = __shfl_xor(cv, off, 64);
          const int orr = __shfl_xor(cr, off, 64);
          const int ow = __shfl_xor(cw, off, 64);
          if (better(ov, orr, cv, cr)) { cv = ov; cr = orr; cw = ow; }
        }
        cv = __shfl(cv, 0, 64);
        cr = __shfl(cr, 0, 64);
        cw = __shfl(cw, 0, 64);
        unsigned long long* rec = a.recs + ((size_t)par * a.gstride + blk) * RW;
        if (cr != INT_MAX) {
          for (int j = lane; j < kb; j += 64)
            TW::put(rec + (size_t)j * NWD, s_wrow[cw][(j - c0 + W) % W], tag32);
        }
        if (blk == 0) {
          unsigned long long* rc = a.rowc + (size_t)par * RW;
          for (int j = lane; j < kb; j += 64)
            TW::put(rc + (size_t)j * NWD, s_rowc[(j - c0 + W) % W], tag32);
        }
        if (lane == 0) {
          const unsigned k32 = (cr != INT_MAX) ? TW::key(cv) : 0u;
          const unsigned row24 = (cr != INT_MAX) ? (unsigned)cr : 0xffffffu;
          st_ag(&a.keys[(size_t)par * a.gstride + blk],
                ((unsigned long long)k32 << 32) | ((unsigned long long)tag8 << 24) | row24);
        }
        // poll the keys of all workgroups and, in the same loads' round trip,
        // row c (published by workgroup 0 with its key): the pivot row when
        // p == c (the common case of a diagonally dominant matrix: no second
        // dependent round trip), the interchange partner of the pivot row's
        // owner otherwise
        constexpr int RQ = (W + 63) / 64;
        const unsigned long long* rcw = a.rowc + (size_t)par * RW;
        T rcv[RQ];
        unsigned long long bk = 0;  // (key << 24) | (0xffffff - row): larger is better
        int bb = 0;
        unsigned long long xk0 = 0, xk1 = 0;  // the key words of workgroups lane, lane + 64
        bool aborted = false;
        for (unsigned spin = 0;; ++spin) {
          bool ok = true;
          unsigned long long lk = 0;
          int lb = 0;
          for (int q = lane; q < G; q += 64) {
            const unsigned long long x = ld_ag(&a.keys[(size_t)par * a.gstride + q]);
            if (q == lane) xk0 = x;
            if (q == lane + 64) xk1 = x;
            if (((unsigned)(x >> 24) & 0xffu) != tag8) ok = false;
            const unsigned long long ck = ((x >> 32) << 24) | (0xffffffull - (x & 0xffffffull));
            if (ck > lk) { lk = ck; lb = q; }
          }
#pragma unroll
          for (int q = 0; q < RQ; ++q) {
            const int j = lane + 64 * q;
            rcv[q] = T(0);
            if (j < kb) ok = TW::get(rcw + (size_t)j * NWD, tag32, rcv[q]) && ok;
          }
          if (__all(ok)) {
            bk = lk;
            bb = lb;
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
              const int j = lane + 64 * q;
              if (j < W) s_rc[(j - c0 + W) % W] = rcv[q];
            }
            break;
          }
          if (spin > kSpinLimit || ld_ag32(a.abort_flag)) {
            if (lane == 0) atomicExch(a.abort_flag, 1u);
            aborted = true;
            break;
          }
        }
        int p = INT_MAX;
        bool nz = false, use_rc = false;
        if (!aborted) {
          // the winner by two DPP wave maxima: the |a| key, then the lowest row
          // (0xffffff - row largest) among the lanes holding that key
          const unsigned lkey = (unsigned)(bk >> 24), linv = (unsigned)(bk & 0xffffffull);
          const unsigned kmax = __ockl_wfred_max_u32(lkey);
          const unsigned imax = __ockl_wfred_max_u32(lkey == kmax ? linv : 0u);
          const unsigned long long hit = __ballot(lkey == kmax && linv == imax);
          const int hl = __builtin_amdgcn_readfirstlane((int)__ffsll((long long)hit) - 1);
          int pb = __builtin_amdgcn_readlane(bb, hl);
          p = (int)(0xffffffu - imax);
          if (sizeof(T) == 8) {
            // workgroups tied on the FP32-rounded key: decide on the exact values
            auto tied = [&](unsigned long long x) {
              return (unsigned)(x >> 32) == kmax && (x & 0xffffffull) != 0xffffffull;
            };
            int ntie = 0;
            if (G <= 128) {  // the key words of the last poll
              ntie = __popcll(__ballot(lane < G && tied(xk0))) +
                     __popcll(__ballot(lane + 64 < G && tied(xk1)));
            } else {
              for (int q = lane; q < G; q += 64)
                if (tied(ld_ag(&a.keys[(size_t)par * a.gstride + q]))) ++ntie;
              for (int off = 32; off >= 1; off >>= 1) ntie += __shfl_xor(ntie, off, 64);
            }
            if (ntie > 1) {
              T ev = T(-1);
              int er = INT_MAX, eb = 0;
              for (int q = lane; q < G && !aborted; q += 64) {
                const unsigned long long x = ld_ag(&a.keys[(size_t)par * a.gstride + q]);
                if ((unsigned)(x >> 32) != kmax || (x & 0xffffffull) == 0xffffffull) continue;
                const unsigned long long* wq =
                    a.recs + ((size_t)par * a.gstride + q) * RW + (size_t)c * NWD;
                T val = T(0);
                unsigned spin = 0;
                while (!TW::get(wq, tag32, val)) {
                  if (++spin > kSpinLimit || ld_ag32(a.abort_flag)) { aborted = true; break; }
                }
                const int rq = (int)(x & 0xffffffull);
                if (better(fabs(val), rq, ev, er)) { ev = fabs(val); er = rq; eb = q; }
              }
              if (__any(aborted)) {
                if (lane == 0) atomicExch(a.abort_flag, 1u);
                aborted = true;
              } else {
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                  const T ov = __shfl_xor(ev, off, 64);
                  const int orr = __shfl_xor(er, off, 64);
                  const int ob = __shfl_xor(eb, off, 64);
                  if (better(ov, orr, ev, er)) { ev = ov; er = orr; eb = ob; }
                }
                p = er;
                pb = eb;
              }
            }
          }
          if (!aborted && p == c) {
            use_rc = true;  // the pivot row is row c, read with the keys
          } else if (!aborted) {
            const unsigned long long* wrec = a.recs + ((size_t)par * a.gstride + pb) * RW;
            aborted = !read_record<T, W>(wrec, tag32, kb, c0, s_pr, a.abort_flag);
          }
          if (!aborted) nz = s_rows[use_rc ? 1 : 0][u] != T(0);
        }
        if (lane == 0) {
          s_p = p;
          s_ok = nz ? 1 : 0;
          s_use_rc = use_rc ? 1 : 0;
          if (aborted) s_abort = 1;
          if (!aborted && blk == 0) {
            a.ipiv[a.k0 + c] = a.k0 + (nz ? p : c);
            s_piv[c] = a.k0 + (nz ? p : c);
            if (!nz) atomicCAS(a.info, 0, a.k0 + c + 1);
          }
        }
      }
      __syncthreads();
      if (s_abort) { stop = true; break; }
      // 4. interchange, multipliers, rank-1 update (zero pivot: nothing, as LAPACK)
      if (s_ok) {
        const int p = s_p;
        if (p != c) {
          if (r == p) {
#pragma unroll
            for (int k = 0; k < W; ++k) v[k] = s_rc[k];
          } else if (r == c) {
#pragma unroll
            for (int k = 0; k < W; ++k) v[k] = s_pr[k];
          }
        }
        const T* const prow = s_rows[s_use_rc];  // the pivot row
        const T piv = prow[u];
        const bool below = valid && r > c;
        // LAPACK's scaling: by the reciprocal, or by division below the safe minimum
        const bool recip = fabs(piv) >= std::numeric_limits<T>::min();
        const T rp = T(1) / piv;
        const T l = below ? (recip ? v[u] * rp : v[u] / piv) : T(0);
        if (below) v[u] = l;
        // later columns: registers u+1 .. W-c0-1, in groups of kPU
#pragma unroll
        for (int g0 = 0; g0 < W; g0 += kPU) {
          if (g0 + kPU <= u) continue;  // compile time: no later column in the group
          if (g0 < W - c0) {            // uniform: the group holds later columns
#pragma unroll
            for (int k = g0; k < g0 + kPU; ++k)
              if (k > u) v[k] -= l * prow[k];
          }
        }
      }
    }
    // rotate by kPU: column c0 + kPU into register 0
    if (!stop) {
      T t[kPU];
#pragma unroll
      for (int k = 0; k < kPU; ++k) t[k] = v[k];
#pragma unroll
      for (int k = 0; k < W - kPU; ++k) v[k] = v[k + kPU];
#pragma unroll
      for (int k = 0; k < kPU; ++k) v[W - kPU + k] = t[k];
      rot += kPU;
    }
  }
  // register k holds column (rot + k) mod W
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int j = (rot + k) % W;
    if (valid && j < kb) Ap[(size_t)r + (size_t)j * lda] = v[k];
  }
  // workgroup 0: the panel's interchanges as a gather list (every s_piv entry
  // was written before one of the loop's barriers); none after an abort
  if (blk == 0 && w == 0 && a.list_len) {
    if (stop) {
      if (lane == 0) *a.list_len = 0;
    } else {
      build_gather_list(lane < kb ? s_piv[lane] : 0, lane + 64 < kb ? s_piv[lane + 64] : 0,
                        a.k0, kb, a.list_dst, a.list_src, a.list_len);
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 panels by tournament pivoting: TSLU, the panel factorization of CALU
// (Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32(4), 2011).
//
// The panel's m rows are cut into leaves of kPR = 512 rows. A leaf workgroup
// runs GEPP on its rows in registers, with no exchange between workgroups, and
// keeps the kb rows it pivoted on as candidates. Candidate sets meet in a tree
// of fan-in kTF = 4 (4 x 128 rows: a leaf's size again): a node runs GEPP on
// the ORIGINAL values of its children's candidates and passes its kb pivot
// rows up. The root's GEPP fixes the panel's kb pivot rows in order; its registers
// then hold their factors L11\U11. The tree needs no grid barrier: a workgroup
// that finishes a node writes the candidate set, waits for its stores to be
// acknowledged and counts itself in at the parent; the last child to arrive
// goes on as the parent's worker and the others leave, so the launch works at
// any residency and nothing waits. Candidate words are written and read with
// relaxed agent-scope atomics (coherent across XCDs); the writer's stores are
// acknowledged (s_waitcnt vmcnt(0)) before it counts itself in, so the worker
// that reads the final count reads complete sets.
//
// Pivot choice: the largest |a| among the rows that have not pivoted yet,
// lowest position on ties (a leaf's positions are its rows in order, an inner
// node's its children's candidates in order), so a panel of one leaf
// (m <= 512) is plain GEPP. The interchanges (ipiv, LAPACK order) reach the
// outer block's columns through perm_build / perm_apply, and tslu_apply then
// forms L21 = A21 U11^-1 row by row with U11 in LDS.
// ---------------------------------------------------------------------------
constexpr int kTF = 4;                 // fan-in of the candidate tree
constexpr int kTW = 128;               // FP32 panel width
constexpr unsigned kNoRow = 0xffffffffu;

struct TsluArgs {
  float* A;
  int lda;
  int m;                      // panel rows (global rows k0 .. k0+m-1)
  int kb;                     // panel width (<= kTW)
  int k0;                     // global index of the panel's first row/column
  int G;                      // leaves
  unsigned* cand;             // [node][kTW][kTW]: word j*kTW + i = column j of candidate i
  unsigned* cid;              // [node][kTW]: panel row of candidate i (kNoRow: none)
  unsigned* count;            // [node] arrivals at an inner node (zero between launches)
  float* top;                 // [kTW][kTW]: row c = L11\U11 row of the step-c pivot
  int* ipiv;                  // ipiv[k0 + c] = global row swapped with row k0 + c
  int* info;                  // first zero pivot: global column + 1 (0 if none)
  unsigned* abort_flag;       // raised if a candidate set reads back inconsistent
};

// Tree level l: n_0 = G nodes, n_{l+1} = ceil(n_l / kTF); its first node slot.
__device__ __forceinline__ void tslu_level(int G, int level, int& size, int& off) {
  int s = G, o = 0;
  for (int l = 0; l < level; ++l) {
    o += s;
    s = (s + kTF - 1) / kTF;
  }
  size = s;
  off = o;
}

// Layout of a node's 512 x 128 block in a workgroup: wave w holds columns
// 16w .. 16w+15, lane l rows 8l .. 8l+7 (a[i][j] = row 8l+i, column 16w+j).
// A step of GEPP then needs, per thread, 8 multipliers and 16 pivot-row
// values: the wave owning the step's column picks the pivot (one wave
// reduction) and publishes the multipliers in LDS (one barrier per step,
// double-buffered); each wave reads its 16 pivot-row values from the lane
// holding the pivot row into scalar registers (v_readlane), where the finished
// columns are zeroed by scalar selects.
constexpr int kTR = 8;    // rows per thread
constexpr int kTC = 16;   // columns per thread

struct TsluShared {
  float mult[2][kPR];          // multipliers of the step, by parity
  int piv_pos[2];              // pivot position of the step (-1: no row left)
  int zero[kTW];               // step c had a zero pivot
  unsigned rid[kTW];           // panel row of step c's pivot
  float stage[kTW][kTW + 1];   // originals of the candidates, row = pivot step
  int flag;
};

// a[i][j] of lane `lane` for wave-uniform i < kTR and lane (a scalar branch
// picks the register; no run-time register index). j: a compile-time index
// after unrolling.
__device__ __forceinline__ float tslu_lane_value(const float (&a)[kTR][kTC], int i, int j,
                                                 int lane) {
  int x;
  switch (i) {
    case 0: x = __builtin_amdgcn_readlane(__float_as_int(a[0][j]), lane); break;
    case 1: x = __builtin_amdgcn_readlane(__float_as_int(a[1][j]), lane); break;
    case 2: x = __builtin_amdgcn_readlane(__float_as_int(a[2][j]), lane); break;
    case 3: x = __builtin_amdgcn_readlane(__float_as_int(a[3][j]), lane); break;
    case 4: x = __builtin_amdgcn_readlane(__float_as_int(a[4][j]), lane); break;
    case 5: x = __builtin_amdgcn_readlane(__float_as_int(a[5][j]), lane); break;
    case 6: x = __builtin_amdgcn_readlane(__float_as_int(a[6][j]), lane); break;
    default: x = __builtin_amdgcn_readlane(__float_as_int(a[7][j]), lane); break;
  }
  return __int_as_float(x);
}

// The 16 values of row i of lane `lane` (wave-uniform i, lane) as scalars.
__device__ __forceinline__ void tslu_lane_row(const float (&a)[kTR][kTC], int i, int lane,
                                              float (&pr)[kTC]) {
#define TSLU_ROW(I)                                                                   \
  case I:                                                                             \
    _Pragma("unroll") for (int j = 0; j < kTC; ++j) pr[j] =                           \
        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[I][j]), lane));      \
    break;
  switch (i) {
    TSLU_ROW(0) TSLU_ROW(1) TSLU_ROW(2) TSLU_ROW(3) TSLU_ROW(4) TSLU_ROW(5) TSLU_ROW(6)
    default:
      TSLU_ROW(7)
  }
#undef TSLU_ROW
}

// GEPP for kb steps over the workgroup's rows (live: bit i = row i of the
// thread still in play). stp[i] = step at which row i became a pivot (-1:
// never). On return a[][] holds, for the pivot rows, their L\U factors.
// Returns the number of steps that found a row.
__device__ __forceinline__ int tslu_gepp(float (&a)[kTR][kTC], unsigned& live, int (&stp)[kTR],
                                         int kb, TsluShared& S) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  int nsel = 0;
  int cb = 0;
  for (; cb * kTC < kb; ++cb) {
    if (w < cb) break;  // wave-uniform: this wave's columns are all finished
#pragma unroll
    for (int jc = 0; jc < kTC; ++jc) {
      const int c = cb * kTC + jc;
      const int par = c & 1;
      if (w == cb) {
        // the wave owning column c: pivot = largest |a| over the live rows,
        // lowest row on ties (|a| bits order like |a|, +1 so a zero beats none)
        unsigned kbest = 0u;
        int ibest = 0;
#pragma unroll
        for (int i = 0; i < kTR; ++i) {
          const unsigned k =
              (((live >> i) & 1u) && c < kb) ? __float_as_uint(fabsf(a[i][jc])) + 1u : 0u;
          if (k > kbest) { kbest = k; ibest = i; }
        }
        const unsigned kmax = __ockl_wfred_max_u32(kbest);
        const unsigned long long hit = __ballot(kbest == kmax && kmax != 0u);
        const int lw = hit ? (int)__ffsll((long long)hit) - 1 : 0;
        const int iw = __builtin_amdgcn_readlane(ibest, lw);
        const float piv = tslu_lane_value(a, iw, jc, lw);
        const bool nz = kmax != 0u && piv != 0.f;
        const float rp = nz ? 1.f / piv : 0.f;
        float mi[kTR];
#pragma unroll
        for (int i = 0; i < kTR; ++i) {
          const bool upd = nz && ((live >> i) & 1u) && !(lane == lw && i == iw);
          mi[i] = upd ? a[i][jc] * rp : 0.f;
        }
        if (nz && !(fabsf(piv) >= 1.17549435e-38f)) {
          // LAPACK divides instead below the safe minimum
#pragma unroll
          for (int i = 0; i < kTR; ++i) mi[i] = mi[i] != 0.f ? a[i][jc] / piv : 0.f;
        }
#pragma unroll
        for (int i = 0; i < kTR; ++i) a[i][jc] = mi[i] != 0.f ? mi[i] : a[i][jc];
        float4* md = reinterpret_cast<float4*>(&S.mult[par][kTR * lane]);
        md[0] = make_float4(mi[0], mi[1], mi[2], mi[3]);
        md[1] = make_float4(mi[4], mi[5], mi[6], mi[7]);
        if (lane == 0) {
          S.piv_pos[par] = kmax ? kTR * lw + iw : -1;
          if (c < kb) S.zero[c] = nz ? 0 : 1;
        }
      }
      __syncthreads();
      const int p = __builtin_amdgcn_readfirstlane(S.piv_pos[par]);
      const float4* md = reinterpret_cast<const float4*>(&S.mult[par][kTR * lane]);
      const float4 m0 = md[0], m1 = md[1];
      const float mi[kTR] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
      float pr[kTC];
      if (p >= 0) {
        const int lw = p >> 3, iw = p & 7;
        nsel += 1;
#pragma unroll
        for (int i = 0; i < kTR; ++i) stp[i] = (lane == lw && i == iw) ? c : stp[i];
        live &= (lane == lw) ? ~(1u << iw) : ~0u;
        tslu_lane_row(a, iw, lw, pr);
      } else {
#pragma unroll
        for (int j = 0; j < kTC; ++j) pr[j] = 0.f;
      }
      // rank-1 update; the owner's columns up to c take pivot-row value 0
      // (scalar selects; waves left of the owner have left the loop)
      const bool owner = w == cb;
#pragma unroll
      for (int j = 0; j < kTC; ++j) {
        const float prj = (owner && j <= jc) ? 0.f : pr[j];
#pragma unroll
        for (int i = 0; i < kTR; ++i) a[i][j] -= mi[i] * prj;
      }
    }
  }
  // A wave whose columns are all finished only keeps step with the others:
  // one barrier per remaining step and the pivot bookkeeping.
  const int nsteps = (kb + kTC - 1) / kTC * kTC;
  for (int c = cb * kTC; c < nsteps; ++c) {
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(S.piv_pos[c & 1]);
    if (p >= 0) {
      const int lw = p >> 3, iw = p & 7;
      nsel += 1;
#pragma unroll
      for (int i = 0; i < kTR; ++i) stp[i] = (lane == lw && i == iw) ? c : stp[i];
      live &= (lane == lw) ? ~(1u << iw) : ~0u;
    }
  }
  __syncthreads();
  return nsel;
}

// Candidate ci of the set at `slot`: its panel row, and its values in this
// wave's columns. The set's words were acknowledged before its writer counted
// itself in at the parent, and this workgroup read that count, so plain
// agent-scope loads see them.
__device__ __forceinline__ unsigned tslu_read_cand(const TsluArgs& a, size_t slot, int ci, int w,
                                                   float (&r)[kTC]) {
  unsigned rid = ld_ag32(a.cid + slot * kTW + ci);
  if (rid != kNoRow && rid >= (unsigned)a.m) {  // never for a consistent set
    atomicExch(a.abort_flag, 1u);
    rid = kNoRow;
  }
  const unsigned* cand = a.cand;
  asm volatile("" : "+s"(cand));  // no addresses kept live across calls
  const unsigned* dw = cand + slot * (size_t)kTW * kTW + ci;
#pragma unroll
  for (int j = 0; j < kTC; ++j) {
    const int col = kTC * w + j;
    r[j] = (rid != kNoRow && col < a.kb) ? __uint_as_float(ld_ag32(dw + (size_t)col * kTW)) : 0.f;
  }
  return rid;
}

// Row i of the thread (position 8 lane + i of node `node` at `level`) in this
// wave's columns: a leaf's row node * 512 + position of the panel Ap, an inner
// node's candidate (position & 127) of child (position >> 7). Returns its
// panel row (kNoRow: none).
__device__ __forceinline__ unsigned tslu_load_row(const TsluArgs& a, const float* Ap, int level,
                                                  int node, int i, float (&r)[kTC]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int pos = kTR * lane + i;
  asm volatile("" : "+s"(Ap));  // a fresh base per call: no addresses kept live
  if (level == 0) {
    const int row = node * kPR + pos;
    const unsigned rr = row < a.m ? (unsigned)row : kNoRow;
    const size_t lda = (size_t)a.lda;
#pragma unroll
    for (int j = 0; j < kTC; ++j) {
      const int col = kTC * w + j;
      r[j] = (rr != kNoRow && col < a.kb) ? Ap[(size_t)row + (size_t)col * lda] : 0.f;
    }
    return rr;
  }
  int np, op;
  tslu_level(a.G, level - 1, np, op);
  const int q = pos >> 7;
  if (kTF * node + q < np)
    return tslu_read_cand(a, (size_t)op + (size_t)kTF * node + q, pos & (kTW - 1), w, r);
#pragma unroll
  for (int j = 0; j < kTC; ++j) r[j] = 0.f;
  return kNoRow;
}

__global__ __launch_bounds__(kPR, 1) void tslu_kernel(TsluArgs a) {
  __shared__ __align__(16) TsluShared S;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const float* Ap = a.A + (size_t)a.k0 + (size_t)a.k0 * (size_t)a.lda;
  float v[kTR][kTC];
  unsigned rid[kTR];

  int level = 0, node = blockIdx.x;
  for (;;) {
    int nlev, off;
    tslu_level(a.G, level, nlev, off);
    unsigned live = 0u;
#pragma unroll
    for (int i = 0; i < kTR; ++i) {
      rid[i] = tslu_load_row(a, Ap, level, node, i, v[i]);
      live |= (rid[i] != kNoRow) ? (1u << i) : 0u;
    }
    int stp[kTR];
#pragma unroll
    for (int i = 0; i < kTR; ++i) stp[i] = -1;
    const int nsel = tslu_gepp(v, live, stp, a.kb, S);
    if (w == 0) {
#pragma unroll
      for (int i = 0; i < kTR; ++i)
        if (stp[i] >= 0) S.rid[stp[i]] = rid[i];
    }
    if (nlev == 1) {
      // root: the pivot rows' L11\U11 (row-major, row c = step c)
#pragma unroll
      for (int i = 0; i < kTR; ++i) {
        if (stp[i] >= 0) {
#pragma unroll
          for (int j = 0; j < kTC; ++j) {
            const int col = kTC * w + j;
            if (col < a.kb) a.top[(size_t)stp[i] * kTW + col] = v[i][j];
          }
        }
      }
      __syncthreads();
      if (tid < 64) {
        // LAPACK-order interchanges: position c takes the step-c pivot row.
        // Positions >= c of the first kb hold original top rows; lane l keeps
        // where[x] (current position of top row x) and topc[p] (top row at
        // position p) for x, p in {l, l + 64}.
        int wh0 = lane, wh1 = lane + 64, tc0 = lane, tc1 = lane + 64;
        if (nsel < a.kb && lane == 0) atomicExch(a.abort_flag, 1u);  // m >= kb rows
        for (int c = 0; c < a.kb; ++c) {
          const int wr = c < nsel ? (int)S.rid[c] : c;
          int pos;
          if (wr >= a.kb) {
            pos = wr;
          } else {
            const int t0 = __shfl(wh0, wr & 63, 64), t1 = __shfl(wh1, wr & 63, 64);
            pos = wr < 64 ? t0 : t1;
          }
          const int x0 = __shfl(tc0, c & 63, 64), x1 = __shfl(tc1, c & 63, 64);
          const int x = c < 64 ? x0 : x1;
          if (lane == 0) a.ipiv[a.k0 + c] = a.k0 + pos;
          if (x < 64) { if (lane == x) wh0 = pos; } else { if (lane == x - 64) wh1 = pos; }
          if (pos < a.kb) {
            if (pos < 64) { if (lane == pos) tc0 = x; } else { if (lane == pos - 64) tc1 = x; }
          }
        }
        if (lane == 0) {
          for (int c = 0; c < a.kb; ++c)
            if (S.zero[c]) {
              atomicCAS(a.info, 0, a.k0 + c + 1);
              break;
            }
        }
      }
      return;
    }
    // the originals of this node's candidates, in pivot order, into LDS
#pragma unroll
    for (int i = 0; i < kTR; ++i) {
      if (stp[i] >= 0) {
        float r[kTC];
        tslu_load_row(a, Ap, level, node, i, r);
#pragma unroll
        for (int j = 0; j < kTC; ++j) S.stage[stp[i]][kTC * w + j] = r[j];
      }
    }
    __syncthreads();
    const size_t slot = (size_t)off + node;
    unsigned* cw = a.cand + slot * (size_t)kTW * kTW;
    for (int wd = tid; wd < kTW * kTW; wd += kPR) {
      const int j = wd >> 7, ci = wd & (kTW - 1);
      if (ci < nsel && j < a.kb) st_ag32(cw + wd, __float_as_uint(S.stage[ci][j]));
    }
    if (tid < kTW) st_ag32(a.cid + slot * kTW + tid, tid < nsel ? S.rid[tid] : kNoRow);
    // every word acknowledged before the arrival is counted
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int nnext, onext;
    tslu_level(a.G, level + 1, nnext, onext);
    const int parent = node / kTF;
    const int nchild = min(kTF, nlev - kTF * parent);
    unsigned* cnt = a.count + onext + parent;
    if (tid == 0)
      S.flag = (int)__hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (S.flag != nchild - 1) return;  // a sibling is still running: it goes on
    if (tid == 0) __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // go on as the parent (its children's candidate sets)
    level += 1;
    node = parent;
  }
}

// Panel rows [kb, m) after the interchanges: L21 = A21 U11^-1 (thread = row,
// U11 and the reciprocals of its diagonal in LDS); workgroup 0 also writes
// L11\U11 into rows [0, kb). Multipliers take LAPACK's scaling (by the
// reciprocal, or by division below the safe minimum); a zero pivot leaves its
// column as is (the factorization reports it through info).
constexpr int kApplyThreads = 256;
__global__ __launch_bounds__(kApplyThreads) void tslu_apply(float* __restrict__ A, int lda, int m,
                                                            int kb, int k0,
                                                            const float* __restrict__ top) {
  __shared__ __align__(16) float sU[kTW][kTW];
  __shared__ float sR[kTW];
  const int tid = threadIdx.x;
  const size_t ld = (size_t)lda;
  float* Ap = A + (size_t)k0 + (size_t)k0 * ld;
  for (int idx = tid; idx < kb * kTW; idx += kApplyThreads) sU[idx / kTW][idx % kTW] = top[idx];
  __syncthreads();
  for (int c = tid; c < kb; c += kApplyThreads) {
    const float piv = sU[c][c];
    sR[c] = fabsf(piv) >= 1.17549435e-38f ? 1.f / piv : 0.f;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int idx = tid; idx < kb * kb; idx += kApplyThreads) {
      const int col = idx / kb, row = idx - col * kb;
      Ap[(size_t)row + (size_t)col * ld] = sU[row][col];
    }
  }
  const int r = kb + blockIdx.x * kApplyThreads + tid;
  if (r >= m) return;
  float v[kTW];
#pragma unroll
  for (int j = 0; j < kTW; ++j) v[j] = (j < kb) ? Ap[(size_t)r + j * ld] : 0.f;
  int rot = 0;  // register k holds column (rot + k) mod kTW
  for (int c0 = 0; c0 < kb; c0 += kPU) {
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int c = c0 + u;
      if (c >= kb) break;
      const float piv = sU[c][c];
      if (piv != 0.f) {
        const float rp = sR[c];
        const float l = rp != 0.f ? v[u] * rp : v[u] / piv;
        v[u] = l;
#pragma unroll
        for (int g0 = 0; g0 < kTW; g0 += kPU) {
          if (g0 + kPU <= u) continue;
          if (g0 < kTW - c0) {
#pragma unroll
            for (int k = g0; k < g0 + kPU; ++k)
              if (k > u) v[k] -= l * sU[c][(c0 + k) & (kTW - 1)];
          }
        }
      }
    }
    float t[kPU];
#pragma unroll
    for (int k = 0; k < kPU; ++k) t[k] = v[k];
#pragma unroll
    for (int k = 0; k < kTW - kPU; ++k) v[k] = v[k + kPU];
#pragma unroll
    for (int k = 0; k < kPU; ++k) v[kTW - kPU + k] = t[k];
    rot += kPU;
  }
#pragma unroll
  for (int k = 0; k < kTW; ++k) {
    const int j = (rot + k) & (kTW - 1);
    if (j < kb) Ap[(size_t)r + (size_t)j * ld] = v[k];
  }
}

// ---------------------------------------------------------------------------
// Row interchanges of one panel as a gather list (dst[t] <- src[t]): built by
// build_gather_list (above), at the end of the GEPP panel kernel or by
// perm_build_fast after a tournament panel.
// ---------------------------------------------------------------------------
constexpr int kMaxPanel = 256;

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
// C -= A B for 16-byte aligned operands: the trailing update of the LU.
//
// Workgroup tile BM x BN, NWM x NWN waves, each wave (BM/NWM) x (BN/NWN) built
// from 16x16 MFMA blocks; K in steps of BK = 16 through a double-buffered LDS
// stage. The accumulators start at zero and the epilogue subtracts them from
// C (read once, written once per launch).
//
// LDS layouts: A as [k][m] (m contiguous: 16-byte stores from the column-major
// source, conflict-free 4-byte fragment reads), B as [n][k] (k contiguous: the
// four k of one MFMA lane come in one 16-byte read). Within a K-step of 16 the
// MFMA k-slot q of a lane (lane / 16) takes source k = 4q + s at sub-step s,
// identically for A and B, so each product term appears once.
//
// Tile order: workgroups are dealt to the 8 XCDs round-robin; block b is
// mapped to tile xcd_chunk(b) so each XCD walks a contiguous range of the
// linear tile order, which runs row by row through groups of 4 tile columns
// (A and B blocks shared in the XCD's L2).
// ---------------------------------------------------------------------------
// SET = false: C -= A B. SET = true: C = A B, where B may be C itself when
// one tile row covers all M rows (each workgroup reads its whole column strip
// of B before it writes the same strip of C).
template <typename T, int BM, int BN, int NWM, int NWN, bool SET = false>
__global__ __launch_bounds__(NWM * NWN * 64) void gemm_tile_kernel(
    int M, int N, int K, const T* A, int lda, const T* B, int ldb, T* C, int ldc, int tiles_m,
    int ntiles) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int NT = NWM * NWN * 64;
  constexpr int BK = 16;
  constexpr int VEC = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  constexpr int WTM = BM / NWM, WTN = BN / NWN;
  constexpr int TI = WTM / 16, TJ = WTN / 16;
  constexpr int PADA = VEC, PADB = VEC;
  constexpr int LA = BM * BK / VEC / NT, LB = BN * BK / VEC / NT;
  static_assert(LA * VEC * NT == BM * BK && LB * VEC * NT == BN * BK, "tile/thread mismatch");
  __shared__ __align__(16) T As[2][BK][BM + PADA];
  __shared__ __align__(16) T Bs[2][BN][BK + PADB];

  // XCD-chunked tile index
  const int bid = blockIdx.x;
  int tile = bid;
  if (ntiles >= 8) {
    const int q = ntiles >> 3, rem = ntiles & 7, x = bid & 7, j = bid >> 3;
    tile = (x < rem) ? x * (q + 1) + j : rem * (q + 1) + (x - rem) * q + j;
  }
  // Linear tile order in groups of kGroupN tile columns, row by row within a
  // group: the ~32 tiles an XCD runs at once cover 8 tile rows x 4 tile
  // columns, so its L2 serves each A block to 4 tiles and each B block to 8
  // (a single tile column would re-read every A block from HBM).
  constexpr int kGroupN = 4;
  const int tiles_n = (ntiles + tiles_m - 1) / tiles_m;
  const int group = tile / (tiles_m * kGroupN);
  const int first_n = group * kGroupN;
  const int gw = min(kGroupN, tiles_n - first_n);
  const int gidx = tile - group * tiles_m * kGroupN;
  const int bm = (gidx / gw) * BM, bn = (first_n + gidx % gw) * BN;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave % NWM) * WTM, wn = (wave / NWM) * WTN;
  const int l16 = lane & 15, q4 = lane >> 4;

  // Lane element (ti, tj, r) of the tile sits at row bm + wm + 16 ti + l16 and
  // column bn + wn + 16 tj + cofs(tj, r): a per-lane base plus a wave-uniform
  // offset.
  const bool interior = (bm + BM <= M) && (bn + BN <= N);
  const int lane_col = (sizeof(T) == 4) ? q4 * 4 : q4;
  T* Cl = C + (size_t)(bm + wm + l16) + (size_t)(bn + wn + lane_col) * ldc;
  auto cofs = [](int tj, int r) { return tj * 16 + ((sizeof(T) == 4) ? r : r * 4); };
  acc_t acc[TI][TJ];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) acc[ti][tj] = acc_t{0, 0, 0, 0};

  // Main loop, instantiated twice: CHK = false for tiles whose every A/B load is
  // in bounds (no per-element tests, so nothing serializes the prefetch of the
  // next K-step behind the current one), CHK = true for edge tiles.
  auto mainloop = [&](auto chk) {
    constexpr bool CHK = decltype(chk)::value;
    vec_t ra[LA], rb[LB];
    auto gload = [&](int kt) {
#pragma unroll
      for (int u = 0; u < LA; ++u) {
        const int v = tid + u * NT;
        const int m4 = (v % (BM / VEC)) * VEC, k = v / (BM / VEC);
        const int r = bm + m4, kk = kt + k;
        const T* src = A + (size_t)r + (size_t)kk * lda;
        if (!CHK || (kk < K && r + VEC <= M)) {
          ra[u] = *reinterpret_cast<const vec_t*>(src);
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e) ra[u][e] = (kk < K && r + e < M) ? src[e] : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < LB; ++u) {
        const int v = tid + u * NT;
        const int k4 = (v % (BK / VEC)) * VEC, n = v / (BK / VEC);
        const int cb = bn + n, kk = kt + k4;
        const T* src = B + (size_t)kk + (size_t)cb * ldb;
        if (!CHK || (cb < N && kk + VEC <= K)) {
          rb[u] = *reinterpret_cast<const vec_t*>(src);
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e) rb[u][e] = (cb < N && kk + e < K) ? src[e] : T(0);
        }
      }
    };
    auto sstore = [&](int buf) {
#pragma unroll
      for (int u = 0; u < LA; ++u) {
        const int v = tid + u * NT;
        const int m4 = (v % (BM / VEC)) * VEC, k = v / (BM / VEC);
        *reinterpret_cast<vec_t*>(&As[buf][k][m4]) = ra[u];
      }
#pragma unroll
      for (int u = 0; u < LB; ++u) {
        const int v = tid + u * NT;
        const int k4 = (v % (BK / VEC)) * VEC, n = v / (BK / VEC);
        *reinterpret_cast<vec_t*>(&Bs[buf][n][k4]) = rb[u];
      }
    };

    gload(0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int kt = 0; kt < K; kt += BK) {
      const bool more = kt + BK < K;
      if (more) gload(kt + BK);
      T fb[TJ][4];
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj) {
        const T* p = &Bs[buf][wn + tj * 16 + l16][q4 * 4];
#pragma unroll
        for (int h = 0; h < 4; h += VEC) {
          const vec_t x = *reinterpret_cast<const vec_t*>(p + h);
#pragma unroll
          for (int e = 0; e < VEC; ++e) fb[tj][h + e] = x[e];
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        T fa[TI];
#pragma unroll
        for (int ti = 0; ti < TI; ++ti) fa[ti] = As[buf][q4 * 4 + s][wm + ti * 16 + l16];
#pragma unroll
        for (int ti = 0; ti < TI; ++ti)
#pragma unroll
          for (int tj = 0; tj < TJ; ++tj)
            acc[ti][tj] = Mfma<T>::op(fb[tj][s], fa[ti], acc[ti][tj]);
      }
      if (more) {
        sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
      }
    }
  };
  if (interior && (K % BK) == 0)
    mainloop(std::false_type{});
  else
    mainloop(std::true_type{});

  // C -= acc: the product is summed on its own and subtracted with one
  // rounding (as the reference GEMM update does), one MFMA row block of C at a
  // time with the next block's loads in flight.
  if constexpr (SET) {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
      const int i = bm + wm + ti * 16 + l16;
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = bn + wn + lane_col + cofs(tj, r);
          if (interior || (i < M && j < N)) Cl[ti * 16 + (size_t)cofs(tj, r) * ldc] = acc[ti][tj][r];
        }
    }
  } else if (interior) {
    T cv[2][TJ][4];
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[0][tj][r] = Cl[(size_t)cofs(tj, r) * ldc];
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
      if (ti + 1 < TI) {
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            cv[(ti + 1) & 1][tj][r] = Cl[(ti + 1) * 16 + (size_t)cofs(tj, r) * ldc];
      }
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Cl[ti * 16 + (size_t)cofs(tj, r) * ldc] = cv[ti & 1][tj][r] - acc[ti][tj][r];
    }
  } else {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
      const int i = bm + wm + ti * 16 + l16;
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = bn + wn + lane_col + cofs(tj, r);
          if (i < M && j < N) Cl[ti * 16 + (size_t)cofs(tj, r) * ldc] -= acc[ti][tj][r];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// C -= A B for FP32 operands on the bf16 matrix cores. Each operand value x is
// split exactly into three bf16 parts, x = h + m + l: h = bf16(x) (round to
// nearest), m = bf16(x - h), l = x - h - m (both differences are exact in FP32,
// and l fits a bf16), so |m| <= 2^-8 |x| and |l| <= 2^-16 |x|. Of the nine
// partial products the six of order down to 2^-16 |a b| are kept,
//   a b ~ hh + hm + mh + hl + lh + mm,
// and the dropped ml + lm + ll stay within 2^-23 (1 + 2^-9) |a b| (each term is
// at most 2^-8 |m_a m_b| <= 2^-24 |a b|) with no bias in sign:
// FP32's unit roundoff is 2^-24 and the matrix cores sum in FP32, so the result
// meets the error bound of an FP32 GEMM. Per 16 x 16 x 16 block that is 3
// v_mfma_f32_16x16x32_bf16 (16 times the flop rate of v_mfma_f32_16x16x4f32)
// in place of 4 FP32 MFMAs: 2.67x less matrix-core time.
//
// The MFMA takes 8 bf16 per lane and operand: lane (l16, g) supplies row l16 at
// k-slots 8g..8g+7, the same slots in both operands. A K-step covers 16 source
// k: lanes g = 0, 1 carry one pair of parts and lanes g = 2, 3 another, each
// over the k-group g & 1 (source k = 8 (g & 1) + e for element e), so each
// instruction sums two partial products (a runs over the columns of C, b over
// its rows, as in gemm_tile_kernel):
//   I1: a = [B_h | B_m], b = [A_h | A_h]  ->  hh + (B_m A_h)
//   I2: a = [B_h | B_l], b = [A_m | A_h]  ->  (B_h A_m) + (B_l A_h)
//   I3: a = [B_h | B_m], b = [A_l | A_m]  ->  (B_h A_l) + mm
//
// LDS per stage and operand: [part 3][k-group 2][row 256] chunks of 8 bf16 (16
// bytes), each (part, k-group) region padded by 128 bytes; a fragment is one
// 16-byte read with 16 lanes on 16 consecutive rows. Global loads: A (column-
// major, rows contiguous) as dwords, one row and one k-group per thread (a wave
// reads 256 contiguous bytes per k); B (k contiguous) as two 16-byte loads, one
// column and one k-group per thread. The parts are formed in registers once per
// element and stored with 16-byte writes, double-buffered (99 KB: one
// workgroup of 8 waves per CU). Tile order, wave tiles (64 x 128) and the
// epilogue are those of gemm_tile_kernel<float, 256, 256, 4, 2>.
constexpr int kSplitRows = 256;
constexpr int kSplitRegion = kSplitRows + 8;  // 16-byte chunks per (part, k-group)
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// x0, x1 -> the packed bf16 pairs of the three parts (x0 in the low half)
__device__ __forceinline__ void split3_pair(float x0, float x1, unsigned& h, unsigned& m,
                                            unsigned& l) {
  const bf16x2_t h2 = {(__bf16)x0, (__bf16)x1};
  const float r0 = x0 - (float)h2[0], r1 = x1 - (float)h2[1];
  const bf16x2_t m2 = {(__bf16)r0, (__bf16)r1};
  const float s0 = r0 - (float)m2[0], s1 = r1 - (float)m2[1];
  const bf16x2_t l2 = {(__bf16)s0, (__bf16)s1};
  h = __builtin_bit_cast(unsigned, h2);
  m = __builtin_bit_cast(unsigned, m2);
  l = __builtin_bit_cast(unsigned, l2);
}

__global__ __launch_bounds__(512) void gemm_split_kernel(int M, int N, int K,
                                                         const float* __restrict__ A, int lda,
                                                         const float* __restrict__ B, int ldb,
                                                         float* __restrict__ C, int ldc,
                                                         int tiles_m, int ntiles) {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  constexpr int BM = kSplitRows, BN = kSplitRows, BK = 16, NWM = 4;
  constexpr int WTM = 64, WTN = 128, TI = WTM / 16, TJ = WTN / 16;
  constexpr int R = kSplitRegion;
  // [buffer][part][k-group][row]
  __shared__ uint4 As[2][3][2][R];
  __shared__ uint4 Bs[2][3][2][R];

  // XCD-chunked tile index, groups of 4 tile columns (as gemm_tile_kernel)
  const int bid = blockIdx.x;
  int tile = bid;
  if (ntiles >= 8) {
    const int q = ntiles >> 3, rem = ntiles & 7, x = bid & 7, j = bid >> 3;
    tile = (x < rem) ? x * (q + 1) + j : rem * (q + 1) + (x - rem) * q + j;
  }
  constexpr int kGroupN = 4;
  const int tiles_n = (ntiles + tiles_m - 1) / tiles_m;
  const int group = tile / (tiles_m * kGroupN);
  const int first_n = group * kGroupN;
  const int gw = min(kGroupN, tiles_n - first_n);
  const int gidx = tile - group * tiles_m * kGroupN;
  const int bm = (gidx / gw) * BM, bn = (first_n + gidx % gw) * BN;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave % NWM) * WTM, wn = (wave / NWM) * WTN;
  const int l16 = lane & 15, q4 = lane >> 4;
  const int kg = q4 & 1, half = q4 >> 1;
  // parts read by this lane: a-operand (B) F1 = h|m, F2 = h|l; b-operand (A)
  // E1 = h|h, E2 = m|h, E3 = l|m (left of the bar: lanes g = 0, 1)
  const int pF1 = half ? 1 : 0, pF2 = half ? 2 : 0;
  const int pE2 = half ? 0 : 1, pE3 = half ? 1 : 2;

  const bool interior = (bm + BM <= M) && (bn + BN <= N);
  const int lane_col = q4 * 4;
  float* Cl = C + (size_t)(bm + wm + l16) + (size_t)(bn + wn + lane_col) * ldc;
  acc_t acc[TI][TJ];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) acc[ti][tj] = acc_t{0, 0, 0, 0};

  // loader roles: A row am, k-group akg; B column bnn, k-group bkg
  const int am = tid & (BM - 1), akg = tid >> 8;
  const int bnn = tid >> 1, bkg = tid & 1;
  const float* Ap = A + (size_t)(bm + am) + (size_t)(8 * akg) * lda;
  const float* Bp = B + (size_t)(8 * bkg) + (size_t)(bn + bnn) * ldb;

  auto mainloop = [&](auto chk) {
    constexpr bool CHK = decltype(chk)::value;
    float ra[8], rb[8];
    auto gload = [&](int kt) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int kk = kt + 8 * akg + e;
        if (!CHK || (kk < K && bm + am < M))
          ra[e] = Ap[(size_t)(kt + e) * lda];
        else
          ra[e] = 0.f;
      }
      const int k0 = kt + 8 * bkg;
      if (!CHK || (bn + bnn < N && k0 + 8 <= K)) {
        const float4 u = *reinterpret_cast<const float4*>(Bp + kt);
        const float4 v = *reinterpret_cast<const float4*>(Bp + kt + 4);
        rb[0] = u.x; rb[1] = u.y; rb[2] = u.z; rb[3] = u.w;
        rb[4] = v.x; rb[5] = v.y; rb[6] = v.z; rb[7] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) rb[e] = (bn + bnn < N && k0 + e < K) ? Bp[kt + e] : 0.f;
      }
    };
    // parts of 8 values (element e in bf16 slot e) into the chunk of row r
    auto put = [&](uint4 (*X)[2][kSplitRegion], int g, int r, const unsigned (&hw)[4],
                   const unsigned (&mw)[4], const unsigned (&lw)[4]) {
      X[0][g][r] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
      X[1][g][r] = make_uint4(mw[0], mw[1], mw[2], mw[3]);
      X[2][g][r] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    };
    auto sstore = [&](int buf) {
      unsigned hw[4], mw[4], lw[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) split3_pair(ra[2 * p], ra[2 * p + 1], hw[p], mw[p], lw[p]);
      put(As[buf], akg, am, hw, mw, lw);
#pragma unroll
      for (int p = 0; p < 4; ++p) split3_pair(rb[2 * p], rb[2 * p + 1], hw[p], mw[p], lw[p]);
      put(Bs[buf], bkg, bnn, hw, mw, lw);
    };

    // One K-step on buffer buf: sweep 1 (I1 and I3, sharing the a-operand F1),
    // then sweep 2 (I2). With STORE, the parts of the next K-step (loaded into
    // ra, rb before the step) are formed and written to buffer nbuf during
    // sweep 2. The step is branch-free, so the scheduling groups at its end
    // order all of it: each a-fragment read one 16-column block ahead of its
    // MFMAs, the split's VALU work spread between sweep 2's MFMAs.
    auto step = [&](int buf, int nbuf, auto store) {
      constexpr bool STORE = decltype(store)::value;
      bf16x8_t e1[TI], e3[TI], e2[TI], f[2];
#pragma unroll
      for (int ti = 0; ti < TI; ++ti) {
        const int r = wm + ti * 16 + l16;
        e1[ti] = __builtin_bit_cast(bf16x8_t, As[buf][0][kg][r]);
        e3[ti] = __builtin_bit_cast(bf16x8_t, As[buf][pE3][kg][r]);
      }
      f[0] = __builtin_bit_cast(bf16x8_t, Bs[buf][pF1][kg][wn + l16]);
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj) {
        if (tj + 1 < TJ)
          f[(tj + 1) & 1] =
              __builtin_bit_cast(bf16x8_t, Bs[buf][pF1][kg][wn + (tj + 1) * 16 + l16]);
#pragma unroll
        for (int ti = 0; ti < TI; ++ti)
          acc[ti][tj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(f[tj & 1], e1[ti], acc[ti][tj], 0, 0, 0);
#pragma unroll
        for (int ti = 0; ti < TI; ++ti)
          acc[ti][tj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(f[tj & 1], e3[ti], acc[ti][tj], 0, 0, 0);
      }
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
        e2[ti] = __builtin_bit_cast(bf16x8_t, As[buf][pE2][kg][wm + ti * 16 + l16]);
      f[0] = __builtin_bit_cast(bf16x8_t, Bs[buf][pF2][kg][wn + l16]);
      // sweep 2, column block tj: its 4 MFMAs, then (STORE) a quarter of one
      // operand's split (tj < 4: A, else B) and, after the last quarter, that
      // operand's 3 LDS writes, fenced so that each slice of VALU work issues
      // while the block's MFMAs run
      unsigned hw[4], mw[4], lw[4];
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj) {
        if (tj + 1 < TJ)
          f[(tj + 1) & 1] =
              __builtin_bit_cast(bf16x8_t, Bs[buf][pF2][kg][wn + (tj + 1) * 16 + l16]);
#pragma unroll
        for (int ti = 0; ti < TI; ++ti)
          acc[ti][tj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(f[tj & 1], e2[ti], acc[ti][tj], 0, 0, 0);
        if constexpr (STORE) {
          const int p = tj & 3;
          if (tj < 4)
            split3_pair(ra[2 * p], ra[2 * p + 1], hw[p], mw[p], lw[p]);
          else
            split3_pair(rb[2 * p], rb[2 * p + 1], hw[p], mw[p], lw[p]);
          if (tj == 3) put(As[nbuf], akg, am, hw, mw, lw);
          if (tj == 7) put(Bs[nbuf], bkg, bnn, hw, mw, lw);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    };

    gload(0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    const int nsteps = (K + BK - 1) / BK;
    for (int st = 0; st + 1 < nsteps; ++st) {
      gload((st + 1) * BK);
      step(buf, buf ^ 1, std::true_type{});
      __syncthreads();
      buf ^= 1;
    }
    step(buf, 0, std::false_type{});
  };
  if (interior && (K % BK) == 0)
    mainloop(std::false_type{});
  else
    mainloop(std::true_type{});

  // C -= acc (as gemm_tile_kernel): one subtraction per entry, the next MFMA
  // row block of C loaded while the current one is written
  if (interior) {
    float cv[2][TJ][4];
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[0][tj][r] = Cl[(size_t)(tj * 16 + r) * ldc];
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
      if (ti + 1 < TI) {
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            cv[(ti + 1) & 1][tj][r] = Cl[(ti + 1) * 16 + (size_t)(tj * 16 + r) * ldc];
      }
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Cl[ti * 16 + (size_t)(tj * 16 + r) * ldc] = cv[ti & 1][tj][r] - acc[ti][tj][r];
    }
  } else {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
      const int i = bm + wm + ti * 16 + l16;
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = bn + wn + lane_col + tj * 16 + r;
          if (i < M && j < N) Cl[ti * 16 + (size_t)(tj * 16 + r) * ldc] -= acc[ti][tj][r];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Triangular solves with FP64 accumulation (factors in T), blocked by kSolveNB
// with the inverses of the diagonal blocks precomputed once per factorization.
// ---------------------------------------------------------------------------
constexpr int kSolveNB = 128;
// Outer block of the LU (K of the trailing update) in 128-column blocks.
constexpr int kOuterBlocks = 4;

__global__ void gather_kernel(int n, const int* __restrict__ perm, const double* __restrict__ b,
                              double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = b[perm[i]];
}

// Inverse of diagonal block b of L (unit lower) or U (upper): kTriInvGroups
// workgroups per block, each building kTriInvCols columns of X = inverse.
// Column c is built row by row by a group of kTriInvLanes lanes (one wave
// holds four columns): lane p sums the terms k = p mod kTriInvLanes of the row's
// dot product in FP64 over two accumulators, the group adds its partial sums
// by lane shuffles, and lane 0 stores the entry in LDS (a wavefront's LDS
// operations complete in order, so the group's next row sees it). A column's
// chain starts at its diagonal: X(i, c) = 0 for i < c (L) or i > c (U). X is
// kept transposed (Xt[c][k], rows padded by 16 bytes); the block of the factor
// is read from an LDS copy for FP32 (row-major) and from memory for FP64.
constexpr int kTriInvThreads = 512;
constexpr int kTriInvLanes = 16;
constexpr int kTriInvCols = kTriInvThreads / kTriInvLanes;  // columns per workgroup
constexpr int kTriInvGroups = kSolveNB / kTriInvCols;       // workgroups per block
constexpr size_t tri_inv_lds(size_t elem) {
  return (size_t)kTriInvCols * (kSolveNB + 16 / elem) * elem +
         (elem == 4 ? (size_t)kSolveNB * kSolveNB * elem : 0);
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(kTriInvThreads) void tri_inv_kernel(int n, const T* __restrict__ F,
                                                                int lda, T* __restrict__ Dinv) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int NB = kSolveNB;
  constexpr int LP = kTriInvLanes;
  constexpr int XP = NB + 16 / (int)sizeof(T);
  constexpr bool kStage = sizeof(T) == 4;
  T* Xt = reinterpret_cast<T*>(smem_raw);  // Xt[cl*XP + k] = X(k, c0 + cl)
  T* Fs = Xt + kTriInvCols * XP;           // FP32: Fs[i*NB + k] = F(s0 + i, s0 + k)
  const int b = blockIdx.x / kTriInvGroups;
  const int cg0 = (blockIdx.x % kTriInvGroups) * kTriInvCols;  // first column of this workgroup
  const int s0 = b * NB;
  const int nb = min(NB, n - s0);
  const int tid = threadIdx.x;
  const int cl = tid / LP, part = tid % LP;
  const int c = cg0 + cl;
  const T* Fb = F + (size_t)s0 + (size_t)s0 * lda;
  if constexpr (kStage) {
    for (int idx = tid; idx < NB * NB; idx += kTriInvThreads) {
      const int i = idx & (NB - 1), k = idx / NB;  // coalesced over i of column k
      Fs[i * NB + k] = (i < nb && k < nb) ? Fb[i + (size_t)k * lda] : T(0);
    }
    __syncthreads();
  }
  auto f = [&](int i, int k) -> double {
    if constexpr (kStage) return (double)Fs[i * NB + k];
    else return (double)Fb[i + (size_t)k * lda];
  };
  T* xc = Xt + cl * XP;
  // this lane's share of sum_{k0 <= k < k1} F(i, k) X(k, c): k = part mod LP
  auto dot = [&](int i, int k0, int k1) -> double {
    double a0 = 0.0, a1 = 0.0;
    int k = k0 + ((part - k0) & (LP - 1));
    for (; k + LP < k1; k += 2 * LP) {
      a0 += f(i, k) * (double)xc[k];
      a1 += f(i, k + LP) * (double)xc[k + LP];
    }
    if (k < k1) a0 += f(i, k) * (double)xc[k];
    double sum = a0 + a1;
#pragma unroll
    for (int off = 1; off < LP; off <<= 1) sum += __shfl_xor(sum, off, 64);
    return sum;
  };
  if (c < nb) {
    if (LOWER) {
      for (int i = part; i < c; i += LP) xc[i] = T(0);
      for (int i = c; i < nb; ++i) {
        const double x = ((i == c) ? 1.0 : 0.0) - dot(i, c, i);
        if (part == 0) xc[i] = (T)x;
      }
    } else {
      for (int i = c + 1 + part; i < nb; i += LP) xc[i] = T(0);
      for (int i = c; i >= 0; --i) {
        const double x = (((i == c) ? 1.0 : 0.0) - dot(i, i + 1, c + 1)) / f(i, i);
        if (part == 0) xc[i] = (T)x;
      }
    }
  }
  __syncthreads();
  T* D = Dinv + (size_t)b * NB * NB;
  // D(i, c) = X(i, c), coalesced over i
  for (int idx = tid; idx < NB * kTriInvCols; idx += kTriInvThreads) {
    const int i = idx & (NB - 1), cc = idx / NB;
    if (i < nb) D[i + (size_t)(cg0 + cc) * NB] = (cg0 + cc < nb) ? Xt[cc * XP + i] : T(0);
  }
}
// X = L^-1 for the unit lower nb x nb FP32 block (nb <= 128) at F, written to
// D (column-major NB x NB: zero above the diagonal and outside nb), by one
// workgroup: the eight 16 x 16 diagonal blocks of X by forward substitution
// (a thread per column, the column in registers), then block rows i = 1..7 of
// X from X_ij = -X_ii sum_{j <= k < i} L_ik X_kj (two passes per block row,
// the block row of L staged in LDS). Sums in FP64, entries rounded to FP32 as
// they are stored, as tri_inv_kernel does; 15 barriers instead of a 128-row
// chain per column.
constexpr int kTiB = 16;
__global__ __launch_bounds__(512) void tri_inv_lower_blocked(int nb, const float* __restrict__ F,
                                                             int lda, float* __restrict__ D) {
  constexpr int NB = kSolveNB;
  constexpr int NBB = NB / kTiB;
  __shared__ float Xs[NB][NB + 1];                  // X(i, c)
  __shared__ float Lr[kTiB][NB + 1];                // block row of L (diagonal blocks first)
  __shared__ double Ss[NBB - 1][kTiB][kTiB + 1];    // S_ij of the current block row
  const int tid = threadIdx.x;
  auto Lg = [&](int i, int k) -> float {  // L(i, k), identity outside nb
    return (i < nb && k < nb) ? F[(size_t)i + (size_t)k * lda] : (i == k ? 1.f : 0.f);
  };
  // the eight diagonal blocks, block bb in Lr[r][bb * 16 + k]
  for (int e = tid; e < NBB * kTiB * kTiB; e += 512) {
    const int bb = e / (kTiB * kTiB), k = (e / kTiB) % kTiB, r = e % kTiB;  // coalesced over r
    Lr[r][bb * kTiB + k] = Lg(bb * kTiB + r, bb * kTiB + k);
  }
  __syncthreads();
  if (tid < NB) {
    const int bb = tid / kTiB, cc = tid % kTiB, r0 = bb * kTiB;
    double x[kTiB];
#pragma unroll
    for (int r = 0; r < kTiB; ++r) {
      double acc = (r == cc) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < r; ++k) acc -= (double)Lr[r][r0 + k] * x[k];
      x[r] = (double)(float)acc;  // x[k] = 0 above the diagonal, so the sum starts at cc
      Xs[r0 + r][r0 + cc] = (float)acc;
    }
  }
  __syncthreads();
  for (int ib = 1; ib < NBB; ++ib) {
    const int i0 = ib * kTiB;
    for (int e = tid; e < kTiB * i0; e += 512) {
      const int k = e / kTiB, r = e % kTiB;  // coalesced over the block row's 16 rows
      Lr[r][k] = Lg(i0 + r, k);
    }
    __syncthreads();
    for (int e = tid; e < ib * kTiB * kTiB; e += 512) {
      const int j = e / (kTiB * kTiB), r = (e / kTiB) % kTiB, cc = e % kTiB;
      double acc = 0.0;
      for (int k = j * kTiB; k < i0; ++k) acc += (double)Lr[r][k] * (double)Xs[k][j * kTiB + cc];
      Ss[j][r][cc] = acc;
    }
    __syncthreads();
    for (int e = tid; e < ib * kTiB * kTiB; e += 512) {
      const int j = e / (kTiB * kTiB), r = (e / kTiB) % kTiB, cc = e % kTiB;
      double acc = 0.0;
      for (int t = 0; t <= r; ++t) acc += (double)Xs[i0 + r][i0 + t] * Ss[j][t][cc];
      Xs[i0 + r][j * kTiB + cc] = (float)(-acc);
    }
    __syncthreads();
  }
  for (int idx = tid; idx < NB * NB; idx += 512) {
    const int i = idx % NB, c = idx / NB;
    D[idx] = (i < nb && c < nb && c <= i) ? Xs[i][c] : 0.f;
  }
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsv_chain_kernel(int n, const T* __restrict__ F, int lda,
                                                         const T* __restrict__ Dinv,
                                                         const double* __restrict__ y,
                                                         unsigned long long* __restrict__ xw,
                                                         unsigned* __restrict__ flags,
                                                         unsigned* __restrict__ ticket,
                                                         unsigned epoch,
                                                         unsigned* __restrict__ abort_flag) {
  constexpr int NB = kSolveNB;
  constexpr int HB = NB / 2;  // columns per half-workgroup
  __shared__ T sD[NB][NB];    // sD[k][i] = Dinv_rb(i, k)
  __shared__ double s_x[NB], s_v[NB], s_z[2][NB];
  __shared__ int s_blk, s_abort;
  const int tid = threadIdx.x, r = tid & (NB - 1), h = tid >> 7;
  const int nblk = (n + NB - 1) / NB;
  if (tid == 0) {
    s_blk = (int)atomicAdd(ticket, 1u);
    s_abort = 0;
  }
  __syncthreads();
  const int t = s_blk;
  if (t >= nblk) return;
  const int rb = LOWER ? t : nblk - 1 - t;
  const int s0 = rb * NB, nb = min(NB, n - s0);
  const int row = s0 + r;
  const bool rok = r < nb;
  {
    const T* Db = Dinv + (size_t)rb * NB * NB;
    for (int idx = tid; idx < NB * NB; idx += 256) sD[idx / NB][idx % NB] = Db[idx];
  }
  double acc = (h == 0 && rok) ? y[row] : 0.0;
  T tv[HB];
  auto load_tile = [&](int q) {
    const int c0 = (LOWER ? q : nblk - 1 - q) * NB;
    const int cn = min(NB, n - c0);
    const T* Fr = F + (size_t)row + (size_t)(c0 + h * HB) * lda;
#pragma unroll
    for (int jj = 0; jj < HB; ++jj)
      tv[jj] = (rok && h * HB + jj < cn) ? Fr[(size_t)jj * lda] : T(0);
  };
  if (t > 0) load_tile(0);
  for (int q = 0; q < t; ++q) {
    const int cbk = LOWER ? q : nblk - 1 - q;
    if (tid == 0) {
      unsigned spin = 0;
      while (ld_ag32(&flags[cbk]) != epoch) {
        if (++spin > kSpinLimit || ld_ag32(abort_flag)) {
          atomicExch(abort_flag, 1u);
          s_abort = 1;
          break;
        }
      }
    }
    __syncthreads();
    if (s_abort) return;
    if (tid < NB) {
      const int j = cbk * NB + tid;
      s_x[tid] = (j < n) ? __longlong_as_double((long long)ld_ag(&xw[j])) : 0.0;
    }
    __syncthreads();
    double part = 0.0;
#pragma unroll
    for (int jj = 0; jj < HB; ++jj) part += (double)tv[jj] * s_x[h * HB + jj];
    acc -= part;
    if (q + 1 < t) load_tile(q + 1);
    __syncthreads();
  }
  s_z[h][r] = acc;
  __syncthreads();
  if (tid < NB) s_v[tid] = (tid < nb) ? s_z[0][tid] + s_z[1][tid] : 0.0;
  __syncthreads();
  double zp = 0.0;
  for (int k = h * HB; k < h * HB + HB; ++k) zp += (double)sD[k][r] * s_v[k];
  __syncthreads();
  s_z[h][r] = zp;
  __syncthreads();
  if (h == 0 && rok)
    st_ag(&xw[row], (unsigned long long)__double_as_longlong(s_z[0][r] + s_z[1][r]));
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&flags[rb], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Transposed triangle as one launch (A^T x = b with A = P^T L U: U^T w = b in
// forward block order, UT = true; L^T z = w in backward order, UT = false), the
// block-chain scheme of trsv_chain_kernel: workgroups take 128-column blocks in
// dependency order from a ticket counter, add the contributions of the blocks
// published before them (epoch flags), apply the transposed inverse of their
// diagonal block and publish. Block cb needs, for each of its columns j,
// s_j = sum over solved rows r of F[r, j] v[r]. Tiles are read as in the
// forward solve (lanes over rows, coalesced): wave q4 takes the tile's columns
// [32 q4, 32 q4 + 32) and lane l rows l and l + 64, and accumulates per column
// across tiles; one transpose-reduction across the wave at the end (63
// shuffles) leaves lane l with the sum for column 32 q4 + (l & 31).
template <typename T, bool UT>
__global__ __launch_bounds__(256) void trsv_chain_trans_kernel(int n, const T* __restrict__ F,
                                                               int lda,
                                                               const T* __restrict__ Dinv,
                                                               const double* __restrict__ y,
                                                               unsigned long long* __restrict__ xw,
                                                               unsigned* __restrict__ flags,
                                                               unsigned* __restrict__ ticket,
                                                               unsigned epoch,
                                                               unsigned* __restrict__ abort_flag) {
  constexpr int NB = kSolveNB;
  constexpr int QC = NB / 4;  // columns per wave
  constexpr int HB = NB / 2;  // rows per lane pair (l, l + 64) span / k-range per half
  __shared__ T sD[NB][NB];    // sD[k][c] = Dinv_cb(k, c)
  __shared__ double s_x[NB], s_v[NB], s_z[2][NB];
  __shared__ int s_blk, s_abort;
  const int tid = threadIdx.x, lane = tid & 63, q4 = tid >> 6;
  const int nblk = (n + NB - 1) / NB;
  if (tid == 0) {
    s_blk = (int)atomicAdd(ticket, 1u);
    s_abort = 0;
  }
  __syncthreads();
  const int t = s_blk;
  if (t >= nblk) return;
  const int cb = UT ? t : nblk - 1 - t;
  const int s0 = cb * NB, nb = min(NB, n - s0);
  {
    const T* Db = Dinv + (size_t)cb * NB * NB;
    for (int idx = tid; idx < NB * NB; idx += 256) sD[idx % NB][idx / NB] = Db[idx];
  }
  double acc[QC];
#pragma unroll
  for (int jj = 0; jj < QC; ++jj) acc[jj] = 0.0;
  T tv[2 * QC];  // [0, QC): row lane, [QC, 2 QC): row lane + 64
  auto load_tile = [&](int q) {
    const int row = (UT ? q : nblk - 1 - q) * NB + lane;
    const bool ra = row < n, rb2 = row + 64 < n;
    const T* Fr = F + (size_t)row + (size_t)(s0 + q4 * QC) * lda;
#pragma unroll
    for (int jj = 0; jj < QC; ++jj) {
      const bool cok = q4 * QC + jj < nb;
      tv[jj] = (ra && cok) ? Fr[(size_t)jj * lda] : T(0);
      tv[QC + jj] = (rb2 && cok) ? Fr[(size_t)jj * lda + 64] : T(0);
    }
  };
  if (t > 0) load_tile(0);
  for (int q = 0; q < t; ++q) {
    const int rbk = UT ? q : nblk - 1 - q;
    if (tid == 0) {
      unsigned spin = 0;
      while (ld_ag32(&flags[rbk]) != epoch) {
        if (++spin > kSpinLimit || ld_ag32(abort_flag)) {
          atomicExch(abort_flag, 1u);
          s_abort = 1;
          break;
        }
      }
    }
    __syncthreads();
    if (s_abort) return;
    if (tid < NB) {
      const int j = rbk * NB + tid;
      s_x[tid] = (j < n) ? __longlong_as_double((long long)ld_ag(&xw[j])) : 0.0;
    }
    __syncthreads();
    const double xa = s_x[lane], xb = s_x[lane + 64];
#pragma unroll
    for (int jj = 0; jj < QC; ++jj) {
      acc[jj] += (double)tv[jj] * xa;
      acc[jj] += (double)tv[QC + jj] * xb;
    }
    if (q + 1 < t) load_tile(q + 1);
    __syncthreads();
  }
  // lanes l and l ^ 32 hold the same columns for other rows: add them, then a
  // transpose-reduction over the remaining offsets leaves lane l with column
  // (l & 31): at offset s, lane l keeps the half of the indices its bit s picks
#pragma unroll
  for (int jj = 0; jj < QC; ++jj) acc[jj] += __shfl_xor(acc[jj], 32, 64);
#pragma unroll
  for (int sft = 16; sft >= 1; sft >>= 1) {
    const bool up = (lane & sft) != 0;
#pragma unroll
    for (int k = 0; k < sft; ++k) {
      const double send = up ? acc[k] : acc[k + sft];
      const double keep = up ? acc[k + sft] : acc[k];
      acc[k] = keep + __shfl_xor(send, sft, 64);
    }
  }
  if (lane < QC) {
    const int c = q4 * QC + lane;
    s_v[c] = (c < nb) ? y[s0 + c] - acc[0] : 0.0;
  }
  __syncthreads();
  // z_c = sum_k Dinv(k, c) v_k; thread (c, h) takes k in half h (rows k >= nb
  // of a partial block's inverse are not written by tri_inv_kernel)
  const int cc = tid & (NB - 1), h = tid >> 7;
  double zp = 0.0;
  const int k1 = min(h * HB + HB, nb);
  for (int k = h * HB; k < k1; ++k) zp += (double)sD[k][cc] * s_v[k];
  s_z[h][cc] = zp;
  __syncthreads();
  if (h == 0 && cc < nb)
    st_ag(&xw[s0 + cc], (unsigned long long)__double_as_longlong(s_z[0][cc] + s_z[1][cc]));
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&flags[cb], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void scatter_kernel(int n, const int* __restrict__ perm, const double* __restrict__ z,
                               double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[perm[i]] = z[i];
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int NWM, int NWN, bool SET = false>
void launch_tile_gemm(bemhip_ctx* c, int M, int N, int K, const T* A, int lda, const T* B,
                      int ldb, T* C, int ldc) {
  const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  const int nt = tm * tn;
  hipLaunchKernelGGL((gemm_tile_kernel<T, BM, BN, NWM, NWN, SET>), dim3(nt),
                     dim3(NWM * NWN * 64), 0, c->stream, M, N, K, A, lda, B, ldb, C, ldc, tm, nt);
}

bool gemm_aligned(const void* A, int lda, const void* B, int ldb, int vec) {
  return ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && (lda % vec == 0) &&
         (ldb % vec == 0);
}

// X := L^-1 X for a unit lower 128 x 128 block L (column-major, ld) and the
// ncols columns of X (column-major, ldx): forward substitution with L in LDS
// and every wave on 8 columns of its own (lane l holds rows l and l + 64), the
// finished row k taken from its lane with v_readlane: no barrier per step.
// 32 columns per workgroup of 4 waves. The in-block U12 of the FP32 LU, so
// that the panel chain does not wait for the block inverse.
constexpr int kTrsmCols = 32;
__global__ __launch_bounds__(256) void trsm_unit_lower128(const float* __restrict__ L, int ld,
                                                          float* __restrict__ X, int ldx,
                                                          int ncols) {
  __shared__ __align__(16) float sL[128 * 128];  // column-major, 64 KB
  for (int t = threadIdx.x; t < 128 * 32; t += 256) {  // float4 per (column, row quad)
    const int k = t >> 5, r = (t & 31) * 4;
    const float* src = L + (size_t)k * ld + r;
    *(float4*)(sL + k * 128 + r) = make_float4(src[0], src[1], src[2], src[3]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j0 = blockIdx.x * kTrsmCols + w * 8;
  float xa[8], xb[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bool ok = j0 + q < ncols;
    xa[q] = ok ? X[(size_t)lane + (size_t)(j0 + q) * ldx] : 0.f;
    xb[q] = ok ? X[(size_t)(lane + 64) + (size_t)(j0 + q) * ldx] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < 127; ++k) {
    float xk[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      xk[q] = __int_as_float(
          __builtin_amdgcn_readlane(__float_as_int(k < 64 ? xa[q] : xb[q]), k & 63));
    if (k < 63) {
      const float la = sL[k * 128 + lane];
      const bool below = lane > k;
#pragma unroll
      for (int q = 0; q < 8; ++q) xa[q] = below ? xa[q] - la * xk[q] : xa[q];
    }
    const float lb = sL[k * 128 + 64 + lane];
    const bool below_b = lane + 64 > k;
#pragma unroll
    for (int q = 0; q < 8; ++q) xb[q] = below_b ? xb[q] - lb * xk[q] : xb[q];
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (j0 + q < ncols) {
      X[(size_t)lane + (size_t)(j0 + q) * ldx] = xa[q];
      X[(size_t)(lane + 64) + (size_t)(j0 + q) * ldx] = xb[q];
    }
  }
}

// X := L11^-1 X for the columns of X (ob <= 512 rows, column-major, ldx), L11
// the unit lower ob x ob diagonal block of the factor (leading dimension lda)
// whose 128-blocks' inverses are linv (kSolveNB x kSolveNB each, leading
// dimension kSolveNB, zero above the diagonal): by 128-blocks b,
// Y = X_b - L[b, <b] X_<b (the product summed, then subtracted once), then
// X_b = Linv_b Y, the solved rows kept in LDS. One workgroup per 8 columns;
// thread (row i of the 128-block, column group cg) holds 4 columns. The
// look-ahead TRSM of the next block's columns in one launch.
constexpr int kTrsmInvCols = 8;
template <typename T>
__global__ __launch_bounds__(256) void trsm_inv_strip(int ob, int nc, const T* __restrict__ L,
                                                      int lda, const T* __restrict__ linv,
                                                      T* __restrict__ X, int ldx) {
  constexpr int CPT = kTrsmInvCols / 2;
  __shared__ __align__(16) T Xs[kOuterBlocks * kSolveNB][kTrsmInvCols];
  __shared__ __align__(16) T Ys[kSolveNB][kTrsmInvCols];
  const int tid = threadIdx.x, i = tid & (kSolveNB - 1), cg = tid >> 7;
  const int j0 = blockIdx.x * kTrsmInvCols;
  for (int jj = 0; jj < kTrsmInvCols; ++jj) {
    const bool in = j0 + jj < nc;
    for (int r = tid; r < ob; r += 256) Xs[r][jj] = in ? X[(size_t)r + (size_t)(j0 + jj) * ldx] : T(0);
  }
  __syncthreads();
  for (int b0 = 0; b0 < ob; b0 += kSolveNB) {
    const int bw = min(kSolveNB, ob - b0);
    const bool row = i < bw;
    T acc[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) acc[q] = T(0);
    const T* Lr = L + (size_t)(b0 + (row ? i : 0));
#pragma unroll 8
    for (int k = 0; k < b0; ++k) {
      const T a = row ? Lr[(size_t)k * lda] : T(0);
#pragma unroll
      for (int q = 0; q < CPT; ++q) acc[q] += a * Xs[k][cg * CPT + q];
    }
    if (row) {
#pragma unroll
      for (int q = 0; q < CPT; ++q) Ys[i][cg * CPT + q] = Xs[b0 + i][cg * CPT + q] - acc[q];
    }
    __syncthreads();
    const T* Lb = linv + (size_t)(b0 / kSolveNB) * kSolveNB * kSolveNB + (row ? i : 0);
#pragma unroll
    for (int q = 0; q < CPT; ++q) acc[q] = T(0);
#pragma unroll 8
    for (int k = 0; k < bw; ++k) {
      const T a = row ? Lb[(size_t)k * kSolveNB] : T(0);
#pragma unroll
      for (int q = 0; q < CPT; ++q) acc[q] += a * Ys[k][cg * CPT + q];
    }
    if (row) {
#pragma unroll
      for (int q = 0; q < CPT; ++q) Xs[b0 + i][cg * CPT + q] = acc[q];
    }
    __syncthreads();
  }
  for (int jj = 0; jj < kTrsmInvCols; ++jj) {
    if (j0 + jj >= nc) break;
    for (int r = tid; r < ob; r += 256) X[(size_t)r + (size_t)(j0 + jj) * ldx] = Xs[r][jj];
  }
}

// X = Linv X in place (M <= 128 rows of X, Linv M x M with leading dimension
// kSolveNB), one 128-row tile row. Requires 16-byte aligned operands.
template <typename T>
void gemm_set_inplace(bemhip_ctx* c, int M, int N, const T* Linv, T* X, int ldx) {
  if (M <= 0 || N <= 0) return;
  launch_tile_gemm<T, 128, 128, 2, 2, true>(c, M, N, M, Linv, kSolveNB, X, ldx, X, ldx);
}

// C -= A B. 16-byte aligned operands (every LU call on the padded matrices)
// take the tile kernel: products with at least one 256-wide tile per CU on
// 256-wide tiles, smaller ones on 128x128 tiles for more workgroups; anything
// else takes the generic kernel.
template <typename T>
int gemm_sub(bemhip_ctx* c, int M, int N, int K, const T* A, int lda, const T* B, int ldb, T* C,
             int ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return BEMHIP_OK;
  constexpr int VEC = 16 / (int)sizeof(T);
  if (gemm_aligned(A, lda, B, ldb, VEC)) {
    const int cus = std::max(1, c->num_cu);
    const long big = (long)((M + 255) / 256) * (long)((N + 255) / 256);
    if (big >= cus) {
      // the large-tile launches (the LU's trailing updates) are timed one by one
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (c->gemm_log) {
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        hipEventRecord(e0, c->stream);
      }
      if constexpr (sizeof(T) == 4) {
        if (c->lu32_gemm == BEMHIP_LU_GEMM_SPLIT_BF16) {
          const int tm = (M + kSplitRows - 1) / kSplitRows, tn = (N + kSplitRows - 1) / kSplitRows;
          hipLaunchKernelGGL(gemm_split_kernel, dim3(tm * tn), dim3(512), 0, c->stream, M, N, K,
                             A, lda, B, ldb, C, ldc, tm, tm * tn);
        } else {
          launch_tile_gemm<T, 256, 256, 4, 2>(c, M, N, K, A, lda, B, ldb, C, ldc);
        }
      } else
        launch_tile_gemm<T, 256, 128, 4, 2>(c, M, N, K, A, lda, B, ldb, C, ldc);
      if (c->gemm_log) {
        hipEventRecord(e1, c->stream);
        c->gemm_log->push_back({e0, e1, 2.0 * M * (double)N * K,
                                sizeof(T) * ((double)M * K + (double)K * N + 2.0 * M * N)});
      }
    } else {
      launch_tile_gemm<T, 128, 128, 2, 2>(c, M, N, K, A, lda, B, ldb, C, ldc);
    }
    return BEMHIP_OK;
  }
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
  // Register-resident panel width: 128 FP32 or 64 FP64 values per thread.
  constexpr int KB = (sizeof(T) == 4) ? 128 : 64;
  constexpr int RW = KB * TagWords<T>::kWords;
  const int Gmax = (n + kPR - 1) / kPR;
  if (n >= (1 << 24)) return fail(c, BEMHIP_ERR_ARG, "lu: n must be below 2^24");
  constexpr int OB = kOuterBlocks * kSolveNB;  // outer block: K of the trailing update
  constexpr int PPB = OB / KB;                   // panels per outer block
  constexpr int LSLOT = 4 * kMaxPanel + 16;      // ints per interchange-list slot
  // FP32 panels by tournament pivoting when asked for
  const bool tslu = sizeof(T) == 4 && c->lu32_pivot == BEMHIP_PIVOT_TOURNAMENT;
  size_t tnodes = 0;  // candidate-tree nodes of the largest panel
  for (int s = Gmax;; s = (s + kTF - 1) / kTF) {
    tnodes += (size_t)s;
    if (s == 1) break;
  }
  const size_t tslu_bytes =
      tslu ? 4 * tnodes * kTW * (size_t)kTW + 4 * tnodes * kTW + 4 * tnodes + 4 * (size_t)kTW * kTW +
                 4 * 256
           : 0;
  // work buffers
  const size_t need = 8 * 4 * (size_t)Gmax + 8 * 2 * (size_t)Gmax * RW + 8 * 2 * (size_t)RW +
                      sizeof(int) * ((size_t)PPB * LSLOT + 16) +
                      sizeof(T) * 2 * kOuterBlocks * kSolveNB * kSolveNB + 8 * 256 + tslu_bytes;
  BEMHIP_HIP_TRY(c, c->lu_work.ensure(need));
  char* wp = c->lu_work.ptr;
  auto carve = [&](size_t bytes) {
    char* p = wp;
    wp += (bytes + 255) & ~size_t(255);
    return p;
  };
  unsigned long long* keys = (unsigned long long*)carve(8 * 4 * (size_t)Gmax);  // 2 buffers
  unsigned long long* recs = (unsigned long long*)carve(8 * 2 * (size_t)Gmax * RW);
  unsigned long long* rowc = (unsigned long long*)carve(8 * 2 * (size_t)RW);
  int* lists = (int*)carve(sizeof(int) * (size_t)PPB * LSLOT);  // one per panel of a block
  int* flags = (int*)carve(sizeof(int) * 16);
  // inverses of the outer block's 128-blocks of L11, by block parity: block j+1's
  // are built on stream B while block j's still serve its TRSM on stream A
  T* linv2 = (T*)carve(sizeof(T) * 2 * kOuterBlocks * kSolveNB * kSolveNB);
  auto linv_of = [&](int k0) {
    return linv2 + (size_t)((k0 / OB) & 1) * kOuterBlocks * kSolveNB * kSolveNB;
  };
  unsigned* t_cand = nullptr;
  unsigned* t_cid = nullptr;
  unsigned* t_count = nullptr;
  float* t_top = nullptr;
  if (tslu) {
    t_cand = (unsigned*)carve(4 * tnodes * kTW * (size_t)kTW);
    t_cid = (unsigned*)carve(4 * tnodes * kTW);
    t_count = (unsigned*)carve(4 * tnodes);
    t_top = (float*)carve(4 * (size_t)kTW * kTW);
  }
  int* info = flags + 4;
  unsigned* abort_flag = (unsigned*)(flags + 8);
  auto l_dst = [&](int q) { return lists + (size_t)q * LSLOT; };
  auto l_src = [&](int q) { return lists + (size_t)q * LSLOT + 2 * kMaxPanel; };
  auto l_len = [&](int q) { return lists + (size_t)q * LSLOT + 4 * kMaxPanel; };
  if ((size_t)(wp - c->lu_work.ptr) > c->lu_work.count)
    return fail(c, BEMHIP_ERR_STATE, "lu: work buffer carve overflow");
  hipStream_t sA = c->stream, sB = c->stream2, sC = c->stream3;
  BEMHIP_HIP_TRY(c, hipMemsetAsync(flags, 0, sizeof(int) * 16, sA));
  BEMHIP_HIP_TRY(c, hipMemsetAsync(keys, 0, 8 * 4 * (size_t)Gmax, sA));
  if (tslu) BEMHIP_HIP_TRY(c, hipMemsetAsync(t_count, 0, 4 * tnodes, sA));
  int key_buf = 0;  // key buffer of the next panel launch
  std::vector<GemmTiming> big_gemms;
  c->gemm_log = &big_gemms;
  struct LogOff {
    bemhip_ctx* c;
    ~LogOff() { c->gemm_log = nullptr; }
  } log_off{c};
  int occ = 0;
  BEMHIP_HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ, (const void*)panel_kernel<T, KB>, kPR, 0));
  if ((int64_t)occ * std::max(1, c->num_cu) < Gmax)
    return fail(c, BEMHIP_ERR_STATE, "lu: panel workgroups cannot all be resident");
  c->dinv_of = nullptr;
  BEMHIP_HIP_TRY(c, hipFuncSetAttribute((const void*)tri_inv_kernel<T, true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)tri_inv_lds(sizeof(T))));
  constexpr int VEC = 16 / (int)sizeof(T);
  const bool aligned = gemm_aligned(linv2, kSolveNB, dA, lda, VEC);

  // Timing: event pairs per phase on the stream that runs it (the phases of
  // the two streams overlap, so their sums exceed the factorization's time).
  enum { kPanel = 0, kSwap = 1, kTrsm = 2, kGemm = 3 };
  struct Seg {
    hipEvent_t a, b;
    int cat;
  };
  std::vector<Seg> segs;
  std::vector<hipEvent_t> joins;
  auto ev_now = [&](hipStream_t st) {
    hipEvent_t e;
    hipEventCreate(&e);
    hipEventRecord(e, st);
    return e;
  };
  // phase timing events (off by default: each event record costs the panel
  // chain time, ~10 ms over an LU at n = 32769)
  const bool timers = c->lu_timers;
  auto tev_now = [&](hipStream_t st) -> hipEvent_t { return timers ? ev_now(st) : nullptr; };
  auto seg_end = [&](hipStream_t st, hipEvent_t a0, int category) {
    if (a0) segs.push_back({a0, ev_now(st), category});
  };
  // the helpers (gemm_sub, trsm_rec, ...) enqueue on c->stream
  struct StreamScope {
    bemhip_ctx* c;
    hipStream_t old;
    StreamScope(bemhip_ctx* cc, hipStream_t st) : c(cc), old(cc->stream) { c->stream = st; }
    ~StreamScope() { c->stream = old; }
  };

  // Factor outer block [k0, k0 + ob) on stream st: register-resident panels,
  // each panel's interchanges applied to the other columns of the block (the
  // list kept in slot q for the columns outside the block), the inverses of
  // the block's 128-blocks of L11, and the rest of the block updated after
  // each panel.
  auto factor_block = [&](hipStream_t st, int k0, int ob) -> int {
    StreamScope scope(c, st);
    T* linv = linv_of(k0);
    for (int s = 0, q = 0; s < ob; s += KB, ++q) {
      const int kc = k0 + s;
      const int kb = std::min(KB, ob - s);
      const int m = n - kc;
      const int G = (m + kPR - 1) / kPR;
      hipEvent_t e0 = tev_now(st);
      bool tournament = false;
      if constexpr (sizeof(T) == 4) {
        if (tslu) {
          tournament = true;
          // tournament: pivots and the winners' L11\U11, the interchanges on
          // every column of the block (the panel's too), then L21
          TsluArgs ta;
          ta.A = (float*)dA; ta.lda = lda; ta.m = m; ta.kb = kb; ta.k0 = kc; ta.G = G;
          ta.cand = t_cand; ta.cid = t_cid; ta.count = t_count; ta.top = t_top;
          ta.ipiv = dipiv; ta.info = info; ta.abort_flag = abort_flag;
          hipLaunchKernelGGL(tslu_kernel, dim3(G), dim3(kPR), 0, st, ta);
          hipLaunchKernelGGL(perm_build_fast, dim3(1), dim3(64), 0, st, dipiv, kc, kb, l_dst(q),
                             l_src(q), l_len(q));
          hipLaunchKernelGGL((perm_apply<T>), dim3((ob + 3) / 4), dim3(256), 0, st, dA, lda, k0,
                             k0 + ob, 0, 0, (const int*)l_dst(q), (const int*)l_src(q),
                             (const int*)l_len(q));
          hipLaunchKernelGGL(tslu_apply,
                             dim3(std::max(1, (m - kb + kApplyThreads - 1) / kApplyThreads)),
                             dim3(kApplyThreads), 0, st, (float*)dA, lda, m, kb, kc,
                             (const float*)t_top);
          BEMHIP_HIP_TRY(c, hipGetLastError());
          seg_end(st, e0, kPanel);
          e0 = tev_now(st);
        }
      }
      if (!tournament) {
        PanelArgs<T> pa;
        pa.A = dA; pa.lda = lda; pa.m = m; pa.kb = kb; pa.k0 = kc; pa.G = G; pa.gstride = Gmax;
        pa.epoch = (++c->panel_epoch) & 0xffffffu;
        pa.ipiv = dipiv; pa.info = info; pa.recs = recs; pa.rowc = rowc;
        pa.keys = keys + (size_t)key_buf * 2 * Gmax;
        pa.keys_next = keys + (size_t)(key_buf ^ 1) * 2 * Gmax;
        key_buf ^= 1;
        pa.abort_flag = abort_flag;
        pa.list_dst = l_dst(q); pa.list_src = l_src(q); pa.list_len = l_len(q);
        void* args[] = {&pa};
        BEMHIP_HIP_TRY(c, hipLaunchCooperativeKernel((const void*)panel_kernel<T, KB>, dim3(G),
                                                     dim3(kPR), args, 0, st));
        seg_end(st, e0, kPanel);
        e0 = tev_now(st);
        const int ncols_blk = ob - kb;  // (the panel kernel built the gather list)
        if (ncols_blk > 0)
          hipLaunchKernelGGL((perm_apply<T>), dim3((ncols_blk + 3) / 4), dim3(256), 0, st, dA, lda,
                             k0, kc, kc + kb, k0 + ob, (const int*)l_dst(q), (const int*)l_src(q),
                             (const int*)l_len(q));
        BEMHIP_HIP_TRY(c, hipGetLastError());
        seg_end(st, e0, kSwap);
        e0 = tev_now(st);
      }
      // inverse of each completed 128-block of this outer block's L11
      if (aligned && ((s + kb) % kSolveNB == 0 || s + kb == ob)) {
        const int jb = (s + kb - 1) / kSolveNB;
        const int d0 = k0 + jb * kSolveNB;
        const int bw = std::min(kSolveNB, ob - jb * kSolveNB);
        const T* Fb = dA + (size_t)d0 + (size_t)d0 * lda;
        T* Db = linv + (size_t)jb * kSolveNB * kSolveNB;
        // on stream C when the block is factored on stream B: only stream A's
        // TRSM needs the inverses (the in-block U12 is solved by substitution
        // in FP32, by trsm_rec in FP64), and it waits for them
        hipStream_t si = st;
        if (st == sB) {
          hipEvent_t e = ev_now(st);
          joins.push_back(e);
          BEMHIP_HIP_TRY(c, hipStreamWaitEvent(sC, e, 0));
          si = sC;
        }
        if constexpr (sizeof(T) == 4)
          hipLaunchKernelGGL(tri_inv_lower_blocked, dim3(1), dim3(512), 0, si, bw,
                             (const float*)Fb, lda, (float*)Db);
        else
          hipLaunchKernelGGL((tri_inv_kernel<T, true>), dim3(kTriInvGroups),
                             dim3(kTriInvThreads), tri_inv_lds(sizeof(T)), si, bw, Fb, lda, Db);
        BEMHIP_HIP_TRY(c, hipGetLastError());
      }
      // the rest of this outer block
      const int rest = k0 + ob - (kc + kb);
      if (rest > 0) {
        T* L11 = dA + (size_t)kc + (size_t)kc * lda;
        T* A12 = dA + (size_t)kc + (size_t)(kc + kb) * lda;
        if constexpr (sizeof(T) == 4) {
          if (kb == 128) {  // U12 = L11^-1 A12 by substitution: no wait for the inverse
            hipLaunchKernelGGL(trsm_unit_lower128, dim3((rest + kTrsmCols - 1) / kTrsmCols),
                               dim3(256), 0, st, (const float*)L11, lda, (float*)A12, lda, rest);
            BEMHIP_HIP_TRY(c, hipGetLastError());
            seg_end(st, e0, kTrsm);
            e0 = tev_now(st);
            T* L21 = dA + (size_t)(kc + kb) + (size_t)kc * lda;
            T* A22 = dA + (size_t)(kc + kb) + (size_t)(kc + kb) * lda;
            gemm_sub(c, n - kc - kb, rest, kb, (const T*)L21, lda, (const T*)A12, lda, A22, lda);
            BEMHIP_HIP_TRY(c, hipGetLastError());
            seg_end(st, e0, kGemm);
            continue;
          }
        }
        if (aligned && KB == kSolveNB) {  // the panel is one 128-block: U12 = L11^-1 A12
          gemm_set_inplace(c, kb, rest,
                           (const T*)(linv + (size_t)(s / kSolveNB) * kSolveNB * kSolveNB), A12,
                           lda);
        } else {
          int stt = trsm_rec(c, kb, rest, (const T*)L11, lda, A12, lda);
          if (stt) return stt;
        }
        seg_end(st, e0, kTrsm);
        e0 = tev_now(st);
        T* L21 = dA + (size_t)(kc + kb) + (size_t)kc * lda;
        T* A22 = dA + (size_t)(kc + kb) + (size_t)(kc + kb) * lda;
        gemm_sub(c, n - kc - kb, rest, kb, (const T*)L21, lda, (const T*)A12, lda, A22, lda);
        BEMHIP_HIP_TRY(c, hipGetLastError());
        seg_end(st, e0, kGemm);
      } else {
        seg_end(st, e0, kTrsm);
      }
    }
    return BEMHIP_OK;
  };
  // The block's interchanges on the columns outside it (stream A).
  auto outer_perms = [&](int k0, int ob) -> int {
    hipEvent_t e0 = tev_now(sA);
    const int ncols = n - ob;
    for (int s = 0, q = 0; s < ob && ncols > 0; s += KB, ++q)
      hipLaunchKernelGGL((perm_apply<T>), dim3((ncols + 3) / 4), dim3(256), 0, sA, dA, lda, 0, k0,
                         k0 + ob, n, (const int*)l_dst(q), (const int*)l_src(q),
                         (const int*)l_len(q));
    BEMHIP_HIP_TRY(c, hipGetLastError());
    seg_end(sA, e0, kSwap);
    return BEMHIP_OK;
  };
  // U12 = L11^-1 A12 for the columns [c0, c1) of block k0 (stream A).
  auto outer_trsm = [&](int k0, int ob, int c0, int c1) -> int {
    const int nc = c1 - c0;
    if (nc <= 0) return BEMHIP_OK;
    const T* linv = linv_of(k0);
    hipEvent_t e0 = tev_now(sA);
    T* L11 = dA + (size_t)k0 + (size_t)k0 * lda;
    T* X = dA + (size_t)k0 + (size_t)c0 * lda;
    if (aligned && nc <= kOuterBlocks * kSolveNB) {
      // the next block's columns (on the panel chain): one launch
      hipLaunchKernelGGL((trsm_inv_strip<T>), dim3((nc + kTrsmInvCols - 1) / kTrsmInvCols),
                         dim3(256), 0, sA, ob, nc, (const T*)L11, lda, linv, X, lda);
    } else if (aligned) {
      // by 128-blocks b: X_b -= L[b, <b] X_<b, then X_b = Linv_b X_b
      for (int b0 = 0; b0 < ob; b0 += kSolveNB) {
        const int bw = std::min(kSolveNB, ob - b0);
        if (b0 > 0)
          gemm_sub(c, bw, nc, b0, (const T*)(L11 + b0), lda, (const T*)X, lda, X + b0, lda);
        gemm_set_inplace(c, bw, nc,
                         (const T*)(linv + (size_t)(b0 / kSolveNB) * kSolveNB * kSolveNB),
                         X + b0, lda);
      }
    } else {
      int st = trsm_rec(c, ob, nc, (const T*)L11, lda, X, lda);
      if (st) return st;
    }
    BEMHIP_HIP_TRY(c, hipGetLastError());
    seg_end(sA, e0, kTrsm);
    return BEMHIP_OK;
  };
  // A[k0+ob.., c0..c1) -= L21 U12[:, c0..c1) (stream A).
  auto outer_gemm = [&](int k0, int ob, int c0, int c1) -> int {
    if (c1 <= c0 || n - k0 - ob <= 0) return BEMHIP_OK;
    hipEvent_t e0 = tev_now(sA);
    gemm_sub(c, n - k0 - ob, c1 - c0, ob, (const T*)(dA + (size_t)(k0 + ob) + (size_t)k0 * lda),
             lda, (const T*)(dA + (size_t)k0 + (size_t)c0 * lda), lda,
             dA + (size_t)(k0 + ob) + (size_t)c0 * lda, lda);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    seg_end(sA, e0, kGemm);
    return BEMHIP_OK;
  };

  // Blocked right-looking LU with look-ahead (the structure of LAPACK getrf
  // over getrf2, plus one block of look-ahead): after block j's U12, block
  // j+1's columns are updated first; block j+1 is then factored on stream B
  // while stream A updates the remaining trailing columns with block j. Block
  // j+1's interchanges reach the columns outside it only after that update:
  // P (A22 - L21 U12) = P A22 - (P L21) U12, and during the overlap the two
  // streams touch disjoint columns.
  int st = factor_block(sA, 0, std::min(OB, n));
  if (st) return st;
  if ((st = outer_perms(0, std::min(OB, n)))) return st;
  for (int k0 = 0; k0 < n; k0 += OB) {
    const int ob = std::min(OB, n - k0);
    const int k1 = k0 + ob;
    if (k1 >= n) break;
    const int ob1 = std::min(OB, n - k1);
    const int k2 = k1 + ob1;
    // the next block's columns first: U12 and the update, then its factorization
    if ((st = outer_trsm(k0, ob, k1, k2))) return st;
    if ((st = outer_gemm(k0, ob, k1, k2))) return st;
    hipEvent_t next_ready = ev_now(sA);
    joins.push_back(next_ready);
    BEMHIP_HIP_TRY(c, hipStreamWaitEvent(sB, next_ready, 0));
    if ((st = factor_block(sB, k1, ob1))) return st;
    hipEvent_t factored = ev_now(sB);
    joins.push_back(factored);
    hipEvent_t inverted = ev_now(sC);  // block k1's L11 inverses (stream C)
    joins.push_back(inverted);
    // meanwhile the remaining trailing columns
    if ((st = outer_trsm(k0, ob, k2, n))) return st;
    if ((st = outer_gemm(k0, ob, k2, n))) return st;
    // block k1's interchanges need only its factorization; its TRSM (the
    // next iteration) needs the inverses too
    BEMHIP_HIP_TRY(c, hipStreamWaitEvent(sA, factored, 0));
    if ((st = outer_perms(k1, ob1))) return st;
    BEMHIP_HIP_TRY(c, hipStreamWaitEvent(sA, inverted, 0));
  }
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(sA));
  double tcat[4] = {0, 0, 0, 0};
  for (auto& sg : segs) {
    float t = 0;
    hipEventElapsedTime(&t, sg.a, sg.b);
    tcat[sg.cat] += t;
    hipEventDestroy(sg.a);
    hipEventDestroy(sg.b);
  }
  for (auto& e : joins) hipEventDestroy(e);
  double big_ms = 0.0, big_flops = 0.0, big_bytes = 0.0;
  for (auto& g : big_gemms) {
    float t = 0;
    hipEventElapsedTime(&t, g.a, g.b);
    big_ms += t;
    big_flops += g.flops;
    big_bytes += g.bytes;
    hipEventDestroy(g.a);
    hipEventDestroy(g.b);
  }
  const double t_panel = tcat[kPanel], t_swap = tcat[kSwap], t_trsm = tcat[kTrsm],
               t_gemm = tcat[kGemm];
  const char* pfx = (sizeof(T) == 8) ? "lu64_" : "lu32_";
  c->ms[std::string(pfx) + "panel"] = t_panel;
  c->ms[std::string(pfx) + "swap"] = t_swap;
  c->ms[std::string(pfx) + "trsm"] = t_trsm;
  c->ms[std::string(pfx) + "gemm"] = t_gemm;
  c->stats[std::string(pfx) + "kb"] = KB;
  c->stats[std::string(pfx) + "ob"] = OB;
  c->stats[std::string(pfx) + "panel_wgs"] = Gmax;
  // large-tile trailing-update launches: count, flops, summed duration
  c->ms[std::string(pfx) + "big_gemm"] = big_ms;
  c->stats[std::string(pfx) + "big_gemm_launches"] = (double)big_gemms.size();
  c->stats[std::string(pfx) + "big_gemm_flops"] = big_flops;
  c->stats[std::string(pfx) + "big_gemm_bytes"] = big_bytes;
  int hflags[16];
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, sA));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(sA));
  if (hflags[8] != 0)
    return fail(c, BEMHIP_ERR_HIP, "lu: panel exchange wait ran out (factorization abandoned)");
  if (info_host) *info_host = hflags[4];
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
    const size_t lds = tri_inv_lds(sizeof(T));
    BEMHIP_HIP_TRY(c, hipFuncSetAttribute((const void*)tri_inv_kernel<T, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    BEMHIP_HIP_TRY(c, hipFuncSetAttribute((const void*)tri_inv_kernel<T, false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((tri_inv_kernel<T, true>), dim3(nblk * kTriInvGroups), dim3(kTriInvThreads),
                       lds, c->stream, n, dLU, lda, DL);
    hipLaunchKernelGGL((tri_inv_kernel<T, false>), dim3(nblk * kTriInvGroups),
                       dim3(kTriInvThreads), lds, c->stream, n, dLU, lda, DU);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    c->dinv_of = dLU;
    c->dinv_n = n;
  }
  return BEMHIP_OK;
}

// A^T x = b in place of d_b, with P A = L U (perm: (P v)[i] = v[perm[i]]):
// U^T w = b, L^T z = w as one block-chain launch each, x = P^T z.
template <typename T>
int lu_solve_trans_t(bemhip_ctx* c, int n, const T* dLU, int lda, const int* dperm,
                     double* d_b) {
  int st = ensure_dinv<T>(c, n, dLU, lda);
  if (st) return st;
  const int nblk = (n + kSolveNB - 1) / kSolveNB;
  const size_t blk_elems = (size_t)kSolveNB * kSolveNB;
  const T* DL = (const T*)c->dinv.ptr;
  const T* DU = DL + nblk * blk_elems;
  BEMHIP_HIP_TRY(c, c->vec64c.ensure(3 * (size_t)n));
  unsigned long long* w = (unsigned long long*)c->vec64c.ptr;
  unsigned long long* z = (unsigned long long*)(c->vec64c.ptr + n);
  const size_t nsync = 2 * (size_t)nblk + 4;
  if (c->solve_sync.count < nsync) {
    BEMHIP_HIP_TRY(c, c->solve_sync.ensure(nsync));
    BEMHIP_HIP_TRY(c, hipMemsetAsync(c->solve_sync.ptr, 0, sizeof(unsigned) * nsync, c->stream));
  }
  unsigned* flags = c->solve_sync.ptr;
  unsigned* tickets = flags + 2 * (size_t)nblk;
  unsigned* abort_flag = tickets + 2;
  if (!c->h_flag) BEMHIP_HIP_TRY(c, hipHostMalloc((void**)&c->h_flag, sizeof(unsigned)));
  const unsigned epoch = ++c->solve_epoch;
  if (epoch == 0) return fail(c, BEMHIP_ERR_STATE, "lu_solve: epoch counter wrapped");
  BEMHIP_HIP_TRY(c, hipMemsetAsync(tickets, 0, sizeof(unsigned) * 3, c->stream));  // + abort
  hipLaunchKernelGGL((trsv_chain_trans_kernel<T, true>), dim3(nblk), dim3(256), 0, c->stream, n,
                     dLU, lda, DU, (const double*)d_b, w, flags, tickets, epoch, abort_flag);
  hipLaunchKernelGGL((trsv_chain_trans_kernel<T, false>), dim3(nblk), dim3(256), 0, c->stream, n,
                     dLU, lda, DL, (const double*)w, z, flags + nblk, tickets + 1, epoch,
                     abort_flag);
  const int tb = 256;
  hipLaunchKernelGGL(scatter_kernel, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n, dperm,
                     (const double*)z, d_b);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->h_flag, abort_flag, sizeof(unsigned), hipMemcpyDeviceToHost,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (*c->h_flag)
    return fail(c, BEMHIP_ERR_HIP, "lu_solve: block-chain wait ran out (solve abandoned)");
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
  BEMHIP_HIP_TRY(c, c->vec64c.ensure(3 * (size_t)n));
  double* y = c->vec64c.ptr;
  unsigned long long* w = (unsigned long long*)(c->vec64c.ptr + n);
  unsigned long long* x = (unsigned long long*)(c->vec64c.ptr + 2 * (size_t)n);
  // flags [2][nblk], tickets [2], abort word; zero once at allocation
  const size_t nsync = 2 * (size_t)nblk + 4;
  if (c->solve_sync.count < nsync) {
    BEMHIP_HIP_TRY(c, c->solve_sync.ensure(nsync));
    BEMHIP_HIP_TRY(c, hipMemsetAsync(c->solve_sync.ptr, 0, sizeof(unsigned) * nsync, c->stream));
  }
  unsigned* flags = c->solve_sync.ptr;
  unsigned* tickets = flags + 2 * (size_t)nblk;
  unsigned* abort_flag = tickets + 2;
  if (!c->h_flag) BEMHIP_HIP_TRY(c, hipHostMalloc((void**)&c->h_flag, sizeof(unsigned)));
  const unsigned epoch = ++c->solve_epoch;
  if (epoch == 0) return fail(c, BEMHIP_ERR_STATE, "lu_solve: epoch counter wrapped");
  BEMHIP_HIP_TRY(c, hipMemsetAsync(tickets, 0, sizeof(unsigned) * 3, c->stream));  // + abort
  const int tb = 256;
  hipLaunchKernelGGL(gather_kernel, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n, dperm,
                     d_b, y);
  hipLaunchKernelGGL((trsv_chain_kernel<T, true>), dim3(nblk), dim3(256), 0, c->stream, n, dLU,
                     lda, DL, (const double*)y, w, flags, tickets, epoch, abort_flag);
  hipLaunchKernelGGL((trsv_chain_kernel<T, false>), dim3(nblk), dim3(256), 0, c->stream, n, dLU,
                     lda, DU, (const double*)w, x, flags + nblk, tickets + 1, epoch, abort_flag);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(d_b, x, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->h_flag, abort_flag, sizeof(unsigned), hipMemcpyDeviceToHost,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (*c->h_flag)
    return fail(c, BEMHIP_ERR_HIP, "lu_solve: block-chain wait ran out (solve abandoned)");
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

// ---------------------------------------------------------------------------
// Re-solve with resident FP64 factors, allocating nothing
// (bemhip_solve_rhs_noalloc). Its device scratch lives in this file-scope
// array, one copy per GPU; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device. Layout, in doubles:
// [0, N) right-hand side, [N, 2N) gathered rhs, [2N, 3N) forward result words,
// [3N, 4N) solution words, then the chain's flags [2][kNoAllocBlocks],
// tickets [2] and abort word (unsigned).
// ---------------------------------------------------------------------------
namespace bemhip {
namespace {

constexpr int kNoAllocMaxN = 65536;
constexpr int kNoAllocBlocks = kNoAllocMaxN / kSolveNB;
constexpr size_t kNoAllocWords = 2 * (size_t)kNoAllocBlocks + 4;
constexpr size_t kNoAllocBytes = 4 * sizeof(double) * (size_t)kNoAllocMaxN +
                                 sizeof(unsigned) * kNoAllocWords;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[256] = {};
unsigned g_scratch_epoch[256] = {};

__global__ void zero_words_kernel(unsigned* __restrict__ w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = 0u;
}

}  // namespace

}  // namespace bemhip

extern "C" int bemhip_solve_rhs_noalloc(bemhip_ctx* c, const double* b, double* x) {
  using namespace bemhip;
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "solve_rhs_noalloc: NULL context");
  if (!b || !x) return fail(c, BEMHIP_ERR_ARG, "solve_rhs_noalloc: NULL b or x");
  const int n = c->nsys();
  if (n > kNoAllocMaxN)
    return fail(c, BEMHIP_ERR_ARG, "solve_rhs_noalloc: the system has " + std::to_string(n) +
                                       " unknowns, above the maximum of " +
                                       std::to_string(kNoAllocMaxN));
  if (!c->factored64 || c->dinv_of != (const void*)c->A64.ptr || c->dinv_n != n)
    return fail(c, BEMHIP_ERR_STATE,
                "solve_rhs_noalloc: needs the FP64 factors of a bemhip_solve(BEMHIP_SOLVE_FP64) "
                "on this system as the last solve");
  if (c->device < 0 || c->device >= 256)
    return fail(c, BEMHIP_ERR_ARG, "solve_rhs_noalloc: device index out of range");
  BEMHIP_HIP_TRY(c, hipSetDevice(c->device));
  unsigned char* base = nullptr;
  unsigned epoch = 0;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(c, BEMHIP_ERR_HIP, std::string("solve_rhs_noalloc: hipGetSymbolAddress: ") +
                                           hipGetErrorString(e));
      g_scratch_ptr[c->device] = p;
    }
    base = (unsigned char*)g_scratch_ptr[c->device];
    epoch = ++g_scratch_epoch[c->device];
  }
  if (epoch == 0) return fail(c, BEMHIP_ERR_STATE, "solve_rhs_noalloc: epoch counter wrapped");
  const int NM = kNoAllocMaxN;
  double* d_b = (double*)base;
  double* d_y = d_b + NM;
  unsigned long long* d_w = (unsigned long long*)(d_b + 2 * (size_t)NM);
  unsigned long long* d_x = (unsigned long long*)(d_b + 3 * (size_t)NM);
  unsigned* flags = (unsigned*)(d_b + 4 * (size_t)NM);
  unsigned* tickets = flags + 2 * (size_t)kNoAllocBlocks;
  unsigned* abort_flag = tickets + 2;
  const int nblk = (n + kSolveNB - 1) / kSolveNB;
  const size_t blk_elems = (size_t)kSolveNB * kSolveNB;
  const double* DL = (const double*)c->dinv.ptr;
  const double* DU = DL + nblk * blk_elems;
  const int tb = 256;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(d_b, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, c->stream, tickets, 3);
  hipLaunchKernelGGL(gather_kernel, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n,
                     c->perm.ptr, (const double*)d_b, d_y);
  hipLaunchKernelGGL((trsv_chain_kernel<double, true>), dim3(nblk), dim3(256), 0, c->stream, n,
                     c->A64.ptr, (int)c->ld64, DL, (const double*)d_y, d_w, flags, tickets, epoch,
                     abort_flag);
  hipLaunchKernelGGL((trsv_chain_kernel<double, false>), dim3(nblk), dim3(256), 0, c->stream, n,
                     c->A64.ptr, (int)c->ld64, DU, (const double*)d_w, d_x,
                     flags + kNoAllocBlocks, tickets + 1, epoch, abort_flag);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  unsigned aborted = 0;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(x, d_x, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(&aborted, abort_flag, sizeof(unsigned), hipMemcpyDeviceToHost,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (aborted)
    return fail(c, BEMHIP_ERR_HIP, "solve_rhs_noalloc: block-chain wait ran out (solve abandoned)");
  return BEMHIP_OK;
}
