// Shared definitions of libbemhip.so: context, device buffers, constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/bemhip.h"
#include "quadconst.hpp"

namespace bemhip {

// Device geometry (structure of arrays).
struct Geo {
  int p, nn, ne, n, ni, nt;
  size_t ld;           // leading dimension of the stored H, G (padded, see padded_ld)
  const double* x;     // nt collocation x
  const double* y;     // nt collocation y
  const double* nx;    // n normals
  const double* ny;
  const double* jw;    // n: J_k * w_k (far lumping weight)
  const double* elen;  // ne element lengths
  const double* gx;    // ne*nn geometric nodes (equispaced order)
  const double* gy;
  // Rows: points p_i = (rx[i], ry[i]), i < nt. The assembly's rows are the
  // collocation points (rx = x, row0 = 0); an evaluation at other points sets
  // rx, ry to them and row0 >= n, so that no row lies on an element. Row i is
  // the source of column j's element when row0 + i is one of its nodes.
  const double* rx;
  const double* ry;
  int row0;
};

// Leading dimension of the device-resident matrices: the smallest multiple of
// `unit` elements >= nt whose column stride is an odd multiple of 256 bytes
// (unit = 64 for FP32, 32 for FP64). Column strides near a power of two
// (ld = nt = 32769 -> 2^17 + 4 bytes) put the columns of a tile on the same
// HBM channels; the odd multiple spreads them.
inline size_t padded_ld(int nt, int unit) {
  size_t m = ((size_t)nt + unit - 1) / unit;
  if ((m & 1) == 0) ++m;
  return m * unit;
}

template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t count = 0;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  hipError_t ensure(size_t n) {
    if (n <= count && ptr) return hipSuccess;
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)&ptr, n * sizeof(T));
    if (e != hipSuccess) { ptr = nullptr; return e; }
    count = n;
    return hipSuccess;
  }
};

// Quadtree plan of the far-field FMM (BEMHIP_FMM assembly): boxes of levels
// 2..L in one global index (level l from lvl_off[l]), Morton-sorted within a
// level; leaves are the level-L boxes.
struct FmmPlan {
  bool built = false;
  int L = 0, p = 0, nbox = 0, nleaf = 0;
  double x0 = 0, y0 = 0, S = 0;       // root square [x0, x0+S) x [y0, y0+S)
  std::vector<int> lvl_off, lvl_n;    // per level (index l), size L+1
  DevBuf<double> center;              // [nbox][2]
  DevBuf<int> child_lo, child_hi;     // children (global indices) of levels < L
  DevBuf<int> parent;                 // parent (global index) of levels > 2
  DevBuf<int> il_ptr, il_idx;         // M2L source boxes (global), CSR over boxes
  DevBuf<int> src_lo, src_hi, tgt_lo, tgt_hi;  // per leaf: ranges of src_order / tgt_order
  DevBuf<int> src_order, tgt_order;   // Morton order -> node / row index
  DevBuf<int> nb_ptr, nb_idx;         // per leaf: neighbour leaves with sources (leaf indices)
  DevBuf<int> cand_ptr, cand_idx;     // per leaf: elements with a node in its 3 x 3 neighbourhood
  DevBuf<double> mpole, local;        // [nbox][p+1] complex (re, im)
  DevBuf<double> binom;               // [2p+1][2p+1]
};

// One timed GEMM launch (event pair on the launching stream, flops).
struct GemmTiming {
  hipEvent_t a, b;
  double flops;
  double bytes;  // algorithmic: A and B read once, C read and written once
};

}  // namespace bemhip

struct bemhip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // second stream (default priority) for concurrent phases
  hipStream_t stream3 = nullptr;  // third stream (default priority): the assembly's on-element
                                  // rule, the LU's block inverses
  std::string err;
  int num_cu = 0;

  // geometry
  bool has_geo = false;
  int p = 0, nn = 0, ne = 0, n = 0, ni = 0, nt = 0;
  size_t ld32 = 0, ld64 = 0;  // leading dimensions of the FP32 / FP64 matrices
  double k = 1.0;
  // problem: dof = 1 Laplace (conductivity k); dof = 2 2-D Kelvin elasticity
  // (shear modulus mu, effective Poisson ratio nu). Node j's component b is
  // unknown dof * j + b; H and A have nsys() rows and columns, G nbnd() columns.
  int dof = 1;
  double mu = 0.0, nu = 0.0;
  int nsys() const { return dof * nt; }
  int nbnd() const { return dof * n; }
  std::vector<double> h_elem_w;
  bemhip::DevBuf<double> x, y, nx, ny, jw, elen, gx, gy;

  // assembly
  bool assembled = false;
  int precision = 0;
  int npg = 0;
  double near_factor = 1.5;
  bemhip::DevBuf<bemhip::QuadConst> qc;
  bemhip::DevBuf<double> H64, G64;
  bemhip::DevBuf<float> H32, G32;
  bemhip::DevBuf<uint32_t> nearbits;   // [ceil(ne/32)][nt]
  bemhip::DevBuf<int> row_cnt, row_ptr, pair_row, pair_elem;
  bemhip::DevBuf<double> pair_h, pair_g;  // [npairs][nn]
  bemhip::DevBuf<int> sing_idx, reg_idx;  // CSR indices of the on-element / other pairs
  bemhip::DevBuf<double> sing_h, sing_g;  // [min(nt, n)][nn] on-element values by row
  int64_t npairs = 0;
  bemhip::DevBuf<double> hpart;   // [ncb][nt] far row sums of H
  bemhip::DevBuf<double> blkbox;  // [ncb][5] node bounding box and cutoff of each column block
  bemhip::DevBuf<double> rowbox;  // [row block][4] node bounding box of each 256-row block
  bemhip::DevBuf<uint8_t> tiles;  // [row block][ncb] 1 where the pair tests ran
  int* h_count = nullptr;          // pinned host word: near-pair count
  int* d_count = nullptr;          // its device address
  bemhip::DevBuf<double> diag64;  // [nt] rigid-body diagonal
  bemhip::DevBuf<char> scan_tmp;

  // boundary conditions / system
  bool has_bc = false;
  std::vector<int64_t> h_bc;
  std::vector<double> h_bv;
  bemhip::DevBuf<int> d_bc;      // n BC codes (0 Dirichlet, 1 Neumann)
  bemhip::DevBuf<double> d_bv;   // n BC values
  bemhip::DevBuf<double> wH, wG, wH2, wG2;  // per-column weights for FP64 operator passes
  bemhip::DevBuf<double> b64;
  bemhip::DevBuf<double> part1, part2;     // [ncb][nt] operator partial sums
  bemhip::DevBuf<double> A64;
  bemhip::DevBuf<float> A32;
  bemhip::DevBuf<int> ipiv, ipiv32;   // pivots of the FP64 / FP32 factorization
  bemhip::DevBuf<int> perm, perm32;   // gather permutations built from them
  bool factored64 = false, factored32 = false;

  // solution of the last bemhip_solve (kept for the adjoint)
  bool has_sol = false;
  int sol_mode = 0;
  bemhip::DevBuf<double> xsol;
  bemhip::DevBuf<double> Jn;        // n |dx/dxi| at each boundary node
  bemhip::DevBuf<double> adj_work;  // adjoint scratch (carved per call)

  // far-field FMM (BEMHIP_FMM assembly)
  bemhip::FmmPlan fmm;

  // interior evaluation (points, weights, its own near-pair list and rule)
  bemhip::DevBuf<bemhip::QuadConst> ev_qc;
  bemhip::DevBuf<double> ev_pts;    // [2][npts] x then y
  bemhip::DevBuf<double> ev_w;      // [4][n] -T, q, 1, 0
  bemhip::DevBuf<double> ev_out;    // [npts]
  bemhip::DevBuf<double> ev_box;    // [ncb][5]
  bemhip::DevBuf<double> ev_rowbox; // [point block][4]
  bemhip::DevBuf<uint32_t> ev_bits; // [ncb][chunk]
  bemhip::DevBuf<int> ev_cnt, ev_ptr, ev_prow, ev_pelem;
  bemhip::DevBuf<uint8_t> ev_tiles;
  bemhip::DevBuf<double> ev_ph, ev_pg;  // [pairs][nn]
  bemhip::DevBuf<double> ev_part;       // [2][ncb][chunk]

  // scratch for LU and solves
  bemhip::DevBuf<char> lu_work;
  unsigned panel_epoch = 0;           // tag base of the LU panel exchange words
  int lu32_pivot = BEMHIP_PIVOT_PARTIAL;  // FP32 panels: partial or tournament pivoting
  int lu32_gemm = BEMHIP_LU_GEMM_SPLIT_BF16;  // FP32 trailing updates: split-bf16 or FP32 MFMA
  bool lu_timers = false;  // HIP events around the LU's phases (the lu32_*/lu64_* phase counters)
  std::vector<bemhip::GemmTiming>* gemm_log = nullptr;  // set during an LU: large-tile GEMM timings
  bemhip::DevBuf<char> dinv;          // inverses of the diagonal blocks of L and U
  const void* dinv_of = nullptr;      // factor the inverses belong to
  int dinv_n = 0;
  bemhip::DevBuf<double> vec64a, vec64b, vec64c;
  bemhip::DevBuf<double> scal;          // device scalars of the norm reductions
  bemhip::DevBuf<double> ir_d;          // GMRES-IR correction (lu_solve_t owns vec64c)
  bemhip::DevBuf<unsigned> solve_sync;  // block-chain solve flags, tickets, abort word
  unsigned solve_epoch = 0;
  unsigned* h_flag = nullptr;           // pinned host word: solve abort flag
  bemhip::DevBuf<float> vec32a;
  bemhip::DevBuf<int> info;

  // timings (ms) of the last calls, keyed by phase
  std::map<std::string, double> ms;
  std::map<std::string, double> stats;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace bemhip {

// System matrices (H, G, A in either precision) stay allocated between calls,
// so a caller that alternates precisions or problem sizes does not re-allocate
// them every time. ensure_matrix(c, buf, n, keep...) ensures one of them; on an
// out-of-memory error it frees the others (except `buf` and the `keep` ones)
// and tries again.
inline void release_spare_matrices(bemhip_ctx* c, const void* k0, const void* k1, const void* k2,
                                   const void* k3) {
  auto rel = [&](auto& b) {
    const void* p = &b;
    if (p != k0 && p != k1 && p != k2 && p != k3) b.release();
  };
  rel(c->H64); rel(c->G64); rel(c->H32); rel(c->G32); rel(c->A64); rel(c->A32);
}
template <typename T>
hipError_t ensure_matrix(bemhip_ctx* c, DevBuf<T>& buf, size_t n, const void* keep1 = nullptr,
                         const void* keep2 = nullptr, const void* keep3 = nullptr) {
  hipError_t e = buf.ensure(n);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    release_spare_matrices(c, &buf, keep1, keep2, keep3);
    e = buf.ensure(n);
  }
  return e;
}

// Scratch check of the dense matrix-free modes, whose column-block partials
// and near bits take `per_pair_block` bytes per (row, 32-element column block):
// O(nt^2 / 32), unlike the O(nt) of the FMM mode.
int check_matrix_free_scratch(bemhip_ctx* c, const char* who, size_t per_pair_block, size_t ncb,
                              size_t rows);

// Error helpers ---------------------------------------------------------------
void set_global_error(const std::string& s);
int fail(bemhip_ctx* c, int code, const std::string& msg);

#define BEMHIP_HIP_TRY(ctx, call)                                                     \
  do {                                                                                \
    hipError_t _e = (call);                                                           \
    if (_e != hipSuccess)                                                             \
      return ::bemhip::fail((ctx), _e == hipErrorOutOfMemory ? BEMHIP_ERR_NOMEM       \
                                                             : BEMHIP_ERR_HIP,        \
                            std::string(#call) + ": " + hipGetErrorString(_e));       \
  } while (0)

// Wave-level reductions (64 lanes) ---------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Host-side launchers (defined in assemble.hip / lu.hip) ----------------------
int assemble_impl(bemhip_ctx* c);
int apply_bc_impl(bemhip_ctx* c);
int operator_apply(bemhip_ctx* c, const double* d_x, double* d_y);
int lu_factor_device(bemhip_ctx* c, int n, int dtype, void* dA, int lda, int* dipiv,
                     int* info_host);
int lu_solve_device(bemhip_ctx* c, int n, int dtype, const void* dLU, int lda,
                    const int* dperm, double* d_b);
int lu_solve_trans_device(bemhip_ctx* c, int n, int dtype, const void* dLU, int lda,
                          const int* dperm, double* d_b);
int build_perm(bemhip_ctx* c, int n, const int* dipiv, int* dperm);
int operator_apply_trans(bemhip_ctx* c, const double* d_v, double* d_y);
int fmm_build(bemhip_ctx* c);
Geo make_geo(bemhip_ctx* c);
hipError_t ensure_count_word(bemhip_ctx* c);  // pinned host word for the near-pair count
int kelvin_assemble_impl(bemhip_ctx* c);
int kelvin_apply_bc_impl(bemhip_ctx* c);
int kelvin_operator_apply(bemhip_ctx* c, const double* d_x, double* d_y);
int kelvin_jacobi_diag(bemhip_ctx* c, double* d);
int fmm_far(bemhip_ctx* c, hipStream_t s, const double* wH, const double* wG, double* out);
int interior_impl(bemhip_ctx* c, int npts, const double* d_px, const double* d_py,
                  const double* d_w, const QuadConst* d_qc, int npg, double near_factor,
                  double* d_out);

// Right preconditioner of the GMRES core (gmres.hip): x = P^-1 u.
// With trans, GMRES solves A^T x = rhs (operator_apply_trans, the FP32 factors
// applied transposed); JACOBI is for A only.
struct Precond {
  enum Kind { JACOBI, LU32 } kind = JACOBI;
  const double* d = nullptr;  // JACOBI: diag A
  bool trans = false;
};
int precond_apply(bemhip_ctx* c, const Precond& P, int n, const double* in, double* out);
int gmres_core(bemhip_ctx* c, int n, const double* rhs, double* dx, bool x_zero, double atol,
               double rtol, int m, int itmax, const Precond& P, int* iters, int* cycles,
               double* beta_out);
// GMRES-IR from the FP32-LU solution, for A x = b64 (trans: A^T x = rhs):
// see solve.hip.
int gmres_ir(bemhip_ctx* c, int nt, bool trans, const double* rhs, double* dx, double* dr,
             double* dy, double tol, int maxit, int* outer, int* inner, bool* converged);
// max |a_i| (and max |b_i|) on the device, read back.
int absmax2(bemhip_ctx* c, int n, const double* a, const double* b, double* ha, double* hb);

// Small host helpers
void gauss_legendre_host(int n, double* x, double* w);

struct ScopedTimer {
  bemhip_ctx* c;
  const char* key;
  ScopedTimer(bemhip_ctx* ctx, const char* k) : c(ctx), key(k) {
    hipEventRecord(c->ev0, c->stream);
  }
  ~ScopedTimer() {
    hipEventRecord(c->ev1, c->stream);
    hipEventSynchronize(c->ev1);
    float t = 0.f;
    hipEventElapsedTime(&t, c->ev0, c->ev1);
    c->ms[key] = t;
  }
};

}  // namespace bemhip
