// Linear solve of the mixed-BC system A x = b (bemhip_solve) and the
// standalone LU entry points. Replaces bem_linsolve
// (src/Core/LinearSolveUtils.jl:44-49) as called by solve (src/Core/Solver.jl:24-64).
//
//   BEMHIP_SOLVE_FP64   FP64 LU of the FP64-assembled A, one forward/backward solve.
//   BEMHIP_SOLVE_MIXED  FP32 LU (MFMA trailing updates) of A32, then FP64
//                       iterative refinement: r = b - A64 x with A64 applied in
//                       FP64 (recomputed far field + FP64 near/singular entries +
//                       FP64 diagonal), d = (LU32)^-1 r, x += d, until
//                       ||d||_inf <= tol ||x||_inf. When the corrections stop
//                       contracting above the FP64 noise floor, GMRES-IR with
//                       the FP32 factors as preconditioner takes over (the
//                       reference's solve falls through to an iterative solve
//                       when its factorization fails, src/Core/Solver.jl:39-50).

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iterator>
#include <vector>

namespace bemhip {

namespace {

__global__ void sub_kernel(int n, const double* __restrict__ b, const double* __restrict__ y,
                           double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

__global__ void axpy_kernel(int n, const double* __restrict__ d, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}

// A32 = FP32 of the column swap of H64, G64 (used once A64 holds FP64 factors).
// Workgroup row y takes columns y, y + gridDim.y, ...
__global__ void swap_round_kernel(int nt, int n, size_t lds, size_t ldd, const int* __restrict__ bc,
                                  const double* __restrict__ H, const double* __restrict__ G,
                                  float* __restrict__ A) {
  for (int j = blockIdx.y; j < nt; j += gridDim.y) {
    const bool dir = (j < n) && bc[j] == 0;
    const double* src = dir ? G + (size_t)j * lds : H + (size_t)j * lds;
    const double sg = dir ? -1.0 : 1.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += gridDim.x * blockDim.x)
      A[(size_t)j * ldd + i] = (float)(sg * src[i]);
  }
}

// A32 = FP32 of A64 (nt x nt; leading dimensions lds, ldd).
__global__ void round_cols_kernel(int nt, size_t lds, size_t ldd, const double* __restrict__ s,
                                  float* __restrict__ d) {
  for (int j = blockIdx.y; j < nt; j += gridDim.y)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += gridDim.x * blockDim.x)
      d[(size_t)j * ldd + i] = (float)s[(size_t)j * lds + i];
}

// max |v| over n entries into out[slot] (one workgroup).
__global__ __launch_bounds__(1024) void absmax_kernel(int n, const double* __restrict__ v,
                                                      double* __restrict__ out, int slot) {
  __shared__ double s[16];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(v[i]));
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s[0];
    for (int w = 1; w < 16; ++w) r = fmax(r, s[w]);
    out[slot] = r;
  }
}

}  // namespace

int absmax2(bemhip_ctx* c, int n, const double* a, const double* b, double* ha, double* hb) {
  BEMHIP_HIP_TRY(c, c->scal.ensure(2));
  double* slots = c->scal.ptr;
  hipLaunchKernelGGL(absmax_kernel, dim3(1), dim3(1024), 0, c->stream, n, a, slots, 0);
  if (b) hipLaunchKernelGGL(absmax_kernel, dim3(1), dim3(1024), 0, c->stream, n, b, slots, 1);
  double h[2] = {0, 0};
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(h, slots, sizeof(double) * (b ? 2 : 1), hipMemcpyDeviceToHost,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  *ha = h[0];
  if (hb) *hb = h[1];
  return BEMHIP_OK;
}

// GMRES-based iterative refinement (Carson and Higham, SIAM J. Sci. Comput.
// 40 (2018) A817-A847): x = (LU32)^-1 rhs, then r = rhs - A x in FP64 and the
// correction A d = r solved by GMRES right-preconditioned by the FP32 factors
// (one cycle of at most 30 products, to ||r - A d||_2 <= 1e-6 ||r||_2), x += d,
// until ||d||_inf <= tol ||x||_inf, or until d stops contracting at the FP64
// noise floor (kappa eps64 <= 1e-6). It converges where classical refinement
// needs kappa u32 < 1 and does not. trans: the same for A^T x = rhs.
int gmres_ir(bemhip_ctx* c, int nt, bool trans, const double* rhs, double* dx, double* dr,
             double* dy, double tol, int maxit, int* outer, int* inner, bool* converged) {
  const int tb = 256, nb = (nt + tb - 1) / tb;
  constexpr int kInner = 30;
  int st = 0;
  BEMHIP_HIP_TRY(c, c->ir_d.ensure(nt));
  double* dd = c->ir_d.ptr;
  Precond P;
  P.kind = Precond::LU32;
  P.trans = trans;
  if ((st = precond_apply(c, P, nt, rhs, dx))) return st;
  double dn_prev = INFINITY;
  int k = 0, total = 0;
  *converged = false;
  for (k = 0; k < maxit; ++k) {
    if ((st = trans ? operator_apply_trans(c, dx, dy) : operator_apply(c, dx, dy))) return st;
    hipLaunchKernelGGL(sub_kernel, dim3(nb), dim3(tb), 0, c->stream, nt, rhs, dy, dr);
    BEMHIP_HIP_TRY(c, hipMemsetAsync(dd, 0, sizeof(double) * nt, c->stream));
    int its = 0;
    st = gmres_core(c, nt, dr, dd, true, 0.0, 1e-6, kInner, kInner, P, &its, nullptr, nullptr);
    total += its;
    if (st && st != BEMHIP_ERR_NOCONV) return st;
    double dn = 0, xn = 0;
    if ((st = absmax2(c, nt, dd, dx, &dn, &xn))) return st;
    if (k < 40) {
      char key[40];
      std::snprintf(key, sizeof(key), "refine_gmres_dx_%02d", k);
      c->stats[key] = xn > 0 ? dn / xn : dn;
      std::snprintf(key, sizeof(key), "refine_gmres_its_%02d", k);
      c->stats[key] = its;
    }
    if (!std::isfinite(dn)) break;
    hipLaunchKernelGGL(axpy_kernel, dim3(nb), dim3(tb), 0, c->stream, nt, dd, dx);
    if (dn <= tol * xn) { *converged = true; ++k; break; }
    if (k >= 1 && dn > 0.5 * dn_prev && dn <= 1e-6 * xn) { *converged = true; ++k; break; }
    dn_prev = dn;
  }
  *outer = k;
  *inner = total;
  return BEMHIP_OK;
}

}  // namespace bemhip

using namespace bemhip;

extern "C" int bemhip_solve(bemhip_ctx* c, int mode, double tol, int maxit, double* x, int* iters,
                            double* relres) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "solve: NULL context");
  if (!c->has_bc) return fail(c, BEMHIP_ERR_STATE, "solve: apply_bc first");
  if (mode != BEMHIP_SOLVE_FP64 && mode != BEMHIP_SOLVE_MIXED && mode != BEMHIP_SOLVE_MIXED_GMRES)
    return fail(c, BEMHIP_ERR_ARG,
                "solve: mode must be BEMHIP_SOLVE_FP64, BEMHIP_SOLVE_MIXED or BEMHIP_SOLVE_MIXED_GMRES");
  if (c->precision == BEMHIP_MATRIX_FREE || c->precision == BEMHIP_FMM)
    return fail(c, BEMHIP_ERR_STATE, "solve: LU solves need stored matrices; use bemhip_solve_gmres");
  if (mode == BEMHIP_SOLVE_FP64 && c->precision != BEMHIP_FP64)
    return fail(c, BEMHIP_ERR_STATE, "solve: FP64 LU needs an FP64 assembly (precision 0)");
  if (mode != BEMHIP_SOLVE_FP64 && (!(tol > 0.0) || maxit < 0))
    return fail(c, BEMHIP_ERR_ARG, "solve: tol must be > 0 and maxit >= 0");
  hipSetDevice(c->device);
  const int nt = c->nsys();  // rows of A: dof * collocation points
  BEMHIP_HIP_TRY(c, c->ipiv.ensure(nt));
  BEMHIP_HIP_TRY(c, c->perm.ensure(nt));
  BEMHIP_HIP_TRY(c, c->ipiv32.ensure(nt));
  BEMHIP_HIP_TRY(c, c->perm32.ensure(nt));
  BEMHIP_HIP_TRY(c, c->vec64a.ensure(nt));  // x
  BEMHIP_HIP_TRY(c, c->vec64b.ensure(std::max(nt, 4)));
  BEMHIP_HIP_TRY(c, c->part2.ensure(2 * (size_t)nt));  // r, y
  double* dx = c->vec64a.ptr;
  double* dr = c->part2.ptr;
  double* dy = c->part2.ptr + nt;
  const int tb = 256, nb = (nt + tb - 1) / tb;
  hipEvent_t t0, t1;
  hipEventCreate(&t0);
  hipEventCreate(&t1);
  hipEventRecord(t0, c->stream);
  int st = 0, info = 0, it = 0;
  double rel = NAN;
  if (mode == BEMHIP_SOLVE_FP64) {
    if (!c->factored64) {
      hipEvent_t f0, f1;
      hipEventCreate(&f0);
      hipEventCreate(&f1);
      hipEventRecord(f0, c->stream);
      st = lu_factor_device(c, nt, 0, c->A64.ptr, (int)c->ld64, c->ipiv.ptr, &info);
      hipEventRecord(f1, c->stream);
      hipEventSynchronize(f1);
      float ms = 0;
      hipEventElapsedTime(&ms, f0, f1);
      c->ms["lu64"] = ms;
      hipEventDestroy(f0);
      hipEventDestroy(f1);
      if (st) return st;
      if (info) return fail(c, BEMHIP_ERR_SINGULAR, "solve: zero pivot in the FP64 LU");
      if ((st = build_perm(c, nt, c->ipiv.ptr, c->perm.ptr))) return st;
      c->factored64 = true;
    }
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(dx, c->b64.ptr, sizeof(double) * nt, hipMemcpyDeviceToDevice,
                                     c->stream));
    if ((st = lu_solve_device(c, nt, 0, c->A64.ptr, (int)c->ld64, c->perm.ptr, dx))) return st;
  } else {
    if (!c->factored32) {
      if (c->precision == BEMHIP_FP64) {
        // FP32 copy of the FP64-assembled system (from H, G once A64 is factored)
        BEMHIP_HIP_TRY(c, ensure_matrix(c, c->A32, c->ld32 * nt, &c->H64, &c->G64, &c->A64));
        const dim3 grid((unsigned)std::min((nt + 255) / 256, 64), (unsigned)std::min(nt, 32768));
        if (c->factored64)
          hipLaunchKernelGGL(swap_round_kernel, grid, dim3(256), 0, c->stream, nt, c->nbnd(), c->ld64,
                             c->ld32, c->d_bc.ptr, c->H64.ptr, c->G64.ptr, c->A32.ptr);
        else
          hipLaunchKernelGGL(round_cols_kernel, grid, dim3(256), 0, c->stream, nt, c->ld64,
                             c->ld32, c->A64.ptr, c->A32.ptr);
      }
      hipEvent_t f0, f1;
      hipEventCreate(&f0);
      hipEventCreate(&f1);
      hipEventRecord(f0, c->stream);
      st = lu_factor_device(c, nt, 1, c->A32.ptr, (int)c->ld32, c->ipiv32.ptr, &info);
      hipEventRecord(f1, c->stream);
      hipEventSynchronize(f1);
      float ms = 0;
      hipEventElapsedTime(&ms, f0, f1);
      c->ms["lu32"] = ms;
      hipEventDestroy(f0);
      hipEventDestroy(f1);
      if (st) return st;
      if (info) return fail(c, BEMHIP_ERR_SINGULAR, "solve: zero pivot in the FP32 LU");
      if ((st = build_perm(c, nt, c->ipiv32.ptr, c->perm32.ptr))) return st;
      c->factored32 = true;
    }
    hipEvent_t r0, r1;
    hipEventCreate(&r0);
    hipEventCreate(&r1);
    hipEventRecord(r0, c->stream);
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(dx, c->b64.ptr, sizeof(double) * nt, hipMemcpyDeviceToDevice,
                                     c->stream));
    if ((st = lu_solve_device(c, nt, 1, c->A32.ptr, (int)c->ld32, c->perm32.ptr, dx))) return st;
    bool converged = false;
    double dn_prev = INFINITY;
    for (const char* key : {"refine_stagnated", "refine_predicted", "refine_gmres_ir",
                            "refine_gmres_outer", "refine_gmres_products"})
      c->stats.erase(key);
    for (auto it2 = c->stats.begin(); it2 != c->stats.end();)
      it2 = (it2->first.rfind("refine_dx_", 0) == 0 || it2->first.rfind("refine_gmres_", 0) == 0)
                ? c->stats.erase(it2)
                : std::next(it2);
    const int classical = (mode == BEMHIP_SOLVE_MIXED_GMRES) ? 0 : maxit;
    for (it = 0; it < classical; ++it) {
      if ((st = operator_apply(c, dx, dy))) return st;
      hipLaunchKernelGGL(sub_kernel, dim3(nb), dim3(tb), 0, c->stream, nt, c->b64.ptr, dy, dr);
      if ((st = lu_solve_device(c, nt, 1, c->A32.ptr, (int)c->ld32, c->perm32.ptr, dr)))
        return st;
      double dn = 0, xn = 0;
      if ((st = absmax2(c, nt, dr, dx, &dn, &xn))) return st;
      if (it < 40) {
        char key[32];
        std::snprintf(key, sizeof(key), "refine_dx_%02d", it);
        c->stats[key] = xn > 0 ? dn / xn : dn;
      }
      if (!std::isfinite(dn)) break;
      hipLaunchKernelGGL(axpy_kernel, dim3(nb), dim3(tb), 0, c->stream, nt, dr, dx);
      if (dn <= tol * xn) { converged = true; ++it; break; }
      // The error left after this correction is about rho * dn, rho = dn / dn_prev
      // the contraction of the step (the refinement converges linearly with the
      // rate ~ kappa u32): stop when that estimate is ten times below tol and the
      // step contracted by at least ten, instead of spending one more operator
      // pass and solve to observe a correction below tol.
      if (it >= 1 && dn < 0.1 * dn_prev && (dn / dn_prev) * dn <= 0.1 * tol * xn) {
        converged = true;
        c->stats["refine_predicted"] = 1.0;
        ++it;
        break;
      }
      // Stagnation at the FP64 noise floor (kappa * eps64): the correction no
      // longer halves and is already below 1e-10 relative.
      if (it >= 1 && dn > 0.5 * dn_prev && dn <= 1e-10 * xn) {
        converged = true;
        c->stats["refine_stagnated"] = 1.0;
        ++it;
        break;
      }
      // Contracting by less than half above that floor (kappa u32 near or
      // above 1): classical refinement would need tens of steps or diverge;
      // GMRES-IR below takes over.
      if (it >= 1 && dn > 0.5 * dn_prev) { ++it; break; }
      dn_prev = dn;
    }
    if (!converged && maxit > 0) {
      int outer = 0, inner = 0;
      if ((st = gmres_ir(c, nt, false, c->b64.ptr, dx, dr, dy, tol, maxit, &outer, &inner,
                         &converged)))
        return st;
      c->stats["refine_gmres_ir"] = 1.0;
      c->stats["refine_gmres_outer"] = outer;
      c->stats["refine_gmres_products"] = inner;
      it += outer;
    }
    hipEventRecord(r1, c->stream);
    hipEventSynchronize(r1);
    float ms = 0;
    hipEventElapsedTime(&ms, r0, r1);
    c->ms["refine"] = ms;
    hipEventDestroy(r0);
    hipEventDestroy(r1);
    c->stats["refine_iters"] = it;
    if (!converged && maxit > 0) {
      hipEventDestroy(t0);
      hipEventDestroy(t1);
      return fail(c, BEMHIP_ERR_NOCONV,
                  "solve: iterative refinement did not converge, nor did GMRES-IR");
    }
  }
  hipEventRecord(t1, c->stream);
  hipEventSynchronize(t1);
  float ms = 0;
  hipEventElapsedTime(&ms, t0, t1);
  c->ms[mode == BEMHIP_SOLVE_FP64 ? "solve64" : "solve_mixed"] = ms;
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  // final FP64 relative residual (reported, outside the timed solve)
  if ((st = operator_apply(c, dx, dy))) return st;
  hipLaunchKernelGGL(sub_kernel, dim3(nb), dim3(tb), 0, c->stream, nt, c->b64.ptr, dy, dr);
  double rn = 0, bn = 0;
  if ((st = absmax2(c, nt, dr, c->b64.ptr, &rn, &bn))) return st;
  rel = bn > 0 ? rn / bn : rn;
  c->stats["relres"] = rel;
  BEMHIP_HIP_TRY(c, c->xsol.ensure(nt));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->xsol.ptr, dx, sizeof(double) * nt, hipMemcpyDeviceToDevice,
                                   c->stream));
  c->has_sol = true;
  c->sol_mode = mode;
  if (x) BEMHIP_HIP_TRY(c, hipMemcpy(x, dx, sizeof(double) * nt, hipMemcpyDeviceToHost));
  if (iters) *iters = it;
  if (relres) *relres = rel;
  return BEMHIP_OK;
}

extern "C" int bemhip_lu_factor(bemhip_ctx* c, int n, int dtype, void* A, int* ipiv) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "lu_factor: NULL context");
  if (n < 1 || !A || !ipiv || (dtype != 0 && dtype != 1))
    return fail(c, BEMHIP_ERR_ARG, "lu_factor: need n >= 1, dtype 0/1 and non-NULL buffers");
  hipSetDevice(c->device);
  // device copy with the padded leading dimension of the resident matrices
  const size_t es = dtype == 0 ? 8 : 4;
  const size_t ld = padded_ld(n, dtype == 0 ? 32 : 64);
  DevBuf<char> dA;
  DevBuf<int> dp;
  BEMHIP_HIP_TRY(c, dA.ensure(es * ld * n));
  BEMHIP_HIP_TRY(c, dp.ensure(n));
  BEMHIP_HIP_TRY(c, hipMemcpy2DAsync(dA.ptr, es * ld, A, es * n, es * n, n, hipMemcpyHostToDevice,
                                     c->stream));
  int info = 0;
  hipEvent_t f0, f1;
  hipEventCreate(&f0);
  hipEventCreate(&f1);
  hipEventRecord(f0, c->stream);
  int st = lu_factor_device(c, n, dtype, dA.ptr, (int)ld, dp.ptr, &info);
  hipEventRecord(f1, c->stream);
  hipEventSynchronize(f1);
  float ms = 0;
  hipEventElapsedTime(&ms, f0, f1);
  c->ms[dtype == 0 ? "lu64" : "lu32"] = ms;
  hipEventDestroy(f0);
  hipEventDestroy(f1);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipMemcpy2DAsync(A, es * n, dA.ptr, es * ld, es * n, n, hipMemcpyDeviceToHost,
                                     c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(ipiv, dp.ptr, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (info) return fail(c, BEMHIP_ERR_SINGULAR, "lu_factor: zero pivot");
  return BEMHIP_OK;
}

extern "C" int bemhip_lu_solve(bemhip_ctx* c, int n, int dtype, const void* LU, const int* ipiv,
                               double* b, int trans) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "lu_solve: NULL context");
  if (n < 1 || !LU || !ipiv || !b || (dtype != 0 && dtype != 1))
    return fail(c, BEMHIP_ERR_ARG, "lu_solve: need n >= 1, dtype 0/1 and non-NULL buffers");
  if (trans != 0 && trans != 1) return fail(c, BEMHIP_ERR_ARG, "lu_solve: trans must be 0 or 1");
  for (int i = 0; i < n; ++i)
    if (ipiv[i] < i || ipiv[i] >= n) return fail(c, BEMHIP_ERR_ARG, "lu_solve: invalid ipiv");
  hipSetDevice(c->device);
  const size_t es = dtype == 0 ? 8 : 4;
  const size_t ld = padded_ld(n, dtype == 0 ? 32 : 64);
  DevBuf<char> dA;
  DevBuf<int> dp, dperm;
  DevBuf<double> db;
  BEMHIP_HIP_TRY(c, dA.ensure(es * ld * n));
  BEMHIP_HIP_TRY(c, dp.ensure(n));
  BEMHIP_HIP_TRY(c, dperm.ensure(n));
  BEMHIP_HIP_TRY(c, db.ensure(n));
  BEMHIP_HIP_TRY(c, hipMemcpy2DAsync(dA.ptr, es * ld, LU, es * n, es * n, n, hipMemcpyHostToDevice,
                                     c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(dp.ptr, ipiv, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(db.ptr, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  int st = build_perm(c, n, dp.ptr, dperm.ptr);
  if (st) return st;
  st = trans ? lu_solve_trans_device(c, n, dtype, dA.ptr, (int)ld, dperm.ptr, db.ptr)
             : lu_solve_device(c, n, dtype, dA.ptr, (int)ld, dperm.ptr, db.ptr);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(b, db.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return BEMHIP_OK;
}
