// Restarted GMRES on the device (bemhip_solve_gmres): the iterative path of
// solve_Hmat (src/Core/Solver.jl:66-86) for a system that is applied, not
// stored. Every product y = A z is the FP64 operator of the mixed-BC system
// (operator_apply: the far field recomputed from the geometry, the near /
// on-element pair values, the rigid-body diagonal).
//
// GMRES(m) with right Jacobi preconditioning, A D^-1 u = b, x = D^-1 u with
// D = diag(A): column j of A is column j of H for Neumann / internal nodes
// (D_j = H_jj, the rigid-body diagonal) and -G's column for Dirichlet nodes
// (D_j = -G_jj, the on-element pair value). The Arnoldi basis V (nt x (m+1),
// FP64) stays on the device; each step orthogonalizes by classical
// Gram-Schmidt applied twice (CGS2: two passes of h = V^T w, w -= V h, one
// kernel each) and normalizes w; the (m+1) x m Hessenberg matrix, its Givens
// rotations and the small triangular solve live on the host. Inner products
// are deterministic: per-workgroup partial sums in a fixed order, then a
// fixed-order sum of the partials.
//
// Stop (as Krylov.gmres, which solve_Hmat calls): ||b - A x||_2 <= atol +
// rtol ||b||_2, tested on the Arnoldi residual estimate inside a cycle and on
// a fresh residual at every restart.
//
// The same GMRES core (gmres_core) also solves the correction equations of
// GMRES-based iterative refinement in the mixed solve (solve.hip), with the
// FP32 LU factors as the right preconditioner instead of Jacobi.

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace bemhip {

namespace {

constexpr int kDotBlocks = 256;   // workgroups of a partial inner-product pass
constexpr int kDotThreads = 256;

// part[blk * k + j] = sum over this workgroup's rows of V[:, j] . w, j < k.
__global__ __launch_bounds__(kDotThreads) void dots_kernel(int n, int k, const double* __restrict__ V,
                                                            size_t ldv, const double* __restrict__ w,
                                                            double* __restrict__ part) {
  __shared__ double s_red[kDotThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  for (int j = 0; j < k; ++j) {
    const double* v = V + (size_t)j * ldv;
    double s = 0.0;
    for (int i = blockIdx.x * kDotThreads + threadIdx.x; i < n; i += gridDim.x * kDotThreads)
      s += v[i] * w[i];
    s = wave_sum(s);
    if (lane == 0) s_red[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kDotThreads / kWave; ++q) t += s_red[q];
      part[(size_t)blockIdx.x * k + j] = t;
    }
    __syncthreads();
  }
}

// h[j] = sum over workgroups of part[blk * k + j] (fixed order).
__global__ void dots_finish(int k, int nblk, const double* __restrict__ part, double* __restrict__ h) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  double t = 0.0;
  for (int b = 0; b < nblk; ++b) t += part[(size_t)b * k + j];
  h[j] = t;
}

// w -= V[:, 0..k) h
__global__ void vsub_kernel(int n, int k, const double* __restrict__ V, size_t ldv,
                            const double* __restrict__ h, double* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += V[(size_t)j * ldv + i] * h[j];
  w[i] -= s;
}

// y = a * x
__global__ void scale_kernel(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

// z = x / d
__global__ void divide_kernel(int n, const double* __restrict__ x, const double* __restrict__ d,
                              double* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x[i] / d[i];
}

// t = V[:, 0..k) y
__global__ void combine_kernel(int n, int k, const double* __restrict__ V, size_t ldv,
                               const double* __restrict__ y, double* __restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += V[(size_t)j * ldv + i] * y[j];
  t[i] = s;
}

// x += z
__global__ void add_kernel(int n, const double* __restrict__ z, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += z[i];
}

// r = b - y
__global__ void resid_kernel(int n, const double* __restrict__ b, const double* __restrict__ y,
                             double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

// Diagonal of A: H_ii (rigid-body diagonal) for Neumann and internal nodes,
// -G_ii (the node's own entry of its on-element pair) for Dirichlet nodes.
__global__ void jacobi_diag_kernel(int nt, int n, int nn, const int* __restrict__ bc,
                                   const double* __restrict__ diag, const int* __restrict__ row_ptr,
                                   const int* __restrict__ pair_elem,
                                   const double* __restrict__ pair_g, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  double v = diag[i];
  if (i < n && bc[i] == 0) {
    const int own = i / nn;
    v = 0.0;
    for (int pr = row_ptr[i]; pr < row_ptr[i + 1]; ++pr)
      if (pair_elem[pr] == own) v = -pair_g[(size_t)pr * nn + (i - own * nn)];
  }
  d[i] = (v != 0.0 && isfinite(v)) ? v : 1.0;
}

}  // namespace

// h[0..k) = V[:, 0..k)^T w on the device, copied to the host.
static int dots(bemhip_ctx* c, int n, int k, const double* V, size_t ldv, const double* w,
                double* part, double* d_h, double* h_host) {
  hipLaunchKernelGGL(dots_kernel, dim3(kDotBlocks), dim3(kDotThreads), 0, c->stream, n, k, V, ldv,
                     w, part);
  hipLaunchKernelGGL(dots_finish, dim3((k + 255) / 256), dim3(256), 0, c->stream, k, kDotBlocks,
                     part, d_h);
  BEMHIP_HIP_TRY(c, hipGetLastError());
  if (h_host) {
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(h_host, d_h, sizeof(double) * k, hipMemcpyDeviceToHost,
                                     c->stream));
    BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return BEMHIP_OK;
}

static int norm2(bemhip_ctx* c, int n, const double* w, double* part, double* d_h, double* out) {
  double h = 0.0;
  int st = dots(c, n, 1, w, (size_t)n, w, part, d_h, &h);
  if (st) return st;
  *out = std::sqrt(h);
  return BEMHIP_OK;
}

// out = P^-1 in: Jacobi (divide by diag A) or the FP32 LU factors of A.
int precond_apply(bemhip_ctx* c, const Precond& P, int n, const double* in, double* out) {
  const int tb = 256, nb = (n + tb - 1) / tb;
  if (P.kind == Precond::JACOBI) {
    hipLaunchKernelGGL(divide_kernel, dim3(nb), dim3(tb), 0, c->stream, n, in, P.d, out);
    BEMHIP_HIP_TRY(c, hipGetLastError());
    return BEMHIP_OK;
  }
  if (!c->factored32) return fail(c, BEMHIP_ERR_STATE, "gmres: no FP32 factors to precondition with");
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(out, in, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                   c->stream));
  if (P.trans) return lu_solve_trans_device(c, n, 1, c->A32.ptr, (int)c->ld32, c->perm32.ptr, out);
  return lu_solve_device(c, n, 1, c->A32.ptr, (int)c->ld32, c->perm32.ptr, out);
}

static int apply_op(bemhip_ctx* c, bool trans, const double* x, double* y) {
  return trans ? operator_apply_trans(c, x, y) : operator_apply(c, x, y);
}

// GMRES(m) on A x = rhs with A P^-1 u = rhs, x = P^-1 u, continuing from the x
// in dx (x_zero: dx holds zeros, the first residual is rhs itself). Stops when
// ||rhs - A x||_2 <= atol + rtol ||r0||_2, r0 the first residual (Arnoldi
// estimate inside a cycle, a fresh residual at each restart), or after itmax
// products. beta_out: the last fresh residual norm.
int gmres_core(bemhip_ctx* c, int n, const double* rhs, double* dx, bool x_zero, double atol,
               double rtol, int m, int itmax, const Precond& P, int* iters, int* cycles_out,
               double* beta_out) {
  const size_t ldv = (size_t)n;
  DevBuf<double> V, work, small;
  BEMHIP_HIP_TRY(c, V.ensure(ldv * (m + 1)));
  BEMHIP_HIP_TRY(c, work.ensure(3 * (size_t)n));  // w (also r), z, t
  BEMHIP_HIP_TRY(c, small.ensure((size_t)kDotBlocks * (m + 1) + 2 * (size_t)(m + 1)));
  double* dw = work.ptr;
  double* dz = work.ptr + n;
  double* dt = work.ptr + 2 * (size_t)n;
  double* part = small.ptr;
  double* d_h = small.ptr + (size_t)kDotBlocks * (m + 1);
  double* d_y = d_h + (m + 1);
  const int tb = 256, nb = (n + tb - 1) / tb;
  int st = 0;
  if (x_zero) {
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(dw, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                     c->stream));
  } else {
    if ((st = apply_op(c, P.trans, dx, dz))) return st;
    hipLaunchKernelGGL(resid_kernel, dim3(nb), dim3(tb), 0, c->stream, n, rhs, dz, dw);
  }
  double beta = 0.0;
  if ((st = norm2(c, n, dw, part, d_h, &beta))) return st;
  const double tol = atol + rtol * beta;
  int it = 0, cycles = 0;
  bool converged = beta <= tol;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h(m + 1), h2(m + 1), y(m);
  while (!converged && it < itmax) {
    ++cycles;
    hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(tb), 0, c->stream, n, 1.0 / beta, dw, V.ptr);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;  // columns of this cycle
    for (; k < m && it < itmax; ++k) {
      double* vk = V.ptr + (size_t)k * ldv;
      double* vk1 = V.ptr + (size_t)(k + 1) * ldv;
      if ((st = precond_apply(c, P, n, vk, dz))) return st;
      if ((st = apply_op(c, P.trans, dz, dw))) return st;
      ++it;
      // CGS2
      if ((st = dots(c, n, k + 1, V.ptr, ldv, dw, part, d_h, h.data()))) return st;
      hipLaunchKernelGGL(vsub_kernel, dim3(nb), dim3(tb), 0, c->stream, n, k + 1, V.ptr, ldv, d_h,
                         dw);
      if ((st = dots(c, n, k + 1, V.ptr, ldv, dw, part, d_h, h2.data()))) return st;
      hipLaunchKernelGGL(vsub_kernel, dim3(nb), dim3(tb), 0, c->stream, n, k + 1, V.ptr, ldv, d_h,
                         dw);
      double hn = 0.0;
      if ((st = norm2(c, n, dw, part, d_h, &hn))) return st;
      for (int j = 0; j <= k; ++j) H[(size_t)k * (m + 1) + j] = h[j] + h2[j];
      H[(size_t)k * (m + 1) + k + 1] = hn;
      if (hn > 0.0)
        hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(tb), 0, c->stream, n, 1.0 / hn, dw, vk1);
      // previous rotations on column k, then a new one zeroing H(k+1, k)
      double* col = &H[(size_t)k * (m + 1)];
      for (int j = 0; j < k; ++j) {
        const double a = col[j], b = col[j + 1];
        col[j] = cs[j] * a + sn[j] * b;
        col[j + 1] = -sn[j] * a + cs[j] * b;
      }
      const double a = col[k], b = col[k + 1];
      const double r = std::hypot(a, b);
      cs[k] = r > 0.0 ? a / r : 1.0;
      sn[k] = r > 0.0 ? b / r : 0.0;
      col[k] = r;
      col[k + 1] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (std::fabs(g[k + 1]) <= tol || hn == 0.0) { ++k; break; }
    }
    // y = R^-1 g (upper triangular k x k), x += P^-1 V y
    for (int i = k - 1; i >= 0; --i) {
      double s = g[i];
      for (int j = i + 1; j < k; ++j) s -= H[(size_t)j * (m + 1) + i] * y[j];
      y[i] = s / H[(size_t)i * (m + 1) + i];
    }
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(d_y, y.data(), sizeof(double) * k, hipMemcpyHostToDevice,
                                     c->stream));
    hipLaunchKernelGGL(combine_kernel, dim3(nb), dim3(tb), 0, c->stream, n, k, V.ptr, ldv, d_y,
                       dt);
    if ((st = precond_apply(c, P, n, dt, dz))) return st;
    hipLaunchKernelGGL(add_kernel, dim3(nb), dim3(tb), 0, c->stream, n, dz, dx);
    // fresh residual
    if ((st = apply_op(c, P.trans, dx, dz))) return st;
    hipLaunchKernelGGL(resid_kernel, dim3(nb), dim3(tb), 0, c->stream, n, rhs, dz, dw);
    if ((st = norm2(c, n, dw, part, d_h, &beta))) return st;
    if (!std::isfinite(beta)) break;
    converged = beta <= tol;
  }
  BEMHIP_HIP_TRY(c, hipGetLastError());
  if (iters) *iters = it;
  if (cycles_out) *cycles_out = cycles;
  if (beta_out) *beta_out = beta;
  return converged ? BEMHIP_OK : BEMHIP_ERR_NOCONV;
}

}  // namespace bemhip

using namespace bemhip;

extern "C" int bemhip_solve_gmres(bemhip_ctx* c, double atol, double rtol, int restart, int itmax,
                                  double* x, int* iters, double* relres) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "solve_gmres: NULL context");
  if (!c->has_bc) return fail(c, BEMHIP_ERR_STATE, "solve_gmres: apply_bc first");
  if (!(atol >= 0.0) || !(rtol >= 0.0) || (atol == 0.0 && rtol == 0.0))
    return fail(c, BEMHIP_ERR_ARG, "solve_gmres: need atol, rtol >= 0, not both zero");
  if (restart < 1 || restart > 512 || itmax < 1)
    return fail(c, BEMHIP_ERR_ARG, "solve_gmres: restart must be in 1..512 and itmax >= 1");
  hipSetDevice(c->device);
  const int n = c->nsys();  // unknowns: dof * collocation points
  DevBuf<double> work, small;
  BEMHIP_HIP_TRY(c, work.ensure(2 * (size_t)n));  // x, D
  BEMHIP_HIP_TRY(c, small.ensure((size_t)kDotBlocks + 2));
  double* dx = work.ptr;
  double* dd = work.ptr + n;
  const int tb = 256, nb = (n + tb - 1) / tb;
  hipEvent_t t0, t1;
  BEMHIP_HIP_TRY(c, hipEventCreate(&t0));
  BEMHIP_HIP_TRY(c, hipEventCreate(&t1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() { hipEventDestroy(a); hipEventDestroy(b); }
  } ev{t0, t1};
  hipEventRecord(t0, c->stream);

  if (c->dof == 2) {
    int sj = kelvin_jacobi_diag(c, dd);
    if (sj) return sj;
  } else {
    hipLaunchKernelGGL(jacobi_diag_kernel, dim3(nb), dim3(tb), 0, c->stream, n, c->n, c->nn,
                       c->d_bc.ptr, c->diag64.ptr, c->row_ptr.ptr, c->pair_elem.ptr,
                       c->pair_g.ptr, dd);
  }
  BEMHIP_HIP_TRY(c, hipMemsetAsync(dx, 0, sizeof(double) * n, c->stream));
  int st = 0;
  double bnorm = 0.0;
  if ((st = norm2(c, n, c->b64.ptr, small.ptr, small.ptr + kDotBlocks, &bnorm))) return st;
  Precond P;
  P.kind = Precond::JACOBI;
  P.d = dd;
  int it = 0, cycles = 0;
  double beta = bnorm;
  st = gmres_core(c, n, c->b64.ptr, dx, true, atol, rtol, restart, itmax, P, &it, &cycles, &beta);
  if (st && st != BEMHIP_ERR_NOCONV) return st;
  const bool converged = st == BEMHIP_OK;
  hipEventRecord(t1, c->stream);
  hipEventSynchronize(t1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, t0, t1);
  c->ms["gmres"] = ms;
  c->stats["gmres_iters"] = it;
  c->stats["gmres_cycles"] = cycles;
  const double rel = bnorm > 0.0 ? beta / bnorm : beta;
  c->stats["gmres_relres"] = rel;
  if (x) {
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  if (iters) *iters = it;
  if (relres) *relres = rel;
  if (!converged)
    return fail(c, BEMHIP_ERR_NOCONV, "solve_gmres: tolerance not reached within itmax products");
  return BEMHIP_OK;
}
