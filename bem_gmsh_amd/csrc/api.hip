// C ABI of libbemhip.so (declared in include/bemhip.h): context management,
// geometry upload, orchestration of the assembly / BC / solve kernels.

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <sstream>
#include <type_traits>

namespace bemhip {

static std::mutex g_err_mu;
static std::string g_err;

void set_global_error(const std::string& s) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = s;
}

int fail(bemhip_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  set_global_error(msg);
  return code;
}

int check_matrix_free_scratch(bemhip_ctx* c, const char* who, size_t per_pair_block, size_t ncb,
                              size_t rows) {
  const double need = (double)per_pair_block * (double)ncb * (double)rows;
  size_t free_b = 0, total_b = 0;
  BEMHIP_HIP_TRY(c, hipMemGetInfo(&free_b, &total_b));
  // buffers this context already holds count as available
  const double held = (double)c->nearbits.count * sizeof(uint32_t) +
                      (double)c->hpart.count * sizeof(double) +
                      (double)c->part1.count * sizeof(double);
  if (need > 0.9 * ((double)free_b + held)) {
    char msg[512];
    snprintf(msg, sizeof(msg),
             "%s: the matrix-free assembly keeps %zu bytes per (row, 32-element block) of near "
             "bits and FP64 partial sums, %.1f GB for %zu rows x %zu blocks, above the %.1f GB "
             "the device has free; BEMHIP_FMM assembles in O(N) memory",
             who, per_pair_block, need * 1e-9, rows, ncb, (0.9 * ((double)free_b + held)) * 1e-9);
    return fail(c, BEMHIP_ERR_NOMEM, msg);
  }
  return BEMHIP_OK;
}

// Gauss-Legendre rule: the n roots of P_n by Newton on the three-term
// recurrence, weights 2 / ((1 - x^2) P_n'(x)^2), ascending order (the rule
// FastGaussQuadrature.gausslegendre returns; src/Core/Assembly_full.jl:719-723).
static void legendre_pn(int n, double x, double* pn, double* dpn) {
  double p0 = 1.0, p1 = x;
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

void gauss_legendre_host(int n, double* x, double* w) {
  if (n == 1) { x[0] = 0.0; w[0] = 2.0; return; }
  const int m = (n + 1) / 2;
  for (int i = 0; i < m; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double pn = 0, dpn = 0;
    for (int it = 0; it < 100; ++it) {
      legendre_pn(n, z, &pn, &dpn);
      const double dz = pn / dpn;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    legendre_pn(n, z, &pn, &dpn);
    const double wi = 2.0 / ((1.0 - z * z) * dpn * dpn);
    x[n - 1 - i] = z;
    x[i] = -z;
    w[n - 1 - i] = wi;
    w[i] = wi;
  }
  if (n % 2 == 1) {
    double pn, dpn;
    x[m - 1] = 0.0;
    legendre_pn(n, 0.0, &pn, &dpn);
    w[m - 1] = 2.0 / (dpn * dpn);
  }
}

// Barycentric bases of the element (src/Core/Interpolation.jl:62-71, :98,
// :127, :193-213).
static void diff_matrix(int n, const double* x, const double* bw, double* D) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double t = (bw[j] / bw[i]) / (x[i] - x[j]);
      D[i * n + j] = t;
      s += t;
    }
    D[i * n + i] = -s;
  }
}

static void build_quadconst(bemhip_ctx* c, QuadConst* q, int npg, double near_factor) {
  std::memset(q, 0, sizeof(*q));
  q->npg = npg;
  q->p = c->p;
  gauss_legendre_host(npg, q->gl_x, q->gl_w);
  gauss_legendre_host(kNinterp, q->g20_x, q->g20_w);
  for (int i = 0; i < kNinterp; ++i) {
    double wi = 1.0;
    for (int j = 0; j < i; ++j) wi *= q->g20_x[i] - q->g20_x[j];
    for (int j = i + 1; j < kNinterp; ++j) wi *= q->g20_x[i] - q->g20_x[j];
    q->g20_bw[i] = 1.0 / wi;
  }
  const int nn = c->nn;
  double w[4];
  gauss_legendre_host(nn, q->lp_x, w);
  for (int i = 0; i < nn; ++i) {
    const double s = ((i + 1) % 2 == 1) ? 1.0 : -1.0;
    q->lp_bw[i] = s * std::sqrt((1.0 - q->lp_x[i] * q->lp_x[i]) * w[i]);
  }
  diff_matrix(nn, q->lp_x, q->lp_bw, q->lp_D);
  const int p = c->p;
  for (int j = 0; j <= p; ++j) {
    q->eq_x[j] = (-1.0 * (p - j) + 1.0 * j) / p;
    const bool odd_n = (p % 2) == 1, even_j = (j % 2) == 0;
    double binom = 1.0;
    for (int t = 1; t <= j; ++t) binom = binom * (p - j + t) / t;
    q->eq_bw[j] = ((odd_n != even_j) ? 1.0 : -1.0) * std::round(binom);
  }
  diff_matrix(nn, q->eq_x, q->eq_bw, q->eq_D);
  for (int j = 0; j < nn && j < (int)c->h_elem_w.size(); ++j) q->ew[j] = c->h_elem_w[j];
  q->k = c->k;
  q->near_factor = near_factor;
}

// Device matrix (rows x cols, leading dimension ld) -> dense host column-major
// (leading dimension rows), converting the element type when TD != TS.
template <typename TD, typename TS>
int copy_out(bemhip_ctx* c, TD* dst, const TS* src, size_t rows, size_t cols, size_t ld) {
  if (std::is_same<TD, TS>::value) {
    BEMHIP_HIP_TRY(c, hipMemcpy2D(dst, rows * sizeof(TS), src, ld * sizeof(TS), rows * sizeof(TS),
                                  cols, hipMemcpyDeviceToHost));
    return BEMHIP_OK;
  }
  std::vector<TS> tmp(rows * cols);
  BEMHIP_HIP_TRY(c, hipMemcpy2D(tmp.data(), rows * sizeof(TS), src, ld * sizeof(TS),
                                rows * sizeof(TS), cols, hipMemcpyDeviceToHost));
  for (size_t t = 0; t < rows * cols; ++t) dst[t] = (TD)tmp[t];
  return BEMHIP_OK;
}

template <typename TD>
int get_HG(bemhip_ctx* c, TD* H, TD* G) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "get_HG: NULL context");
  if (!c->assembled) return fail(c, BEMHIP_ERR_STATE, "get_HG: assemble first");
  if (c->precision == BEMHIP_MATRIX_FREE || c->precision == BEMHIP_FMM)
    return fail(c, BEMHIP_ERR_STATE, "get_HG: a matrix-free assembly stores no H, G");
  hipSetDevice(c->device);
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  const size_t nt = c->nsys(), n = c->nbnd();
  int st = 0;
  if (c->precision == BEMHIP_FP64) {
    if (H && (st = copy_out(c, H, c->H64.ptr, nt, nt, c->ld64))) return st;
    if (G && (st = copy_out(c, G, c->G64.ptr, nt, n, c->ld64))) return st;
  } else {
    if (H && (st = copy_out(c, H, c->H32.ptr, nt, nt, c->ld32))) return st;
    if (G && (st = copy_out(c, G, c->G32.ptr, nt, n, c->ld32))) return st;
  }
  return BEMHIP_OK;
}

}  // namespace bemhip

using namespace bemhip;

static const char* kStatusNames[] = {"ok",           "bad argument",   "HIP runtime error",
                                     "invalid state", "singular matrix", "no convergence",
                                     "out of memory", "no device"};

extern "C" {

int bemhip_version(void) { return 10000; }

const char* bemhip_strerror(int status) {
  if (status < 0 || status > 7) return "unknown status";
  return kStatusNames[status];
}

const char* bemhip_last_error(const bemhip_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  static thread_local std::string copy;
  {
    std::lock_guard<std::mutex> lk(g_err_mu);
    copy = g_err;
  }
  return copy.c_str();
}

int bemhip_gauss_legendre(int n, double* x, double* w) {
  if (n < 1 || !x || !w) return fail(nullptr, BEMHIP_ERR_ARG, "gauss_legendre: n must be >= 1");
  gauss_legendre_host(n, x, w);
  return BEMHIP_OK;
}

int bemhip_device_count(int* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = (e == hipSuccess) ? n : 0;
  if (e != hipSuccess || n == 0)
    return fail(nullptr, BEMHIP_ERR_NODEV, "no HIP device visible");
  return BEMHIP_OK;
}

int bemhip_create(bemhip_ctx** out, int device_id) {
  if (!out) return fail(nullptr, BEMHIP_ERR_ARG, "create: out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(nullptr, BEMHIP_ERR_NODEV, "create: no HIP device visible");
  if (device_id < 0 || device_id >= n)
    return fail(nullptr, BEMHIP_ERR_ARG, "create: device_id out of range");
  auto* c = new bemhip_ctx();
  c->device = device_id;
  if (hipSetDevice(device_id) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipStreamCreate(&c->stream2) != hipSuccess || hipStreamCreate(&c->stream3) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->stream2) hipStreamDestroy(c->stream2);
    if (c->stream3) hipStreamDestroy(c->stream3);
    delete c;
    return fail(nullptr, BEMHIP_ERR_HIP, "create: HIP stream/event creation failed");
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_id) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  *out = c;
  return BEMHIP_OK;
}

void bemhip_destroy(bemhip_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  hipStreamSynchronize(c->stream2);
  hipStreamSynchronize(c->stream3);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  hipStreamDestroy(c->stream);
  hipStreamDestroy(c->stream2);
  hipStreamDestroy(c->stream3);
  if (c->h_count) hipHostFree(c->h_count);
  if (c->h_flag) hipHostFree(c->h_flag);
  delete c;
}

int bemhip_set_geometry(bemhip_ctx* c, int p, int ne, int n, int ni, const double* coll_xy,
                        const double* nrm_xy, const double* J, const double* elem_w,
                        const double* geo_xy, const double* elen, double k) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "set_geometry: NULL context");
  if (p != 1 && p != 2) return fail(c, BEMHIP_ERR_ARG, "set_geometry: element order must be 1 or 2");
  if (ne < 1 || n != ne * (p + 1) || ni < 0)
    return fail(c, BEMHIP_ERR_ARG, "set_geometry: need ne >= 1, n == ne*(p+1), ni >= 0");
  if (!coll_xy || !nrm_xy || !J || !elem_w || !geo_xy || !elen)
    return fail(c, BEMHIP_ERR_ARG, "set_geometry: NULL array");
  if (!(k > 0.0) || !std::isfinite(k)) return fail(c, BEMHIP_ERR_ARG, "set_geometry: k must be > 0");
  const int nn = p + 1, nt = n + ni;
  if (nt > (1 << 26)) return fail(c, BEMHIP_ERR_ARG, "set_geometry: nt too large");
  for (int64_t t = 0; t < (int64_t)nt * 2; ++t)
    if (!std::isfinite(coll_xy[t])) return fail(c, BEMHIP_ERR_ARG, "set_geometry: non-finite coordinate");
  for (int e = 0; e < ne; ++e)
    if (!(elen[e] >= 0.0) || !std::isfinite(elen[e]))
      return fail(c, BEMHIP_ERR_ARG, "set_geometry: invalid element length");
  hipSetDevice(c->device);
  c->has_geo = false;
  c->assembled = false;
  c->has_bc = false;
  c->factored32 = c->factored64 = false;
  c->has_sol = false;
  c->p = p; c->nn = nn; c->ne = ne; c->n = n; c->ni = ni; c->nt = nt; c->k = k;
  c->dof = 1;  // Laplace until bemhip_set_elasticity
  c->mu = c->nu = 0.0;
  c->ld32 = padded_ld(nt, 64);
  c->ld64 = padded_ld(nt, 32);
  c->h_elem_w.assign(elem_w, elem_w + nn);
  std::vector<double> hx(nt), hy(nt), hnx(n), hny(n), hjw(n), hgx((size_t)ne * nn),
      hgy((size_t)ne * nn);
  for (int i = 0; i < nt; ++i) { hx[i] = coll_xy[2 * (size_t)i]; hy[i] = coll_xy[2 * (size_t)i + 1]; }
  for (int j = 0; j < n; ++j) {
    hnx[j] = nrm_xy[2 * (size_t)j];
    hny[j] = nrm_xy[2 * (size_t)j + 1];
    hjw[j] = J[j] * elem_w[j % nn];  // el.Jacobian[k] * dad.elem_weight[k]
  }
  for (size_t t = 0; t < (size_t)ne * nn; ++t) { hgx[t] = geo_xy[2 * t]; hgy[t] = geo_xy[2 * t + 1]; }
  auto up = [&](DevBuf<double>& d, const std::vector<double>& h) -> int {
    BEMHIP_HIP_TRY(c, d.ensure(h.size()));
    BEMHIP_HIP_TRY(c, hipMemcpyAsync(d.ptr, h.data(), sizeof(double) * h.size(),
                                     hipMemcpyHostToDevice, c->stream));
    return BEMHIP_OK;
  };
  int st;
  if ((st = up(c->x, hx)) || (st = up(c->y, hy)) || (st = up(c->nx, hnx)) ||
      (st = up(c->ny, hny)) || (st = up(c->jw, hjw)) || (st = up(c->gx, hgx)) ||
      (st = up(c->gy, hgy)))
    return st;
  std::vector<double> hl(elen, elen + ne);
  if ((st = up(c->elen, hl))) return st;
  std::vector<double> hJ(J, J + n);
  if ((st = up(c->Jn, hJ))) return st;
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->has_geo = true;
  return BEMHIP_OK;
}

int bemhip_set_elasticity(bemhip_ctx* c, double mu, double nu) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "set_elasticity: NULL context");
  if (!c->has_geo) return fail(c, BEMHIP_ERR_STATE, "set_elasticity: set_geometry first");
  if (!(mu > 0.0) || !std::isfinite(mu))
    return fail(c, BEMHIP_ERR_ARG, "set_elasticity: mu must be finite and > 0");
  if (!(nu > -1.0 && nu < 0.5))
    return fail(c, BEMHIP_ERR_ARG, "set_elasticity: nu must be in (-1, 0.5)");
  if ((int64_t)2 * c->nt > (int64_t)1 << 26)
    return fail(c, BEMHIP_ERR_ARG, "set_elasticity: nt too large");
  c->assembled = false;
  c->has_bc = false;
  c->factored32 = c->factored64 = false;
  c->has_sol = false;
  c->dof = 2;
  c->mu = mu;
  c->nu = nu;
  c->ld32 = padded_ld(2 * c->nt, 64);
  c->ld64 = padded_ld(2 * c->nt, 32);
  return BEMHIP_OK;
}

int bemhip_set_lu_pivoting(bemhip_ctx* c, int mode) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "set_lu_pivoting: NULL context");
  if (mode != BEMHIP_PIVOT_TOURNAMENT && mode != BEMHIP_PIVOT_PARTIAL)
    return fail(c, BEMHIP_ERR_ARG,
                "set_lu_pivoting: mode must be BEMHIP_PIVOT_TOURNAMENT or BEMHIP_PIVOT_PARTIAL");
  c->lu32_pivot = mode;
  c->factored32 = false;
  return BEMHIP_OK;
}

int bemhip_set_lu_timers(bemhip_ctx* c, int on) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "set_lu_timers: NULL context");
  c->lu_timers = on != 0;
  return BEMHIP_OK;
}

int bemhip_set_lu_gemm(bemhip_ctx* c, int mode) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "set_lu_gemm: NULL context");
  if (mode != BEMHIP_LU_GEMM_FP32 && mode != BEMHIP_LU_GEMM_SPLIT_BF16)
    return fail(c, BEMHIP_ERR_ARG,
                "set_lu_gemm: mode must be BEMHIP_LU_GEMM_FP32 or BEMHIP_LU_GEMM_SPLIT_BF16");
  c->lu32_gemm = mode;
  c->factored32 = false;
  return BEMHIP_OK;
}

int bemhip_assemble(bemhip_ctx* c, int npg, double near_factor, int precision) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "assemble: NULL context");
  if (!c->has_geo) return fail(c, BEMHIP_ERR_STATE, "assemble: set_geometry first");
  if (npg < 1 || npg > kMaxNpg) return fail(c, BEMHIP_ERR_ARG, "assemble: npg must be in 1..64");
  if (c->dof == 2 && precision == BEMHIP_FMM)
    return fail(c, BEMHIP_ERR_ARG,
                "assemble: elasticity has no FMM far field (BEMHIP_FP64, BEMHIP_MIXED or "
                "BEMHIP_MATRIX_FREE)");
  if (!(near_factor >= 0.0)) return fail(c, BEMHIP_ERR_ARG, "assemble: near_factor must be >= 0");
  if (precision != BEMHIP_FP64 && precision != BEMHIP_MIXED && precision != BEMHIP_MATRIX_FREE &&
      precision != BEMHIP_FMM)
    return fail(c, BEMHIP_ERR_ARG,
                "assemble: precision must be BEMHIP_FP64, BEMHIP_MIXED, BEMHIP_MATRIX_FREE or "
                "BEMHIP_FMM");
  if ((precision == BEMHIP_FP64 || precision == BEMHIP_MIXED) &&
      (int64_t)c->nsys() * (int64_t)c->nsys() > (int64_t)1 << 40)
    return fail(c, BEMHIP_ERR_ARG,
                "assemble: nt too large for stored H, G (use BEMHIP_MATRIX_FREE or BEMHIP_FMM)");
  hipSetDevice(c->device);
  c->assembled = false;
  c->has_bc = false;
  c->factored32 = c->factored64 = false;
  c->has_sol = false;
  c->npg = npg;
  c->near_factor = near_factor;
  c->precision = precision;
  QuadConst q;
  build_quadconst(c, &q, npg, near_factor);
  BEMHIP_HIP_TRY(c, c->qc.ensure(1));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->qc.ptr, &q, sizeof(q), hipMemcpyHostToDevice, c->stream));
  hipEventRecord(c->ev0, c->stream);
  hipEvent_t t0, t1;
  hipEventCreate(&t0);
  hipEventCreate(&t1);
  hipEventRecord(t0, c->stream);
  int st = (c->dof == 2) ? kelvin_assemble_impl(c) : assemble_impl(c);
  hipEventRecord(t1, c->stream);
  hipEventSynchronize(t1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, t0, t1);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->ms["assemble"] = ms;
  c->stats["npairs"] = (double)c->npairs;
  c->assembled = true;
  return BEMHIP_OK;
}

int bemhip_interior_laplace(bemhip_ctx* c, int npts, const double* pts_xy, const double* T,
                            const double* q, int npg, double near_factor, double* u) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "interior_laplace: NULL context");
  if (!c->has_geo) return fail(c, BEMHIP_ERR_STATE, "interior_laplace: set_geometry first");
  if (c->dof != 1) return fail(c, BEMHIP_ERR_STATE, "interior_laplace: Laplace problems only");
  if (npts < 0) return fail(c, BEMHIP_ERR_ARG, "interior_laplace: npts must be >= 0");
  if (npg < 1 || npg > kMaxNpg)
    return fail(c, BEMHIP_ERR_ARG, "interior_laplace: npg must be in 1..64");
  if (!(near_factor >= 0.0))
    return fail(c, BEMHIP_ERR_ARG, "interior_laplace: near_factor must be >= 0");
  if (npts == 0) return BEMHIP_OK;
  if (!pts_xy || !T || !q || !u) return fail(c, BEMHIP_ERR_ARG, "interior_laplace: NULL buffer");
  const int n = c->n;
  std::vector<double> hp(2 * (size_t)npts), hw(4 * (size_t)n);
  for (int i = 0; i < npts; ++i) {
    const double x = pts_xy[2 * (size_t)i], y = pts_xy[2 * (size_t)i + 1];
    if (!std::isfinite(x) || !std::isfinite(y))
      return fail(c, BEMHIP_ERR_ARG, "interior_laplace: non-finite point coordinate");
    hp[i] = x;
    hp[(size_t)npts + i] = y;
  }
  for (int j = 0; j < n; ++j) {
    hw[j] = -T[j];
    hw[(size_t)n + j] = q[j];
    hw[2 * (size_t)n + j] = 1.0;
    hw[3 * (size_t)n + j] = 0.0;
  }
  hipSetDevice(c->device);
  QuadConst qh;
  build_quadconst(c, &qh, npg, near_factor);
  BEMHIP_HIP_TRY(c, c->ev_qc.ensure(1));
  BEMHIP_HIP_TRY(c, c->ev_pts.ensure(hp.size()));
  BEMHIP_HIP_TRY(c, c->ev_w.ensure(hw.size()));
  BEMHIP_HIP_TRY(c, c->ev_out.ensure((size_t)npts));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->ev_qc.ptr, &qh, sizeof(qh), hipMemcpyHostToDevice,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->ev_pts.ptr, hp.data(), hp.size() * sizeof(double),
                                   hipMemcpyHostToDevice, c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->ev_w.ptr, hw.data(), hw.size() * sizeof(double),
                                   hipMemcpyHostToDevice, c->stream));
  hipEvent_t t0, t1;
  BEMHIP_HIP_TRY(c, hipEventCreate(&t0));
  BEMHIP_HIP_TRY(c, hipEventCreate(&t1));
  hipEventRecord(t0, c->stream);
  int st = interior_impl(c, npts, c->ev_pts.ptr, c->ev_pts.ptr + npts, c->ev_w.ptr,
                         c->ev_qc.ptr, npg, near_factor, c->ev_out.ptr);
  hipEventRecord(t1, c->stream);
  hipEventSynchronize(t1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, t0, t1);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipMemcpy(u, c->ev_out.ptr, sizeof(double) * npts, hipMemcpyDeviceToHost));
  c->ms["interior"] = ms;
  c->stats["interior_points"] = (double)npts;
  return BEMHIP_OK;
}

int bemhip_get_HG(bemhip_ctx* c, double* H, double* G) { return get_HG<double>(c, H, G); }
int bemhip_get_HG_f32(bemhip_ctx* c, float* H, float* G) { return get_HG<float>(c, H, G); }

int bemhip_near_pairs(bemhip_ctx* c, int64_t* count) {
  if (!c || !count) return fail(c, BEMHIP_ERR_ARG, "near_pairs: NULL argument");
  if (!c->assembled) return fail(c, BEMHIP_ERR_STATE, "near_pairs: assemble first");
  *count = c->npairs;
  return BEMHIP_OK;
}

int bemhip_get_near_pairs(bemhip_ctx* c, int* row, int* elem, double* h, double* g) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "get_near_pairs: NULL context");
  if (!c->assembled) return fail(c, BEMHIP_ERR_STATE, "get_near_pairs: assemble first");
  hipSetDevice(c->device);
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  const size_t np = (size_t)c->npairs;
  const size_t per = (size_t)c->nn * c->dof * c->dof;  // values per pair
  if (np == 0) return BEMHIP_OK;
  if (row) BEMHIP_HIP_TRY(c, hipMemcpy(row, c->pair_row.ptr, sizeof(int) * np, hipMemcpyDeviceToHost));
  if (elem) BEMHIP_HIP_TRY(c, hipMemcpy(elem, c->pair_elem.ptr, sizeof(int) * np, hipMemcpyDeviceToHost));
  if (h) BEMHIP_HIP_TRY(c, hipMemcpy(h, c->pair_h.ptr, sizeof(double) * np * per, hipMemcpyDeviceToHost));
  if (g) BEMHIP_HIP_TRY(c, hipMemcpy(g, c->pair_g.ptr, sizeof(double) * np * per, hipMemcpyDeviceToHost));
  return BEMHIP_OK;
}

int bemhip_apply_bc(bemhip_ctx* c, const int64_t* BC, const double* BV, double* A, double* b) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "apply_bc: NULL context");
  if (!c->assembled) return fail(c, BEMHIP_ERR_STATE, "apply_bc: assemble first");
  if (!BC || !BV) return fail(c, BEMHIP_ERR_ARG, "apply_bc: NULL BC/BV");
  if (A && (c->precision == BEMHIP_MATRIX_FREE || c->precision == BEMHIP_FMM))
    return fail(c, BEMHIP_ERR_STATE, "apply_bc: a matrix-free assembly stores no A");
  const int nbc = c->nbnd();
  std::vector<int> hbc(nbc);
  for (int j = 0; j < nbc; ++j) {
    if (BC[j] != 0 && BC[j] != 1)
      return fail(c, BEMHIP_ERR_ARG, "apply_bc: BC codes must be 0 (Dirichlet) or 1 (Neumann)");
    if (!std::isfinite(BV[j])) return fail(c, BEMHIP_ERR_ARG, "apply_bc: non-finite BV");
    hbc[j] = (int)BC[j];
  }
  hipSetDevice(c->device);
  c->has_bc = false;
  c->factored32 = c->factored64 = false;
  c->has_sol = false;
  BEMHIP_HIP_TRY(c, c->d_bc.ensure(nbc));
  BEMHIP_HIP_TRY(c, c->d_bv.ensure(nbc));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->d_bc.ptr, hbc.data(), sizeof(int) * nbc,
                                   hipMemcpyHostToDevice, c->stream));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->d_bv.ptr, BV, sizeof(double) * nbc,
                                   hipMemcpyHostToDevice, c->stream));
  c->h_bc.assign(BC, BC + nbc);
  c->h_bv.assign(BV, BV + nbc);
  int st = apply_bc_impl(c);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->has_bc = true;
  const size_t nt = c->nsys();
  if (b) BEMHIP_HIP_TRY(c, hipMemcpy(b, c->b64.ptr, sizeof(double) * nt, hipMemcpyDeviceToHost));
  if (A) {
    if (c->precision == BEMHIP_FP64) {
      if ((st = copy_out(c, A, c->A64.ptr, nt, nt, c->ld64))) return st;
    } else {
      if ((st = copy_out(c, A, c->A32.ptr, nt, nt, c->ld32))) return st;
    }
  }
  return BEMHIP_OK;
}

int bemhip_apply_operator(bemhip_ctx* c, const double* x, double* y) {
  if (!c || !x || !y) return fail(c, BEMHIP_ERR_ARG, "apply_operator: NULL argument");
  if (!c->has_bc) return fail(c, BEMHIP_ERR_STATE, "apply_operator: apply_bc first");
  hipSetDevice(c->device);
  const size_t nt = c->nsys();
  BEMHIP_HIP_TRY(c, c->vec64a.ensure(nt));
  BEMHIP_HIP_TRY(c, c->vec64b.ensure(nt));
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(c->vec64a.ptr, x, sizeof(double) * nt, hipMemcpyHostToDevice,
                                   c->stream));
  int st = operator_apply(c, c->vec64a.ptr, c->vec64b.ptr);
  if (st) return st;
  BEMHIP_HIP_TRY(c, hipMemcpyAsync(y, c->vec64b.ptr, sizeof(double) * nt, hipMemcpyDeviceToHost,
                                   c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return BEMHIP_OK;
}

int bemhip_counters(bemhip_ctx* c, char* json, size_t cap) {
  if (!c || !json || cap == 0) return fail(c, BEMHIP_ERR_ARG, "counters: NULL argument");
  std::ostringstream os;
  os.precision(9);
  os << "{\"device\":" << c->device << ",\"num_cu\":" << c->num_cu << ",\"p\":" << c->p
     << ",\"ne\":" << c->ne << ",\"n\":" << c->n << ",\"nt\":" << c->nt << ",\"dof\":" << c->dof
     << ",\"precision\":" << c->precision << ",\"npg\":" << c->npg << ",\"ms\":{";
  bool first = true;
  for (auto& kv : c->ms) {
    os << (first ? "" : ",") << "\"" << kv.first << "\":" << kv.second;
    first = false;
  }
  os << "},\"stats\":{";
  first = true;
  for (auto& kv : c->stats) {
    os << (first ? "" : ",") << "\"" << kv.first << "\":" << kv.second;
    first = false;
  }
  os << "}}";
  const std::string s = os.str();
  if (s.size() + 1 > cap) return fail(c, BEMHIP_ERR_ARG, "counters: buffer too small");
  std::memcpy(json, s.c_str(), s.size() + 1);
  return BEMHIP_OK;
}

int bemhip_synchronize(bemhip_ctx* c) {
  if (!c) return fail(nullptr, BEMHIP_ERR_ARG, "synchronize: NULL context");
  hipSetDevice(c->device);
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream2));
  BEMHIP_HIP_TRY(c, hipStreamSynchronize(c->stream3));
  return BEMHIP_OK;
}

}  // extern "C"
