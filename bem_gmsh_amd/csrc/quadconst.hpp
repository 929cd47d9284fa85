// Compile-time limits and the quadrature constants of the near/singular
// rules, shared by the assembly kernels and the adjoint. Plain C++ (no HIP
// types), so the rule templates that read it compile for any target.
#pragma once

namespace bemhip {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxNpg = 64;        // largest near-field Gauss rule
constexpr int kNinterp = 20;       // Guiggiani interpolation nodes (reference default)
constexpr int kEB = 32;            // elements per far-field column block (one bit word)
constexpr int kFarThreads = 256;   // rows per far-field workgroup

// Quadrature constants shared by the near/singular kernels (device buffer).
struct QuadConst {
  int npg;
  int p;
  double gl_x[kMaxNpg];     // near-field Gauss-Legendre rule
  double gl_w[kMaxNpg];
  double g20_x[kNinterp];   // interpolated-Guiggiani rule (ninterp = 20)
  double g20_w[kNinterp];
  double g20_bw[kNinterp];  // barycentric weights 1/prod(x_i - x_j)
  double lp_x[4];           // Legendre(p) collocation nodes
  double lp_bw[4];          // Legendre(p) barycentric weights
  double lp_D[16];          // Legendre(p) differentiation matrix, row-major
  double eq_x[4];           // Equispaced(p) nodes
  double eq_bw[4];          // Equispaced(p) barycentric weights
  double eq_D[16];          // Equispaced(p) differentiation matrix, row-major
  double ew[4];             // element Gauss weights (dad.elem_weight)
  double k;                 // conductivity
  double near_factor;
};

}  // namespace bemhip
