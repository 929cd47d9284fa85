/*
 * bemhip.h — C ABI of libbemhip.so, the MI355X (gfx950) implementation of the
 * dense 2-D Laplace boundary-element path of BEM_gmsh:
 *
 *     format2d  ->  assemble!(dad, npg)  ->  solve(dad)
 *
 * Plain C types only (pointers + sizes); every call returns a status code
 * (BEMHIP_OK on success) and records a message readable with
 * bemhip_last_error(). No exception crosses this boundary.
 *
 * Memory: the caller owns host buffers; the context owns device buffers
 * (geometry, H, G, A, LU factors, pivots), which stay resident in HBM between
 * calls. All host matrices are column-major with leading dimension nt, exactly
 * the Julia layout of dad.H (nt x nt) and dad.G (nt x n); on the device the
 * library pads the leading dimension.
 *
 * Threading: a context is bound to one device and two HIP streams (default
 * priority; the second runs the far field during assembly and the panels of the
 * LU's look-ahead). Each call orders its work on both and returns when it is
 * done; calls on a context are serialized and not re-entrant. Separate contexts
 * are independent.
 *
 * Reference interfaces (paths relative to the BEM_gmsh repository root) are
 * cited on each declaration.
 */
#ifndef BEMHIP_H
#define BEMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bemhip_ctx bemhip_ctx;

/* Status codes. */
#define BEMHIP_OK 0
#define BEMHIP_ERR_ARG 1        /* invalid argument (reference: ArgumentError)    */
#define BEMHIP_ERR_HIP 2        /* HIP runtime failure                            */
#define BEMHIP_ERR_STATE 3      /* call out of order (reference: missing cache)   */
#define BEMHIP_ERR_SINGULAR 4   /* zero pivot in the LU                           */
#define BEMHIP_ERR_NOCONV 5     /* iterative refinement did not converge          */
#define BEMHIP_ERR_NOMEM 6      /* device allocation failed                       */
#define BEMHIP_ERR_NODEV 7      /* no usable GPU                                  */

/* Assembly precision. */
#define BEMHIP_FP64 0           /* H, G stored and computed in FP64               */
#define BEMHIP_MIXED 1          /* far field FP32, near/singular/diagonal FP64    */
#define BEMHIP_MATRIX_FREE 2    /* no stored H, G: near pairs and FP64 diagonal   */
                                /* only; the system is applied on the fly (GMRES) */
#define BEMHIP_FMM 3            /* as MATRIX_FREE, the far field by a fast        */
                                /* multipole method (O(N) per product)            */

/* Solve modes. */
#define BEMHIP_SOLVE_FP64 0     /* FP64 LU (MFMA trailing updates)                */
#define BEMHIP_SOLVE_MIXED 1    /* FP32 LU (MFMA) + FP64 iterative refinement     */
                                /* (GMRES-IR when refinement stops contracting)   */
#define BEMHIP_SOLVE_MIXED_GMRES 2  /* FP32 LU + GMRES-IR directly (no classical  */
                                /* refinement): for systems known to be ill-     */
                                /* conditioned (kappa * 6e-8 near or above 1)    */

/* Library version (major*10000 + minor*100 + patch). */
int bemhip_version(void);

/* Short name of a status code ("ok", "bad argument", ...). */
const char* bemhip_strerror(int status);

/* Last error message of a context, or of the library when ctx is NULL. */
const char* bemhip_last_error(const bemhip_ctx* ctx);

/* Gauss-Legendre nodes/weights on [-1,1], ascending (host only).
 * Replaces FastGaussQuadrature.gausslegendre as called from
 * src/Core/Assembly_full.jl:719-723 and src/Core/Input.jl:86-89. */
int bemhip_gauss_legendre(int n, double* x, double* w);

/* Number of visible GPUs (0 and BEMHIP_ERR_NODEV when none). */
int bemhip_device_count(int* count);

/* Create / destroy a context on device_id. Replaces the backend selection of
 * _ka_backend (src/Laplace/Assembly_GPU.jl:152-167). */
int bemhip_create(bemhip_ctx** out, int device_id);
void bemhip_destroy(bemhip_ctx* ctx);

/* Upload the collocation geometry built by format2d (src/Core/Input.jl:136-197,
 * packed like _pack_boundary2d, src/Laplace/Assembly_GPU.jl:169-233).
 *   p        element order (1 or 2); element e owns nodes e*(p+1) .. e*(p+1)+p
 *   ne, n    elements and boundary nodes (n == ne*(p+1)); ni internal nodes
 *   coll_xy  (nt x 2) collocation points, boundary first     (dad.collocation)
 *   nrm_xy   (n x 2) outward unit normals                    (dad.Normal)
 *   J        (n) |dx/dxi| at each boundary node              (Element.Jacobian)
 *   elem_w   (p+1) Gauss-Legendre weights of the element rule (dad.elem_weight)
 *   geo_xy   (ne x (p+1) x 2) geometric nodes, xi = -1..1 equispaced (Element.geo)
 *   elen     (ne) element lengths                             (Element.Length)
 *   k        conductivity (Laplace.k)
 * Invalidates any assembled matrices and factors. */
int bemhip_set_geometry(bemhip_ctx* ctx, int p, int ne, int n, int ni,
                        const double* coll_xy, const double* nrm_xy, const double* J,
                        const double* elem_w, const double* geo_xy, const double* elen,
                        double k);

/* Switch the context's problem to 2-D isotropic elasticity (the Kelvin kernel
 * of fundamental(props::Elasticity, r, n), src/Elasticity/Fundamental.jl:28-57),
 * as format2d(mesh, Elasticity(E, nu, rho)) selects it for the reference
 * (src/Core/Structures.jl:149-263). Call after bemhip_set_geometry (which
 * resets the problem to Laplace).
 *   mu   shear modulus E / (2 (1 + nu)) (Elasticity.mu)
 *   nu   effective Poisson ratio (effective_nu, src/Core/Structures.jl:260-263:
 *        nu for plane strain, nu / (1 + nu) for plane stress), in (-1, 0.5)
 * Every system size doubles: node j's component b is unknown 2j + b; H and A
 * are 2nt x 2nt, G is 2nt x 2n, BC and BV have 2n entries (per dof), x has
 * 2nt. bemhip_assemble then builds H_G_full_direct for a Vectorial problem
 * (src/Core/Assembly_full.jl:120-175; BEMHIP_FP64, or BEMHIP_MIXED: FP32
 * storage of FP64 values, FP64 right-hand side and refinement operator
 * recomputed from the geometry), bemhip_apply_bc the
 * per-dof applyBC (src/Core/Boundary_conditions.jl:153-175) and bemhip_solve
 * the solve of src/Core/Solver.jl:115-135 (split_sol!, :137-146, is left to
 * the caller). BEMHIP_MATRIX_FREE assembly and bemhip_solve_gmres work as for
 * Laplace; BEMHIP_FMM is Laplace-only (BEMHIP_ERR_ARG); the adjoint and
 * interior evaluation return BEMHIP_ERR_STATE. Invalidates any assembled
 * matrices and factors. */
int bemhip_set_elasticity(bemhip_ctx* ctx, double mu, double nu);

/* Assemble H (nt x nt) and G (nt x n) on the device. Replaces
 * H_G_full_direct (src/Core/Assembly_full.jl:61-85) behind
 * assemble!(dad; method, npg) (src/Core/Assembly_full.jl:107-118): far
 * nodal lumping (:801-810), near classification (:736-754), tan-p3c near
 * quadrature (:649-695, src/Core/Integration.jl:68-93, :1117-1365),
 * interpolated-Guiggiani on-element integrals (:397-424,
 * src/Core/Integration.jl:1597-1845) and the rigid-body diagonal (:384-390).
 *   npg          Gauss points of the near rule (>= 1, <= 64)
 *   near_factor  far-lumping cutoff factor (reference default 1.5; > 1e20 = all near)
 *   precision    BEMHIP_FP64, BEMHIP_MIXED, BEMHIP_MATRIX_FREE (near pairs and
 *                the FP64 rigid-body diagonal only, no H, G; the far field is
 *                recomputed in every product, for bemhip_solve_gmres;
 *                bemhip_get_HG and the LU solves then return BEMHIP_ERR_STATE),
 *                or BEMHIP_FMM (as MATRIX_FREE, the far field by a 2-D fast
 *                multipole method, O(N) per product; BEMHIP_ERR_ARG when the
 *                boundary is too coarse for a two-level quadtree) */
int bemhip_assemble(bemhip_ctx* ctx, int npg, double near_factor, int precision);

/* Copy the assembled H, G to host buffers (either may be NULL), converting to
 * double (FP32 storage in MIXED mode), or to float with the _f32 variant. */
int bemhip_get_HG(bemhip_ctx* ctx, double* H, double* G);
int bemhip_get_HG_f32(bemhip_ctx* ctx, float* H, float* G);

/* Near/on-element pair list of the last assembly (row-sorted CSR): count, and
 * per pair its row, element and the (p+1) FP64 values added to H and G.
 * Replaces the host mask of _near_skip_mask (src/Laplace/Assembly_GPU.jl:236-253)
 * and the near correction of _correct_near_cpu! (:322-342). */
int bemhip_near_pairs(bemhip_ctx* ctx, int64_t* count);
int bemhip_get_near_pairs(bemhip_ctx* ctx, int* row, int* elem, double* h, double* g);

/* Mixed-BC column swap: A = H with Dirichlet (BC==0) columns replaced by -G,
 * b = -sum_{BC=0} H[:,j] BV_j + sum_{BC=1} G[:,j] BV_j.
 * Replaces applyBC (src/Core/Boundary_conditions.jl:20-75). BC, BV have n
 * entries. A (nt x nt, FP64) and b (nt) are optional host copies; the system
 * stays on the device for bemhip_solve. */
int bemhip_apply_bc(bemhip_ctx* ctx, const int64_t* BC, const double* BV, double* A,
                    double* b);

/* Solve A x = b on the device and return x (nt). Replaces bem_linsolve
 * (src/Core/LinearSolveUtils.jl:44-49) as called by solve
 * (src/Core/Solver.jl:24-64); split_sol! (:88-98) is left to the caller.
 *   mode     BEMHIP_SOLVE_FP64 (needs FP64 assembly), BEMHIP_SOLVE_MIXED or
 *            BEMHIP_SOLVE_MIXED_GMRES. MIXED refines classically (d = (LU32)^-1 r,
 *            r = b - A x in FP64) and, once the correction stops halving above
 *            1e-10 ||x||, switches to GMRES-IR: A d = r by GMRES with the FP32
 *            factors as right preconditioner (Carson and Higham 2018); this
 *            stands where the reference falls through to an iterative solve
 *            (src/Core/Solver.jl:39-50). MIXED_GMRES starts with GMRES-IR.
 *   tol      refinement stop: ||dx||_inf <= tol * ||x||_inf (MIXED modes); GMRES-IR
 *            also stops when its corrections stagnate at the FP64 noise floor
 *            (kappa * eps64 <= 1e-6)
 *   maxit    cap on the classical steps and, separately, the GMRES-IR steps
 *   iters    (out, may be NULL) refinement steps taken (classical + GMRES-IR)
 *   relres   (out, may be NULL) final ||b - A x||_inf / ||b||_inf in FP64
 * The factors stay resident (re-used by later solves on the same system). */
int bemhip_solve(bemhip_ctx* ctx, int mode, double tol, int maxit, double* x, int* iters,
                 double* relres);

/* Solve A x = b for a new right-hand side b (host, nsys entries) with the FP64
 * factors of the last bemhip_solve(BEMHIP_SOLVE_FP64) on this context, and
 * allocate nothing: no hipMalloc, hipFree or allocating library call. No
 * counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, tight serving or time-stepping loops re-solving one
 * factored system). Its device scratch is a fixed 2.1 MB array per GPU, so
 * the system may have at most 65536 unknowns (BEMHIP_ERR_ARG beyond);
 * BEMHIP_ERR_STATE if the last solve was not an FP64 one of this system.
 * Every call on a GPU shares that GPU's scratch: calls that overlap in time
 * on one GPU (other threads, other contexts) race. Same arithmetic as the
 * triangular solves of bemhip_solve (no refinement, no residual report). */
int bemhip_solve_rhs_noalloc(bemhip_ctx* ctx, const double* b, double* x);

/* Restarted GMRES on the device for the current system, the iterative path of
 * solve_Hmat (src/Core/Solver.jl:66-86) for operators that are not stored:
 * every product y = A x is the FP64 operator of bemhip_apply_operator (far
 * field recomputed, near / on-element pairs, rigid-body diagonal), with right
 * Jacobi preconditioning (the diagonal of A). Works after any assembly; the
 * only solve after a BEMHIP_MATRIX_FREE assembly.
 *   atol, rtol  stop when ||b - A x||_2 <= atol + rtol ||b||_2 (reference
 *               defaults 1e-10, 1e-8)
 *   restart     Krylov basis size m (1..512): GMRES(m)
 *   itmax       cap on the products with A (> 0)
 *   x           (out, nt) solution
 *   iters       (out, may be NULL) products with A taken
 *   relres      (out, may be NULL) final ||b - A x||_2 / ||b||_2 from a fresh
 *               product
 * Returns BEMHIP_ERR_NOCONV (x still written) when itmax is reached first. */
int bemhip_solve_gmres(bemhip_ctx* ctx, double atol, double rtol, int restart, int itmax,
                       double* x, int* iters, double* relres);

/* y = A x in FP64 for the current system (stored A, or recomputed on the fly
 * from the FP64 kernels in MIXED mode). x, y have nt entries. Used for the
 * refinement residual; exposed for verification. */
int bemhip_apply_operator(bemhip_ctx* ctx, const double* x, double* y);

/* Pivoting of the FP32 LU (the factorization of BEMHIP_SOLVE_MIXED /
 * MIXED_GMRES and of bemhip_lu_factor with dtype 1). No counterpart in the
 * reference, whose LU is LAPACK's FP64 getrf (src/Core/LinearSolveUtils.jl:44-49).
 *   BEMHIP_PIVOT_PARTIAL (default): GEPP, LAPACK's pivots (ties aside)
 *   BEMHIP_PIVOT_TOURNAMENT: each 128-column panel chooses its pivot rows by
 *     tournament pivoting (TSLU of CALU: GEPP on 512-row leaves, the leaves'
 *     pivot rows reduced in a tree of GEPPs, no exchange between workgroups
 *     per column); |L| entries may exceed 1
 * The FP64 LU always pivots partially. Discards FP32 factors held by the
 * context. */
#define BEMHIP_PIVOT_PARTIAL 0
#define BEMHIP_PIVOT_TOURNAMENT 1
int bemhip_set_lu_pivoting(bemhip_ctx* ctx, int mode);

/* Arithmetic of the FP32 LU's large trailing updates (C -= A B with at least
 * one 256 x 256 tile per CU). No counterpart in the reference, whose LU is
 * LAPACK's FP64 getrf (src/Core/LinearSolveUtils.jl:44-49).
 *   BEMHIP_LU_GEMM_FP32: FP32 matrix cores (v_mfma_f32_16x16x4f32)
 *   BEMHIP_LU_GEMM_SPLIT_BF16 (default): every FP32 operand split exactly into three
 *     bf16 parts (x = h + m + l), the six partial products down to 2^-16 |a b|
 *     on the bf16 matrix cores with FP32 sums; the dropped terms stay within
 *     2^-23 (1 + 2^-9) |a b| per product (FP32's unit roundoff is 2^-24), at 2.67x less
 *     matrix-core time
 * Discards FP32 factors held by the context. */
#define BEMHIP_LU_GEMM_FP32 0
#define BEMHIP_LU_GEMM_SPLIT_BF16 1
int bemhip_set_lu_gemm(bemhip_ctx* ctx, int mode);

/* Phase timers of the LU (no counterpart in the reference): on != 0 records
 * HIP events around its panels, interchanges, triangular solves and updates,
 * reported by bemhip_counters as lu32_panel / lu32_swap / lu32_trsm /
 * lu32_gemm (lu64_* for FP64). Off by default: the event records lengthen
 * the panel chain (about 10 ms of a 280 ms FP32 LU at n = 32769). The
 * large-update timings (lu32_big_gemm) are always recorded. */
int bemhip_set_lu_timers(bemhip_ctx* ctx, int on);

/* Standalone dense LU of a host matrix (column-major n x n), the LU seam of
 * bem_linsolve (src/Core/LinearSolveUtils.jl:44-49): partial pivoting in FP64,
 * FP32 pivoting as set by bemhip_set_lu_pivoting.
 *   dtype 0: A is double*, 1: A is float*. A is overwritten by L\U (unit L)
 *   and ipiv (n, 0-based: row i was swapped with row ipiv[i]).
 * bemhip_lu_solve solves A x = b (trans = 0) or A^T x = b (trans = 1) with
 * those factors, in place of b (nrhs = 1). */
int bemhip_lu_factor(bemhip_ctx* ctx, int n, int dtype, void* A, int* ipiv);
int bemhip_lu_solve(bemhip_ctx* ctx, int n, int dtype, const void* LU, const int* ipiv,
                    double* b, int trans);

/* Adjoint shape gradient of a boundary-flux misfit. No counterpart in the
 * reference: it differentiates assemble -> applyBC -> solve
 * (src/Core/Assembly_full.jl:61-85, src/Core/Boundary_conditions.jl:20-75,
 * src/Core/Solver.jl:24-64) for the system of the last bemhip_solve:
 *   J = 1/2 sum_{j<n} w_j (q_j - q_target_j)^2
 *   q_target (n) target fluxes
 *   w        (n) weights, or NULL for the boundary quadrature weights
 *            J_j * elem_w[k] (then differentiated as well)
 *   dJ_dgeo  (out, ne x (p+1) x 2) dJ/d geo_xy, per element copy of each
 *            geometric node (an end node shared by two elements gets the sum
 *            of both entries)
 *   J        (out, may be NULL) the objective
 * BC values and internal points are held fixed. One transposed solve with the
 * factors of the last solve (FP64, or FP32 with FP64 refinement), then one
 * pass over the element kernels recomputed with their derivatives (near and
 * on-element rules in forward-mode dual numbers). */
int bemhip_adjoint(bemhip_ctx* ctx, const double* q_target, const double* w, double* dJ_dgeo,
                   double* J);

/* Potential at points inside the domain from a boundary solution. Replaces
 * interior_laplace (src/Laplace/Lubrication.jl:399-417), whose formula the
 * reference's tests use as _interior_T (test/laplace.jl:36-53):
 *   u(p) = (sum_j g_j(p) q_j - sum_j h_j(p) T_j) / c(p),   c(p) = -sum_j h_j(p)
 * over the n boundary nodes, h_j, g_j the integrals of T*, U* times the shape
 * function of node j over its element, seen from p.
 *   npts, pts_xy  points (npts x 2, interleaved like coll_xy); a point must not
 *                 lie on an element (there the reference switches to its
 *                 singular rule; this call does not)
 *   T, q          (n) boundary potential and flux (dad.T[1:n], dad.q[1:n])
 *   npg           Gauss points of the near rule (1..64, as bemhip_assemble)
 *   near_factor   > 1e20: every (point, element) pair through the tan-p3c near
 *                 rule, as interior_laplace does (integrate_element,
 *                 src/Core/Assembly_full.jl:262-313); otherwise the assembly's
 *                 rules with that cutoff (far nodal lumping), i.e. the value
 *                 of an internal row of H, G (src/Core/Assembly_full.jl:61-85)
 *   u             (out, npts)
 * Needs the geometry only; leaves the assembled system, factors and solution
 * of the context untouched. */
int bemhip_interior_laplace(bemhip_ctx* ctx, int npts, const double* pts_xy, const double* T,
                            const double* q, int npg, double near_factor, double* u);

/* Timings (ms, HIP events on the context's streams) and sizes of the last calls
 * as a JSON object. */
int bemhip_counters(bemhip_ctx* ctx, char* json, size_t cap);

/* Wait for all work queued on the context's streams. */
int bemhip_synchronize(bemhip_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* BEMHIP_H */
