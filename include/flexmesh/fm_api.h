/*
 * fm_api.h -- C ABI of the MI355X quasi-static Newton path (P1-tet
 * Neo-Hookean residual and tangent, fixed-CSR assembly, ILU(0)-PCG, Newton).
 *
 * Plain pointers and sizes only; no exceptions cross this boundary.  Every call
 * returns an int status (FM_OK = 0, < 0 on error) and fm_last_error() gives the
 * message.  Host arrays are caller-owned and copied; device memory is owned by
 * the context.  One HIP stream per context; every call is synchronous on
 * return.  A context is not thread-safe; distinct contexts may be used
 * concurrently.
 *
 * Conventions (those of the reference engine, FEM/fem_types.hpp):
 *   DOF numbering node*3 + d; node coordinates SoA (x[nn], y[nn], z[nn]);
 *   connectivity SoA (conn[a*ne + e], a = 0..3); CSR with uint32 row_ptr and
 *   col_idx; FP64 values at the host boundary.
 */
#ifndef FLEXMESH_FM_API_H
#define FLEXMESH_FM_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FM_OK 0
#define FM_EINVAL (-1)      /* bad argument or size */
#define FM_ESTATE (-2)      /* call out of order (e.g. no mesh) */
#define FM_EHIP (-3)        /* HIP runtime error */
#define FM_ENOCONV (-4)     /* solver did not converge */
#define FM_EDEVICE (-5)     /* a device-side bound (spin limit) was hit */

/* precision of the linear algebra: FM_FP64 = plain FP64 drop-in path;
   FM_MIXED = FP32 assembly/SpMV/ILU(0)/PCG under an FP64 Newton residual;
   FM_MIXED_PC = FP64 tangent, SpMV and PCG vectors with the FP64 ILU(0)
   factor rounded to FP32 and applied in FP32 as preconditioner (for tangents
   whose condition number makes FP32 directions inaccurate: slender bodies) */
#define FM_FP64 0
#define FM_MIXED 1
#define FM_MIXED_PC 2

typedef struct fm_ctx fm_ctx;

/* ---- context ---------------------------------------------------------------- */
int fm_create(int device, fm_ctx** out);
int fm_destroy(fm_ctx* ctx);
const char* fm_last_error(const fm_ctx* ctx);
const char* fm_version(void);

/* ---- synthetic inputs (host side, no context) ---------------------------------
 * Structured box of nx*ny*nz hexes, 6 Kuhn tets each, node id
 * k(nx+1)(ny+1) + j(nx+1) + i.  Replaces generate_unit_cube_mesh
 * (FEM/adaptive_fem_engine.hpp:2219-2294), whose split is non-conforming.
 * X: 3*(nx+1)(ny+1)(nz+1) doubles SoA; conn: 4*6*nx*ny*nz uint32 SoA. */
int fm_mesh_box(int nx, int ny, int nz, double Lx, double Ly, double Lz, double* X, uint32_t* conn);
/* Moves interior nodes by U(-amp, +amp) per coordinate (counter-based RNG keyed by seed). */
int fm_mesh_jitter(uint32_t n_nodes, double* X, const uint8_t* is_interior, double amp, uint64_t seed);
/* out[e] = median * exp(sigma * z_e), z_e standard normal (counter-based Box-Muller keyed by seed). */
int fm_field_lognormal(uint32_t n, double median, double sigma, uint64_t seed, double* out);

/* ---- mesh, material, boundary conditions ---------------------------------------
 * Replace MeshTopology / ProblemDescriptor (FEM/fem_types.hpp:168-256); the
 * material is per element (mu_e, lam_e) instead of one (E, nu) per problem
 * (ProblemDescriptor.E/.nu, used at FEM/adaptive_fem_engine.hpp:1034, 1187). */
int fm_set_mesh(fm_ctx* ctx, uint32_t n_nodes, const double* X, uint32_t n_elems, const uint32_t* conn);
/* New node coordinates for the mesh already set (3*n_nodes doubles SoA, the same
   node count and connectivity): the pattern, colouring, schedules, material,
   boundary conditions, load and u are kept; the assembled tangent and its
   factor are dropped and the predictor history restarts.  The adapt loop
   moves nodes in place this way (laplacian_smooth after edge_split,
   FEM/adaptive_fem_engine.hpp:2662-2663). */
int fm_set_coords(fm_ctx* ctx, uint32_t n_nodes, const double* X);
int fm_set_material(fm_ctx* ctx, const double* mu_e, const double* lam_e);
/* Dirichlet DOFs and their values at load factor 1 (BoundaryConditions::dirichlet_bcs,
   FEM/fem_types.hpp:258-264); vals may be NULL (all zero). */
int fm_set_bcs(fm_ctx* ctx, uint32_t n_dir, const uint32_t* dofs, const double* vals);
/* External nodal force at load factor 1 (3*nn doubles) or NULL (no force);
   replaces apply_neumann_traction (FEM/adaptive_fem_engine.hpp:733-742). */
int fm_set_load(fm_ctx* ctx, const double* f_ext);

/* ---- pattern and colouring ----------------------------------------------------
 * build_csr_sparsity_pattern_rigorous (FEM/adaptive_fem_engine.hpp:1069-1120)
 * with the intended column rule; outputs may be NULL.  row_ptr: 3nn+1. */
int fm_build_pattern(fm_ctx* ctx, uint32_t* row_ptr_out, uint32_t* col_idx_out, uint64_t* nnz_out);
/* MeshColoring::compute (FEM/adaptive_fem_engine.hpp:1549-1602); colors_out may be NULL. */
int fm_color(fm_ctx* ctx, uint32_t* colors_out, uint32_t* n_colors_out);

/* ---- assembly ------------------------------------------------------------------
 * assemble_global_threadlocal (FEM/adaptive_fem_engine.hpp:1126-1233):
 * f_int(u) and the tangent K on the pattern, kept on the device (fm_use_tangent)
 * and copied to K_vals_out when it is not NULL.  prec selects the K precision
 * on the device (K_vals_out is FP64 either way).  No boundary conditions are
 * applied.  u may be NULL (zero); f_int_out may be NULL. */
int fm_assemble(fm_ctx* ctx, const double* u, double* f_int_out, double* K_vals_out, int prec);
/* assemble_global_residual_only (:990-1059): f_int(u) alone (3nn doubles), FP64,
   no tangent, no boundary conditions; the context's state is not changed. */
int fm_assemble_residual_only(fm_ctx* ctx, const double* u, double* f_int_out);
/* assemble_global_residual_only (:990-1059) and the Newton residual
   R = load*f_ext - f_int(u) with constrained rows zeroed. R_out may be NULL;
   norm_out (||R||_2) may be NULL. */
int fm_residual(fm_ctx* ctx, const double* u, double load, double* R_out, double* norm_out);

/* No counterpart in the reference.  fm_residual without any allocation, for
   callers that cannot allocate (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc): the same kernels and results as
   fm_residual, with the reduction's scratch (block partials of R.R and the
   reduce state, 256 KiB) in a static device array that each GPU holds until
   the process ends.  Limit: n_dofs <= 8,388,608 (FM_EINVAL beyond).  The
   context must be set up (FM_ESTATE otherwise: call fm_residual once first).
   Every call on one GPU shares that GPU's array: calls that overlap on one GPU,
   from any contexts or threads, race. */
int fm_residual_noalloc(fm_ctx* ctx, const double* u, double load, double* R_out, double* norm_out);

/* ---- sparse linear algebra -----------------------------------------------------
 * A general CSR matrix (sorted or unsorted columns) for the solver entries
 * below; replaces SparseCSR (FEM/fem_types.hpp:304-320).  Rows whose groups of
 * three share one column set made of whole 3x3 blocks are solved node-blocked. */
int fm_set_matrix(fm_ctx* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals);
/* use the assembled tangent as the matrix (Dirichlet rows/columns eliminated
   when eliminate != 0, FEM/nonlinear_solver.hpp:388-396) */
int fm_use_tangent(fm_ctx* ctx, int eliminate);
/* SparseCSR::matvec (FEM/fem_types.hpp:310-319) in the given precision */
int fm_spmv(fm_ctx* ctx, int prec, const double* x, double* y);
/* ILU0Preconditioner::factorize (FEM/adaptive_fem_engine.hpp:305-470): L and U at
   A's positions, u_ii at the diagonal positions (LU_out may be NULL) */
int fm_ilu0_factor(fm_ctx* ctx, int prec, double* LU_out);
/* ILU0Preconditioner::solve (:473-504) with the last factorization */
int fm_ilu0_apply(fm_ctx* ctx, int prec, const double* r, double* z);
/* pcg_solve_ilu0 (:1283-1341): factorizes, then PCG from x (in/out) with the
   stopping rule sqrt(r.z / b.b) < tol.  *iters follows the reference return
   value (max_it = not converged, 0 = trivially converged). */
int fm_pcg(fm_ctx* ctx, int prec, const double* b, double* x, double tol, int max_it, int* iters, double* relres);

/* ---- Newton ---------------------------------------------------------------------
 * One load step (1-based) of the Newton loop of AdaptiveFEMEngine::run
 * (FEM/adaptive_fem_engine.hpp:2456-2592) behind the NonlinearNewtonSolver
 * interface (FEM/nonlinear_solver.hpp:23-64, 87-238). */
typedef struct {
    uint32_t n_steps;     /* load steps to reach load factor 1 */
    int prec;             /* FM_FP64, FM_MIXED or FM_MIXED_PC */
    double newton_rtol;   /* ||R_k|| <= rtol ||R_0|| */
    int max_newton;
    double pcg_tol;       /* pcg_solve_ilu0 tol of the FP64 path */
    int pcg_max;
    double inner_tol;     /* FP32 PCG tol of the mixed path (inexact-Newton forcing term; default 3e-3,
                             which suits lift = 1; 1e-2 suits lift = 2 alone, 1e-3 with recycle and
                             inner_tol_max, where it is the smallest forcing term) */
    int ls_max;           /* Armijo backtracks (armijo_line_search_true, :933-985) */
    double armijo_c;
    int lift;             /* 1: predictor solve with the prescribed increment; 2: start from the
                             Lagrange extrapolation of the previous converged steps of the
                             device-resident state (quadratic through three, linear through two;
                             the predictor solve until two are known; any fm_set_* or a host u
                             restarts the history); from a quadratic start, later Newton
                             iterations keep the first tangent and factor while each iteration
                             cuts ||R|| by 100x or more */
    int inner_max;        /* FP32 PCG iteration cap of the mixed path */
    int ls_mode;          /* 0: Armijo on ||R|| only (reference); 1: a trial is also accepted when the
                             potential energy sum V psi - lf f_ext.u decreases by c alpha R.du */
    /* NewtonSolverConfig members (FEM/nonlinear_solver.hpp:23-34) */
    double tol_absolute;  /* also converged when ||R|| < tol_absolute (:25, test at :127); 0 = off */
    double ls_alpha0;     /* first trial step of the line search (line_search_alpha, :27); default 1 */
    double ls_beta;       /* backtracking factor alpha <- beta alpha, 0 < beta < 1 (line_search_beta, :28) */
    int reuse_tangent;    /* 1: every iteration of a step solves with the step's first tangent and
                             factor (reuse_tangent, :30, :141-143) */
    double energy_tol;    /* also converged when |Pi_k - Pi_(k-1)| < energy_tol after an iteration
                             (energy_convergence_tol, :32, :215-227); 0 = off */
    int record_energy;    /* 1: evaluate the potential energy at the start and after every iteration
                             for fm_newton_history (it is evaluated anyway when energy_tol > 0) */
    int inexact;          /* FM_FP64 only: 1 = the FP64 inner solves take the mixed path's forcing
                             term (inner_tol, inner_max, normalised right-hand side) instead of the
                             reference's pcg_tol / pcg_max, so both precisions run one algorithm */
    /* inner solves of FM_MIXED, FM_MIXED_PC and inexact FM_FP64 (no counterpart in the
       reference, whose pcg_solve_ilu0 takes its start x from the caller, :1283-1288);
       ritz applies to FM_MIXED and inexact FM_FP64 only */
    int recycle;          /* lift = 2: the first Newton solve of a load step starts from the
                             Galerkin projection onto the latest 2 recycle (<= 12) Newton
                             corrections of the load path (the A-norm-best start in their span;
                             a step that takes k Newton iterations adds k of them); 0 = x0 = 0 */
    double inner_tol_max; /* > inner_tol: target-aware forcing term, each inner solve to
                             clamp(0.5 newton_rtol R_0 / ||R_k||, inner_tol, inner_tol_max) (a solve
                             that needs only a small reduction stops early); 0 = inner_tol throughout */
    int coarse_degree;    /* 1..4: every inner solve (and the predictor solve) starts from the
                             Galerkin projection onto the polynomial displacement fields of total
                             degree <= coarse_degree in the node coordinates (12, 30, 60 or 105
                             fields, zero on constrained DOFs), together with the recycled
                             corrections in a step's first solve: the smooth modes ILU(0) resolves
                             slowly; 0 = off */
    int coarse_additive;  /* with coarse_degree > 0: those fields Z also enter every PCG iteration
                             as an additive coarse correction, z = M^{-1} r + Z (Z^T K Z)^{-1} Z^T r
                             (a two-level preconditioner with the ILU(0) solve M^{-1}); the PCG
                             stopping rule sqrt(r.z / b.b) then measures r in that
                             preconditioner's norm; 0 = off */
    int lag_precond;      /* 1: the later Newton iterations of a load step re-assemble the tangent
                             but keep the preconditioner of the step's first iteration: its ILU(0)
                             factor and, with coarse_additive, its coarse operator Z^T K Z, whose
                             inverse also gives those solves' Galerkin start (a lagged
                             preconditioner: the Krylov operator is the current tangent); 0 = off */
    int extrap_order;     /* lift = 2: degree of the extrapolation through the previous converged
                             states, 1..4 (3: cubic through four); 0 = the default, quadratic */
    int slope_start;      /* lift = 2: the predictor solve from the unloaded state (the first step of
                             a load path) is solved to 1e-5 and its increment over the load factor
                             kept as the path's slope there; while fewer states are known than
                             extrap_order needs, the extrapolation takes the known states and
                             that slope (Hermite data: quadratic for the second step, cubic for
                             the third); 0 = off */
    int factor_steps;     /* with lag_precond: a load step that starts from an extrapolation of
                             degree >= 2 re-assembles the tangent but keeps the ILU(0) factor of
                             an earlier step while that factor has served fewer than factor_steps
                             steps (a lagged preconditioner across steps; the Galerkin start and
                             the coarse operator use the current tangent); 0 or 1 = a new factor
                             every step */
    int keep_coarse;      /* with factor_steps: such a step also keeps the coarse operator
                             Z^T K Z of the step that formed the factor (its Galerkin start
                             forms only the recycled corrections' rows); 0 = off */
    int ritz;             /* with coarse_additive: each inner solve keeps its first 48 Lanczos
                             vectors, and the eigenvectors of its Lanczos matrix with the ritz
                             (<= 8) smallest eigenvalues combine them into approximations of the
                             preconditioned operator's slowest modes; the next load step's first
                             solve starts from the Galerkin projection onto them too, when it
                             also has recycled corrections (not at a path's first step); 0 = off */
} fm_newton_cfg;

typedef struct {
    int converged;
    int newton_iters;
    int pcg_iters;
    int assemblies;
    int residuals;
    int factorizations;
    double r0;
    double rel_res;
    double seconds;
    int pcg_cap_hits;     /* linear solves that stopped at their iteration cap */
} fm_newton_stats;

void fm_newton_cfg_default(fm_newton_cfg* cfg);
/* u: 3nn doubles in/out (NULL: use and keep the context's device-resident state) */
int fm_newton_step(fm_ctx* ctx, uint32_t step, double* u, const fm_newton_cfg* cfg, fm_newton_stats* stats);
/* the same Newton solve to an arbitrary load factor (LoadStep::load_factor,
   FEM/fem_types.hpp:266-271): prescribed values load_factor*dir_vals, external
   force load_factor*f_ext */
int fm_newton_solve_to(fm_ctx* ctx, double load_factor, double* u, const fm_newton_cfg* cfg, fm_newton_stats* stats);
/* Per-iteration history of the last fm_newton_step / fm_newton_solve_to
 * (NewtonSolverResult::residual_history, step_size_history, energy_history,
 * FEM/nonlinear_solver.hpp:42-44): residual[0] = ||R_0||, then ||R|| after the
 * load-step predictor when one ran (step size 1), then after every Newton
 * iteration; step_size[i] is the line-search step that led to residual[i+1]
 * (len - 1 entries); energy[i] is the potential energy at the state of
 * residual[i] (NaN unless record_energy or energy_tol was set).  At most cap
 * entries are written to each non-NULL array; *len is the full length. */
int fm_newton_history(fm_ctx* ctx, uint32_t cap, double* residual, double* step_size, double* energy,
                      uint32_t* len);
/* device-resident displacement (3nn doubles); fm_set_u(ctx, NULL) returns the
   state to the unloaded configuration (u = 0 at load factor 0, the start of a
   load path, with the predictor history restarted there) */
int fm_get_u(fm_ctx* ctx, double* u);
int fm_set_u(fm_ctx* ctx, const double* u);

/* ---- adjoint ----------------------------------------------------------------------
 * No reference counterpart (new): at the converged final-load state u, solve
 * K lambda = dL/du (FP64 PCG) and return dL/dmu_e, dL/dlam_e and dL/d(load
 * factor of f_ext and of the Dirichlet values).  Outputs may be NULL. */
int fm_adjoint(fm_ctx* ctx, const double* dLdu, double tol, int max_it, double* dL_dmu, double* dL_dlam,
               double* dL_dload, int* iters);

/* ---- rotation part (FP64) -----------------------------------------------------------
 * Replaces sl3_retraction / exponential_map / matrix_logarithm
 * (CORE/lie_operator.cpp:94-312) and transport_stress_frame_objective
 * (FEM/state_transport.hpp:201-249) for the rotation part: R = polar(F),
 * w = log R (angle-axis), Q = exp(alpha [w]x) (Rodrigues).  F, R, Q are n
 * row-major 3x3; w is 3n.  R, w, Q may be NULL. */
int fm_rot_logexp_batch(fm_ctx* ctx, uint32_t n, const double* F, double alpha, double* R, double* w, double* Q);
/* per-element rotation vector of the polar factor of F at the device-resident u
   (3*ne doubles, element order) */
int fm_element_rotations(fm_ctx* ctx, double* w);

/* ---- element fields (FP64) -----------------------------------------------------------
 * Replaces StressExtractor::extract_element_fields (FEM/stress_extraction.hpp:187-240)
 * and the element-state update update_element_states_rigorous
 * (FEM/adaptive_fem_engine.hpp:2716-2794) for the displacement u (3*nn, NULL: the
 * device-resident state).  Per element, in element order: det F (signed),
 * Cauchy stress sigma = P F^T / max(|det F|, 0.01) as six SoA rows
 * (xx, yy, zz, yz, xz, xy; 6*ne), von Mises sqrt(3/2 dev:dev), pressure tr/3, the
 * largest principal stress, the strain-energy density, the maximum stretch
 * sqrt(max(eig_max(F^T F), 1)), and F as nine SoA rows (F_ij, row-major; 9*ne).
 * F is the exact P1 gradient, not projected onto det F = 1.  Degenerate elements
 * (volume < 1e-30) give zeros.  Any output may be NULL. */
typedef struct fm_element_fields_out {
    double* det_F;
    double* cauchy;
    double* von_mises;
    double* pressure;
    double* max_principal;
    double* energy;
    double* max_stretch;
    double* F;
} fm_element_fields_out;
int fm_element_fields(fm_ctx* ctx, const double* u, const fm_element_fields_out* out);

/* ---- ZZ recovered-gradient error estimate (FP64) --------------------------------------
 * Replaces ZZErrorEstimator::estimate_with_bounds / estimate
 * (FEM/error_estimator.hpp:128-232) for the displacement u (3*nn, NULL: the
 * device-resident state): raw element samples g = (F11 - 1, F22 - 1, F33 - 1),
 * their volume-weighted node averages (node_avg, 3*nn), SPR linear fits over
 * each node's patch of at most 32 incident elements evaluated at the node
 * (recovered, 3*nn), element indicators sqrt(sum |recovered - avg|^2 / 4) cbrt(V_e)
 * (elem_errors, ne, element order) and the global estimate sqrt(sum V_e sum / 4).
 * A patch whose fit is undetermined (fewer than four non-coplanar centroids)
 * takes the patch mean.  Any output may be NULL. */
int fm_zz_estimate(fm_ctx* ctx, const double* u, double* elem_errors, double* node_avg, double* recovered,
                   double* global_estimate);

/* ---- checkpoint ----------------------------------------------------------------------
 * Versioned, checksummed load-step checkpoint, a superset of CheckpointFile
 * (FEM/adaptive_fem_engine.hpp:1347-1452): "FMCKPT02", u32 version 2, u32
 * flags, u64 length L; then L bytes exactly as CheckpointFile::save writes them
 * (u32 load_step, adapt_iter, n_nodes, n_elements, n_dofs; per node x, y, z, u,
 * v, w; per element 4 x u32 nodes, volume, jacobian_det, quality_metric; per
 * element state F_data[9] (F_ij at i*3+j), the Mandel stress of
 * neo_hookean_response (M = C S, M11, M22, M33, sqrt2 M12, sqrt2 M13, sqrt2 M23)
 * at J = |det F|, eps_plastic[6] = 0, kappa = 0, gamma_p = 0; u); then mu_e
 * and lam_e per element; then the FNV-1a hash of all preceding bytes.  F is
 * the exact P1 gradient (not projected onto det F = 1).
 * fm_checkpoint_save writes the device-resident state.  fm_checkpoint_load
 * checks the hash and, on a context with a mesh, that the file's node
 * coordinates, connectivity and material are the context's, bit for bit
 * (FM_EINVAL otherwise), then restores u; on an empty context it sets the
 * mesh, the material and u from the file. */
int fm_checkpoint_save(fm_ctx* ctx, const char* path, uint32_t load_step);
int fm_checkpoint_load(fm_ctx* ctx, const char* path, uint32_t* load_step);

/* ---- measurement ------------------------------------------------------------------------
 * Per-kernel-class HIP-event timing on the context's stream. */
#define FM_K_SPMV 0
#define FM_K_ILU_FWD 1
#define FM_K_ILU_BWD 2
#define FM_K_ILU_FACTOR 3
#define FM_K_ASSEMBLE 4
#define FM_K_RESIDUAL 5
#define FM_K_PCG_OTHER 6
#define FM_K_FIELDS 7
#define FM_K_GALERKIN 8  /* Galerkin start: products with the basis, Gram matrix, combination */
#define FM_K_COARSE 9    /* additive coarse correction: moments Z^T r and the coarse solve */
#define FM_K_COUNT 10
/* on = 0: off; 1: every launch of the timed kernel classes is bracketed by
   HIP events on the context's stream; k > 1: of the classes launched once per
   PCG iteration (SpMV, the two sweeps, the other PCG kernels, the coarse
   correction) one launch in
   every k (launches 0, k, 2k, ... counted from fm_profile_reset), since each
   pair of events puts a bubble of a few microseconds into the stream */
int fm_profile_enable(fm_ctx* ctx, int on);
int fm_profile_reset(fm_ctx* ctx);
/* launches[k]: launches of kernel class k while profiling was on; ms[k]: their
   total milliseconds (the mean of the timed launches times launches[k]) */
int fm_profile_get(fm_ctx* ctx, double* ms, uint64_t* launches);
/* No counterpart in the reference.  The moments Z^T r of r (3 nn doubles) over
   the polynomial displacement fields of total degree <= degree (1..4) of the
   Galerkin start and the additive coarse level, computed by the PCG's own
   moment kernel in the precision prec (FM_FP64 or FM_MIXED): out[c nmono + k]
   = sum over nodes v of mono_k(xs_v) r[3v + c], zero on constrained DOFs,
   with xs the node coordinates mapped to [-1, 1] over the bounding box and
   rounded to FP32, and the monomials mono_k by degree (1; x, y, z; x^2, y^2,
   z^2, xy, yz, xz; x^3, y^3, z^3, x^2 y, x^2 z, x y^2, y^2 z, x z^2, y z^2,
   xyz; x^4, y^4, z^4, x^3 y, x^3 z, x y^3, y^3 z, x z^3, y z^3, x^2 y^2,
   y^2 z^2, x^2 z^2, x^2 y z, x y^2 z, x y z^2); nmono = 4, 10, 20, 35 */
int fm_coarse_moments(fm_ctx* ctx, int prec, int degree, const double* r, double* out);
/* No counterpart in the reference.  alpha_j and beta_j of every iteration of
   the last PCG solve of the FP32 or FP64 Krylov path (the first min(cap, len)
   of len; beta of the last iteration 0): the Lanczos coefficients from which
   the preconditioned operator's Ritz values follow (T_jj = 1/alpha_j +
   beta_(j-1)/alpha_(j-1), T_j,j+1 = sqrt(beta_j)/alpha_j).  At most 4096
   iterations are recorded; len is 0 after an FM_MIXED_PC solve (its flexible
   beta has no Lanczos relation). */
int fm_pcg_history(fm_ctx* ctx, uint32_t cap, double* alpha, double* beta, uint32_t* len);
/* sizes for roofline bookkeeping: n (rows), nnz, n_elems, n_nodes, n_colors */
int fm_info(fm_ctx* ctx, uint64_t* out5);
/* kernels chosen for the current matrix: ILU(0) sweep kind (0 level-scheduled,
   1 chain bundles), forward levels, backward levels, chains, bundles, forward
   wave-steps, lanes per node, bundle extent gj, bundle extent gk, SpMV kind
   (0 CSR, 1 node-blocked) */
int fm_sweep_info(fm_ctx* ctx, uint64_t* out10);

#ifdef __cplusplus
}
#endif
#endif
