/*
 * mhx.h — C ABI of the MI355X-native assembly + linear-solve engine.
 *
 * Plain C types only (no torch, no HIP types): a Trilinos-side integration
 * (a custom AssemblyManager / LinearAlgebraInterface) or a ctypes/cgo stub can
 * bind it directly.  Every entry point returns MHX_OK (0) or a negative
 * status; the message is available from mhx_last_error().  Nothing throws
 * across this boundary.
 *
 * Conventions kept from MrHyDE (reference paths relative to the MrHyDE tree):
 *   - residual returned is -F(u)           src/managers/assembly/assemblyManager_scatter.hpp:227
 *   - Jacobian is +dF/du                   src/managers/assembly/assemblyManager_scatter.hpp:261
 *   - strong Dirichlet rows: residual 0,   src/managers/assembly/assemblyManager_scatter.hpp:208,253
 *     Jacobian row = e_d (diag 1); the     src/managers/assembly/assemblyManager_constraints.hpp:125-141
 *     Dirichlet columns are NOT zeroed
 *   - u carries the Dirichlet values       src/managers/solver/solverManager_util.hpp:60-146
 *   - J du = res, u <- u + du              src/managers/solver/solverManager_solvers.hpp:523,602
 */
#ifndef MHX_H
#define MHX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MHX_OK                 0
#define MHX_ERR_ARG           -1   /* bad argument / inconsistent description     */
#define MHX_ERR_DEVICE        -2   /* HIP runtime error                            */
#define MHX_ERR_UNSUPPORTED   -3   /* element/physics combination not built        */
#define MHX_ERR_NOT_CONVERGED -4   /* Krylov/Newton hit its iteration cap          */
#define MHX_ERR_NONFINITE     -5   /* NaN/Inf detected in a residual norm          */
#define MHX_ERR_NOMEM         -6   /* device allocation failed / over budget       */
#define MHX_ERR_STATE         -7   /* call order (e.g. solve before assembly)      */

/* ---- physics modules (reference registry src/physics/physicsImporter.cpp) */
#define MHX_PHYS_THERMAL        1  /* src/physics/thermal.cpp                      */
#define MHX_PHYS_NAVIER_STOKES  2  /* src/physics/navierstokes.cpp                 */

/* ---- box sides, used for Dirichlet side sets ("left", "right", ...) ---- */
#define MHX_SIDE_XMIN  1u   /* "left"   */
#define MHX_SIDE_XMAX  2u   /* "right"  */
#define MHX_SIDE_YMIN  4u   /* "bottom" */
#define MHX_SIDE_YMAX  8u   /* "top"    */
#define MHX_SIDE_ZMIN 16u   /* "back"   */
#define MHX_SIDE_ZMAX 32u   /* "front"  */
#define MHX_SIDE_ALL  63u   /* "all boundaries" */

#define MHX_MAX_MODES 8
#define MHX_MAX_NQ_MODES 4

/*
 * Problem description: the subset of a MrHyDE input deck this path needs.
 * Mesh: structured box, NX x NY (x NZ) quads/hexes (Mesh: NX/NY/NZ, xmin..).
 * Discretization: "order" (HGRAD_{QUAD,HEX}_Cn, equispaced nodes) and the
 * cubature "quadrature" degree (Gauss, n = degree/2 + 1 points per axis),
 * src/interfaces/discretization/discretizationInterface_basis.hpp:219-350,
 * discretizationInterface_integration.hpp:91-106.
 *
 * Thermal coefficients (src/physics/thermal.cpp:48-65 function names):
 *   "thermal diffusion" k(T,x) = k0_e * (1 + diff_u2 * T*T)
 *   "thermal source"    f(x)   = src_e + sum_m amp_m * prod_d s(freq_md * x_d)
 *                               with s(w x) = sin(w x) for w != 0 and 1 for w == 0.
 * Arrays are host pointers, copied at create time; NULL means k0 = 1, src = 0.
 */
typedef struct mhx_problem_desc {
  int32_t dim;            /* 2 or 3 */
  int32_t order;          /* Lagrange order p >= 1 */
  int32_t quadrature;     /* cubature degree (reference key "quadrature") */
  int32_t nx, ny, nz;     /* elements per axis (nz ignored in 2D) */
  double  xmin, xmax, ymin, ymax, zmin, zmax;
  double  perturb;        /* interior-vertex jitter, fraction of h (0 = affine mesh) */
  uint64_t perturb_seed;
  int32_t physics;        /* MHX_PHYS_* */
  int32_t workset_size;   /* reference "workset size" (CPU restatement only) */

  /* thermal */
  const double* k0;       /* [n_elem] or NULL */
  double  diff_u2;
  const double* src_elem; /* [n_elem] or NULL */
  int32_t n_modes;
  double  mode_amp[MHX_MAX_MODES];
  double  mode_freq[MHX_MAX_MODES][3];
  /* "density" and "specific heat" (thermal.cpp:83-84), constants; 0 means
   * the reference default 1.0.  They scale the time derivative only. */
  double  th_density, th_specific_heat;

  /* Neumann conditions (thermal boundaryResidual, "Neumann",
   * src/physics/thermal.cpp:210-222): on the sides in neumann_sides the
   * outward flux g = q(x) . n enters as res += int_side g phi, with
   * q_c(x) = sum_m nq_amp[m][c] * prod_d sin(nq_freq[m][d] x_d + nq_phase[m][d]).
   * Navier-Stokes (navierstokes boundaryResidual, traction, "Neumann u<c>",
   * src/physics/navierstokes.cpp:853-1000): q is the traction vector and
   * velocity component c gets res_c += int_side q_c phi on those sides
   * (where component c is not Dirichlet); a phase of pi/2 with frequency 0
   * gives a constant. */
  uint32_t neumann_sides;
  int32_t  n_nq_modes;    /* <= MHX_MAX_NQ_MODES */
  double  nq_amp[MHX_MAX_NQ_MODES][3];
  double  nq_freq[MHX_MAX_NQ_MODES][3];
  double  nq_phase[MHX_MAX_NQ_MODES][3];

  /* weak Dirichlet sides (thermal boundaryResidual "weak Dirichlet",
   * src/physics/thermal.cpp:226-281, Nitsche): on the sides in wdbc_sides
   *   res += int_side (10/h) k (T - g) phi - k (grad T . n) phi
   *                   - s k (T - g) (grad phi . n),
   * h = (side measure)^(1/(dim-1)) per element side, k the thermal
   * diffusion at T, s the physics' "form_param" (wdbc_form 0: s = 1, the
   * reference default; 1: s = -1; 2: s = 0), and the data
   * g(x) = wd_const + wd_lin . x + sum_m wd_amp[m] prod_d sin(wd_freq[m][d] x_d + wd_phase[m][d]). */
  uint32_t wdbc_sides;
  int32_t  wdbc_form;
  double  wd_const;
  double  wd_lin[3];
  int32_t n_wd_modes;     /* <= MHX_MAX_NQ_MODES */
  int32_t wd_reserved;
  double  wd_amp[MHX_MAX_NQ_MODES];
  double  wd_freq[MHX_MAX_NQ_MODES][3];
  double  wd_phase[MHX_MAX_NQ_MODES][3];

  /* strong Dirichlet data (reference: "Dirichlet conditions", use strong DBCs) */
  uint32_t dbc_sides;     /* MHX_SIDE_* mask */
  double  dbc_const;      /* value when dbc_values == NULL */
  const double* dbc_values; /* [n_dof] nodal values (only Dirichlet entries read) or NULL */

  /* Navier-Stokes (src/physics/navierstokes.cpp), steady, Q1-Q1: unknowns
   * per node (ux, uy, [uz], pr), DOF = node * (dim+1) + field.  Velocity
   * component c is Dirichlet on dbc_sides | ns_lid_sides | ns_comp_sides[c]
   * (the reference's per-variable "Dirichlet conditions" lists): ns_lid_u[c]
   * on the lid sides (edges shared with other sides included), 0 elsewhere. */
  double  ns_rho;         /* "density"   (> 0) */
  double  ns_mu;          /* "viscosity" (> 0) */
  int32_t ns_supg, ns_pspg;  /* "useSUPG", "usePSPG" */
  uint32_t ns_lid_sides;
  double  ns_lid_u[3];
  double  ns_force[3];    /* "source ux", "source uy", "source uz" (constants) */
  int32_t ns_pin_pressure;/* 1: pr = 0 at the corner node (xmin, ymin[, zmin]) */
  uint32_t ns_comp_sides[3];

  /* Coefficient functions: the deck's "Functions" block, each entry
   * registered by name as FunctionManager::addFunction(name, expression,
   * "ip") does (src/managers/function/functionManager_create.hpp:14-39,
   * src/interfaces/physics/physicsInterface_functions.hpp:94-150).  The
   * thermal module reads "thermal diffusion", "thermal source",
   * "specific heat", "density", "advection x|y|z" (src/physics/
   * thermal.cpp:55-61; defaults 1, 0, 1, 1, 0); a given one replaces the
   * closed form above (k0, diff_u2, src_elem, modes, th_density,
   * th_specific_heat).  Other names are functions the expressions may use.
   * Expressions follow the reference interpreter (operators + - * / ^ < <=
   * > >=, sin cos exp log tan abs max min mean sqrt sinh cosh, numbers, pi)
   * over "T", "grad(T)[x|y|z]", "T_t", "x", "y", "z", "t" (fn_time) and the
   * element data "k0" and "src" where given.  Compiled at mhx_ctx_create;
   * MHX_ERR_ARG with a message for what cannot be compiled. */
  int32_t n_functions;
  int32_t fn_reserved;
  const char* const* function_names;   /* [n_functions] */
  const char* const* function_exprs;   /* [n_functions] */
  double  fn_time;                     /* value of "t" (set the stage time with mhx_ctx_update) */
} mhx_problem_desc;

typedef struct mhx_ctx mhx_ctx;

/* ---- assembly flags ---------------------------------------------------- */
#define MHX_RES         1   /* compute residual  (assembleRes)               */
#define MHX_JAC         2   /* compute Jacobian  (assembleJacRes)            */
#define MHX_ACCUMULATE  4   /* add into res/jac (reference semantics: caller zeroes) */
#define MHX_DEVICE_PTRS 8   /* pointer arguments are device pointers         */

/* ---- Krylov options (reference Solver keys) ---------------------------- */
#define MHX_KRYLOV_CG     1   /* "Pseudo Block CG" / SolverManager::PCG        */
#define MHX_KRYLOV_GMRES  2   /* "Block GMRES" (default)                       */
#define MHX_PREC_NONE     0
#define MHX_PREC_JACOBI   1   /* Ifpack2 RELAXATION "Jacobi"                   */
#define MHX_PREC_BLOCK_JACOBI 2  /* nodal block Jacobi (multi-field)           */
#define MHX_PREC_AMG      3   /* smoothed-aggregation AMG, Chebyshev(2) V-cycle:
                                 the reference's default MueLu preconditioner,
                                 linearAlgebraInterface_solvers.hpp:201-253   */
#define MHX_PREC_FP32     0x100  /* inner Krylov + preconditioner in FP32, FP64 refinement */

typedef struct mhx_krylov_opts {
  int32_t solver;        /* MHX_KRYLOV_* */
  int32_t precond;       /* MHX_PREC_* (| MHX_PREC_FP32) */
  int32_t max_iters;     /* "max linear iters" */
  int32_t restart;       /* "Num Blocks" (GMRES) */
  double  tol;           /* "linear TOL": ||b - A x||_2 / ||b||_2 <= tol */
  int32_t flags;         /* MHX_DEVICE_PTRS */
  int32_t inner_iters;   /* FP32 inner iterations per refinement sweep (0 = auto: 10 restart cycles, >= 500) */
  double  inner_tol;     /* FP32 inner relative tolerance (0 = auto: 1e-4, 1e-3 with AMG) */
} mhx_krylov_opts;

typedef struct mhx_newton_opts {
  int32_t max_iters;       /* "max nonlinear iters" (default 10) */
  double  nl_tol;          /* "nonlinear TOL" (relative, default 1e-6) */
  int32_t backtracking;    /* "allow backtracking" (default true) */
  int32_t flags;           /* MHX_DEVICE_PTRS */
  mhx_krylov_opts linear;
  /* Convergence tests (solverManager_solvers.hpp:430-445, options of
   * solverManager_construct.hpp:59-63): the relative test (scaled inf-norm
   * < nl_tol, or inf-norm < 1e-100) unless relative_off; with relative_off
   * and use_abs_tol, inf-norm < abs_tol; with neither, max_iters steps. */
  int32_t relative_off;    /* "use relative TOL" = false (default: the relative test) */
  int32_t use_abs_tol;     /* "use absolute TOL" (default false) */
  double  abs_tol;         /* "absolute nonlinear TOL" (0 -> min(nl_tol, 1e-6)) */
} mhx_newton_opts;

typedef struct mhx_solve_stats {
  int32_t iters;           /* Krylov iterations (sum over Newton steps) */
  int32_t nl_iters;        /* Newton iterations */
  double  rel_res;         /* final FP64 true relative residual ||b-Ax||/||b|| */
  double  res_norm0;       /* first nonlinear residual inf-norm */
  double  res_norm;        /* last nonlinear residual inf-norm */
  int32_t refinements;     /* FP64 refinement sweeps (mixed precision) */
  int32_t status;
  double  t_assemble_ms, t_solve_ms;
} mhx_solve_stats;

/* ---- context lifecycle -------------------------------------------------- */
/* Builds mesh, DOF numbering, colouring, CRS graph and device storage.
 * Replaces: MeshInterface/DiscretizationInterface::buildDOFManagers
 * (src/interfaces/discretization/discretizationInterface_dof.hpp:14-241),
 * AssemblyManager::createGroups (assemblyManager_groups.hpp:14-476),
 * LinearAlgebraInterface::setupLinearAlgebra (linearAlgebraInterface_construct.hpp:140-265). */
int  mhx_ctx_create(const mhx_problem_desc* desc, mhx_ctx** out);
void mhx_ctx_destroy(mhx_ctx* ctx);
const char* mhx_last_error(const mhx_ctx* ctx);  /* ctx may be NULL */

/* n_dof, nnz, n_elem, number of element colours */
int  mhx_sizes(const mhx_ctx* ctx, int64_t* n_dof, int64_t* nnz, int64_t* n_elem, int32_t* n_colors);

/* Host copies of the CRS pattern (rowptr[n_dof+1], colidx[nnz], sorted columns).
 * Replaces: Tpetra graph from insertGlobalIndices + fillComplete
 * (src/interfaces/linear_algebra/linearAlgebraInterface_construct.hpp:212-265). */
int  mhx_crs_pattern(const mhx_ctx* ctx, int32_t* rowptr, int32_t* colidx);

/* mask[n_dof]: 1 for strong-Dirichlet DOFs (AssemblyManager::isFixedDOF,
 * src/managers/assembly/assemblyManager_constraints.hpp:14-95). */
int  mhx_dirichlet_mask(const mhx_ctx* ctx, int32_t* mask);

/* u[dbc] = g.  Replaces SolverManager::setDirichlet
 * (src/managers/solver/solverManager_util.hpp:60-146). */
int  mhx_set_dirichlet(mhx_ctx* ctx, double* u, int32_t flags);

/* Physical coordinates of every DOF node, xyz[n_dof*3] (host). */
int  mhx_dof_coords(const mhx_ctx* ctx, double* xyz);

/* Residual and/or Jacobian assembly over all elements.
 *   u[n_dof] in, res[n_dof] = -F(u) out, jac_vals[nnz] in CRS order out.
 * Replaces AssemblyManager::assembleJacRes / assembleRes
 * (src/managers/assembly/assemblyManager.hpp:237-255, :330-345;
 *  assemblyManager_jacres.hpp:121-630, :638-877) including gather, workset
 * seeding, physics volumeResidual, scatter and dofConstraints. */
int  mhx_assemble(mhx_ctx* ctx, const double* u, int32_t flags, double* res, double* jac_vals);

/* y = J x with J in the context's CRS pattern (jac_vals == NULL: the
 * context's Jacobian).  Replaces Tpetra::CrsMatrix::apply as used by
 * SolverManager::PCG (src/managers/solver/solverManager_solvers.hpp:845,877). */
int  mhx_spmv(mhx_ctx* ctx, const double* jac_vals, const double* x, double* y, int32_t flags);

/* Solve J x = r.  jac_vals == NULL uses the Jacobian held by the context
 * (the last internal assembly).  x is used as the initial guess.
 * Replaces LinearAlgebraInterface::linearSolver
 * (src/interfaces/linear_algebra/linearAlgebraInterface_solvers.hpp:14-152)
 * and SolverManager::PCG (src/managers/solver/solverManager_solvers.hpp:812-905). */
int  mhx_linear_solve(mhx_ctx* ctx, const double* jac_vals, const double* r, double* x,
                      const mhx_krylov_opts* opts, mhx_solve_stats* stats);

/* Allocation-free variant of mhx_linear_solve.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It runs the same kernels in the same order as
 * mhx_linear_solve with solver MHX_KRYLOV_CG and precond MHX_PREC_NONE or
 * MHX_PREC_JACOBI (FP64), so x, the iteration count and rel_res are the
 * same bits; the Krylov vectors, the Jacobi scaling and the reduction
 * partials live in one file-scope device array per GPU of
 * 5 * MHX_NOALLOC_MAX_DOF + 1280 doubles (40 MiB, held by every GPU the
 * library runs on until the process ends) instead of context buffers.
 * It calls no hipMalloc/hipFree and records no phase timers.
 *   - n_dof <= MHX_NOALLOC_MAX_DOF, else MHX_ERR_ARG;
 *   - opts->flags must carry MHX_DEVICE_PTRS (r, x and jac_vals are device
 *     pointers), else MHX_ERR_ARG; jac_vals == NULL: the context's Jacobian;
 *   - any other solver or preconditioner: MHX_ERR_UNSUPPORTED.
 * Every call on a GPU shares that GPU's array and leaves its contents
 * behind: calls that overlap on one GPU (any contexts, any threads) race. */
#define MHX_NOALLOC_MAX_DOF 1048576
int  mhx_linear_solve_noalloc(mhx_ctx* ctx, const double* jac_vals, const double* r, double* x,
                              const mhx_krylov_opts* opts, mhx_solve_stats* stats);

/* MatrixMarket files.  Replaces LinearAlgebraInterface::writeToFile
 * (src/interfaces/linear_algebra/linearAlgebraInterface_util.hpp:316-331,
 * Tpetra::MatrixMarket::Writer::writeSparseFile / writeDenseFile): J as
 * "coordinate real general" with 1-based DOF indices in the engine's
 * numbering, vectors as "array real general" (n x 1), values with 17
 * significant digits (a read returns the same bits).  jac_vals == NULL: the
 * context's Jacobian.  The readers place a file's entries into the context's
 * CRS pattern (coordinate general or symmetric; duplicates summed); an entry
 * outside the pattern is MHX_ERR_ARG.  MHX_DEVICE_PTRS in flags: the value
 * pointers are device pointers. */
int  mhx_write_matrix_market(mhx_ctx* ctx, const char* path, const double* jac_vals, int32_t flags);
int  mhx_write_vector_market(mhx_ctx* ctx, const char* path, const double* vec, int32_t flags);
int  mhx_read_matrix_market(mhx_ctx* ctx, const char* path, double* jac_vals, int32_t flags);
int  mhx_read_vector_market(mhx_ctx* ctx, const char* path, double* vec, int32_t flags);

/* Build the smoothed-aggregation hierarchy of J (jac_vals == NULL: the
 * context's Jacobian; MHX_DEVICE_PTRS in flags: device pointer) and keep it
 * in the context; a solve with MHX_PREC_AMG rebuilds it from its own J.
 * Returns the number of levels.  Replaces
 * LinearAlgebraInterface::buildAMGPreconditioner
 * (src/interfaces/linear_algebra/linearAlgebraInterface_solvers.hpp:201-253). */
int  mhx_amg_setup(mhx_ctx* ctx, const double* jac_vals, int32_t flags);

/* One operator of the hierarchy held by the context, copied to host arrays
 * (any of rp/ci/vals may be NULL to query sizes only): which = 0 the level's
 * operator A_l (CRS), 1 its prolongator P_l, 2 its restriction R_l = P_l^T,
 * 3 the dense inverse of the coarsest operator (row-major in vals).
 * info[0..3] = rows, cols, nnz, number of levels; *lambda (if not NULL) =
 * the level's power-iteration estimate of rho(D^-1 A_l).
 * Introspection of MueLu::Hierarchy levels; no single reference call. */
int  mhx_amg_level(mhx_ctx* ctx, int32_t level, int32_t which, int64_t info[4], double* lambda, int32_t* rp,
                   int32_t* ci, double* vals);

/* Time-integration state used by every following assembly and Newton solve
 * (thermal): the residual gains rho cp dT/dt phi with, per node,
 *   dT/dt = (BDF_0 u + sum_{k=1..order} BDF_k u_prev[k-1]) / dt,
 * the backward-difference weights of `bdf_order` 1..6
 * (SolverManager::setBackwardDifference, solverManager_setup.hpp:443-500)
 * and the single-stage BWE tableau (Workset::computeSolnTransientSeeded,
 * src/tools/workset.cpp:601-680); the Jacobian gains BDF_0/dt rho cp times
 * the mass matrix.  u_prev: bdf_order vectors of n_dof values, most recent
 * first (MHX_DEVICE_PTRS in flags: device pointer).  The source is evaluated
 * from the descriptor's current amplitudes (update them per step with
 * mhx_ctx_update).  bdf_order = 0 returns to steady assembly. */
int  mhx_set_time_state(mhx_ctx* ctx, double dt, int32_t bdf_order, const double* u_prev, int32_t flags);

/* One stage of a diagonally implicit Runge-Kutta step (SolverManager::
 * transientSolver's stage loop, src/managers/solver/solverManager_solvers.hpp
 * :125-151, with Workset::computeSolnTransientSeeded, src/tools/workset.cpp
 * :601-680).  The following assemblies and Newton solves take the unknown U
 * as stage `stage` of the tableau (A row-major n_stages x n_stages, b):
 *   T      = alpha_u U + (1 - alpha_u) u_prev[0]
 *            + sum_{j<stage} A_sj / b_j (u_stages[j] - u_prev[0]),
 *   dT/dt  = (BDF_0 U + sum_{k>=1} BDF_k u_prev[k-1]) / (dt b_stage),
 * alpha_u = A_ss / b_s.  After the last stage the step's solution is
 * u_prev[0] + sum_s (U_s - u_prev[0]) (solverManager_solvers.hpp:145-150);
 * coefficients at the stage time t_n + c_s dt come from mhx_ctx_update.
 * u_stages: `stage` vectors (MHX_DEVICE_PTRS: device pointers).  Weak
 * Dirichlet sides are refused for multi-stage tableaus.  mhx_set_time_state
 * is this call with the one-stage tableau A = b = 1. */
int  mhx_set_stage_state(mhx_ctx* ctx, double dt, int32_t bdf_order, const double* u_prev, int32_t n_stages,
                         const double* butcher_A, const double* butcher_b, int32_t stage, const double* u_stages,
                         int32_t flags);

/* Newton on u in place (u must already carry the Dirichlet values, see
 * mhx_set_dirichlet).  Replaces SolverManager::nonlinearSolver
 * (src/managers/solver/solverManager_solvers.hpp:290-643); status 0 on
 * convergence or when the iteration cap is reached, as the reference does. */
int  mhx_nonlinear_solve(mhx_ctx* ctx, double* u, const mhx_newton_opts* opts, mhx_solve_stats* stats);

/* L2 errors against a manufactured solution T = sum_m amp_m prod_d s(freq_md x_d)
 * at the volume cubature points: returns sqrt(sum w (T_h - T)^2) and the
 * gradient norm.  Replaces PostprocessManager::computeError L2/GRAD
 * (src/managers/postprocess/postprocessManager_error_estimation.hpp:173-330). */
int  mhx_l2_error(mhx_ctx* ctx, const double* u, int32_t n_modes, const double* amp,
                  const double* freq /* [n_modes*3] */, int32_t flags,
                  double* l2, double* l2_grad);

/* ---- tuning options -----------------------------------------------------
 * Named integer options of a context (the role of the reference's
 * "workset size" / "assembly partitioning" keys, scripts/input-files/
 * input-defaults.yaml:124).  Known names:
 *   "assembly.colouring"  0 = element colouring (one launch per element
 *                         colour), 1 = block colouring (one launch per
 *                         block colour, element colours as in-workgroup
 *                         sub-phases), 2 = x-segments (a workgroup merges a
 *                         run of x-adjacent elements in LDS and writes each
 *                         row's share as contiguous runs; Navier-Stokes:
 *                         segments of 8 (3-D) / 16 (2-D) elements; block
 *                         colouring runs element colouring there),
 *                         3 = pencil streaming (a wavefront walks an
 *                         x-segment of an element pencil two elements per
 *                         step and writes each row's x-run once; 3-D Q2
 *                         thermal with 3x3x3 Gauss on an unperturbed box
 *                         mesh only, MHX_ERR_ARG elsewhere).  Default: 3
 *                         where it applies, 2 for Navier-Stokes, else 0.
 *   "assembly.segment"    pencil streaming: elements per x-segment, even,
 *                         2 .. 4096 (default 12).
 *   "krylov.spmv_lanes"   lanes that share one CRS row in the sparse
 *                         matrix-vector product: 4, 8, 16, 32 or 64 (default
 *                         chosen from the mean row length at context build).
 *   "krylov.lattice_columns"  1 (default): sparse products on the
 *                         context's lattice pattern compute each entry's
 *                         column from the pattern's closed form instead of
 *                         reading the column index (same sums); 0: read
 *                         the column indices.
 *   "amg.smoother"        smoother of the AMG V-cycle: -1 = automatic
 *                         (default: 0 for one unknown per node, 1 for
 *                         coupled fields), 0 = degree-2
 *                         Chebyshev on the point diagonal, 1 =
 *                         damped Jacobi sweeps on the nodal n_field x
 *                         n_field diagonal blocks, 2 = one multicolour
 *                         nodal block Gauss-Seidel sweep each side; 1 and
 *                         2 are for coupled multi-field systems (the reference's MueLu
 *                         "smoother: type" key, CHEBYSHEV or RELAXATION,
 *                         src/interfaces/linear_algebra/
 *                         linearAlgebraInterface_solvers.hpp:225).
 *   "amg.reuse"           0 (default, the reference's MueLu "reuse: none"):
 *                         every solve builds its hierarchy; 1: later solves
 *                         on the context keep the coarse levels of the last
 *                         full set-up and refresh only the fine level's
 *                         smoother data from the new operator (diagonal or
 *                         nodal block inverses; "amg.sweeps" and
 *                         "amg.omega_pct" take their current values).  The
 *                         Chebyshev eigenvalue bound of every level stays the
 *                         one estimated at that full set-up; a change of
 *                         precision, smoother or unknowns per node forces a
 *                         full set-up.
 *   "amg.sweeps"          block-Jacobi smoother sweeps before and after
 *                         the coarse correction, 1 .. 16 (default 3).
 *   "amg.omega_pct"       block-Jacobi damping in percent, 1 .. 200
 *                         (default 40).
 * Returns MHX_ERR_ARG for an unknown name or value. */
int  mhx_set_option(mhx_ctx* ctx, const char* name, int64_t value);

/* FunctionManager::evaluate(name, "point") on the host
 * (src/managers/function/functionManager_evaluate.hpp:14-50): compiles the
 * descriptor's function `name` as mhx_ctx_create does and evaluates it at n
 * points.  pts[11*i ..]: T, T_t, grad(T)[x], [y], [z], x, y, z, t, k0, src.
 * out[6*i ..]: value, then its derivatives by T, grad(T)[x], [y], [z] and
 * T_t.  Needs no GPU.  MHX_ERR_ARG with mhx_expr_last_error() for an
 * expression that does not compile. */
int  mhx_expr_eval(const mhx_problem_desc* desc, const char* name, int64_t n, const double* pts, double* out);
const char* mhx_expr_last_error(void);
int  mhx_get_option(const mhx_ctx* ctx, const char* name, int64_t* value);

/* ---- device-resident state --------------------------------------------
 * Context-owned device vectors, so a caller can keep u, res and J in HBM
 * between calls and pass them back with MHX_DEVICE_PTRS (the role of the
 * reference's overlapped Tpetra vectors/matrix created once and reused,
 * src/managers/solver/solverManager_solvers.hpp:322-333). */
#define MHX_BUF_U    0   /* [n_dof] */
#define MHX_BUF_RES  1   /* [n_dof] */
#define MHX_BUF_JAC  2   /* [nnz]   */
#define MHX_BUF_DU   3   /* [n_dof] */
#define MHX_BUF_AUX  4   /* [n_dof] */
int  mhx_state_ptr(mhx_ctx* ctx, int32_t which, double** dev_ptr, int64_t* count);
int  mhx_state_upload(mhx_ctx* ctx, int32_t which, const double* host, int64_t count);
int  mhx_state_download(mhx_ctx* ctx, int32_t which, double* host, int64_t count);

/* ---- profiling: per-phase HIP-event timers (reference Teuchos timers) -- */
int  mhx_profile_enable(mhx_ctx* ctx, int32_t on);
/* Returns phase count; fills up to cap names (each <= 47 chars + NUL in a
 * 48-byte slot), accumulated milliseconds and launch counts. */
int  mhx_profile_read(mhx_ctx* ctx, int32_t cap, char* names48, double* ms, int64_t* counts);
int  mhx_profile_reset(mhx_ctx* ctx);
/* Drain the context's HIP stream. */
int  mhx_synchronize(mhx_ctx* ctx);

/* Run the context's work on the caller's stream (a hipStream_t passed as
 * void*; NULL = the legacy default stream), so device buffers produced on
 * that stream are ordered before the engine reads them.  The context's own
 * stream is drained first. */
int  mhx_set_stream(mhx_ctx* ctx, void* hip_stream);

/* ---- seeded inputs ------------------------------------------------------
 * Counter-based generator (splitmix64 of seed ^ index) so every consumer
 * produces identical fields.  With z ~ N(0,1) from Box-Muller of draws 2i
 * and 2i+1: kind 0 uniform a + (b-a) u on [a,b); kind 1 lognormal
 * exp(a + b z) (log-mean a, log-sd b); kind 2 normal a + b z. */
#define MHX_RNG_UNIFORM   0
#define MHX_RNG_LOGNORMAL 1
#define MHX_RNG_NORMAL    2
int  mhx_rng_fill(uint64_t seed, int64_t n, int32_t kind, double a, double b, double* out);

/* ---- batched multiscale subgrid solves ----------------------------------
 * Many independent thermal subproblems, each a box (the unit cube unless
 * extent is given) cut into n^3 Q1 hexahedra (2 <= n <= 16, 2x2x2 Gauss),
 * element conductivities k, optional element sources, strong Dirichlet data
 * on every boundary node; each is assembled and
 * solved with Jacobi-PCG to ||b - A x|| / ||b|| <= tol inside ONE persistent
 * launch (optionally replayed from a captured hipGraph).  FP32_REFINED runs
 * the inner PCG in FP32 inside FP64 iterative refinement (FP64 residuals).
 * Replaces the sequential per-macro-group subgrid loop:
 * SubGridDtN_Solver::nonlinearSolver (src/subgrid/subgridDtN_solver.cpp:903-1043)
 * called from MultiscaleManager::evaluateMacroMicroMacroMap
 * (src/managers/multiscale/multiscaleManager.cpp:866-1188). */
#define MHX_BATCH_FP64          0
#define MHX_BATCH_FP32_REFINED  1
typedef struct mhx_batch mhx_batch;
typedef struct mhx_batch_desc {
  int32_t n_inst;
  const int32_t* n_axis;   /* [n_inst] elements per axis */
  const double* k;         /* concatenated n^3 element conductivities (host) */
  const double* g;         /* concatenated (n+1)^3 nodal arrays; boundary entries = Dirichlet data (host) */
  double  tol;             /* default 1e-10 */
  int32_t max_iters;       /* per instance, default 10000 */
  int32_t precision;       /* MHX_BATCH_* */
  double  inner_tol;       /* FP32 inner relative tolerance, default 1e-5 */
  int32_t max_refine;      /* default 20 */
  int32_t wg_size;         /* 512 or 1024; 0 = autotune on this batch */
  int32_t use_graph;       /* 1 = capture the launch once and replay it */
  const double* extent;    /* [n_inst][3] box side lengths (Lx, Ly, Lz); NULL = unit cubes (host) */
  const double* src;       /* concatenated n^3 element sources f_e; NULL = none (host) */
  int32_t reserved[4];
} mhx_batch_desc;
int  mhx_batch_create(const mhx_batch_desc* desc, mhx_batch** out);
int  mhx_batch_solve(mhx_batch* b, int32_t flags);
/* Host copies: x (concatenated nodal solutions), iterations, FP64 relative
 * residual and refinement sweeps per instance (any pointer may be NULL). */
int  mhx_batch_results(mhx_batch* b, double* x, int32_t* iters, double* relres, int32_t* refinements);
int  mhx_batch_info(const mhx_batch* b, int64_t* total_nodes, int32_t* wg_size, double* last_ms);
int  mhx_batch_set_wg(mhx_batch* b, int32_t wg_size);
void mhx_batch_destroy(mhx_batch* b);

/* ---- sizes and coefficient updates -------------------------------------
 * n_dof, nnz and the device bytes a context needs for assembly and one
 * Krylov solve with the given options (opts may be NULL: CG), without
 * building it.  No reference counterpart (admission control of the server). */
int  mhx_footprint(const mhx_problem_desc* desc, const mhx_krylov_opts* opts, int64_t* n_dof, int64_t* nnz,
                   int64_t* device_bytes);
/* New coefficients on an existing context: element k0 / source, source modes,
 * Dirichlet values, Navier-Stokes parameters.  Mesh, discretisation, physics
 * and Dirichlet sides must equal the context's (MHX_ERR_ARG otherwise); the
 * mesh, CRS graph and solver buffers stay resident.  The role of re-running
 * a sample with new parameters in the reference's UQ loop
 * (AnalysisManager::forwardModel, src/managers/analysis/analysisManager_solve.hpp). */
int  mhx_ctx_update(mhx_ctx* ctx, const mhx_problem_desc* desc);

/* ---- serving: continuous batching of independent solves -----------------
 * Each request is one steady solve (mhx_nonlinear_solve with the server's
 * Newton options, from u = Dirichlet data, 0 elsewhere).  Up to max_active
 * requests run at once, each on its own host worker and HIP stream; a queued
 * request is admitted (first fit) when a worker is free and its footprint
 * (mhx_footprint) fits hbm_budget next to the contexts held; an idle context
 * with the same mesh is reused (max_cached idle contexts are kept; while
 * requests are queued the cache may hold more, within hbm_budget, since a
 * context's destruction synchronises the device).  Results
 * do not depend on the mix: reductions run in a fixed order per request.
 * No counterpart in the reference (its UQ loop runs samples one by one,
 * src/managers/analysis/analysisManager_solve.hpp). */
typedef struct mhx_server mhx_server;
typedef struct mhx_server_desc {
  int64_t hbm_budget;       /* bytes of device memory for contexts; 0 = half the free memory */
  int32_t max_active;       /* concurrent requests (default 8) */
  int32_t max_cached;       /* idle contexts kept for reuse (0 = none) */
  mhx_newton_opts newton;   /* per request; zeros -> 2 Newton steps, nl_tol 1e-8, CG + linear.precond to 1e-10 */
  int32_t reserved[4];
} mhx_server_desc;
int  mhx_server_create(const mhx_server_desc* desc, mhx_server** out);
/* Copies the description and its arrays; never blocks on the device.
 * MHX_ERR_NOMEM when the request alone exceeds hbm_budget. */
int  mhx_server_submit(mhx_server* s, const mhx_problem_desc* desc, int64_t* ticket);
/* 0: still queued/running (wait = 0); 1: done, x[n_dof] (host, may be NULL),
 * stats and the latency from submit to completion filled in and the ticket
 * released; < 0: the request failed (message: mhx_server_last_error). */
int  mhx_server_poll(mhx_server* s, int64_t ticket, int32_t wait, double* x, mhx_solve_stats* stats,
                     double* latency_ms);
/* done, active, queued, bytes held, contexts reused, contexts created, idle contexts, budget */
int  mhx_server_info(mhx_server* s, int64_t* out8);
const char* mhx_server_last_error(mhx_server* s);
void mhx_server_destroy(mhx_server* s);

/* Plain device copy y = x of n_doubles FP64 values, repeated reps times on a
 * private stream; returns the mean milliseconds per copy.  A calibration
 * kernel of known traffic (8 B read + 8 B written per value) for HBM
 * counters and the achievable-bandwidth figure; not part of the reference. */
int  mhx_copy_probe(int64_t n_doubles, int32_t reps, double* ms_per_copy);

/* The dense B^T D B form of the Q2 hexahedron's element contraction (27 x
 * 108 x 27 per element) on the FP64 matrix cores, v_mfma_f64_16x16x4f64,
 * over n_elem synthetic elements: mean milliseconds per pass over them, and
 * the largest relative error of the computed matrices (the first 64 in
 * full, every element's entry sum) against the host.  variant 0 reads the
 * basis table from LDS for every instruction, variant 1 holds it in
 * registers.  The measured alternative to the sum-factorised column stage
 * of the Q2 assembly kernels; no counterpart in the reference. */
int  mhx_mfma_probe(int64_t n_elem, int32_t reps, int32_t variant, double* ms_per_pass, double* max_rel_err);

/* Library version / build string. */
const char* mhx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MHX_H */
