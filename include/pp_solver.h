/*
 * pp_solver.h — C ABI of the MI355X pressure-projection engine.
 *
 * The engine replaces the body of the reference's in-process pressure backends
 *   Solver::solve_pressure_correction_icpcg          (src/core/coupled/pressure.cpp:302-321)
 *   Solver::solve_pressure_correction_ildlt_pcg      (pressure.cpp:323-342)
 *   Solver::solve_pressure_correction_liu_split_icpcg (pressure.cpp:344-365)
 *   Solver::solve_pressure_correction_liu_split_ildlt_pcg (pressure.cpp:470-491)
 * reached from the scheme-string switch Solver::solve_pressure_correction (pressure.cpp:1031-1089).
 * One call assembles the 5-point variable-coefficient operator and the right-hand side from the
 * Solver's fields, factors it (ICC(0) or ILDLT(0), natural ordering), runs projected PCG from x0 = 0
 * and returns the interior of pressure_correction_ after the mean subtraction the reference wrappers
 * apply; ghost filling (apply_pressure_bc, operators.cpp:143-181) stays with the caller.
 *
 * Conventions
 *   - Arrays are double, i-major with j contiguous, like Field2D (src/core/base/field.hpp:23-33).
 *     Host entry points take a leading dimension (elements between consecutive i) so a caller can
 *     pass &field(0,0) of a ghosted Field2D with ld = ny + 2*ghost without copying.
 *   - Return value 0 = success. Non-zero = error; pp_report.message carries the reference's error
 *     text ("ICC preconditioner construction failed", "pcg breakdown: p^T A p is invalid", ...),
 *     which the C++ host turns into std::runtime_error like the reference.
 *   - Not converging within max_iterations is not an error (ch_sparse_krylov.hpp:659).
 *   - One context per host thread; calls on one context are serialised by the caller.
 */
#ifndef PP_SOLVER_H
#define PP_SOLVER_H

#ifdef __cplusplus
extern "C" {
#endif

/* Scheme codes; names of the new scheme strings in parentheses. */
enum pp_scheme {
  PP_ICPCG = 1,           /* hip_icpcg        <- icpcg            */
  PP_ILDLT_PCG = 2,       /* hip_ildlt_pcg    <- ildlt_pcg        */
  PP_SPLIT_ICPCG = 3,     /* hip_split_icpcg  <- split_icpcg, liu_split_icpcg, paper_split_icpcg */
  PP_SPLIT_ILDLT_PCG = 4  /* hip_split_ildlt_pcg <- split_ildlt_pcg, ...                         */
};

/* Arithmetic of the inner Krylov loop. */
enum pp_precision {
  PP_FP64 = 0,  /* the reference's arithmetic, operation for operation                          */
  PP_MIXED = 1  /* FP32 SpMV/sweeps/updates with FP64-accumulated dots inside FP64 refinement    */
};

/* Boundary types (config.hpp:23-27). Index order of bc arrays: left, right, bottom, top. */
enum pp_bc_type { PP_BC_UNSET = 0, PP_BC_DIRICHLET = 1, PP_BC_NEUMANN = 2 };

typedef struct pp_problem {
  int nx, ny;             /* cells; nx 1..512, ny 1..65535 (1..1024 if periodic in x), >= 4     */
                          /* along a periodic direction                                         */
  double dx, dy;          /* the Solver's dx_, dy_ (lx / nx, ly / ny)                            */
  double dt;              /* cfg_.dt                                                              */
  int periodic_x, periodic_y;
  int bc_type[4];         /* pressure_bc_left/right/bottom/top .type                             */
  double bc_value[4];     /* ... .value                                                           */
  int scheme;             /* enum pp_scheme                                                       */
  int precision;          /* enum pp_precision                                                    */
  int max_iterations;     /* cfg_.poisson_iterations (values < 1 act as 1)                        */
  double tolerance;       /* cfg_.pressure_tolerance (relative or absolute residual)              */
  double density_ratio;   /* split schemes: rho_ref = min(1, density_ratio) (operators.cpp:411)   */
  int step_index;         /* split schemes: 0 -> p_ext = p^n, else 2 p^n - p^{n-1} (operators.cpp:413-422) */
  int inner_max_iterations; /* PP_MIXED: cap per inner FP32 solve (<= 0: engine default)         */
  double inner_tolerance;   /* PP_MIXED: inner relative reduction (<= 0: engine default)         */
  int ring_depth;           /* sweep wave roles: 0 = the context's tuning table (pp_set_tuning)
                               or the built-in default; a power of two >= 8 with
                               ceil(nx/64) * ring_depth <= 64: loader waves stage that many
                               levels per sweep wave; -1: direct sweeps, no loader waves (grids
                               wider than 384 always). Does not change results.                  */
} pp_problem;

typedef struct pp_report {
  int iterations;            /* last_pressure_iterations_ (mixed: total inner iterations)         */
  int outer_cycles;          /* PP_MIXED: refinement cycles; PP_FP64: 1                           */
  double relative_residual;  /* the backend's return value                                        */
  double absolute_residual;
  double iterate_residual;   /* LinearSolveReport::iterate_residual                              */
  double kernel_ms;          /* device time of the solve kernel(s), HIP events on the launch stream */
  double phase_ms[10];       /* device time per phase (setup, SpMV, x/r update, |r|, forward sweep,
                                backward sweep, mean(z), r.z, p update, output)                   */
  int status;                /* 0 ok; see pp_status                                               */
  char message[256];
} pp_report;

enum pp_status {
  PP_OK = 0,
  PP_ERR_FACTOR = 1,        /* "ICC preconditioner construction failed" / "ILDLT ..."            */
  PP_ERR_PAP = 2,           /* "pcg breakdown: p^T A p is invalid"                               */
  PP_ERR_RZ0 = 3,           /* "pcg preconditioner breakdown: non-positive r^T M^{-1} r"         */
  PP_ERR_RZ = 4,            /* "pcg preconditioner breakdown: non-positive new r^T M^{-1} r"     */
  PP_ERR_NODIAG = 5,        /* "sparse matrix assembly failed to produce a diagonal entry"       */
  PP_ERR_SWEEP = 6,         /* internal: wavefront hand-off did not complete                     */
  PP_ERR_BICG_RHO = 7,      /* "bicgstab breakdown: invalid rho"                                 */
  PP_ERR_BICG_RHAT_V = 8,   /* "bicgstab breakdown: invalid r_hat^T A M^{-1} p"                  */
  PP_ERR_BICG_TT = 9,       /* "bicgstab breakdown: invalid t^T t"                               */
  PP_ERR_BICG_OMEGA = 10,   /* "bicgstab breakdown: invalid omega"                               */
  PP_ERR_ARG = 100,         /* invalid argument (sizes, scheme, pointers)                       */
  PP_ERR_HIP = 101          /* HIP runtime failure                                               */
};

typedef struct pp_ctx pp_ctx;

/* Context bound to one device; owns device buffers, layout tables and a stream. */
int pp_create(pp_ctx **ctx, int device);
void pp_destroy(pp_ctx *ctx);
/* Text of the last error on this context ("" if none). */
const char *pp_last_error(const pp_ctx *ctx);
/* Device time in ms of the kernels the last call on this context ran (HIP events on that call's
 * stream): the solve kernel(s), the whole batch, or the velocity correction kernel. The solve
 * entries also report it in pp_report.kernel_ms; the reference has no counterpart (its backends
 * run on the host). */
double pp_last_kernel_ms(const pp_ctx *ctx);

/*
 * One pressure correction from host fields (replaces pressure.cpp:302-365, 470-491 between the
 * zero-fill of pressure_correction_ and apply_pressure_bc). Synchronous.
 *   rho    nx x ny      rho_mid_   (ld_rho >= ny)
 *   u      (nx+1) x ny  u_star_    (ld_u >= ny)
 *   v      nx x (ny+1)  v_star_    (ld_v >= ny+1)
 *   p_n, p_nm1  nx x ny pressure_, pressure_previous_step_ (split schemes; NULL = zero)
 *   p_out  nx x ny      interior of pressure_correction_ (ld_out >= ny)
 */
int pp_solve(pp_ctx *ctx, const pp_problem *pb, const double *rho, int ld_rho, const double *u, int ld_u,
             const double *v, int ld_v, const double *p_n, int ld_pn, const double *p_nm1, int ld_pnm1,
             double *p_out, int ld_out, pp_report *report);

/*
 * Same solve on device-resident, contiguous arrays (rho nx*ny, u (nx+1)*ny, v nx*(ny+1), p_n/p_nm1
 * nx*ny or NULL, p_out nx*ny). Work is queued on `stream` (a hipStream_t, NULL = the context's
 * stream); the call waits for completion to fill `report`.
 */
int pp_solve_device(pp_ctx *ctx, const pp_problem *pb, const double *d_rho, const double *d_u,
                    const double *d_v, const double *d_p_n, const double *d_p_nm1, double *d_p_out,
                    pp_report *report, void *stream);

/*
 * The reference backend's whole contract on device-resident Field2D buffers (pressure.cpp:302-365,
 * 470-491 from the zero-fill of pressure_correction_ through apply_pressure_bc,
 * operators.cpp:143-181). Every field is a whole ghosted buffer with `ghost` layers, element (i, j)
 * of an a x b field at (i + ghost) * (b + 2 ghost) + j + ghost: rho_mid (nx x ny), u_star
 * ((nx+1) x ny), v_star (nx x (ny+1)), pressure / pressure_prev (nx x ny, read by split schemes,
 * may be NULL otherwise). d_p_corr (nx x ny) receives pressure_correction_: zero-filled, the
 * interior solved, ghosts filled as apply_pressure_bc fills them (all zeros after a failure). The
 * output feeds pp_velocity_correction_device directly. Synchronous on `stream`.
 */
int pp_solve_device_field(pp_ctx *ctx, const pp_problem *pb, int ghost, const double *d_rho_mid,
                          const double *d_u_star, const double *d_v_star, const double *d_pressure,
                          const double *d_pressure_prev, double *d_p_corr, pp_report *report, void *stream);

/*
 * Per-iteration PCG log: the scalar fields of one call of the reference's CGMonitor
 * (CGMonitorData, ch_sparse_krylov.hpp:47-58; the calls at 547-555, 569-579, 612-621 and
 * 636-648), the vectors left out. Record k belongs to iteration k: record 0 is the initial
 * residual (preconditioned_residual = sqrt(r.z) once z exists), record k >= 1 follows PCG
 * iteration k (beta and preconditioned_residual stay 0 on the converged last iteration, as in the
 * reference). Fields the reference leaves at their defaults are 0.
 */
typedef struct pp_monitor_record {
  int iteration;
  int pad;
  double alpha;
  double beta;
  double absolute_residual;
  double relative_residual;
  double iterate_residual;
  double preconditioned_residual;
} pp_monitor_record;

/*
 * pp_solve / pp_solve_device with the PCG log: the kernel writes one record per monitor call into
 * `records` (host memory for pp_solve_monitored, device memory for pp_solve_device_monitored), up
 * to `capacity` records; *count receives the number of monitor calls the solve made (report
 * iterations + 1 after a normal end), which may exceed capacity. Replaces passing a CGMonitor to
 * solve_preconditioned_cg (the online analysis hook, pressure.cpp:1078-1087; the frozen replay,
 * pressure_analysis.cpp:255-294) for the scalar part. FP64 solves only (PP_ERR_ARG for PP_MIXED).
 */
int pp_solve_monitored(pp_ctx *ctx, const pp_problem *pb, const double *rho, int ld_rho, const double *u, int ld_u,
                       const double *v, int ld_v, const double *p_n, int ld_pn, const double *p_nm1, int ld_pnm1,
                       double *p_out, int ld_out, pp_monitor_record *records, int capacity, int *count,
                       pp_report *report);
int pp_solve_device_monitored(pp_ctx *ctx, const pp_problem *pb, const double *d_rho, const double *d_u,
                              const double *d_v, const double *d_p_n, const double *d_p_nm1, double *d_p_out,
                              pp_monitor_record *d_records, int capacity, int *count, pp_report *report,
                              void *stream);

/*
 * Coefficient-level solve (no counterpart in the reference; the differentiable form of the
 * non-split operator of pressure.cpp:123-196). Inputs are face inverse densities and a source:
 *   k_u  (nx+1) x ny   k_f = 1/rho_f on x-faces (face i lies between cells i-1 and i)
 *   k_v  nx x (ny+1)   on y-faces
 *   src  nx x ny       right-hand side before boundary terms (the reference's -div/dt)
 * The operator has c_f = k_f / h^2 on every face; wall faces contribute the reference's boundary
 * terms (Dirichlet 2c to the diagonal and 2c*value to b, Neumann c*value*h to b). Along a periodic
 * direction face nx (ny) is face 0 and its entry is not read. The rest is the FP64 path: regularize,
 * factor (scheme PP_ICPCG or PP_ILDLT_PCG), PCG from 0, mean subtraction. Arrays contiguous.
 */
int pp_solve_coeffs(pp_ctx *ctx, const pp_problem *pb, const double *k_u, const double *k_v, const double *src,
                    double *p_out, pp_report *report);
int pp_solve_coeffs_device(pp_ctx *ctx, const pp_problem *pb, const double *d_k_u, const double *d_k_v,
                           const double *d_src, double *d_p_out, pp_report *report, void *stream);

/*
 * Adjoint of pp_solve_coeffs (no counterpart in the reference). Given the solution p and dL/dp,
 * solves A lam = P dL/dp with the same PCG (P the zero-mean projection when no edge is Dirichlet)
 * and returns
 *   dL/dsrc = lam                                                        (nx x ny)
 *   dL/dk_f = -(lam_a - lam_b)(p_a - p_b) / h^2 on interior and periodic faces,
 *             -2 lam_c (p_c - value) / h^2 on Dirichlet walls, lam_c value h / h^2 on Neumann walls
 *   into dL/dk_u ((nx+1) x ny) and dL/dk_v (nx x (ny+1)); the duplicate periodic face gets 0.
 * report describes the adjoint solve.
 */
int pp_adjoint(pp_ctx *ctx, const pp_problem *pb, const double *k_u, const double *k_v, const double *p,
               const double *dLdp, double *dLdsrc, double *dLdk_u, double *dLdk_v, pp_report *report);
int pp_adjoint_device(pp_ctx *ctx, const pp_problem *pb, const double *d_k_u, const double *d_k_v,
                      const double *d_p, const double *d_dLdp, double *d_dLdsrc, double *d_dLdk_u,
                      double *d_dLdk_v, pp_report *report, void *stream);

/*
 * Batched solve (no counterpart in the reference, whose solver handles one system per call):
 * `count` independent pressure corrections of any sizes, schemes and precisions (problems[k].
 * precision: PP_FP64 or PP_MIXED), one launch per kernel family (FP64 ICC, FP64 ILDLT, mixed ICC,
 * mixed ILDLT) with one workgroup per system, dispatched largest first so that CUs freed by small
 * systems pick up the remaining ones. Each system's result is the one pp_solve_device gives for
 * it. Device arrays are contiguous as for pp_solve_device; the
 * pointer arrays are host arrays of `count` device pointers (d_p_n / d_p_nm1 may be NULL, or hold
 * NULL entries). reports[k] gets system k's report; kernel_ms is the duration of the whole batch.
 * Returns PP_OK, or the status of
 * the first system (in index order) that failed.
 */
int pp_solve_batched_device(pp_ctx *ctx, int count, const pp_problem *problems, const double *const *d_rho,
                            const double *const *d_u, const double *const *d_v, const double *const *d_p_n,
                            const double *const *d_p_nm1, double *const *d_p_out, pp_report *reports,
                            void *stream);

/*
 * Velocity boundary conditions of the correction step: the reference Config's u_bc_* / v_bc_*
 * (config.hpp:116-123) per edge in left, right, bottom, top order, and the wall speeds an unset u
 * bottom / top edge takes (bottom_wall_velocity_x, top_wall_velocity_x, config.hpp:105-106).
 * Unset edges behave as in the reference (operators.cpp:62-112): no-slip Dirichlet 0, u bottom/top
 * at the wall speed.
 */
typedef struct pp_velocity_bc {
  int u_type[4];
  double u_value[4];
  int v_type[4];
  double v_value[4];
  double bottom_wall_velocity_x;
  double top_wall_velocity_x;
} pp_velocity_bc;

/*
 * Velocity correction and pressure-history shift after a solve: replaces
 * Solver::apply_pressure_velocity_correction (operators.cpp:480-534), the step solver.cpp:272 runs
 * after solve_pressure_correction. All fields are whole Field2D buffers with `ghost` layers
 * (element (i, j) of an a x b field at (i + ghost) * (b + 2 ghost) + j + ghost):
 *   read:       rho_mid, p_corr (nx x ny; p_corr = pressure_correction_ with its ghosts filled, as
 *               pp_solve's caller leaves it), u_star ((nx+1) x ny), v_star (nx x (ny+1));
 *   read/write: u ((nx+1) x ny), v (nx x (ny+1)), pressure (pressure_), pressure_prev
 *               (pressure_previous_step_): on entry the current fields (the reference reads old
 *               ghost cells of u, v and pressure, and the split extrapolation reads both
 *               pressures), on return u^{n+1}, v^{n+1}, the new pressure and the previous one.
 * pb supplies the grid, dt, periodicity, pressure BCs, scheme (split schemes add the explicit
 * term with rho_ref = min(1, density_ratio)) and step_index. speeds[0] / speeds[1] receive
 * last_boundary_speed_pre_correction_ / _post_correction_ (operators.cpp:463-478). Results are
 * the reference's bit for bit.
 */
int pp_velocity_correction_device(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost,
                                  const double *d_rho_mid, const double *d_u_star, const double *d_v_star,
                                  const double *d_p_corr, double *d_u, double *d_v, double *d_pressure,
                                  double *d_pressure_prev, double *speeds, void *stream);
int pp_velocity_correction(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost,
                           const double *rho_mid, const double *u_star, const double *v_star, const double *p_corr,
                           double *u, double *v, double *pressure, double *pressure_prev, double *speeds);

/*
 * The pressure part of one timestep from host Field2D buffers, the two calls solver.cpp:271-272
 * makes (solve_pressure_correction, then apply_pressure_velocity_correction) with one upload and
 * one download: the fields go to the device once, pp_solve_device_field and the device correction
 * run back to back, and p_corr (pressure_correction_ with ghosts), u, v, pressure and
 * pressure_prev come back. Same buffers and meanings as pp_velocity_correction, with p_corr an
 * output. On a solve failure the status and report carry the reference's error, p_corr is all
 * zeros and u, v, pressure, pressure_prev are untouched (the reference throws before correcting).
 */
int pp_pressure_step(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost, const double *rho_mid,
                     const double *u_star, const double *v_star, double *p_corr, double *u, double *v,
                     double *pressure, double *pressure_prev, double *speeds, pp_report *report);

/*
 * The momentum predictor's linear solve (momentum.cpp:469-472): build_diagonal_preconditioner
 * (ch_sparse_krylov.hpp:421-431) and solve_preconditioned_bicgstab (ch_sparse_krylov.hpp:839-963)
 * on an n x n CSR matrix in the form finalize_row_maps produces (ch_sparse_krylov.hpp:105-133):
 * row_ptr[n + 1], col_idx / values[row_ptr[n]] with ascending columns and one diagonal entry per
 * row. x is the initial guess on entry (the reference passes zeros) and the solution on return.
 * use_relative_tolerance as the reference's argument (momentum.cpp passes false). The report's
 * iterations and absolute / relative / iterate residuals are LinearSolveReport's; kernel_ms is the
 * device time of the whole solve, host round trips between its passes included. Errors carry the
 * reference's texts: PP_ERR_ARG "cannot solve an empty linear system" (n <= 0), PP_ERR_NODIAG,
 * PP_ERR_BICG_*; x then holds the iterate at the breakdown, as the reference's vector does.
 * Dot products are summed in 512-row chunks (differing from the reference's sequential sums by
 * rounding only); everything else is the reference's arithmetic term for term. pp_bicgstab checks
 * row_ptr and the column indices (PP_ERR_ARG); pp_bicgstab_device trusts them (every col_idx in
 * [0, n), row_ptr non-decreasing from 0).
 */
int pp_bicgstab_device(pp_ctx *ctx, int n, const int *d_row_ptr, const int *d_col_idx, const double *d_values,
                       const double *d_rhs, double *d_x, int max_iterations, double tolerance,
                       int use_relative_tolerance, pp_report *report, void *stream);
int pp_bicgstab(pp_ctx *ctx, int n, const int *row_ptr, const int *col_idx, const double *values, const double *rhs,
                double *x, int max_iterations, double tolerance, int use_relative_tolerance, pp_report *report);

/*
 * Tuning table of the context (written by the project's autotuner, ding_nsch_core_solver_amd/
 * autotune.py): the sweep wave roles (ring depth, or -1 for direct sweeps) to use for grids of
 * exactly nx[k] x ny[k] when pp_problem.ring_depth is 0. Replaces the previous table; entries with an invalid depth are
 * rejected. count = 0 clears it.
 */
int pp_set_tuning(pp_ctx *ctx, int count, const int *nx, const int *ny, const int *ring_depth);

/*
 * The momentum predictor (Solver::compute_momentum_fluxes + Solver::solve_momentum_predictor,
 * momentum.cpp:55-85, 157-502; called at solver.cpp:266-267): the convection fields H(u), H(v), the
 * implicit full-viscous-tensor u-v system with its right-hand side (momentum.cpp:157-468), assembled
 * on the device into the CSR form finalize_row_maps produces, BiCGSTAB with the diagonal
 * preconditioner from x0 = u^n (as pp_bicgstab_device), and the scatter into u_star_ / v_star_ with
 * apply_u_velocity_bc / apply_v_bc. Fields are whole Field2D buffers with `ghost` layers (ghost
 * cells are read, as the reference reads them): u, u_star, u_rhs_prev, surface_fx_u are
 * (nx+1) x ny faces; v, v_star, v_rhs_prev, surface_fy_v are nx x (ny+1); rho_mid, eta (the
 * Solver's eta_), eta_prev (eta_previous_step_) are nx x ny cells. u_rhs_prev / v_rhs_prev hold
 * momentum_u_rhs_prev_ / momentum_v_rhs_prev_ on entry (read when step_index > 0) and this step's
 * convection fields on return. The report carries the BiCGSTAB's iterations and residuals (the
 * reference's return value is absolute_residual); kernel_ms is the device time of the whole step.
 * Errors carry the reference's texts (BiCGSTAB breakdowns, a dropped diagonal; explicit u/v BCs on
 * a periodic axis, as validate_boundary_configuration).
 */
typedef struct pp_momentum_problem {
  int nx, ny;
  double dx, dy, dt;
  double re;                  /* the viscous weight is alpha = 0.5 dt / re (momentum.cpp:162) */
  double body_force_x, body_force_y;
  int periodic_x, periodic_y;
  int step_index;             /* 0: convection H^n; else 1.5 H^n - 0.5 H^{n-1} */
  int max_iterations;         /* momentum_iterations */
  double tolerance;           /* momentum_tolerance (absolute, as momentum.cpp:470-471 passes it) */
} pp_momentum_problem;
int pp_momentum_predictor_device(pp_ctx *ctx, const pp_momentum_problem *mp, const pp_velocity_bc *vbc, int ghost,
                                 const double *d_u, const double *d_v, const double *d_rho_mid, const double *d_eta,
                                 const double *d_eta_prev, const double *d_surface_fx_u, const double *d_surface_fy_v,
                                 double *d_u_rhs_prev, double *d_v_rhs_prev, double *d_u_star, double *d_v_star,
                                 pp_report *report, void *stream);
int pp_momentum_predictor(pp_ctx *ctx, const pp_momentum_problem *mp, const pp_velocity_bc *vbc, int ghost,
                          const double *u, const double *v, const double *rho_mid, const double *eta,
                          const double *eta_prev, const double *surface_fx_u, const double *surface_fy_v,
                          double *u_rhs_prev, double *v_rhs_prev, double *u_star, double *v_star, pp_report *report);

/*
 * The Cahn-Hilliard semi-implicit linear solve (Solver::solve_phase_linear_system_eq25,
 * phase_field.cpp:437-503, called from solve_cahn_hilliard_semi_implicit at 690): the operator
 * A = alpha0 I - beta1 Lap + beta2 Lap^2 on the grid (build_laplacian_matrix, multiply,
 * build_cahn_hilliard_operator, ch_sparse_krylov.hpp:135-233), its ILDLT(0) preconditioner when it
 * builds, else the diagonal one (build_preconditioner, 252-335, 421-442), and PCG without projection
 * from x0 = c (515-660) to ch_tolerance within ch_inner_iterations. A PCG breakdown with the ILDLT
 * preconditioner switches to the diagonal one and solves again from the iterate the failed attempt
 * left, as the reference's catch does; that choice is kept with the cached operator. The context
 * caches the operator, its level schedule and its factor, and reuses them while the grid and the
 * coefficients stay the same (alpha0 within 1e-14 max(1, |alpha0|), as ensure_ch_linear_system):
 * a run rebuilds them at its first two steps (alpha0 = 1/dt, then 3/(2 dt)). rhs and c are
 * natural-order nx * ny arrays (row i * ny + j, the Field2D interiors); c holds the initial guess on
 * entry and the solution on return. beta1 = a1 / Pe, beta2 = a2 / Pe with a1 = 3 sqrt(2) / (16 Cn),
 * a2 = 3 sqrt(2) Cn / 4 (config.hpp:9-15).
 */
enum pp_ch_preconditioner { PP_CH_DIAGONAL = 0, PP_CH_ILDLT = 1, PP_CH_DIAGONAL_FALLBACK = 2 };
typedef struct pp_ch_problem {
  int nx, ny;
  double dx, dy;
  int periodic_x, periodic_y;
  double alpha0, beta1, beta2;
  int max_iterations;         /* ch_inner_iterations */
  double tolerance;           /* ch_tolerance */
} pp_ch_problem;
typedef struct pp_ch_report {
  int iterations;             /* last_ch_iterations_ */
  int preconditioner;         /* pp_ch_preconditioner: SparsePCG[ILDLT] / [Diagonal] / [DiagonalFallback] */
  double iterate_residual;    /* the reference's return value (report.iterate_residual) */
  double absolute_residual, relative_residual;
  double equation_residual;   /* |rhs - A c| / |rhs| of the returned c (last_ch_equation_residual_) */
  double kernel_ms;           /* device time of the solve (the factorization of a new operator excluded) */
  int levels;                 /* levels of the triangular sweeps */
  int status;
  char message[256];
} pp_ch_report;
int pp_ch_solve_device(pp_ctx *ctx, const pp_ch_problem *cp, const double *d_rhs, double *d_c, pp_ch_report *report,
                       void *stream);
int pp_ch_solve(pp_ctx *ctx, const pp_ch_problem *cp, const double *rhs, double *c, pp_ch_report *report);

/*
 * Allocation-free form of pp_solve_device, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in
 * the reference. Same problem, inputs, output, report and error texts as pp_solve_device, on the
 * current device and the given stream (null: the default stream), with the built-in sweep ring depth
 * and wait budget (no context, so no tuning table or wait limit). Every buffer the kernels need beyond
 * the inputs and the output is a fixed file-scope device array of PP_NOALLOC_SCRATCH_BYTES bytes that
 * each GPU the library runs on holds until the process ends; grids of more than PP_NOALLOC_MAX_CELLS
 * cells are rejected with PP_ERR_ARG. The call makes no hipMalloc, hipFree or other allocation; it
 * synchronizes the stream to read the report (kernel_ms is 0). The first call on a device looks up the
 * array's address there and zeroes it (one fill kernel and a device synchronization). All calls on one GPU share that GPU's
 * array: calls that overlap on one GPU (several host threads or streams at once) race and must be
 * serialized by the caller.
 */
#define PP_NOALLOC_MAX_CELLS 131072
#define PP_NOALLOC_SCRATCH_BYTES 20635392
int pp_solve_device_noalloc(const pp_problem *pb, const double *d_rho, const double *d_u, const double *d_v,
                            const double *d_p_n, const double *d_p_nm1, double *d_p_out, pp_report *report,
                            void *stream);

/*
 * Watchdog of the wavefront sweeps (no counterpart in the reference, whose sweeps are sequential).
 * The waves of a solve hand boundary values and staged operands to each other through LDS counters;
 * each wait is bounded, and a wait that spins past `sleeps` sleeps fails the solve with
 * PP_ERR_SWEEP ("wavefront hand-off did not complete") instead of hanging. Once one wait has given
 * up, every later wait of that launch returns at once and the solve loop leaves at its next
 * iteration, so a failure costs one wait budget. sleeps = 0 restores the built-in budget (2^24
 * sleeps, about a second); a small value makes an ordinary solve fail, which is how the error path
 * is tested. Applies to later calls on the context.
 */
int pp_set_wait_limit(pp_ctx *ctx, int sleeps);

/*
 * Summation order of the pressure mean in the velocity correction (pp_velocity_correction[_device],
 * pp_pressure_step): subtract_mean (operators.cpp:447-460, called at 526-532) sums the correction
 * sequentially, i outer and j inner, one dependent addition per cell. exact = 1 keeps that order
 * and the reference's bits (a single lane runs the chain: about 180 us of the 210 us correction at
 * 192x192). exact = 0, the default, sums in a 256-thread tree (thread t adds cells t, t + 256, ...
 * of the i-outer order, then the 64-lane and 4-wave trees), which moves the mean by rounding only
 * (measured within 1e-12 of the reference's pressure, u and v relative to their maxima) and takes a
 * few microseconds. Applies to later calls on the context.
 */
int pp_set_exact_mean(pp_ctx *ctx, int exact);

#ifdef __cplusplus
}
#endif

#endif /* PP_SOLVER_H */
