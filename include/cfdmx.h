/*
 * cfdmx.h -- C-ABI of the MI355X (gfx950) time-step library for incompressible
 * flow past a cylinder.
 *
 * The reference (jimmy3769tw/CFD_NSsolver_forHPC_flowPassingCylinder) has no
 * library or FFI layer: its hot path is a set of header-defined C++ functions that
 * runSeqOmp (src/RUN_CPU.cpp:9-440) calls on host std::vectors.  Each entry point
 * below replaces one of those functions (cited per declaration) and operates on
 * device-resident state owned by a context; the C++ adapter
 * cfdmx_reference_api.hpp re-exposes them under the reference's own names and
 * signatures.
 *
 * Conventions
 *   - Plain C types only; every function returns an int status (CFDMX_OK or a
 *     negative CFDMX_ERR_*) and never throws.  cfdmx_last_error() gives the text.
 *   - Host arrays use the reference layouts: ghosted fields of G = nx*ny*nz doubles,
 *     nx = nxc + 4 (two ghost layers, structureGrid.hpp:34,86-90), index
 *     (i*ny + j)*nz + k (structureGrid.hpp:153-156); pressure unknowns of
 *     N = nxc*nyc*nzc doubles indexed ((i-2)*nyc + (j-2))*nzc + (k-2)
 *     (structureGrid.hpp:176-181).
 *   - Calls on one context are ordered on that context's HIP stream; a context is
 *     not re-entrant.  Functions returning host values synchronise the stream.
 */
#ifndef CFDMX_H
#define CFDMX_H

/* the library is built with hidden visibility; only these entry points are exported */
#define CFDMX_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

#define CFDMX_OK 0
#define CFDMX_ERR_ARG (-1)    /* invalid argument or configuration */
#define CFDMX_ERR_HIP (-2)    /* HIP runtime failure (device, memory, launch) */
#define CFDMX_ERR_STATE (-3)  /* call not valid in the current context state */

#define CFDMX_GRID_UNIFORM 0 /* grid::init_uniform, structureGrid.hpp:336-367 */
#define CFDMX_GRID_SML 1     /* grid::init_nonuniform_SML, structureGrid.hpp:250-329 */

#define CFDMX_SOLVER_BICGSTAB 0 /* solver::bicgstabRe2, bicgstabRe2.hpp:82-180 */
#define CFDMX_SOLVER_RBSOR 1    /* red-black ordering of SorPipeLine_seq's update */

#define CFDMX_PRECOND_NONE 0   /* SPE 7-slot operator, SPE_sparseMatrix.hpp:42-106 */
#define CFDMX_PRECOND_JACOBI 1 /* row-scaled CSR operator, structureGrid.hpp:557-654 */

#define CFDMX_PRECISION_FP64 0
#define CFDMX_PRECISION_MIXED 1 /* FP32 inner Krylov, FP64 outer refinement */
/* the mixed mode's algorithm (multigrid-preconditioned inner BiCGSTAB, FP64 refinement)
   with the inner solve in FP64 as well: the same-algorithm FP64 comparison */
#define CFDMX_PRECISION_FP64_MG 2

typedef struct cfdmx_config {
    /* grid (class grid, structureGrid.hpp:20-40, made runtime) */
    int nx, ny, nz; /* interior cells: grid::nxCal, nyCal, nzCal */
    double lx, ly, lz;
    int grid_kind;
    /* physics and time (simpulationVariable, 0_1_SimulationVariable.hpp:26-31,95-106) */
    double re, dt;
    /* boundary conditions: the Num/Dir table of FLOWpassing.hpp:6-40, face x component */
    double bc_num[6][3];
    double bc_dir[6][3];
    int z_periodic;          /* faces 4/5 periodic (no reference counterpart) */
    int outflow_convective;  /* face 1 u convective (no reference counterpart) */
    double convective_speed; /* advection speed of the convective outflow */
    int flux_correction;     /* shift the outlet face so its flux equals the inlet's */
    /* pressure (bicgstabRe2.hpp:47-52, 0_1_SimulationVariable.hpp:51-67) */
    int solver;
    int precond;
    double tol;
    int maxit;
    double sor_omega;
    int skip_first_pressure; /* RUN_CPU.cpp:303: no pressure solve on step 1 */
    int precision;
    double inner_tol; /* multigrid modes: relative tolerance of each inner solve;
                         0 (the default) takes the mode's fastest measured value,
                         1e-5 in both modes.  The FP64 refinement stops
                         at tol, or unconverged (cfdmx_step_info.unconverged) after 40
                         passes, at maxit inner iterations in all, on a pass that would
                         raise the residual (dropped: x stays as it was) or on an inner
                         divergence */
    /* launch shaping */
    int xcd_remap; /* map consecutive logical blocks to one XCD (0 = hardware order) */
    int mg_graphs; /* multigrid modes: replay each V-cycle from a captured HIP graph
                      (default 1; 0 launches its kernels one by one, same results) */
    /* Smagorinsky LES (SmagorinskyModel.hpp:38, built with TERBULENCE_SMAGORINSKY): the
     * constant Cs of the eddy viscosity (Cs delta)^2 |S| that each step adds to 1/Re in the
     * diffusion; 0 = off (the reference's default build; its constant is 0.18) */
    double smagorinsky_cs;
    /* CheckSteadyStatebyMaxVal_omp (CheckSteadystate.hpp:247-307, called every step at
     * RUN_CPU.cpp:380-384): 1 = each step also reports max(T0 - T3) per component in
     * cfdmx_step_info.maxdiff, folded into the projection (it then reads T0 as well:
     * 24 more bytes per cell); 0 (the default) = off */
    int steady_check;
    /* multigrid modes (no reference counterpart): 1 (the default) starts each pressure
     * solve (and each pressure solve of the adjoint's reverse pass, from its last two) from
     * 2 x_n - x_(n-1), the linear extrapolation of the last two solutions
     * (x_(n-1) kept in the solver's history vector, which checkpoints store), instead of x_n; the
     * solve still ends at tol.  0 = start from x_n.  The FP64 path always starts from x_n,
     * as bicgstabRe2 does */
    int mg_extrapolate;
} cfdmx_config;

typedef struct cfdmx_ctx cfdmx_ctx;

/* per-step report (the stdout of RUN_CPU.cpp:346-435 and Time_cDcL.dat) */
typedef struct cfdmx_step_info {
    double cd, cl;    /* CD_CL, virtualForceIntergrator.hpp:92-111 */
    double time;      /* simulation time at the call (simpulationVariable::time) */
    int iters;        /* pressure iterations (simu.iters) */
    double err;       /* pressure residual (simu.error) */
    double cfl;       /* get_Cfl, cfl.hpp:3-79 */
    int solved;       /* 1 if a pressure solve ran this step */
    int restarts;     /* cumulative BiCGSTAB rho restarts */
    int resets;       /* cumulative BiCGSTAB NaN resets */
    /* steady_check: max(T0 - T3) of u, v, w over the faces the projection writes,
     * starting at 0 as the reference's loop does (CheckSteadystate.hpp:253-301); the
     * faces it skips are left out, because the reference's T3 holds its initial values
     * there (copyVel copies T3 into T0, never back).  0 when steady_check is off. */
    double maxdiff[3];
    /* cumulative pressure solves that ended with err not below tol: the iteration
     * limit, and in the multigrid modes also the 40-pass limit of the refinement, a
     * refinement pass that would raise the residual (dropped) or an inner solve that
     * diverged; err then holds the residual reached */
    int unconverged;
} cfdmx_step_info;

/* Defaults: FLOWpassing table, reference solver defaults (tol 1e-3, 5000 iterations,
 * omega 1.8), uniform 20 x 10 x 0.5 domain. */
CFDMX_API int cfdmx_config_default(cfdmx_config *cfg);

/* Context lifecycle; replaces the allocations of runSeqOmp (RUN_CPU.cpp:18-28,64,120,
 * 158-163) and generateGride (Generator.hpp:14-46).  Builds the grid metrics. */
CFDMX_API int cfdmx_create(const cfdmx_config *cfg, cfdmx_ctx **out);
CFDMX_API int cfdmx_destroy(cfdmx_ctx *ctx);
CFDMX_API const char *cfdmx_last_error(const cfdmx_ctx *ctx); /* ctx may be NULL */
CFDMX_API int cfdmx_sizes(const cfdmx_ctx *ctx, long long *G, long long *N);

/* Grid metrics as built by the reference generators: X,Y,Z (n+1 each), then
 * Dx,Dy,Dz,Dxs,Dys,Dzs (n each).  structureGrid.hpp:93,336-367,433-486. */
CFDMX_API int cfdmx_grid_metrics(const cfdmx_ctx *ctx, double *out);

/* Field movement.  which: 0 = T0 (u^n), 1 = T1 (u*), 2 = T3 (u^{n+1}).
 * Pressure p is ghosted (pressure::p), x is X_result (MxClass::X_result). */
CFDMX_API int cfdmx_upload_velocity(cfdmx_ctx *ctx, int which, const double *u, const double *v,
                          const double *w);
CFDMX_API int cfdmx_download_velocity(cfdmx_ctx *ctx, int which, double *u, double *v, double *w);
CFDMX_API int cfdmx_upload_pressure(cfdmx_ctx *ctx, const double *p);
CFDMX_API int cfdmx_download_pressure(cfdmx_ctx *ctx, double *p);
CFDMX_API int cfdmx_upload_xresult(cfdmx_ctx *ctx, const double *x);
CFDMX_API int cfdmx_download_xresult(cfdmx_ctx *ctx, double *x);
CFDMX_API int cfdmx_download_rhs(cfdmx_ctx *ctx, double *b);
CFDMX_API int cfdmx_upload_rhs(cfdmx_ctx *ctx, const double *b); /* MxClass::matB, N doubles */
CFDMX_API int cfdmx_upload_eta(cfdmx_ctx *ctx, const double *eta);
CFDMX_API int cfdmx_download_eta(cfdmx_ctx *ctx, double *eta);
CFDMX_API int cfdmx_download_force(cfdmx_ctx *ctx, double *f); /* DfibArray::f, 3*G */
CFDMX_API int cfdmx_upload_force(cfdmx_ctx *ctx, const double *f);
CFDMX_API int cfdmx_set_inflow(cfdmx_ctx *ctx, const double *profile); /* u inflow per ghosted j, ny */

/* Plot3D output: OutputPLOT3D_Qfile (qfile.hpp:6-108) and OutputPlot3D_Xfile
 * (xfileWrite.hpp:6-64), single-block files in the reference's layout.
 * q file: int32 1, nx, ny, nz (interior); float32 0, 0, Re, time; then five float32
 * interior arrays with i fastest: p; u, v, w averaged to cell centres from the faces
 * on either side (the average formed in FP64, then rounded); eta.  The velocity is T0
 * (u^n with its boundary faces set) where the reference reads T3; the two differ only
 * on faces its T3 never refreshes (the outlet face and the ghost faces keep their
 * initial values there).
 * cfdmx_plot3d_q returns once the fields are staged on the device: the copy to host
 * memory runs on a second stream, and the file is written by the next cfdmx_plot3d_q,
 * by cfdmx_plot3d_flush or by cfdmx_destroy, so time steps proceed meanwhile.
 * x file: int32 1, nx, ny, nz; the cell-centre coordinates Xc, Yc, Zc
 * (structureGrid.hpp:216-229) as float32 arrays with i fastest.  A file that cannot be
 * written returns CFDMX_ERR_ARG (for q, from the call that writes it). */
CFDMX_API int cfdmx_plot3d_q(cfdmx_ctx *ctx, const char *path, double time);
CFDMX_API int cfdmx_plot3d_flush(cfdmx_ctx *ctx);
CFDMX_API int cfdmx_plot3d_x(cfdmx_ctx *ctx, const char *path);

/* FP64 checkpoint / resume.  No reference counterpart: it stands where the reference
 * would resume from a float32 q file (QfileRead, qfileRead.hpp:6-100, called only from
 * commented-out code in RUN_CPU.cpp:89).  cfdmx_checkpoint_save writes every array a
 * later step reads (T0, T1, T3, p, eta, the forcing, X_result and its two history
 * copies, the inflow profile) with the step counter, time, dt, Re, the mask parameters
 * and the solver states; cfdmx_checkpoint_load restores them into a context created
 * with the same configuration (size, grid kind, domain, z periodicity, precision,
 * solver, operator, boundary-condition table, outflow and flux-correction settings and
 * Smagorinsky constant: CFDMX_ERR_ARG otherwise), after which steps continue bit for
 * bit as in the saved run as long as the tolerances (tol, maxit, sor_omega, inner_tol)
 * and mg_extrapolate are also the saved run's, which the load does not check.  Both synchronise the
 * context's stream. */
CFDMX_API int cfdmx_checkpoint_save(cfdmx_ctx *ctx, const char *path);
CFDMX_API int cfdmx_checkpoint_load(cfdmx_ctx *ctx, const char *path);

/* Seeded initial field on the device: u = u0 + amp*(2U-1) on interior cells with U
 * from splitmix64(seed ^ interior index), v = w = 0, p = 0; then the initial
 * BC_staggered_main + BC_staggered_copy (RUN_CPU.cpp:66-109).  It starts a new run: the
 * forcing, X_result and its history and the solver state are reset as for a freshly
 * built solver (bicgstabRe2.hpp:17-52, RUN_CPU.cpp:120). */
CFDMX_API int cfdmx_init_flow(cfdmx_ctx *ctx, double u0, double amp, unsigned long long seed);
/* Initial T1 = T3 = T0 copy + boundary conditions (RUN_CPU.cpp:104-109) for fields
 * uploaded by the caller. */
CFDMX_API int cfdmx_start(cfdmx_ctx *ctx);

/* Immersed-boundary mask.  DFIB_CylinderZ setEta.hpp:136-204 (centre, radius and
 * sub-cell count parameters; the reference fixes 6.5, ly/2, 0.5, 100). */
CFDMX_API int cfdmx_set_eta_cylinder_z(cfdmx_ctx *ctx, double cx, double cy, double radius, int nsub);
/* Smoothed mask of half-width eps for differentiation (no reference counterpart). */
CFDMX_API int cfdmx_set_eta_smoothed(cfdmx_ctx *ctx, double cx, double cy, double radius, double eps);

/* simpulationVariable::set_Re / set_dt (0_1_SimulationVariable.hpp:26-31,95-106) after
 * creation */
CFDMX_API int cfdmx_set_physics(cfdmx_ctx *ctx, double re, double dt);
/* bicgstabRe2::setTolerance (bicgstabRe2.hpp:38) and the iteration limit
 * (iters_max_, :47); also the SOR tolerance/limit (simu.p_criteria, p_sor_iter_max) */
CFDMX_API int cfdmx_set_tolerance(cfdmx_ctx *ctx, double tol, int maxit);

/* ---- stages of one time step (RUN_CPU.cpp:257-436) ---- */
/* ConvectionDifussion, Convection_Difussion.hpp:174-450: T0 -> T1, QUICK, explicit Euler */
CFDMX_API int cfdmx_convdiff(cfdmx_ctx *ctx);
/* SmagorinskyModel, SmagorinskyModel.hpp:30-79 (GradientGauss :6-27): the eddy viscosity
 * Viseff of every interior cell from T0, which cfdmx_convdiff then adds to 1/Re
 * (Convection_Difussion.hpp:251-255,334-338,414-418).  Requires smagorinsky_cs > 0;
 * cfdmx_step runs it at the start of each step (RUN_CPU.cpp:261-263).  The reference's
 * source does not compile (SmagorinskyModel.hpp:61) and is switched off; this computes
 * the formulas it states. */
CFDMX_API int cfdmx_smagorinsky(cfdmx_ctx *ctx);
CFDMX_API int cfdmx_download_viseff(cfdmx_ctx *ctx, double *viseff); /* N values, icelCal order */
/* BC_staggered_main, Boundary_Condition.hpp:37-176, on T0 and p */
CFDMX_API int cfdmx_bc_main(cfdmx_ctx *ctx);
/* BC_staggered_copy, updateBC.hpp:6-110: ghosts T0 -> T1 */
CFDMX_API int cfdmx_bc_copy(cfdmx_ctx *ctx);
/* createBMatrix_omp, b_matrix.hpp:35-98: b from T1 (kept on the device) */
CFDMX_API int cfdmx_pressure_rhs(cfdmx_ctx *ctx);
/* bicgstabRe2::solve, bicgstabRe2.hpp:82-180, on the device b and X_result; returns
 * iterations and relative residual like std::pair<int,double> */
CFDMX_API int cfdmx_solve_bicgstab(cfdmx_ctx *ctx, int *iters, double *residual);
/* Red-black SOR with SorPipeLine_seq's coefficients and update
 * (SorPipeLine.seq.hpp:31-123) on p, RHS from T1; stop on max |dp| < tol */
CFDMX_API int cfdmx_solve_rbsor(cfdmx_ctx *ctx, int *iters, double *err);
/* Pressure_transform_X_result, Transform.hpp:8-25 */
CFDMX_API int cfdmx_transform(cfdmx_ctx *ctx);
/* update_UandF_omp, updateUandF.hpp:90-180: T1, p, eta -> T3, f */
CFDMX_API int cfdmx_project_ib(cfdmx_ctx *ctx);
/* copyVel(T3, T0), copyVel.hpp:4-12 (a device pointer swap) */
CFDMX_API int cfdmx_swap_t0_t3(cfdmx_ctx *ctx);
/* CD_CL, virtualForceIntergrator.hpp:92-111 */
CFDMX_API int cfdmx_cd_cl(cfdmx_ctx *ctx, double *cd, double *cl);
/* CD_CL as above, for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  No reference counterpart.
 * Reads a caller's force field f_dev (device memory, 3*G doubles in the DfibArray::f
 * layout) and writes cd, cl to cdcl_dev[0..1] (device memory), ordered on the context's
 * stream; it neither synchronises nor allocates: the reduction's block partials live in
 * a file-scope device array (1 MiB per GPU, held until the process ends), which limits
 * the grid to 64 Mi interior cells (CFDMX_ERR_ARG beyond).  Every call on one GPU shares
 * that array: calls that overlap on one GPU (contexts on different streams) race.  The
 * result equals cfdmx_cd_cl's on the same field bit for bit. */
CFDMX_API int cfdmx_cd_cl_noalloc(cfdmx_ctx *ctx, const double *f_dev, double *cdcl_dev);
/* get_Cfl, cfl.hpp:3-79, on T1 */
CFDMX_API int cfdmx_cfl(cfdmx_ctx *ctx, double *cfl);

/* Pressure operator on host vectors of N doubles: SPE_matrix::multiply_omp
 * (SPE_sparseMatrix.hpp:92-106) or CSR_matrix::multiply_omp (CSR_sparseMatrix.hpp:231-245). */
CFDMX_API int cfdmx_apply_operator(cfdmx_ctx *ctx, const double *x, double *r);
/* Its transpose, r = A^T y, on host vectors of N doubles.  No counterpart in the reference:
 * the adjoint of the pressure solve needs it wherever A is not symmetric, i.e. the SPE
 * operator on the SML grid (stretched widths and the y-slot swap, SPE_sparseMatrix.hpp:
 * 79,83) and the row-scaled CSR operator.  Column n sums, in a row's slot order (-x, -y,
 * -z, diag, +z, +y, +x), the entries the neighbouring rows hold for it.  Like
 * cfdmx_apply_operator it overwrites two solver work vectors that every solve sets up. */
CFDMX_API int cfdmx_apply_operator_transpose(cfdmx_ctx *ctx, const double *y, double *r);

/* n whole time steps (the loop body of runSeqOmp, RUN_CPU.cpp:257-436, without file
 * output).  info (may be NULL) receives the last step's report. */
CFDMX_API int cfdmx_step(cfdmx_ctx *ctx, int nsteps, cfdmx_step_info *info);

CFDMX_API int cfdmx_synchronize(cfdmx_ctx *ctx);

/* ---- adjoint of the time-averaged drag (no reference counterpart: the reference
 * has no adjoint).  Callers: gradient-based design and its checks (the PyTorch
 * autograd wrapper in the Python package).  Requires the uniform grid, the SPE
 * operator, BiCGSTAB and no Smagorinsky model; any precision (the multigrid modes
 * solve the pressure adjoint with their own solver).  The objective is
 * J = (1/nsteps) sum of Cd over nsteps steps from the current state, whose start
 * boundary conditions are applied first.  The context is returned to its state at
 * entry. */
typedef struct cfdmx_adjoint_info {
    double forward_ms;  /* forward pass: a checkpoint every K steps, the last segment
                           taped as it runs */
    double reverse_ms;  /* recompute of every K-step segment but the last, plus the
                           adjoint steps */
    int forward_iters;  /* pressure iterations of the forward pass */
    int adjoint_iters;  /* iterations of the pressure adjoints */
    int nsteps, checkpoints;
} cfdmx_adjoint_info;

/* J only, by the same forward path. */
CFDMX_API int cfdmx_drag_objective(cfdmx_ctx *ctx, int nsteps, double *J);
/* J and dJ/d(inflow profile) (grad_inflow: ny ghosted entries, zero where the profile is
 * not read; all zero without cfdmx_set_inflow) and dJ/d(cx, cy, R) of the smoothed mask
 * set by cfdmx_set_eta_smoothed (grad_mask[3]); either may be NULL.  checkpoint_every
 * = K bounds the tape to K steps (memory: about (4G + N) * 8 bytes per tape step and
 * (4G + 3N) * 8 per checkpoint, plus 63 G * 8 for the conv-diff rows).  Where the device
 * has room, the forward pass also keeps the pressure solution of every step before the
 * last segment (N * 8 bytes per step), and the recompute of the segments then copies it
 * instead of solving.  The results are the same bits for every K and either way;
 * K >= nsteps tapes every step and recomputes none. */
CFDMX_API int cfdmx_adjoint_drag(cfdmx_ctx *ctx, int nsteps, int checkpoint_every, double *J,
                       double *grad_inflow, double *grad_mask, cfdmx_adjoint_info *info);

/* ---- measurement (no reference counterpart: the reference times phases with
 * stopWatch, STL_clock.hpp:8-42) ---- */
#define CFDMX_K_CONVDIFF 0
#define CFDMX_K_BC 1
#define CFDMX_K_RHS 2
#define CFDMX_K_BICG_INIT 3
#define CFDMX_K_BICG_A 4
#define CFDMX_K_BICG_B 5
#define CFDMX_K_BICG_C 6
#define CFDMX_K_BICG_CHECK 7 /* initial fold and end-of-batch convergence check */
#define CFDMX_K_SOR 8
#define CFDMX_K_PROJECT 9
#define CFDMX_K_MG 10 /* mixed precision: multigrid V-cycle kernels */
#define CFDMX_K_PB 11 /* mixed precision: FP32 BiCGSTAB products and updates */
#define CFDMX_K_COUNT 12
/* Enable (1) or disable (0) HIP-event timing of every launch of the kernels above on
 * the context's stream; resets the totals. */
CFDMX_API int cfdmx_profile(cfdmx_ctx *ctx, int enable);
/* Accumulated device time and launch count of one kernel class. */
CFDMX_API int cfdmx_profile_read(cfdmx_ctx *ctx, int kernel, double *total_ms, long long *launches);

#ifdef __cplusplus
}
#endif
#endif /* CFDMX_H */
