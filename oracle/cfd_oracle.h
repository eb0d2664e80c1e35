/*
 * cfd_oracle.h -- CPU oracle for the flow-past-a-cylinder time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference solver's hot path (jimmy3769tw/CFD_NSsolver_forHPC_flowPassingCylinder),
 * used as the checker for the HIP library.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Every function cites the reference file:line it restates.  Arrays use the
 * reference layout: ghosted fields icel = (i*ny + j)*nz + k with two ghost
 * layers (gC = 2, structureGrid.hpp:34,153-156); pressure unknowns use the
 * interior layout icelCal = ((i-2)*nyc + (j-2))*nzc + (k-2) (structureGrid.hpp:176-181).
 *
 * Reductions come in two orders:
 *   ORC_RED_SERIAL  -- index order, exactly math::dot / L2Norm at one OpenMP
 *                      thread (math.hpp:13-36) and virtualF_Int (virtualForceIntergrator.hpp:79-86);
 *   ORC_RED_BLOCKED -- the device's canonical orders (blocks of 256 threads x 4 cells
 *                      over the linear cell index for the initial residual, or the
 *                      plane-marching order of the BiCGSTAB phase kernels for the
 *                      (Ap,r0), (As,s), (As,As), (r,r), (r,r0) sums;
 *                      per-wave tree, fixed two-level combine), so that the HIP
 *                      library can be checked bit-for-bit.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_GRID_UNIFORM 0
#define ORC_GRID_SML 1

#define ORC_RED_SERIAL 0
#define ORC_RED_BLOCKED 1

#define ORC_SOLVER_BICGSTAB 0
#define ORC_SOLVER_RBSOR 1
#define ORC_SOLVER_SOR_SEQ 2

#define ORC_PRECOND_NONE 0   /* SPE 7-slot operator (SPE_sparseMatrix.hpp:42-106) */
#define ORC_PRECOND_JACOBI 1 /* row-scaled CSR operator (structureGrid.hpp:557-654) */

/* canonical device reduction shapes */
#define ORC_RB_THREADS 256
#define ORC_RB_CELLS 4
#define ORC_MARCH_TILE 1024 /* plane cells per tile of the marching stencil kernels */
#define ORC_MARCH_CHUNK 8   /* planes per marching block */

typedef struct orc_grid {
    int nxc, nyc, nzc; /* interior cells (grid::nxCal, nyCal, nzCal) */
    int nx, ny, nz;    /* ghosted sizes, n = nCal + 4 */
    double lx, ly, lz;
    double *X, *Y, *Z;          /* node coordinates, n+1 each */
    double *Dx, *Dy, *Dz;       /* cell widths, n each */
    double *Dxs, *Dys, *Dzs;    /* centre-to-centre spacings, n each */
} orc_grid;

typedef struct orc_bc {
    double num[6][3]; /* Num table (FLOWpassing.hpp:6-40): face x component */
    double dir[6][3]; /* Dir table */
    const double *inflow; /* optional u inflow per ghosted j (length ny); NULL = dir[0][0] */
    int z_periodic;       /* faces 4/5 periodic (no reference counterpart) */
    int outflow_convective; /* face 1 u: convective outflow (no reference counterpart) */
    double conv_speed;
    int flux_correction;  /* rescale the outlet flux to the inlet flux */
} orc_bc;

typedef struct orc_params {
    double dt, nu;
    int solver, precond;
    double tol;
    int maxit;
    double omega;
    int reduction;
    int skip_first_pressure; /* RUN_CPU.cpp:303: no pressure solve on step 1 */
    int swap_buffers;        /* 0: copyVel T3->T0 (copyVel.hpp:4-12); 1: pointer swap */
    double smag_cs;          /* Smagorinsky constant (SmagorinskyModel.hpp:38: 0.18); 0: off */
} orc_params;

/* ---- grid (structureGrid.hpp) ---- */
int orc_grid_build(orc_grid *g, int kind);

/* ---- reductions ---- */
double orc_blocked_sum(const double *v, long n);
double orc_march_sum(const orc_grid *g, const double *v);
double orc_serial_sum(const double *v, long n);

/* ---- stages ---- */
void orc_convdiff(const orc_grid *g, double dt, double nu, int z_periodic,
                  const double *u, const double *v, const double *w,
                  double *us, double *vs, double *ws);
/* the same with the eddy viscosity viseff (per interior cell, icelCal order) added to nu */
void orc_convdiff_les(const orc_grid *g, double dt, double nu, const double *viseff,
                      int z_periodic, const double *u, const double *v, const double *w,
                      double *us, double *vs, double *ws);
/* SmagorinskyModel (SmagorinskyModel.hpp:6-79): coef[icelCal] = (cs delta)^2 with
 * delta = cbrt-by-pow(Dx Dy Dz), then viseff = coef * sqrt(|strain sum|) */
void orc_smagorinsky_coef(const orc_grid *g, double cs, double *coef);
void orc_smagorinsky(const orc_grid *g, const double *coef, const double *u, const double *v,
                     const double *w, double *viseff);
void orc_bc_main(const orc_grid *g, const orc_bc *bc, int advance_outflow, double dt,
                 double *u, double *v, double *w, double *p);
void orc_bc_copy(const orc_grid *g, const double *u0, const double *v0, const double *w0,
                 double *u1, double *v1, double *w1);
void orc_rhs(const orc_grid *g, double dt, int precond, int z_periodic,
             const double *us, const double *vs, const double *ws, double *b);
void orc_apply_A(const orc_grid *g, int precond, int z_periodic, const double *x, double *r);
void orc_apply_AT(const orc_grid *g, int precond, int z_periodic, const double *y, double *r);
void orc_spe_build(const orc_grid *g, double *val);
void orc_spe_apply_vals(const orc_grid *g, const double *val, const double *x, double *r);

typedef struct orc_bicg orc_bicg;
orc_bicg *orc_bicg_new(long n);
void orc_bicg_free(orc_bicg *s);
int orc_bicg_timestep(const orc_bicg *s);
int orc_bicg_solve(orc_bicg *s, const orc_grid *g, int precond, int z_periodic, int reduction,
                   double tol, int maxit, const double *b, double *x, double *res_out);

void orc_transform(const orc_grid *g, const double *x, double *p);
int orc_sor(const orc_grid *g, int redblack, int z_periodic, double dt, double omega, double tol,
            int itmax, const double *us, const double *vs, const double *ws, double *p,
            double *err_out);
void orc_update_uf(const orc_grid *g, double dt, int z_periodic, int outflow_convective,
                   const double *p, const double *eta,
                   const double *us, const double *vs, const double *ws,
                   double *u, double *v, double *w, double *f);
void orc_eta_cylinder_z(const orc_grid *g, double cx, double cy, double radius, int nsub,
                        double *eta);
void orc_eta_smoothed(const orc_grid *g, double cx, double cy, double radius, double eps,
                      double *eta);
void orc_cd_cl(const orc_grid *g, int reduction, const double *f, double *cd, double *cl);
double orc_cfl(const orc_grid *g, double dt, int z_periodic,
               const double *u, const double *v, const double *w);
void orc_maxdiff(const orc_grid *g, const double *u0, const double *v0, const double *w0,
                 const double *u3, const double *v3, const double *w3, double out[3]);
void orc_maxdiff_faces(const orc_grid *g, int zp, const double *u0, const double *v0,
                       const double *w0, const double *u3, const double *v3, const double *w3,
                       double out[3]);
void orc_init_velocity(const orc_grid *g, double u0, double amp, unsigned long long seed,
                       double *u, double *v, double *w);

/* ---- whole time loop (RUN_CPU.cpp:66-109,257-436) ---- */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_new(const orc_grid *g, const orc_bc *bc, const orc_params *prm);
void orc_sim_free(orc_sim *s);
/* which: 0..2 T0 u,v,w; 3..5 T1; 6..8 T3; 9 p; 10 eta; 11 f (3G); 12 X_result (N) */
double *orc_sim_field(orc_sim *s, int which);
void orc_sim_start(orc_sim *s); /* initial BC_main + BC_copy, T1 = T3 = T0 */
/* out: cd, cl, time, iters, err, cfl, restarts, solved, max(T0 - T3) of u, v, w */
int orc_sim_step(orc_sim *s, double out[11]);
/* out: cd, cl, time, iters, err, cfl, restarts, solve_flag */

/* ---- adjoint of the time-averaged drag (cfd_adjoint.c; no reference counterpart) ----
 * Uniform grid, SPE operator.  J = (1/nsteps) sum_n Cd^n from U^0 = BC_start(U_init);
 * the mask is orc_eta_smoothed(cx, cy, R, eps); bc->inflow (ghosted j) is the inflow
 * profile.  grad_inflow (length ny, ghosted j) and grad_mask (d/dcx, d/dcy, d/dR) may be
 * NULL.  The pressure adjoint is solved to adj_tol. */
double orc_adjoint_drag(const orc_grid *g, const orc_bc *bc, const orc_params *prm,
                        const double *u_init, const double *v_init, const double *w_init,
                        double cx, double cy, double R, double eps, int nsteps,
                        double adj_tol, double *grad_inflow, double grad_mask[3]);
double orc_drag_objective(const orc_grid *g, const orc_bc *bc, const orc_params *prm,
                          const double *u_init, const double *v_init, const double *w_init,
                          double cx, double cy, double R, double eps, int nsteps);
/* stage transposes, for dot-product tests */
void orc_adj_convdiff(const orc_grid *g, double dt, double nu, int zp, const double *u,
                      const double *v, const double *w, const double *lus, const double *lvs,
                      const double *lws, double *lu, double *lv, double *lw);
void orc_adj_bc_main(const orc_grid *g, const orc_bc *bc, int advance_outflow, double dt,
                     double *lu, double *lv, double *lw, double *l_inflow);

#ifdef __cplusplus
}
#endif
#endif
