// cfdmx_internal.hpp -- structures shared by the kernels and the context code.
#pragma once
#include "cfdmx_device.hpp"

namespace cfdmx {

// Num/Dir boundary table (FLOWpassing.hpp:6-40) and the extra outflow options
struct BCParams {
    double num[6][3];
    double dir[6][3];
    const double *inflow; // device, ghosted j, or nullptr
    int convective;
    double conv_coef; // convective_speed*dt/Dx[nx-3]
    int flux_correction;
};

// BiCGSTAB scalars kept on the device (bicgstabRe2.hpp:89-176)
struct KState {
    double rho[2]; // r1r0 of iteration m in slot m & 1
    double alpha; // a
    double omega; // w
    double beta;  // b
    double norm0;
    double res;
    double tol;
    double restart_factor; // pRestartFactor_
    int iter;              // pIter
    int maxit;
    int done; // 0 running, 1 converged, 2 iteration limit
    int mode; // next phase A: 0 update p, 1 restart, 2 first iteration
    int timestep;
    int restarts;
};

// SorPipeLine's six face coefficients and -1/(their sum), formed once on the host when
// every metric table is a constant (constant = 1), else per cell in the kernel
struct SorCoef {
    double c[7];
    int constant;
};

struct SorState {
    double residual;
    double tol;
    int iters;
    int itmax;
    int done;
    int pad;
};

// Krylov vectors of one BiCGSTAB solve.  q holds p - omega*Ap between phase C and the
// next phase A.  p is written by phase A for phase C only on grids whose vectors fit the
// caches; above that C forms it again from r and q (k_march_C).
template <typename T>
struct BicgBuf {
    T *r0, *r, *as, *q, *p, *ap;
};
using BicgBuffers = BicgBuf<double>;

void launch_convdiff(const Geo &g, hipStream_t s, const double *u, const double *v,
                     const double *w, double *us, double *vs, double *ws, double dt,
                     double nu, double *cfl_part, const double *viseff = nullptr);
void launch_smagorinsky(const Geo &g, hipStream_t s, const double *u, const double *v,
                        const double *w, const double *coef, double *viseff);
void launch_bc(const Geo &g, hipStream_t s, const BCParams &bc, int advance, double *u,
               double *v, double *w, double *p, double *u1, double *v1, double *w1,
               double *flux_part);
void launch_bc_copy(const Geo &g, hipStream_t s, const double *u0, const double *v0,
                    const double *w0, double *u1, double *v1, double *w1);
void launch_rhs(const Geo &g, hipStream_t s, int pre, const double *us, const double *vs,
                const double *ws, double dt, double *b);
void launch_apply(const Geo &g, hipStream_t s, int pre, const double *x, double *r);
void launch_apply_t(const Geo &g, hipStream_t s, int pre, const double *y, double *r);
void launch_bicg_init(const Geo &g, hipStream_t s, int pre, const double *b, const double *x,
                      double *r0, double *r, double *px1, double *part, KState *st);
// one BiCGSTAB iteration m = three kernels; each folds the previous phase's partials
// (part[0] <- A, part[1] <- B, part[2] <- C)
template <typename T>
void launch_bicg_A(const Geo &g, hipStream_t s, int pre, BicgBuf<T> &B, double *const part[3],
                   KState *st, int m);
template <typename T>
void launch_bicg_B(const Geo &g, hipStream_t s, int pre, BicgBuf<T> &B, double *const part[3],
                   KState *st, int m);
template <typename T>
void launch_bicg_C(const Geo &g, hipStream_t s, BicgBuf<T> &B, T *x, double *const part[3],
                   KState *st, int m);
void launch_bicg_check(const Geo &g, hipStream_t s, double *const part[3], KState *st, int m);
// Multigrid-preconditioned solve (mixed precision: T = float; the same algorithm in
// FP64: T = double).  One multigrid level (interior cells, n = (i*ny + j)*nz + k) and
// the scalars of the preconditioned inner BiCGSTAB.
template <typename T>
struct MGLevelT {
    int nx, ny, nz, n;
    T cx, cy, cz;       // 1/h^2 per direction
    int fx, fy, fz;     // coarsening factors from the finer level (1 or 2)
    int zper;
    FastDiv dz, dy, dzh; // divisions by nz, ny and nz/2
    // u and f of the level.  split = 1 (levels with even nz): stored by colour, the red
    // cells ((i + j + k) even) in [0, n/2) and the black ones in [n/2, n), cell k of row
    // (i, j) at (i*ny + j)*(nz/2) + k/2 of its colour's half.  split = 0: natural order.
    T *u, *f;
    int split;
    // level 0 when split: the natural-order right-hand side and result vectors of the
    // Krylov solve (read by the first sweep, written by the last)
    const T *fin;
    T *uout;
    // coarsest level (stored in natural order): the orthonormal eigenvectors of the 1-D
    // operators, Qx (nx x nx, row m = mode m), Qy, Qz, then their eigenvalues (nx, ny, nz
    // values); the level's operator is their Kronecker sum
    const T *fd;
};

// the hierarchy stops at the first level of at most this many cells, solved exactly
constexpr int kMGCoarseCells = 512;
constexpr int kMGCoarseAxis = 512; // largest axis length of that level
// a level above the coarsest of at most this many cells (stored by colour) runs its whole
// V-cycle, the coarse solve included, in one workgroup with its u in LDS (k_mg_tail)
constexpr int kMGTailCells = 4096;

struct PBState {
    double rho, alpha, omega, beta, norm0, res, tol;
    int iter, maxit, done;
};

template <typename T>
void launch_mg_vcycle(hipStream_t s, const MGLevelT<T> *L, int nlev, int l);
template <typename T>
void launch_pb_spmv(const Geo &g, hipStream_t s, const MGLevelT<T> &L0, int ns, const T *in,
                    T *out, const T *w, double *part, const PBState *st);
template <typename T>
void launch_pb_update_p(const Geo &g, hipStream_t s, int first, const T *r, const T *v, T *p,
                        const PBState *st);
template <typename T>
void launch_pb_update_s(const Geo &g, hipStream_t s, const double *part, const T *r, const T *v,
                        T *sv, PBState *st);
template <typename T>
void launch_pb_update_xr(const Geo &g, hipStream_t s, int first, const double *part_in,
                         const T *ph, const T *sh, const T *sv, const T *t, const T *r0,
                         double *e, T *r, double *part_out, PBState *st);
void launch_pb_end(const Geo &g, hipStream_t s, const double *part, PBState *st);

// iterative refinement: the FP64 residual into r, r0 (the correction e, FP64 in both
// modes, starts at 0 through the inner solve's first update), x += e with x kept in xb,
// fold of the three sums
template <typename T>
void launch_mixed_resid(const Geo &g, hipStream_t s, int pre, const double *b, const double *x,
                        T *r, T *r0, double *part);
void launch_mixed_update(const Geo &g, hipStream_t s, double *x, const double *e, double *xb);
void launch_mixed_fold(const Geo &g, hipStream_t s, const double *part, double *out);

void launch_bicg_fin0(const Geo &g, hipStream_t s, double *part, KState *st);
void launch_nanfix(hipStream_t s, int n, double *x, const double *px1, int *count);
void launch_copy(hipStream_t s, int n, const double *a, double *b);
void launch_extrapolate(const Geo &g, hipStream_t s, double *x, double *xp);
void launch_fill(hipStream_t s, int n, double val, double *a);
void launch_sor_rhs(const Geo &g, hipStream_t s, const double *us, const double *vs,
                    const double *ws, double dt, double *cf6);
void launch_sor_half(const Geo &g, hipStream_t s, int colour, double omega, const double *cf6,
                     double *p, double *part, SorState *st);
void launch_sor_fin(const Geo &g, hipStream_t s, double *part, SorState *st);
// red-black SOR on colour-split storage (even nzc only)
int sor_rb_blocks(const Geo &g);
void launch_sor_split(const Geo &g, hipStream_t s, const double *us, const double *vs,
                      const double *ws, double dt, const double *p, double *cf6c, double *pc);
void launch_sor_merge(const Geo &g, hipStream_t s, const double *pc, double *p);
void launch_sor_rb(const Geo &g, hipStream_t s, int colour, double omega, const SorCoef &sc,
                   const double *cf6c, double *pc, const double *p, double *part, SorState *st);
void launch_sor_rb_fin(const Geo &g, hipStream_t s, double *part, SorState *st);
// prev (T0 of the step) and dpart (3 * nblocks doubles) add the steady-state check;
// both null leaves it out.  ghosts (grids with nzc a multiple of 4): 0 leaves every z
// ghost untouched; 1 also stores the z ghosts of u, v, w and p in the rows written, for a
// caller whose boundary conditions rewrite them next; 2 as 1 and f's z ghosts as zeros,
// for a caller that knows they are zero
void launch_project(const Geo &g, hipStream_t s, int from_x, double dt, int convective,
                    const double *xres, double *p, const double *eta, const double *us,
                    const double *vs, const double *ws, double *u, double *v, double *w,
                    double *f, double *part, const double *const *prev = nullptr,
                    double *dpart = nullptr, int ghosts = 0);
void launch_max3_fin(const Geo &g, hipStream_t s, const double *dpart, double *out);
void launch_cdcl_fin(const Geo &g, hipStream_t s, const double *part, double lz, double *out);
void launch_force_partial(const Geo &g, hipStream_t s, const double *f, double *part);
void launch_max_fin(const Geo &g, hipStream_t s, const double *part, double *out);
void launch_eta_cylz(const Geo &g, hipStream_t s, const double *Xn, const double *Yn, double cx,
                     double cy, double R, int nsub, double inv_nsub2, double *eta);
void launch_init_flow(const Geo &g, hipStream_t s, double u0, double amp,
                      unsigned long long seed, double *u, double *v, double *w, double *p);

void launch_transform(const Geo &g, hipStream_t s, const double *x, double *p);
void launch_plot3d_q(const Geo &g, hipStream_t s, const double *u, const double *v,
                     const double *w, const double *p, const double *eta, float *out);
void launch_set_int(hipStream_t s, int *a, int val);

// adjoint of the drag (cfdmx_adjoint.hip)
void launch_adj_bc(const Geo &g, hipStream_t s, const BCParams &bc, int advance, double *lu,
                   double *lv, double *lw, double *tin, double *lin, double *part, double rarea);
void launch_adj_project(const Geo &g, hipStream_t s, double dt, int use_x, int convective,
                        double wf, const double *x, const double *p, const double *eta,
                        double *const us[3], double *const l3[3], double *const ls[3],
                        double *const le[3], double *lx, double *leta);
void launch_adj_rhs(const Geo &g, hipStream_t s, double dt, const double *lb,
                    double *const ls[3]);
void launch_adj_remove_mean(const Geo &g, hipStream_t s, double *a, double *part);
void launch_adj_periodic_interior(const Geo &g, hipStream_t s, double *const l[3]);
void launch_adj_convdiff(const Geo &g, hipStream_t s, double dt, double nu,
                         const double *const U[3], double *Jr, double *const ls[3],
                         double *const lu[3]);
void launch_adj_mask(const Geo &g, hipStream_t s, const double *Xn, const double *Yn, double cx,
                     double cy, double R, double eps, const double *leta, double *part,
                     double *out);

} // namespace cfdmx
