/*
 * cfd_adjoint.c -- CPU reference adjoint of the time-averaged drag coefficient
 * (test infrastructure only, like cfd_oracle.c; no reference counterpart: the
 * reference solver has no adjoint).
 *
 * Forward map, one time step n (orc_sim_step with swap_buffers = 1):
 *   S1  U* = convdiff(U^{n-1}) on the computed faces; the entries orc_bc_copy lists
 *       are copies of U^{n-1}; with periodic z the z-ghosts of the interior columns
 *       are images of U*                              (Convection_Difussion.hpp:174-450)
 *   S2  b = -(1/dt) D U*                               (b_matrix.hpp:47-93)
 *   S3  x = A^+ b  (BiCGSTAB on the SPE operator, treated as an exact solve)
 *   S4  T2 = U* - dt G x on the projected faces, U^n = (1 - eta_f) T2,
 *       f = (U^n - T2) / dt; the outlet face carries U* when the outflow is
 *       convective                                     (updateUandF.hpp:90-180)
 *   S5  Cd^n = -2/lz sum f_x dV                        (virtualForceIntergrator.hpp:92-111)
 *   S6  U^n <- BC(U^n)  (table faces, inflow profile, convective outflow, flux
 *       correction, periodic z)                        (Boundary_Condition.hpp:37-176)
 * and U^0 = BC_start(U_init).  Objective J = (1/N) sum_n Cd^n.
 *
 * The adjoint runs the transposes in reverse: BC^T, projection^T (which also
 * yields the eta and x adjoints), A^+ (A is symmetric on uniform grids, so the
 * pressure adjoint reuses the forward BiCGSTAB; its solution is made mean-free),
 * D^T, then the transposed Jacobian of the QUICK update, linearised at U^{n-1}
 * (the upwind branch of each face value is frozen at the forward state), plus the
 * copies.  Each stage is written in scatter form over the forward loops; the HIP
 * library computes the same transposes in gather form.  Gradients: the inflow
 * profile u_in(j) of face 0 and the smoothed mask's centre and radius
 * (orc_eta_smoothed).
 */
#include "cfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GC 2

static inline long IC(const orc_grid *g, long i, long j, long k)
{
    return (i * g->ny + j) * g->nz + k;
}

/* --------------------------------------------------------------------- */
/* QUICK face-value weights over the stencil [f(+2), f(+1), f(0), f(-1), f(-2)]
 * of QuickUnitA / QuickUnitB (Convection_Difussion.hpp:20-90), for the branch
 * chosen by the sign of the transport velocity.                          */
/* --------------------------------------------------------------------- */
static void quick_weights(int unitA, double Up, double Un, const double *Di, const double *Ds,
                          int I, double wp[5], double wn[5])
{
    for (int m = 0; m < 5; ++m)
        wp[m] = wn[m] = 0.0;
    if (unitA) {
        if (Up > 0.0) {
            const double C = 0.125 * Di[I + 1] * Di[I + 1] / Ds[I];
            wp[1] = 0.5 - C / Di[I + 1];
            wp[2] = 0.5 + C / Di[I + 1] + C / Di[I];
            wp[3] = -C / Di[I];
        } else {
            const double C = 0.125 * Di[I + 1] * Di[I + 1] / Ds[I + 1];
            wp[0] = -C / Di[I + 2];
            wp[1] = 0.5 + C / Di[I + 2] + C / Di[I + 1];
            wp[2] = 0.5 - C / Di[I + 1];
        }
        if (Un > 0.0) {
            const double C = 0.125 * Di[I] * Di[I] / Ds[I - 1];
            wn[2] = 0.5 - C / Di[I];
            wn[3] = 0.5 + C / Di[I] + C / Di[I - 1];
            wn[4] = -C / Di[I - 1];
        } else {
            const double C = 0.125 * Di[I] * Di[I] / Ds[I];
            wn[1] = -C / Di[I + 1];
            wn[2] = 0.5 + C / Di[I + 1] + C / Di[I];
            wn[3] = 0.5 - C / Di[I];
        }
    } else {
        if (Up > 0.0) {
            const double C = 0.125 * Ds[I] * Ds[I] / Di[I];
            wp[1] = 0.5 - C / Ds[I];
            wp[2] = 0.5 + C / Ds[I] + C / Ds[I - 1];
            wp[3] = -C / Ds[I - 1];
        } else {
            const double C = 0.125 * Ds[I] * Ds[I] / Di[I + 1];
            wp[0] = -C / Ds[I + 1];
            wp[1] = 0.5 + C / Ds[I + 1] + C / Ds[I];
            wp[2] = 0.5 - C / Ds[I];
        }
        if (Un > 0.0) {
            const double C = 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I - 1];
            wn[2] = 0.5 - C / Ds[I - 1];
            wn[3] = 0.5 + C / Ds[I - 1] + C / Ds[I - 2];
            wn[4] = -C / Ds[I - 2];
        } else {
            const double C = 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I];
            wn[1] = -C / Ds[I];
            wn[2] = 0.5 + C / Ds[I] + C / Ds[I - 1];
            wn[3] = 0.5 - C / Ds[I - 1];
        }
    }
}

/* L[c] += J^T lout for the update of component d at its computed faces */
static void adj_component(const orc_grid *g, int d, double dt, double nu, int zp,
                          const double *const U[3], const double *lout, double *const L[3])
{
    const int n[3] = {g->nx, g->ny, g->nz};
    const long s[3] = {(long)g->ny * g->nz, g->nz, 1};
    const double *D[3] = {g->Dx, g->Dy, g->Dz};
    const double *Ds[3] = {g->Dxs, g->Dys, g->Dzs};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = GC;
        hi[a] = n[a] - GC;
    }
    if (!(d == 2 && zp)) hi[d] = n[d] - GC - 1; /* skip of the last face (d = 1) */
    const double *f = U[d];
    double *Lf = L[d];
    int I[3];
    for (I[0] = lo[0]; I[0] < hi[0]; ++I[0])
        for (I[1] = lo[1]; I[1] < hi[1]; ++I[1])
            for (I[2] = lo[2]; I[2] < hi[2]; ++I[2]) {
                const long ic = IC(g, I[0], I[1], I[2]);
                const double lam = lout[ic];
                if (lam == 0.0) continue;
                const int Id = I[d];
                const long sd = s[d];
                const double IT = (Ds[d][Id] * 0.5) / D[d][Id];
                /* identity */
                Lf[ic] += lam;
                /* own direction: QUICK-A, (fP^2 - fN^2) / Ds */
                {
                    const double Up = 0.5 * (f[ic + sd] + f[ic]);
                    const double Un = 0.5 * (f[ic - sd] + f[ic]);
                    double wp[5], wn[5];
                    quick_weights(1, Up, Un, D[d], Ds[d], Id, wp, wn);
                    double fP = 0.0, fN = 0.0;
                    for (int m = 0; m < 5; ++m) {
                        fP += wp[m] * f[ic + (2 - m) * sd];
                        fN += wn[m] * f[ic + (2 - m) * sd];
                    }
                    const double gP = -dt * 2.0 * fP / Ds[d][Id] * lam;
                    const double gN = dt * 2.0 * fN / Ds[d][Id] * lam;
                    for (int m = 0; m < 5; ++m)
                        Lf[ic + (2 - m) * sd] += wp[m] * gP + wn[m] * gN;
                    /* diffusion along d */
                    const double cp = nu * dt / (D[d][Id + 1] * Ds[d][Id]);
                    const double cm = nu * dt / (D[d][Id] * Ds[d][Id]);
                    Lf[ic + sd] += cp * lam;
                    Lf[ic - sd] += cm * lam;
                    Lf[ic] -= (cp + cm) * lam;
                }
                /* cross directions: QUICK-B, (fP*Bp - fN*Bn) / D */
                for (int b = 0; b < 3; ++b) {
                    if (b == d) continue;
                    const int Ib = I[b];
                    const long sb = s[b];
                    const double *B = U[b];
                    double *LB = L[b];
                    const double Bp = B[ic] + (B[ic + sd] - B[ic]) * IT;
                    const double Bn = B[ic - sb] + (B[ic - sb + sd] - B[ic - sb]) * IT;
                    double wp[5], wn[5];
                    quick_weights(0, Bp, Bn, D[b], Ds[b], Ib, wp, wn);
                    double fP = 0.0, fN = 0.0;
                    for (int m = 0; m < 5; ++m) {
                        fP += wp[m] * f[ic + (2 - m) * sb];
                        fN += wn[m] * f[ic + (2 - m) * sb];
                    }
                    const double gP = -dt * Bp / D[b][Ib] * lam;
                    const double gN = dt * Bn / D[b][Ib] * lam;
                    for (int m = 0; m < 5; ++m)
                        Lf[ic + (2 - m) * sb] += wp[m] * gP + wn[m] * gN;
                    const double gBp = -dt * fP / D[b][Ib] * lam;
                    const double gBn = dt * fN / D[b][Ib] * lam;
                    LB[ic] += (1.0 - IT) * gBp;
                    LB[ic + sd] += IT * gBp;
                    LB[ic - sb] += (1.0 - IT) * gBn;
                    LB[ic - sb + sd] += IT * gBn;
                    /* diffusion along b */
                    const double cp = nu * dt / (Ds[b][Ib] * D[b][Ib]);
                    const double cm = nu * dt / (Ds[b][Ib - 1] * D[b][Ib]);
                    Lf[ic + sb] += cp * lam;
                    Lf[ic - sb] += cm * lam;
                    Lf[ic] -= (cp + cm) * lam;
                }
            }
}

void orc_adj_convdiff(const orc_grid *g, double dt, double nu, int zp, const double *u,
                      const double *v, const double *w, const double *lus, const double *lvs,
                      const double *lws, double *lu, double *lv, double *lw)
{
    const double *U[3] = {u, v, w};
    double *const L[3] = {lu, lv, lw};
    adj_component(g, 0, dt, nu, zp, U, lus, L);
    adj_component(g, 1, dt, nu, zp, U, lvs, L);
    adj_component(g, 2, dt, nu, zp, U, lws, L);
}

/* transpose of a[dst] = a[src] (dst is overwritten) */
static inline void adj_copy(double *l, long dst, long src)
{
    const double t = l[dst];
    l[dst] = 0.0;
    l[src] += t;
}

/* transpose of periodic_z_fill (cfd_oracle.c): the four ghost images per column */
static void adj_periodic(const orc_grid *g, double *l, int all_columns)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nzc = g->nzc;
    const int lo = all_columns ? 0 : GC;
    for (int i = lo; i < nx - lo; ++i)
        for (int j = lo; j < ny - lo; ++j) {
            const long c = IC(g, i, j, 0);
            adj_copy(l, c + nz - 1, c + nz - 1 - nzc);
            adj_copy(l, c + nz - 2, c + nz - 2 - nzc);
            adj_copy(l, c + 0, c + 0 + nzc);
            adj_copy(l, c + 1, c + 1 + nzc);
        }
}

/* transpose of a[g1] = a[g2] = cst + wgt * a[src] */
static inline double adj_ghost2(double *l, long g1, long g2, long src, double wgt)
{
    const double t = l[g1] + l[g2];
    l[g1] = l[g2] = 0.0;
    l[src] += wgt * t;
    return t;
}

/* transpose of the velocity part of orc_bc_main; l_inflow (ghosted j) accumulates
 * the adjoint of the face-0 inflow values when bc->inflow is set */
void orc_adj_bc_main(const orc_grid *g, const orc_bc *bc, int advance_outflow, double dt,
                     double *lu, double *lv, double *lw, double *l_inflow)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double(*Num)[3] = bc->num;
    if (bc->z_periodic) {
        adj_periodic(g, lw, 1);
        adj_periodic(g, lv, 1);
        adj_periodic(g, lu, 1);
    } else {
        for (int i = GC; i < nx - GC; ++i)
            for (int j = GC; j < ny - GC; ++j) {
                /* face 5 (forward order w, u, v) */
                adj_ghost2(lv, IC(g, i, j, nz - 1), IC(g, i, j, nz - 2), IC(g, i, j, nz - 3),
                           2.0 * Num[5][1] - 1.0);
                adj_ghost2(lu, IC(g, i, j, nz - 1), IC(g, i, j, nz - 2), IC(g, i, j, nz - 3),
                           2.0 * Num[5][0] - 1.0);
                const long a = IC(g, i, j, nz - 1), b = IC(g, i, j, nz - 2), c = IC(g, i, j, nz - 3);
                const double t = lw[a] + lw[b] + lw[c];
                lw[a] = lw[b] = lw[c] = 0.0;
                lw[IC(g, i, j, nz - 4)] += Num[5][2] * t;
            }
        for (int i = GC; i < nx - GC; ++i)
            for (int j = GC; j < ny - GC; ++j) {
                /* face 4 (u, v, w) */
                adj_ghost2(lw, IC(g, i, j, 0), IC(g, i, j, 1), IC(g, i, j, 2), Num[4][2]);
                adj_ghost2(lv, IC(g, i, j, 0), IC(g, i, j, 1), IC(g, i, j, 2),
                           2.0 * Num[4][1] - 1.0);
                adj_ghost2(lu, IC(g, i, j, 0), IC(g, i, j, 1), IC(g, i, j, 2),
                           2.0 * Num[4][0] - 1.0);
            }
    }
    /* face 3 (v, u, w) */
    for (int i = GC; i < nx - GC; ++i)
        for (int k = GC; k < nz - GC; ++k) {
            adj_ghost2(lw, IC(g, i, ny - 1, k), IC(g, i, ny - 2, k), IC(g, i, ny - 3, k),
                       2.0 * Num[3][2] - 1.0);
            adj_ghost2(lu, IC(g, i, ny - 1, k), IC(g, i, ny - 2, k), IC(g, i, ny - 3, k),
                       2.0 * Num[3][0] - 1.0);
            const long a = IC(g, i, ny - 1, k), b = IC(g, i, ny - 2, k), c = IC(g, i, ny - 3, k);
            const double t = lv[a] + lv[b] + lv[c];
            lv[a] = lv[b] = lv[c] = 0.0;
            lv[IC(g, i, ny - 4, k)] += Num[3][1] * t;
        }
    /* face 2 (u, v, w) */
    for (int i = GC; i < nx - GC; ++i)
        for (int k = GC; k < nz - GC; ++k) {
            adj_ghost2(lw, IC(g, i, 0, k), IC(g, i, 1, k), IC(g, i, 2, k), 2.0 * Num[2][2] - 1.0);
            adj_ghost2(lv, IC(g, i, 0, k), IC(g, i, 1, k), IC(g, i, 2, k), Num[2][1]);
            adj_ghost2(lu, IC(g, i, 0, k), IC(g, i, 1, k), IC(g, i, 2, k), 2.0 * Num[2][0] - 1.0);
        }
    /* flux correction: u_b = u_b + (qin - qout) / area on the outlet face */
    if (bc->flux_correction) {
        double area = 0.0, lcorr = 0.0;
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                area += g->Dy[j] * g->Dz[k];
                const long a = IC(g, nx - 1, j, k), b = IC(g, nx - 2, j, k), c = IC(g, nx - 3, j, k);
                const double t = lu[a] + lu[b] + lu[c];
                lu[a] = lu[b] = 0.0;
                lu[c] = t; /* u_b + corr: identity in u_b */
                lcorr += t;
            }
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const double w = g->Dy[j] * g->Dz[k] / area;
                lu[IC(g, 1, j, k)] += lcorr * w;
                lu[IC(g, nx - 3, j, k)] -= lcorr * w;
            }
    }
    /* face 1 (u, v, w) */
    for (int j = GC; j < ny - GC; ++j)
        for (int k = GC; k < nz - GC; ++k) {
            adj_ghost2(lw, IC(g, nx - 1, j, k), IC(g, nx - 2, j, k), IC(g, nx - 3, j, k),
                       2.0 * Num[1][2] - 1.0);
            adj_ghost2(lv, IC(g, nx - 1, j, k), IC(g, nx - 2, j, k), IC(g, nx - 3, j, k),
                       2.0 * Num[1][1] - 1.0);
            const long a = IC(g, nx - 1, j, k), b = IC(g, nx - 2, j, k), c = IC(g, nx - 3, j, k);
            const double t = lu[a] + lu[b] + lu[c];
            lu[a] = lu[b] = lu[c] = 0.0;
            if (bc->outflow_convective && advance_outflow) {
                const double cc = bc->conv_speed * dt / g->Dx[nx - 3];
                lu[c] += (1.0 - cc) * t;
                lu[IC(g, nx - 4, j, k)] += cc * t;
            } else {
                lu[IC(g, nx - 4, j, k)] += Num[1][0] * t;
            }
        }
    /* face 0 (u, v, w) */
    for (int j = GC; j < ny - GC; ++j)
        for (int k = GC; k < nz - GC; ++k) {
            adj_ghost2(lw, IC(g, 0, j, k), IC(g, 1, j, k), IC(g, 2, j, k), 2.0 * Num[0][2] - 1.0);
            adj_ghost2(lv, IC(g, 0, j, k), IC(g, 1, j, k), IC(g, 2, j, k), 2.0 * Num[0][1] - 1.0);
            const double t = adj_ghost2(lu, IC(g, 0, j, k), IC(g, 1, j, k), IC(g, 2, j, k),
                                        Num[0][0]);
            if (bc->inflow && l_inflow) l_inflow[j] += t;
        }
}

/* transpose of the orc_bc_copy assignments T1[e] = T0[e], for the entries of T1 that
 * the update of the next step does not overwrite (the computed faces and, with
 * periodic z, the z-ghosts of the interior columns) */
static void adj_bc_copy(const orc_grid *g, int zp, const double *lus, const double *lvs,
                        const double *lws, double *lu, double *lv, double *lw)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const long G = (long)nx * ny * nz;
    char *mu = (char *)calloc((size_t)G, 1), *mv = (char *)calloc((size_t)G, 1),
         *mw = (char *)calloc((size_t)G, 1);
    /* the copy set, as in orc_bc_copy */
    for (int j = GC; j < ny - GC; ++j)
        for (int k = GC; k < nz - GC; ++k) {
            for (int a = 0; a < 2; ++a) {
                const long c = IC(g, a, j, k);
                mu[c] = mv[c] = mw[c] = 1;
            }
            const int xe[3] = {nx - 1, nx - 3, nx - 2};
            for (int a = 0; a < 3; ++a) {
                const long c = IC(g, xe[a], j, k);
                mu[c] = 1;
                if (a != 1) mv[c] = mw[c] = 1;
            }
        }
    for (int i = GC; i < nx - GC; ++i)
        for (int k = GC; k < nz - GC; ++k) {
            for (int a = 0; a < 2; ++a) {
                const long c = IC(g, i, a, k);
                mu[c] = mv[c] = mw[c] = 1;
            }
            for (int jj = ny - 2; jj < ny; ++jj) mu[IC(g, i, jj, k)] = mw[IC(g, i, jj, k)] = 1;
            for (int jj = ny - 3; jj < ny; ++jj) mv[IC(g, i, jj, k)] = 1;
        }
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j) {
            for (int a = 0; a < 2; ++a) {
                const long c = IC(g, i, j, a);
                mu[c] = mv[c] = mw[c] = 1;
            }
            for (int kk = nz - 2; kk < nz; ++kk) mu[IC(g, i, j, kk)] = mv[IC(g, i, j, kk)] = 1;
            for (int kk = nz - 3; kk < nz; ++kk) mw[IC(g, i, j, kk)] = 1;
        }
    /* minus what convdiff writes afterwards */
    if (zp)
        for (int i = GC; i < nx - GC; ++i)
            for (int j = GC; j < ny - GC; ++j) {
                const long c = IC(g, i, j, 0);
                const long e[4] = {c, c + 1, c + nz - 2, c + nz - 1};
                for (int q = 0; q < 4; ++q) mu[e[q]] = mv[e[q]] = mw[e[q]] = 0;
                mw[IC(g, i, j, nz - 3)] = 0; /* the w face nz-3 is computed */
            }
    for (long e = 0; e < G; ++e) {
        if (mu[e]) lu[e] += lus[e];
        if (mv[e]) lv[e] += lvs[e];
        if (mw[e]) lw[e] += lws[e];
    }
    free(mu);
    free(mv);
    free(mw);
}

/* transpose of orc_rhs (SPE path): lus += -(1/dt) D^T lb */
static void adj_rhs(const orc_grid *g, double dt, const double *lb, double *lus, double *lvs,
                    double *lws)
{
    const double ddt = 1.0 / dt;
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k) {
                const long ic = IC(g, i, j, k);
                const long row = ((long)(i - GC) * g->nyc + (j - GC)) * g->nzc + (k - GC);
                const double t = -ddt * lb[row];
                lus[ic] += t / g->Dx[i];
                lus[IC(g, i - 1, j, k)] -= t / g->Dx[i];
                lvs[ic] += t / g->Dy[j];
                lvs[IC(g, i, j - 1, k)] -= t / g->Dy[j];
                lws[ic] += t / g->Dz[k];
                lws[IC(g, i, j, k - 1)] -= t / g->Dz[k];
            }
}

/* transpose of orc_update_uf with the drag weight: inputs the adjoint of the new
 * velocity (lu3..) and of f_x (dJ/df_x = wf * dV on interior cells); outputs
 * lus.. (adjoint of U*, overwritten), lx (adjoint of the pressure unknowns,
 * overwritten) and leta (accumulated) */
static void adj_update(const orc_grid *g, double dt, int zp, int convective, const double *x,
                       int use_x, const double *p, const double *eta, const double *us,
                       const double *vs, const double *ws, const double *lu3,
                       const double *lv3, const double *lw3, double wf, double *lus,
                       double *lvs, double *lws, double *lx, double *leta)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const long G = (long)nx * ny * nz;
    const long n = (long)g->nxc * g->nyc * g->nzc;
    memset(lus, 0, sizeof(double) * G);
    memset(lvs, 0, sizeof(double) * G);
    memset(lws, 0, sizeof(double) * G);
    memset(lx, 0, sizeof(double) * n);
    /* pressure at interior cells: X_result (solved step) or the ghosted p */
#define PV(I, J, K)                                                                            \
    (use_x ? x[((long)((I)-GC) * g->nyc + ((J)-GC)) * g->nzc + ((K)-GC)] : p[IC(g, I, J, K)])
#define LX(I, J, K) lx[((long)((I)-GC) * g->nyc + ((J)-GC)) * g->nzc + ((K)-GC)]
    for (int d = 0; d < 3; ++d) {
        const int hi_i = d == 0 ? nx - GC - 1 : nx - GC;
        const int hi_j = d == 1 ? ny - GC - 1 : ny - GC;
        const int hi_k = d == 2 ? (zp ? nz - GC : nz - GC - 1) : nz - GC;
        const double *S = d == 0 ? us : d == 1 ? vs : ws;
        const double *L3 = d == 0 ? lu3 : d == 1 ? lv3 : lw3;
        double *LS = d == 0 ? lus : d == 1 ? lvs : lws;
        const double *Dd = d == 0 ? g->Dx : d == 1 ? g->Dy : g->Dz;
        const double *Dsd = d == 0 ? g->Dxs : d == 1 ? g->Dys : g->Dzs;
        for (int i = GC; i < hi_i; ++i)
            for (int j = GC; j < hi_j; ++j)
                for (int k = GC; k < hi_k; ++k) {
                    int ip = i, jp = j, kp = k;
                    if (d == 0) ip = i + 1;
                    if (d == 1) jp = j + 1;
                    if (d == 2) kp = (zp && k == nz - GC - 1) ? GC : k + 1;
                    const int Id = d == 0 ? i : d == 1 ? j : k;
                    const long ic = IC(g, i, j, k), icP = IC(g, ip, jp, kp);
                    const double T2 = S[ic] - dt * (PV(ip, jp, kp) - PV(i, j, k)) / Dsd[Id];
                    const double IT = (Dsd[Id] * 0.5) / Dd[Id];
                    const double e = eta[ic] + (eta[icP] - eta[ic]) * IT;
                    /* u = (1 - e) T2, f = (u - T2) / dt = -e T2 / dt */
                    const double lf = (d == 0) ? wf * g->Dx[i] * g->Dy[j] * g->Dz[k] : 0.0;
                    const double lT2 = (1.0 - e) * L3[ic] - e / dt * lf;
                    const double le = -T2 * L3[ic] - T2 / dt * lf;
                    LS[ic] += lT2;
                    if (use_x) {
                        LX(ip, jp, kp) += -dt / Dsd[Id] * lT2;
                        LX(i, j, k) += dt / Dsd[Id] * lT2;
                    }
                    leta[ic] += (1.0 - IT) * le;
                    leta[icP] += IT * le;
                }
    }
#undef PV
#undef LX
    if (convective)
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k)
                lus[IC(g, nx - 3, j, k)] += lu3[IC(g, nx - 3, j, k)];
}

/* derivative of the smoothed mask value at plane cell (i, j) w.r.t. cx, cy, R */
static void eta_smoothed_grad(const orc_grid *g, int i, int j, double cx, double cy, double R,
                              double eps, double out[3])
{
    const double pi = 3.14159265358979323846;
    const double xc = (g->X[i] + g->X[i + 1]) * 0.5 - cx;
    const double yc = (g->Y[j] + g->Y[j + 1]) * 0.5 - cy;
    const double dist = sqrt(xc * xc + yc * yc);
    const double phi = R - dist;
    out[0] = out[1] = out[2] = 0.0;
    if (phi < -eps || phi > eps || dist == 0.0) return;
    const double h = 0.5 * (1.0 / eps + cos(pi * phi / eps) / eps); /* dH/dphi */
    out[0] = h * xc / dist; /* dphi/dcx = xc/dist */
    out[1] = h * yc / dist;
    out[2] = h;
}

double orc_adjoint_drag(const orc_grid *g, const orc_bc *bc, const orc_params *prm,
                        const double *u_init, const double *v_init, const double *w_init,
                        double cx, double cy, double R, double eps, int nsteps,
                        double adj_tol, double *grad_inflow, double grad_mask[3])
{
    const long G = (long)g->nx * g->ny * g->nz;
    const long n = (long)g->nxc * g->nyc * g->nzc;
    const int zp = bc->z_periodic;
    orc_params P = *prm;
    P.swap_buffers = 1;
    orc_sim *sim = orc_sim_new(g, bc, &P);
    double *eta = orc_sim_field(sim, 10);
    orc_eta_smoothed(g, cx, cy, R, eps, eta);
    memcpy(orc_sim_field(sim, 0), u_init, sizeof(double) * G);
    memcpy(orc_sim_field(sim, 1), v_init, sizeof(double) * G);
    memcpy(orc_sim_field(sim, 2), w_init, sizeof(double) * G);
    orc_sim_start(sim);
    /* tape: U^{n-1} (after BC) and the pressure used by step n */
    double *tapeU = (double *)malloc(sizeof(double) * 3 * G * (size_t)nsteps);
    double *tapeX = (double *)malloc(sizeof(double) * n * (size_t)nsteps);
    double *tapeP = (double *)malloc(sizeof(double) * G * (size_t)nsteps);
    int *solved = (int *)malloc(sizeof(int) * (size_t)nsteps);
    double J = 0.0;
    for (int s = 0; s < nsteps; ++s) {
        for (int c = 0; c < 3; ++c)
            memcpy(tapeU + (3 * (size_t)s + c) * G, orc_sim_field(sim, c), sizeof(double) * G);
        memcpy(tapeP + (size_t)s * G, orc_sim_field(sim, 9), sizeof(double) * G);
        double out[11];
        orc_sim_step(sim, out);
        J += out[0];
        solved[s] = (int)out[7];
        memcpy(tapeX + (size_t)s * n, orc_sim_field(sim, 12), sizeof(double) * n);
    }
    J /= nsteps;
    const double wf = (-2.0 / g->lz) / nsteps; /* dJ/d(f_x dV) */

    double *lu = (double *)calloc((size_t)G, sizeof(double));
    double *lv = (double *)calloc((size_t)G, sizeof(double));
    double *lw = (double *)calloc((size_t)G, sizeof(double));
    double *lus = (double *)calloc((size_t)G, sizeof(double));
    double *lvs = (double *)calloc((size_t)G, sizeof(double));
    double *lws = (double *)calloc((size_t)G, sizeof(double));
    double *us = (double *)calloc((size_t)G, sizeof(double));
    double *vs = (double *)calloc((size_t)G, sizeof(double));
    double *ws = (double *)calloc((size_t)G, sizeof(double));
    double *leta = (double *)calloc((size_t)G, sizeof(double));
    double *lx = (double *)calloc((size_t)n, sizeof(double));
    double *mu = (double *)calloc((size_t)n, sizeof(double));
    double *l_inflow = (double *)calloc((size_t)g->ny, sizeof(double));
    orc_bicg *adj = orc_bicg_new(n);
    const double nu = P.nu;
    for (int s = nsteps - 1; s >= 0; --s) {
        const double *U0 = tapeU + 3 * (size_t)s * G;
        /* S6^T: BC of this step */
        orc_adj_bc_main(g, bc, 1, P.dt, lu, lv, lw, l_inflow);
        /* recompute U* of this step from U^{n-1} (convdiff + the copies) */
        memcpy(us, U0, sizeof(double) * G);
        memcpy(vs, U0 + G, sizeof(double) * G);
        memcpy(ws, U0 + 2 * G, sizeof(double) * G);
        orc_convdiff(g, P.dt, nu, zp, U0, U0 + G, U0 + 2 * G, us, vs, ws);
        /* S4^T, S5^T */
        adj_update(g, P.dt, zp, bc->outflow_convective, tapeX + (size_t)s * n, solved[s],
                   tapeP + (size_t)s * G, eta, us, vs, ws, lu, lv, lw, wf, lus, lvs, lws, lx,
                   leta);
        if (solved[s]) {
            /* S3^T: A mu = lx (A symmetric), mean-free; S2^T */
            double res;
            orc_bicg_solve(adj, g, ORC_PRECOND_NONE, zp, ORC_RED_SERIAL, adj_tol, P.maxit, lx,
                           mu, &res);
            double mean = 0.0;
            for (long q = 0; q < n; ++q) mean += mu[q];
            mean /= (double)n;
            for (long q = 0; q < n; ++q) mu[q] -= mean;
            adj_rhs(g, P.dt, mu, lus, lvs, lws);
        }
        /* S1^T: periodic images of U*, the QUICK update, the copies */
        if (zp) {
            adj_periodic(g, lus, 0);
            adj_periodic(g, lvs, 0);
            adj_periodic(g, lws, 0);
        }
        memset(lu, 0, sizeof(double) * G);
        memset(lv, 0, sizeof(double) * G);
        memset(lw, 0, sizeof(double) * G);
        orc_adj_convdiff(g, P.dt, nu, zp, U0, U0 + G, U0 + 2 * G, lus, lvs, lws, lu, lv, lw);
        adj_bc_copy(g, zp, lus, lvs, lws, lu, lv, lw);
    }
    /* U^0 = BC_start(U_init) */
    orc_adj_bc_main(g, bc, 0, P.dt, lu, lv, lw, l_inflow);
    if (grad_inflow)
        memcpy(grad_inflow, l_inflow, sizeof(double) * g->ny);
    if (grad_mask) {
        grad_mask[0] = grad_mask[1] = grad_mask[2] = 0.0;
        for (int i = GC; i < g->nx - GC; ++i)
            for (int j = GC; j < g->ny - GC; ++j) {
                double le = 0.0;
                for (int k = 0; k < g->nz; ++k) le += leta[IC(g, i, j, k)];
                double d3[3];
                eta_smoothed_grad(g, i, j, cx, cy, R, eps, d3);
                for (int q = 0; q < 3; ++q) grad_mask[q] += le * d3[q];
            }
    }
    free(lu); free(lv); free(lw); free(lus); free(lvs); free(lws);
    free(us); free(vs); free(ws); free(leta); free(lx); free(mu); free(l_inflow);
    free(tapeU); free(tapeX); free(tapeP); free(solved);
    orc_bicg_free(adj);
    orc_sim_free(sim);
    return J;
}

/* J alone, by the same forward path (for finite differences) */
double orc_drag_objective(const orc_grid *g, const orc_bc *bc, const orc_params *prm,
                          const double *u_init, const double *v_init, const double *w_init,
                          double cx, double cy, double R, double eps, int nsteps)
{
    const long G = (long)g->nx * g->ny * g->nz;
    orc_params P = *prm;
    P.swap_buffers = 1;
    orc_sim *sim = orc_sim_new(g, bc, &P);
    orc_eta_smoothed(g, cx, cy, R, eps, orc_sim_field(sim, 10));
    memcpy(orc_sim_field(sim, 0), u_init, sizeof(double) * G);
    memcpy(orc_sim_field(sim, 1), v_init, sizeof(double) * G);
    memcpy(orc_sim_field(sim, 2), w_init, sizeof(double) * G);
    orc_sim_start(sim);
    double J = 0.0;
    for (int s = 0; s < nsteps; ++s) {
        double out[11];
        orc_sim_step(sim, out);
        J += out[0];
    }
    orc_sim_free(sim);
    return J / nsteps;
}
