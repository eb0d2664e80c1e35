/*
 * cfd_oracle.c -- CPU oracle (test infrastructure only; see cfd_oracle.h).
 *
 * Plain C restatement of the reference hot path.  Arithmetic follows the
 * reference expression by expression (same operands, same association, no FMA
 * contraction: build with -ffp-contract=off), so that at one OpenMP thread the
 * reference and this file agree bit for bit (checked against oracle/_ref).
 */
#include "cfd_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GC 2

static inline long IC(const orc_grid *g, long i, long j, long k)
{
    return (i * g->ny + j) * g->nz + k;
}
static inline long ICAL(const orc_grid *g, long i, long j, long k)
{
    return ((i - GC) * g->nyc + (j - GC)) * g->nzc + (k - GC);
}

static void periodic_z_fill(const orc_grid *g, double *a, int all_columns);

/* ------------------------------------------------------------------------ */
/* grid: init_uniform structureGrid.hpp:336-367; unit_SML :250-310;          */
/* init_nonuniform_SML :312-329; init_gC_dx :433-486                         */
/* ------------------------------------------------------------------------ */
static void gc_dx(double *X, double *D, double *Ds, int n)
{
    /* structureGrid.hpp:435-483, one axis */
    X[1] = 2 * X[2] - X[3];
    X[0] = 2 * X[1] - X[2];
    X[n - 1] = 2 * X[n - 2] - X[n - 3];
    X[n] = 2 * X[n - 1] - X[n - 2];
    for (int i = GC; i < n - GC - 1; ++i)
        Ds[i] = (X[i + 2] - X[i]) / 2.0;
    for (int i = GC; i < n - GC - 1; ++i)
        D[i] = (X[i + 1] - X[i]);
    D[0] = D[1] = D[2];
    Ds[0] = Ds[1] = Ds[2];
    D[n - 1] = D[n - 2] = D[n - 3] = X[n - 2] - X[n - 3];
    Ds[n - 1] = Ds[n - 2] = Ds[n - 3] = Ds[n - 4];
}

static int unit_sml(double *a, double lsml, int ns, int off_ns, double l, int n)
{
    /* structureGrid.hpp:250-305 */
    int ncal = n - 2 * GC;
    const double dS = lsml / ns, d = (l - lsml) / (ncal - ns);
    a[GC] = 0.0;
    if (ns != 0) {
        if ((n - ns) % 2 != 0)
            return -1;
        int nlEnd0 = (ncal - ns) / 2;
        int nsEnd0 = nlEnd0 + ns;
        int nlEnd1 = nsEnd0 + nlEnd0;
        nlEnd0 += off_ns;
        nsEnd0 += off_ns;
        if (nlEnd0 <= 0)
            return -2;
        for (int i = 0; i < nlEnd0; i++)
            a[i + GC + 1] = a[i + GC] + d;
        for (int i = nlEnd0; i < nsEnd0; ++i)
            a[i + GC + 1] = a[i + GC] + dS;
        for (int i = nsEnd0; i < nlEnd1; i++)
            a[i + GC + 1] = a[i + GC] + d;
    } else {
        for (int i = GC; i < n - GC; i++)
            a[i + 1] = a[i] + l / ncal;
    }
    return 0;
}

int orc_grid_build(orc_grid *g, int kind)
{
    g->nx = g->nxc + 2 * GC;
    g->ny = g->nyc + 2 * GC;
    g->nz = g->nzc + 2 * GC;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (kind == ORC_GRID_UNIFORM) {
        /* structureGrid.hpp:336-367 */
        double dx = g->lx / (nx - 2.0 * GC);
        double dy = g->ly / (ny - 2.0 * GC);
        double dz = g->lz / (nz - 2.0 * GC);
        g->X[0] = 0.0 - 2.0 * dx;
        for (int i = 1; i < nx + 1; ++i)
            g->X[i] = g->X[i - 1] + dx;
        g->Y[0] = 0.0 - 2.0 * dy;
        for (int j = 1; j < ny + 1; ++j)
            g->Y[j] = g->Y[j - 1] + dy;
        g->Z[0] = 0.0 - 2.0 * dz;
        for (int k = 1; k < nz + 1; ++k)
            g->Z[k] = g->Z[k - 1] + dz;
        for (int i = 0; i < nx; ++i) g->Dx[i] = g->Dxs[i] = dx;
        for (int j = 0; j < ny; ++j) g->Dy[j] = g->Dys[j] = dy;
        for (int k = 0; k < nz; ++k) g->Dz[k] = g->Dzs[k] = dz;
        return 0;
    }
    if (kind == ORC_GRID_SML) {
        /* structureGrid.hpp:312-329: 30% of x and y refined with 100 cells, z uniform;
         * every unit_SML call runs init_gC_dx on all three axes (:307), which is
         * idempotent once the interior nodes are final. */
        memset(g->X, 0, sizeof(double) * (nx + 1));
        memset(g->Y, 0, sizeof(double) * (ny + 1));
        memset(g->Z, 0, sizeof(double) * (nz + 1));
        const double lxS = 0.3 * g->lx, lyS = 0.3 * g->ly, lzS = 0.5 * g->lz;
        int rc = unit_sml(g->X, lxS, 100, -4, g->lx, nx);
        if (rc) return rc;
        rc = unit_sml(g->Y, lyS, 100, 0, g->ly, ny);
        if (rc) return rc;
        rc = unit_sml(g->Z, lzS, 0, 0, g->lz, nz);
        if (rc) return rc;
        gc_dx(g->X, g->Dx, g->Dxs, nx);
        gc_dx(g->Y, g->Dy, g->Dys, ny);
        gc_dx(g->Z, g->Dz, g->Dzs, nz);
        return 0;
    }
    return -10;
}

/* ------------------------------------------------------------------------ */
/* reductions                                                                */
/* ------------------------------------------------------------------------ */
double orc_serial_sum(const double *v, long n)
{
    double r = 0.0;
    for (long i = 0; i < n; ++i)
        r += v[i];
    return r;
}

static double tree256(double *a)
{
    /* per 64-lane wave: a[l] += a[l+s] for s = 32..1; then (w0 + w1) + (w2 + w3) */
    for (int w = 0; w < 4; ++w) {
        double *x = a + 64 * w;
        for (int s = 32; s >= 1; s >>= 1)
            for (int l = 0; l < s; ++l)
                x[l] = x[l] + x[l + s];
    }
    return (a[0] + a[64]) + (a[128] + a[192]);
}

double orc_blocked_sum(const double *v, long n)
{
    const long per = (long)ORC_RB_THREADS * ORC_RB_CELLS;
    const long nb = (n + per - 1) / per;
    double *part = (double *)malloc(sizeof(double) * (nb > 0 ? nb : 1));
    double acc[ORC_RB_THREADS];
    for (long b = 0; b < nb; ++b) {
        for (int t = 0; t < ORC_RB_THREADS; ++t) {
            double s = 0.0;
            for (int c = 0; c < ORC_RB_CELLS; ++c) {
                long m = b * per + (long)c * ORC_RB_THREADS + t;
                if (m < n)
                    s += v[m];
            }
            acc[t] = s;
        }
        part[b] = tree256(acc);
    }
    for (int t = 0; t < ORC_RB_THREADS; ++t) {
        double s = 0.0;
        for (long m = t; m < nb; m += ORC_RB_THREADS)
            s += part[m];
        acc[t] = s;
    }
    double r = tree256(acc);
    free(part);
    return r;
}

/* march order of the device's plane-marching stencil kernels: block b = chunk*ntile +
 * tile covers planes [8*chunk, 8*chunk+8) of the plane-index tile [1024*tile,
 * 1024*tile+1024), downwards for even chunks and upwards for odd ones; thread t
 * accumulates plane by plane the cells q = 1024*tile + 256c + t, c = 0..3; then the
 * same block fold and partial fold as orc_blocked_sum. */
double orc_march_sum(const orc_grid *g, const double *v)
{
    const long cyz = (long)g->nyc * g->nzc;
    const int ntile = (int)((cyz + ORC_MARCH_TILE - 1) / ORC_MARCH_TILE);
    const int nchunk = (g->nxc + ORC_MARCH_CHUNK - 1) / ORC_MARCH_CHUNK;
    const long nb = (long)ntile * nchunk;
    double *part = (double *)malloc(sizeof(double) * nb);
    double acc[ORC_RB_THREADS];
    for (long b = 0; b < nb; ++b) {
        const int tile = (int)(b % ntile), chunk = (int)(b / ntile);
        const int i0 = chunk * ORC_MARCH_CHUNK;
        const int i1 = i0 + ORC_MARCH_CHUNK < g->nxc ? i0 + ORC_MARCH_CHUNK : g->nxc;
        const long q0 = (long)tile * ORC_MARCH_TILE;
        for (int t = 0; t < ORC_RB_THREADS; ++t) {
            double s = 0.0;
            for (int step = 0; step < i1 - i0; ++step) {
                const int i = (chunk & 1) ? i0 + step : i1 - 1 - step;
                for (int c = 0; c < ORC_RB_CELLS; ++c) {
                    const long q = q0 + (long)c * ORC_RB_THREADS + t;
                    if (q < cyz)
                        s += v[(long)i * cyz + q];
                }
            }
            acc[t] = s;
        }
        part[b] = tree256(acc);
    }
    for (int t = 0; t < ORC_RB_THREADS; ++t) {
        double s = 0.0;
        for (long m = t; m < nb; m += ORC_RB_THREADS)
            s += part[m];
        acc[t] = s;
    }
    const double r = tree256(acc);
    free(part);
    return r;
}

static double reduce(int mode, const double *v, long n)
{
    return mode == ORC_RED_SERIAL ? orc_serial_sum(v, n) : orc_blocked_sum(v, n);
}

static double dot_mode(int mode, const double *a, const double *b, long n, double *tmp)
{
    if (mode == ORC_RED_SERIAL) {
        double r = 0.0; /* math.hpp:13-23 at one thread */
        for (long i = 0; i < n; ++i)
            r += a[i] * b[i];
        return r;
    }
    for (long i = 0; i < n; ++i)
        tmp[i] = a[i] * b[i];
    return orc_blocked_sum(tmp, n);
}

/* dot product in the march order (device) or index order (reference) */
static double dot_march(int mode, const orc_grid *g, const double *a, const double *b, long n,
                        double *tmp)
{
    if (mode == ORC_RED_SERIAL)
        return dot_mode(mode, a, b, n, tmp);
    for (long i = 0; i < n; ++i)
        tmp[i] = a[i] * b[i];
    return orc_march_sum(g, tmp);
}

/* ------------------------------------------------------------------------ */
/* QUICK: QuickUnitA/B Convection_Difussion.hpp:20-90                        */
/* ------------------------------------------------------------------------ */
static inline void quickA(double Up, double Un, double fpp, double fp, double fc, double fn,
                          double fnn, const double *Di, const double *Ds, int I,
                          double *P, double *N)
{
    if (Up > 0.0)
        *P = 0.5 * (fc + fp) - 0.125 * Di[I + 1] * Di[I + 1] / Ds[I] *
                                   ((fp - fc) / Di[I + 1] - (fc - fn) / Di[I]);
    else
        *P = 0.5 * (fc + fp) - 0.125 * Di[I + 1] * Di[I + 1] / Ds[I + 1] *
                                   ((fpp - fp) / Di[I + 2] - (fp - fc) / Di[I + 1]);
    if (Un > 0.0)
        *N = 0.5 * (fn + fc) - 0.125 * Di[I] * Di[I] / Ds[I - 1] *
                                   ((fc - fn) / Di[I] - (fn - fnn) / Di[I - 1]);
    else
        *N = 0.5 * (fn + fc) - 0.125 * Di[I] * Di[I] / Ds[I] *
                                   ((fp - fc) / Di[I + 1] - (fc - fn) / Di[I]);
}

static inline void quickB(double Up, double Un, double fpp, double fp, double fc, double fn,
                          double fnn, const double *Di, const double *Ds, int I,
                          double *P, double *N)
{
    if (Up > 0.0)
        *P = 0.5 * (fc + fp) - 0.125 * Ds[I] * Ds[I] / Di[I] *
                                   ((fp - fc) / Ds[I] - (fc - fn) / Ds[I - 1]);
    else
        *P = 0.5 * (fc + fp) - 0.125 * Ds[I] * Ds[I] / Di[I + 1] *
                                   ((fpp - fp) / Ds[I + 1] - (fp - fc) / Ds[I]);
    if (Un > 0.0)
        *N = 0.5 * (fn + fc) - 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I - 1] *
                                   ((fc - fn) / Ds[I - 1] - (fn - fnn) / Ds[I - 2]);
    else
        *N = 0.5 * (fn + fc) - 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I] *
                                   ((fp - fc) / Ds[I] - (fc - fn) / Ds[I - 1]);
}

/* ConvectionDifussion Convection_Difussion.hpp:174-450 (first-order time, QUICK) */
void orc_convdiff(const orc_grid *g, double dt, double nu, int z_periodic,
                  const double *u, const double *v, const double *w,
                  double *us, double *vs, double *ws)
{
    orc_convdiff_les(g, dt, nu, NULL, z_periodic, u, v, w, us, vs, ws);
}

/* with TERBULENCE_SMAGORINSKY the diffusion uses nu = simu.nu + Viseff[icelCal(i,j,k)] at
 * each face (Convection_Difussion.hpp:251-255,334-338,414-418) */
void orc_convdiff_les(const orc_grid *g, double dt, double nu0, const double *viseff,
                      int z_periodic, const double *u, const double *v, const double *w,
                      double *us, double *vs, double *ws)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const long sx = (long)ny * nz, sy = nz, sz = 1;
    const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz;
    const double *Dxs = g->Dxs, *Dys = g->Dys, *Dzs = g->Dzs;

    /* u: :198-274, faces i in [2, nx-3) */
    for (int i = GC; i < nx - GC - 1; ++i)
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const long ic = IC(g, i, j, k);
                const long xm = ic - sx, xp = ic + sx, ym = ic - sy, yp = ic + sy, zm = ic - sz,
                           zp = ic + sz;
                const double IT = (Dxs[i] * 0.5) / Dx[i];
                const double Up = 0.5 * (u[xp] + u[ic]);
                const double Un = 0.5 * (u[xm] + u[ic]);
                const double Vp = v[ic] + (v[xp] - v[ic]) * IT;
                const double Vn = v[ym] + (v[ym + xp - ic] - v[ym]) * IT;
                const double Wp = w[ic] + (w[xp] - w[ic]) * IT;
                const double Wn = w[zm] + (w[zm + xp - ic] - w[zm]) * IT;
                double uPx, uNx, uPy, uNy, uPz, uNz;
                quickA(Up, Un, u[ic + 2 * sx], u[xp], u[ic], u[xm], u[ic - 2 * sx], Dx, Dxs, i,
                       &uPx, &uNx);
                quickB(Vp, Vn, u[ic + 2 * sy], u[yp], u[ic], u[ym], u[ic - 2 * sy], Dy, Dys, j,
                       &uPy, &uNy);
                quickB(Wp, Wn, u[ic + 2 * sz], u[zp], u[ic], u[zm], u[ic - 2 * sz], Dz, Dzs, k,
                       &uPz, &uNz);
                const double convection = -dt * ((uPx * uPx - uNx * uNx) / Dxs[i] +
                                                 (uPy * Vp - uNy * Vn) / Dy[j] +
                                                 (uPz * Wp - uNz * Wn) / Dz[k]);
                const double nu = viseff ? nu0 + viseff[ICAL(g, i, j, k)] : nu0;
                const double difussion =
                    nu * dt *
                    (((u[xp] - u[ic]) / Dx[i + 1] - (u[ic] - u[xm]) / Dx[i]) / Dxs[i] +
                     ((u[yp] - u[ic]) / Dys[j] - (u[ic] - u[ym]) / Dys[j - 1]) / Dy[j] +
                     ((u[zp] - u[ic]) / Dzs[k] - (u[ic] - u[zm]) / Dzs[k - 1]) / Dz[k]);
                us[ic] = u[ic] + convection + difussion;
            }

    /* v: :278-355, faces j in [2, ny-3) */
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC - 1; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const long ic = IC(g, i, j, k);
                const long xm = ic - sx, xp = ic + sx, ym = ic - sy, yp = ic + sy, zm = ic - sz,
                           zp = ic + sz;
                const double IT = (Dys[j] * 0.5) / Dy[j];
                const double Vp = 0.5 * (v[yp] + v[ic]);
                const double Vn = 0.5 * (v[ym] + v[ic]);
                const double Up = u[ic] + (u[yp] - u[ic]) * IT;
                const double Un = u[xm] + (u[xm + yp - ic] - u[xm]) * IT;
                const double Wp = w[ic] + (w[yp] - w[ic]) * IT;
                const double Wn = w[zm] + (w[zm + yp - ic] - w[zm]) * IT;
                double vPx, vNx, vPy, vNy, vPz, vNz;
                quickB(Up, Un, v[ic + 2 * sx], v[xp], v[ic], v[xm], v[ic - 2 * sx], Dx, Dxs, i,
                       &vPx, &vNx);
                quickA(Vp, Vn, v[ic + 2 * sy], v[yp], v[ic], v[ym], v[ic - 2 * sy], Dy, Dys, j,
                       &vPy, &vNy);
                quickB(Wp, Wn, v[ic + 2 * sz], v[zp], v[ic], v[zm], v[ic - 2 * sz], Dz, Dzs, k,
                       &vPz, &vNz);
                const double convection = -dt * ((Up * vPx - Un * vNx) / Dx[i] +
                                                 (vPy * vPy - vNy * vNy) / Dys[j] +
                                                 (Wp * vPz - Wn * vNz) / Dz[k]);
                const double nu = viseff ? nu0 + viseff[ICAL(g, i, j, k)] : nu0;
                const double difussion =
                    nu * dt *
                    (((v[xp] - v[ic]) / Dxs[i] - (v[ic] - v[xm]) / Dxs[i - 1]) / Dx[i] +
                     ((v[yp] - v[ic]) / Dy[j + 1] - (v[ic] - v[ym]) / Dy[j]) / Dys[j] +
                     ((v[zp] - v[ic]) / Dzs[k] - (v[ic] - v[zm]) / Dzs[k - 1]) / Dz[k]);
                vs[ic] = v[ic] + convection + difussion;
            }

    /* w: :361-438, faces k in [2, nz-3); periodic z also computes face nz-3 */
    const int kend = z_periodic ? nz - GC : nz - GC - 1;
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < kend; ++k) {
                const long ic = IC(g, i, j, k);
                const long xm = ic - sx, xp = ic + sx, ym = ic - sy, yp = ic + sy, zm = ic - sz,
                           zp = ic + sz;
                const double IT = (Dzs[k] * 0.5) / Dz[k];
                const double Wp = 0.5 * (w[zp] + w[ic]);
                const double Wn = 0.5 * (w[zm] + w[ic]);
                const double Up = u[ic] + (u[zp] - u[ic]) * IT;
                const double Un = u[xm] + (u[xm + zp - ic] - u[xm]) * IT;
                const double Vp = v[ic] + (v[zp] - v[ic]) * IT;
                const double Vn = v[ym] + (v[ym + zp - ic] - v[ym]) * IT;
                double wPx, wNx, wPy, wNy, wPz, wNz;
                quickB(Up, Un, w[ic + 2 * sx], w[xp], w[ic], w[xm], w[ic - 2 * sx], Dx, Dxs, i,
                       &wPx, &wNx);
                quickB(Vp, Vn, w[ic + 2 * sy], w[yp], w[ic], w[ym], w[ic - 2 * sy], Dy, Dys, j,
                       &wPy, &wNy);
                quickA(Wp, Wn, w[ic + 2 * sz], w[zp], w[ic], w[zm], w[ic - 2 * sz], Dz, Dzs, k,
                       &wPz, &wNz);
                const double convection = -dt * ((Up * wPx - Un * wNx) / Dx[i] +
                                                 (Vp * wPy - Vn * wNy) / Dy[j] +
                                                 (wPz * wPz - wNz * wNz) / Dzs[k]);
                const double nu = viseff ? nu0 + viseff[ICAL(g, i, j, k)] : nu0;
                const double difussion =
                    nu * dt *
                    (((w[xp] - w[ic]) / Dxs[i] - (w[ic] - w[xm]) / Dxs[i - 1]) / Dx[i] +
                     ((w[yp] - w[ic]) / Dys[j] - (w[ic] - w[ym]) / Dys[j - 1]) / Dy[j] +
                     ((w[zp] - w[ic]) / Dz[k + 1] - (w[ic] - w[zm]) / Dz[k]) / Dzs[k]);
                ws[ic] = w[ic] + convection + difussion;
            }
    if (z_periodic) {
        /* u* z-ghosts are images of the new interior (the divergence at k = 2 reads
         * the face k = 1 == nz-3); no reference counterpart */
        periodic_z_fill(g, us, 0);
        periodic_z_fill(g, vs, 0);
        periodic_z_fill(g, ws, 0);
    }
}

/* SmagorinskyModel.hpp:30-79 with GradientGauss :6-27.  The reference's source does not
 * compile (`auot`, :61) and is disabled (TERBULENCE_SMAGORINSKY undefined); this restates
 * its formulas.  Per interior cell: central differences 0.5/Dm * (phi[m] - phi[p]) of u
 * over (Dxs, Dy, Dz), v over (Dx, Dys, Dz) and w over (Dx, Dy, Dzs); temp = |2 ux^2 + vy^2
 * + wz^2 + (uy + vx)^2 + (uz + wx)^2 + (vz + wy)^2| (the factor 2 on the first term only,
 * as written); Viseff = (Cs delta)^2 sqrt(temp), delta = pow(Dx Dy Dz, 1/3). */
void orc_smagorinsky_coef(const orc_grid *g, double cs, double *coef)
{
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k) {
                const double delta = pow(g->Dx[i] * g->Dy[j] * g->Dz[k], 1.0 / 3.0);
                const double cd = cs * delta;
                coef[ICAL(g, i, j, k)] = cd * cd;
            }
}

void orc_smagorinsky(const orc_grid *g, const double *coef, const double *u, const double *v,
                     const double *w, double *viseff)
{
    const long sx = (long)g->ny * g->nz, sy = g->nz;
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k) {
                const long c = IC(g, i, j, k);
                const double ux = 0.5 / g->Dxs[i] * (u[c - sx] - u[c + sx]);
                const double uy = 0.5 / g->Dy[j] * (u[c - sy] - u[c + sy]);
                const double uz = 0.5 / g->Dz[k] * (u[c - 1] - u[c + 1]);
                const double vx = 0.5 / g->Dx[i] * (v[c - sx] - v[c + sx]);
                const double vy = 0.5 / g->Dys[j] * (v[c - sy] - v[c + sy]);
                const double vz = 0.5 / g->Dz[k] * (v[c - 1] - v[c + 1]);
                const double wx = 0.5 / g->Dx[i] * (w[c - sx] - w[c + sx]);
                const double wy = 0.5 / g->Dy[j] * (w[c - sy] - w[c + sy]);
                const double wz = 0.5 / g->Dzs[k] * (w[c - 1] - w[c + 1]);
                const double a = uy + vx, b = uz + wx, d = vz + wy;
                const double temp = fabs(2.0 * (ux * ux) + vy * vy + wz * wz + a * a + b * b + d * d);
                const long n = ICAL(g, i, j, k);
                viseff[n] = coef[n] * sqrt(temp);
            }
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions: BC_staggered_main Boundary_Condition.hpp:37-176      */
/* (faces applied in order 0..5, then p), FLOWpassing table FLOWpassing.hpp. */
/* Periodic z and convective outflow have no reference counterpart.         */
/* ------------------------------------------------------------------------ */
/* periodic images along z: ghost k <- k + nzc (k < 2), k - nzc (k >= nz-2).  The
 * boundary-condition fill covers every (i,j) column, so x/y ghost columns also get
 * images (the w face k = nz-3, computed when z is periodic, reads them). */
static void periodic_z_fill(const orc_grid *g, double *a, int all_columns)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nzc = g->nzc;
    const int lo = all_columns ? 0 : GC;
    for (int i = lo; i < nx - lo; ++i)
        for (int j = lo; j < ny - lo; ++j) {
            const long c = IC(g, i, j, 0);
            a[c + 1] = a[c + 1 + nzc];
            a[c + 0] = a[c + 0 + nzc];
            a[c + nz - 2] = a[c + nz - 2 - nzc];
            a[c + nz - 1] = a[c + nz - 1 - nzc];
        }
}

void orc_bc_main(const orc_grid *g, const orc_bc *bc, int advance_outflow, double dt,
                 double *u, double *v, double *w, double *p)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double(*Num)[3] = bc->num;
    const double(*Dir)[3] = bc->dir;
    /* face 0: :72-80 */
    for (int j = GC; j < ny - GC; ++j)
        for (int k = GC; k < nz - GC; ++k) {
            const int b = 0;
            const double d0 = bc->inflow ? bc->inflow[j] : Dir[b][0];
            u[IC(g, 0, j, k)] = u[IC(g, 1, j, k)] = d0 + Num[b][0] * u[IC(g, 2, j, k)];
            v[IC(g, 0, j, k)] = v[IC(g, 1, j, k)] =
                Dir[b][1] * 2. + (2.0 * Num[b][1] - 1.0) * v[IC(g, 2, j, k)];
            w[IC(g, 0, j, k)] = w[IC(g, 1, j, k)] =
                Dir[b][2] * 2. + (2.0 * Num[b][2] - 1.0) * w[IC(g, 2, j, k)];
        }
    /* face 1: :84-94 (u optionally convective) */
    for (int j = GC; j < ny - GC; ++j)
        for (int k = GC; k < nz - GC; ++k) {
            const int b = 1;
            double ub;
            if (bc->outflow_convective && advance_outflow) {
                const double uo = u[IC(g, nx - 3, j, k)];
                const double c = bc->conv_speed * dt / g->Dx[nx - 3];
                ub = uo - c * (uo - u[IC(g, nx - 4, j, k)]);
            } else {
                ub = Dir[b][0] + Num[b][0] * u[IC(g, nx - 4, j, k)];
            }
            u[IC(g, nx - 1, j, k)] = u[IC(g, nx - 2, j, k)] = u[IC(g, nx - 3, j, k)] = ub;
            v[IC(g, nx - 1, j, k)] = v[IC(g, nx - 2, j, k)] =
                Dir[b][1] * 2. + (2.0 * Num[b][1] - 1.0) * v[IC(g, nx - 3, j, k)];
            w[IC(g, nx - 1, j, k)] = w[IC(g, nx - 2, j, k)] =
                Dir[b][2] * 2. + (2.0 * Num[b][2] - 1.0) * w[IC(g, nx - 3, j, k)];
        }
    if (bc->flux_correction) {
        /* outlet flux shifted to the inlet flux (no reference counterpart): sums of
         * u*Dy*Dz over the interior (j,k), m = (j-2)*nzc + (k-2), in the device's
         * canonical blocked order */
        const long m2 = (long)g->nyc * g->nzc;
        double *t = (double *)calloc((size_t)(3 * m2), sizeof(double));
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const long m = (long)(j - GC) * g->nzc + (k - GC);
                const double a = g->Dy[j] * g->Dz[k];
                t[m] = u[IC(g, 1, j, k)] * a;
                t[m2 + m] = u[IC(g, nx - 3, j, k)] * a;
                t[2 * m2 + m] = a;
            }
        const double qin = orc_blocked_sum(t, m2), qout = orc_blocked_sum(t + m2, m2),
                     area = orc_blocked_sum(t + 2 * m2, m2);
        free(t);
        const double corr = (qin - qout) / area;
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const double ub = u[IC(g, nx - 3, j, k)] + corr;
                u[IC(g, nx - 1, j, k)] = u[IC(g, nx - 2, j, k)] = u[IC(g, nx - 3, j, k)] = ub;
            }
    }
    /* face 2: :97-105 */
    for (int i = GC; i < nx - GC; ++i)
        for (int k = GC; k < nz - GC; ++k) {
            const int b = 2;
            u[IC(g, i, 0, k)] = u[IC(g, i, 1, k)] =
                Dir[b][0] * 2. + (2.0 * Num[b][0] - 1.0) * u[IC(g, i, 2, k)];
            v[IC(g, i, 0, k)] = v[IC(g, i, 1, k)] = Dir[b][1] + Num[b][1] * v[IC(g, i, 2, k)];
            w[IC(g, i, 0, k)] = w[IC(g, i, 1, k)] =
                Dir[b][2] * 2. + (2.0 * Num[b][2] - 1.0) * w[IC(g, i, 2, k)];
        }
    /* face 3: :107-117 */
    for (int i = GC; i < nx - GC; ++i)
        for (int k = GC; k < nz - GC; ++k) {
            const int b = 3;
            v[IC(g, i, ny - 1, k)] = v[IC(g, i, ny - 2, k)] = v[IC(g, i, ny - 3, k)] =
                Dir[b][1] + Num[b][1] * v[IC(g, i, ny - 4, k)];
            u[IC(g, i, ny - 1, k)] = u[IC(g, i, ny - 2, k)] =
                Dir[b][0] * 2. + (2.0 * Num[b][0] - 1.0) * u[IC(g, i, ny - 3, k)];
            w[IC(g, i, ny - 1, k)] = w[IC(g, i, ny - 2, k)] =
                Dir[b][2] * 2. + (2.0 * Num[b][2] - 1.0) * w[IC(g, i, ny - 3, k)];
        }
    if (bc->z_periodic) {
        periodic_z_fill(g, u, 1);
        periodic_z_fill(g, v, 1);
        periodic_z_fill(g, w, 1);
    } else {
        /* face 4: :121-129 */
        for (int i = GC; i < nx - GC; ++i)
            for (int j = GC; j < ny - GC; ++j) {
                const int b = 4;
                u[IC(g, i, j, 0)] = u[IC(g, i, j, 1)] =
                    Dir[b][0] * 2. + (2.0 * Num[b][0] - 1.0) * u[IC(g, i, j, 2)];
                v[IC(g, i, j, 0)] = v[IC(g, i, j, 1)] =
                    Dir[b][1] * 2. + (2.0 * Num[b][1] - 1.0) * v[IC(g, i, j, 2)];
                w[IC(g, i, j, 0)] = w[IC(g, i, j, 1)] = Dir[b][2] + Num[b][2] * w[IC(g, i, j, 2)];
            }
        /* face 5: :133-143 */
        for (int i = GC; i < nx - GC; ++i)
            for (int j = GC; j < ny - GC; ++j) {
                const int b = 5;
                w[IC(g, i, j, nz - 1)] = w[IC(g, i, j, nz - 2)] = w[IC(g, i, j, nz - 3)] =
                    Dir[b][2] + Num[b][2] * w[IC(g, i, j, nz - 4)];
                u[IC(g, i, j, nz - 1)] = u[IC(g, i, j, nz - 2)] =
                    Dir[b][0] * 2. + (2.0 * Num[b][0] - 1.0) * u[IC(g, i, j, nz - 3)];
                v[IC(g, i, j, nz - 1)] = v[IC(g, i, j, nz - 2)] =
                    Dir[b][1] * 2. + (2.0 * Num[b][1] - 1.0) * v[IC(g, i, j, nz - 3)];
            }
    }
    /* pressure, Neumann copy: :149-172 */
    for (int j = GC; j < ny - GC; ++j)
        for (int k = GC; k < nz - GC; ++k) {
            p[IC(g, 0, j, k)] = p[IC(g, 1, j, k)] = p[IC(g, 2, j, k)];
            p[IC(g, nx - 1, j, k)] = p[IC(g, nx - 2, j, k)] = p[IC(g, nx - 3, j, k)];
        }
    for (int i = GC; i < nx - GC; ++i)
        for (int k = GC; k < nz - GC; ++k) {
            p[IC(g, i, 0, k)] = p[IC(g, i, 1, k)] = p[IC(g, i, 2, k)];
            p[IC(g, i, ny - 1, k)] = p[IC(g, i, ny - 2, k)] = p[IC(g, i, ny - 3, k)];
        }
    if (bc->z_periodic) {
        periodic_z_fill(g, p, 1);
    } else {
        for (int i = GC; i < nx - GC; ++i)
            for (int j = GC; j < ny - GC; ++j) {
                p[IC(g, i, j, 0)] = p[IC(g, i, j, 1)] = p[IC(g, i, j, 2)];
                p[IC(g, i, j, nz - 1)] = p[IC(g, i, j, nz - 2)] = p[IC(g, i, j, nz - 3)];
            }
    }
}

/* BC_staggered_copy updateBC.hpp:6-110: ghosts and the outlet/top/back faces T0 -> T1 */
void orc_bc_copy(const orc_grid *g, const double *u0, const double *v0, const double *w0,
                 double *u1, double *v1, double *w1)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int j = GC; j < ny - GC; ++j)
        for (int k = GC; k < nz - GC; ++k) {
            const int xs[2] = {0, 1};
            for (int a = 0; a < 2; ++a) {
                const long c = IC(g, xs[a], j, k);
                u1[c] = u0[c];
                v1[c] = v0[c];
                w1[c] = w0[c];
            }
            const int xe[3] = {nx - 1, nx - 3, nx - 2};
            for (int a = 0; a < 3; ++a) {
                const long c = IC(g, xe[a], j, k);
                u1[c] = u0[c];
                if (a != 1) {
                    v1[c] = v0[c];
                    w1[c] = w0[c];
                }
            }
        }
    for (int i = GC; i < nx - GC; ++i)
        for (int k = GC; k < nz - GC; ++k) {
            for (int a = 0; a < 2; ++a) {
                const long c = IC(g, i, a, k);
                u1[c] = u0[c];
                v1[c] = v0[c];
                w1[c] = w0[c];
            }
            for (int jj = ny - 2; jj < ny; ++jj) {
                const long c = IC(g, i, jj, k);
                u1[c] = u0[c];
                w1[c] = w0[c];
            }
            for (int jj = ny - 3; jj < ny; ++jj) {
                const long c = IC(g, i, jj, k);
                v1[c] = v0[c];
            }
        }
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j) {
            for (int a = 0; a < 2; ++a) {
                const long c = IC(g, i, j, a);
                u1[c] = u0[c];
                v1[c] = v0[c];
                w1[c] = w0[c];
            }
            for (int kk = nz - 2; kk < nz; ++kk) {
                const long c = IC(g, i, j, kk);
                u1[c] = u0[c];
                v1[c] = v0[c];
            }
            for (int kk = nz - 3; kk < nz; ++kk) {
                const long c = IC(g, i, j, kk);
                w1[c] = w0[c];
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Poisson right-hand side: createBMatrix_omp b_matrix.hpp:35-98             */
/* ------------------------------------------------------------------------ */
static double jacobi_diag(const orc_grid *g, int z_periodic, int i, int j, int k);

void orc_rhs(const orc_grid *g, double dt, int precond, int z_periodic,
             const double *us, const double *vs, const double *ws, double *b)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double ddt = 1.0 / dt;
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const long ic = IC(g, i, j, k);
                const long row = ICAL(g, i, j, k);
                const double div = (us[ic] - us[IC(g, i - 1, j, k)]) / g->Dx[i] +
                                   (vs[ic] - vs[IC(g, i, j - 1, k)]) / g->Dy[j] +
                                   (ws[ic] - ws[IC(g, i, j, k - 1)]) / g->Dz[k];
                if (precond == ORC_PRECOND_JACOBI) {
                    const double jp = 1.0 / jacobi_diag(g, z_periodic, i - GC, j - GC, k - GC);
                    b[row] = -jp * ddt * div;
                } else {
                    b[row] = -ddt * div;
                }
            }
}

/* ------------------------------------------------------------------------ */
/* Pressure operators                                                        */
/* ------------------------------------------------------------------------ */
/* coefficient lambdas, SPE_sparseMatrix.hpp:52-58 and structureGrid.hpp:560-566
 * (index shifted by gC) */
static inline double cxp(const orc_grid *g, int i) { return 1.0 / g->Dx[i + GC] / g->Dxs[i + GC]; }
static inline double cyp(const orc_grid *g, int j) { return 1.0 / g->Dy[j + GC] / g->Dys[j + GC]; }
static inline double czp(const orc_grid *g, int k) { return 1.0 / g->Dz[k + GC] / g->Dzs[k + GC]; }
static inline double cxm(const orc_grid *g, int i) { return 1.0 / g->Dx[i + GC] / g->Dxs[i + GC - 1]; }
static inline double cym(const orc_grid *g, int j) { return 1.0 / g->Dy[j + GC] / g->Dys[j + GC - 1]; }
static inline double czm(const orc_grid *g, int k) { return 1.0 / g->Dz[k + GC] / g->Dzs[k + GC - 1]; }

/* SPE_matrix::setupPressure diagonal, SPE_sparseMatrix.hpp:68-76 */
static double spe_diag(const orc_grid *g, int z_periodic, int i, int j, int k)
{
    double ijk = cxm(g, i) + cym(g, j) + czm(g, k) + cxp(g, i) + cyp(g, j) + czp(g, k);
    if (i == 0) ijk -= cxm(g, i);
    if (i == g->nxc - 1) ijk -= cxp(g, i);
    if (j == 0) ijk -= cym(g, j);
    if (j == g->nyc - 1) ijk -= cyp(g, j);
    if (!z_periodic) {
        if (k == 0) ijk -= czm(g, k);
        if (k == g->nzc - 1) ijk -= czp(g, k);
    }
    return ijk;
}

/* createPressureMatrixCSR diagonal before scaling, structureGrid.hpp:590-601 */
static double jacobi_diag(const orc_grid *g, int z_periodic, int i, int j, int k)
{
    double aa = 0.0, bb = 0.0, cc = 0.0;
    if (i == 0) aa = -cxm(g, i);
    if (i == g->nxc - 1) aa = -cxp(g, i);
    if (j == 0) bb = -cym(g, j);
    if (j == g->nyc - 1) bb = -cyp(g, j);
    if (!z_periodic) {
        if (k == 0) cc = -czm(g, k);
        if (k == g->nzc - 1) cc = -czp(g, k);
    }
    return cxm(g, i) + cym(g, j) + czm(g, k) + cxp(g, i) + cyp(g, j) + czp(g, k) + aa + bb + cc;
}

/* SPE_matrix::setupPressure value table, SPE_sparseMatrix.hpp:42-88 (val_ is zero-filled) */
void orc_spe_build(const orc_grid *g, double *val)
{
    const int nxc = g->nxc, nyc = g->nyc, nzc = g->nzc;
    memset(val, 0, sizeof(double) * 7 * (long)nxc * nyc * nzc);
    for (int i = 0; i < nxc; ++i)
        for (int j = 0; j < nyc; ++j)
            for (int k = 0; k < nzc; ++k) {
                const long ic7 = (((long)i * nyc + j) * nzc + k) * 7;
                if (i != 0) val[ic7 + 0] = -cxm(g, i);
                if (j != 0) val[ic7 + 1] = -cyp(g, j); /* sic: +y coefficient in the -y slot */
                if (k != 0) val[ic7 + 2] = -czm(g, k);
                val[ic7 + 3] = spe_diag(g, 0, i, j, k);
                if (k != nzc - 1) val[ic7 + 4] = -czp(g, k);
                if (j != nyc - 1) val[ic7 + 5] = -cym(g, j); /* sic */
                if (i != nxc - 1) val[ic7 + 6] = -cxp(g, i);
            }
}

/* SPE_matrix::multiply_omp SPE_sparseMatrix.hpp:92-106 with the stored table */
void orc_spe_apply_vals(const orc_grid *g, const double *val, const double *x, double *r)
{
    const long nz = g->nzc, nynz = (long)g->nyc * g->nzc, n = nynz * g->nxc;
    for (long i = 0; i < n; ++i) {
        double sum = 0.0;
        const long i7 = i * 7;
        if (i - nynz >= 0) sum += val[i7 + 0] * x[i - nynz];
        if (i - nz >= 0) sum += val[i7 + 1] * x[i - nz];
        if (i - 1 >= 0) sum += val[i7 + 2] * x[i - 1];
        sum += val[i7 + 3] * x[i];
        if (i + 1 < n) sum += val[i7 + 4] * x[i + 1];
        if (i + nz < n) sum += val[i7 + 5] * x[i + nz];
        if (i + nynz < n) sum += val[i7 + 6] * x[i + nynz];
        r[i] = sum;
    }
}

/* Matrix-free application.  precond == NONE: the SPE operator (slot order and
 * boundary folding of SPE_sparseMatrix.hpp:42-106; zero-valued boundary slots
 * contribute nothing).  precond == JACOBI: the row-scaled CSR operator of
 * createPressureMatrixCSR (structureGrid.hpp:557-654) applied by
 * CSR_matrix::multiply_omp (CSR_sparseMatrix.hpp:231-245).  z_periodic wraps the
 * z coupling (no reference counterpart). */
void orc_apply_A(const orc_grid *g, int precond, int z_periodic, const double *x, double *r)
{
    const int nxc = g->nxc, nyc = g->nyc, nzc = g->nzc;
    const long sy = nzc, sx = (long)nyc * nzc;
    for (int i = 0; i < nxc; ++i)
        for (int j = 0; j < nyc; ++j)
            for (int k = 0; k < nzc; ++k) {
                const long n = ((long)i * nyc + j) * nzc + k;
                const int kfirst = (k == 0), klast = (k == nzc - 1);
                const long zm_n = kfirst ? n + (nzc - 1) : n - 1;
                const long zp_n = klast ? n - (nzc - 1) : n + 1;
                const int has_zm = !kfirst || z_periodic;
                const int has_zp = !klast || z_periodic;
                if (precond == ORC_PRECOND_JACOBI) {
                    const double jp = 1.0 / jacobi_diag(g, z_periodic, i, j, k);
                    const double ijk = jacobi_diag(g, z_periodic, i, j, k);
                    double s = 0.0;
                    if (i != 0) s += (-cxm(g, i) * jp) * x[n - sx];
                    if (j != 0) s += (-cym(g, j) * jp) * x[n - sy];
                    if (has_zm) s += (-czm(g, k) * jp) * x[zm_n];
                    s += (ijk * jp) * x[n];
                    if (has_zp) s += (-czp(g, k) * jp) * x[zp_n];
                    if (j != nyc - 1) s += (-cyp(g, j) * jp) * x[n + sy];
                    if (i != nxc - 1) s += (-cxp(g, i) * jp) * x[n + sx];
                    r[n] = s;
                } else {
                    double s = 0.0;
                    if (i != 0) s += (-cxm(g, i)) * x[n - sx];
                    if (j != 0) s += (-cyp(g, j)) * x[n - sy];
                    if (has_zm) s += (-czm(g, k)) * x[zm_n];
                    s += spe_diag(g, z_periodic, i, j, k) * x[n];
                    if (has_zp) s += (-czp(g, k)) * x[zp_n];
                    if (j != nyc - 1) s += (-cym(g, j)) * x[n + sy];
                    if (i != nxc - 1) s += (-cxp(g, i)) * x[n + sx];
                    r[n] = s;
                }
            }
}

/* The seven slot values (-x, -y, -z, diag, +z, +y, +x) of row (i, j, k) of the operator
 * orc_apply_A applies, each formed by the expression it uses there (a slot without a
 * neighbour gets its value anyway; callers skip it). */
static void row_vals(const orc_grid *g, int precond, int z_periodic, int i, int j, int k,
                     double v[7])
{
    if (precond == ORC_PRECOND_JACOBI) {
        const double jp = 1.0 / jacobi_diag(g, z_periodic, i, j, k);
        const double ijk = jacobi_diag(g, z_periodic, i, j, k);
        v[0] = -cxm(g, i) * jp;
        v[1] = -cym(g, j) * jp;
        v[2] = -czm(g, k) * jp;
        v[3] = ijk * jp;
        v[4] = -czp(g, k) * jp;
        v[5] = -cyp(g, j) * jp;
        v[6] = -cxp(g, i) * jp;
    } else {
        v[0] = -cxm(g, i);
        v[1] = -cyp(g, j); /* the y-slot swap, as in orc_spe_build */
        v[2] = -czm(g, k);
        v[3] = spe_diag(g, z_periodic, i, j, k);
        v[4] = -czp(g, k);
        v[5] = -cym(g, j);
        v[6] = -cxp(g, i);
    }
}

/* The transpose of orc_apply_A: r = A^T y (no reference counterpart).  The adjoint of the
 * pressure solve needs it wherever A is not symmetric: the SPE operator on the SML grid
 * (stretched widths, and the y-slot swap) and the row-scaled CSR operator.  Every coupling
 * is mutual in position: row n has a neighbour in slot q exactly when that row has n in
 * slot 6 - q (also across a periodic z wrap).  So column n takes, in a row's slot order,
 * slot 6 - q of the row in its slot q, starting from +0.0 and skipping absent slots. */
void orc_apply_AT(const orc_grid *g, int precond, int z_periodic, const double *y, double *r)
{
    const int nxc = g->nxc, nyc = g->nyc, nzc = g->nzc;
    const long sy = nzc, sx = (long)nyc * nzc;
    double v[7];
    for (int i = 0; i < nxc; ++i)
        for (int j = 0; j < nyc; ++j)
            for (int k = 0; k < nzc; ++k) {
                const long n = ((long)i * nyc + j) * nzc + k;
                const int kfirst = (k == 0), klast = (k == nzc - 1);
                const int km = kfirst ? nzc - 1 : k - 1, kp = klast ? 0 : k + 1;
                double s = 0.0;
                if (i != 0) {
                    row_vals(g, precond, z_periodic, i - 1, j, k, v);
                    s += v[6] * y[n - sx];
                }
                if (j != 0) {
                    row_vals(g, precond, z_periodic, i, j - 1, k, v);
                    s += v[5] * y[n - sy];
                }
                if (!kfirst || z_periodic) {
                    row_vals(g, precond, z_periodic, i, j, km, v);
                    s += v[4] * y[n + (km - k)];
                }
                row_vals(g, precond, z_periodic, i, j, k, v);
                s += v[3] * y[n];
                if (!klast || z_periodic) {
                    row_vals(g, precond, z_periodic, i, j, kp, v);
                    s += v[2] * y[n + (kp - k)];
                }
                if (j != nyc - 1) {
                    row_vals(g, precond, z_periodic, i, j + 1, k, v);
                    s += v[1] * y[n + sy];
                }
                if (i != nxc - 1) {
                    row_vals(g, precond, z_periodic, i + 1, j, k, v);
                    s += v[0] * y[n + sx];
                }
                r[n] = s;
            }
}

/* ------------------------------------------------------------------------ */
/* BiCGSTAB: solver::bicgstabRe2 bicgstabRe2.hpp:17-180                      */
/* ------------------------------------------------------------------------ */
struct orc_bicg {
    long n;
    double *r0, *r1, *p1, *s1, *ap, *as, *px1, *px2, *tmp;
    int timestep, iter_total, reset_members, resets, restarts;
};

orc_bicg *orc_bicg_new(long n)
{
    orc_bicg *s = (orc_bicg *)calloc(1, sizeof(orc_bicg));
    s->n = n;
    double **v[] = {&s->r0, &s->r1, &s->p1, &s->s1, &s->ap, &s->as, &s->px1, &s->px2, &s->tmp};
    for (unsigned q = 0; q < sizeof(v) / sizeof(v[0]); ++q)
        *v[q] = (double *)calloc((size_t)n, sizeof(double));
    return s;
}

void orc_bicg_free(orc_bicg *s)
{
    if (!s) return;
    free(s->r0); free(s->r1); free(s->p1); free(s->s1); free(s->ap); free(s->as);
    free(s->px1); free(s->px2); free(s->tmp);
    free(s);
}

int orc_bicg_timestep(const orc_bicg *s) { return s->timestep; }

static double l2norm_mode(int mode, const double *a, long n, double *tmp)
{
    return sqrt(dot_mode(mode, a, a, n, tmp)); /* math.hpp:27-36 */
}

int orc_bicg_solve(orc_bicg *S, const orc_grid *g, int precond, int zp, int mode, double tol,
                   int maxit, const double *rhs, double *x, double *res_out)
{
    const long n = S->n;
    const double restart_factor = 1e-7;
    double r1r0 = 0, r1r0l = 0, a = 0, w = 0, b = 0;
    S->timestep++;
    const double norm0 = l2norm_mode(mode, rhs, n, S->tmp);
    orc_apply_A(g, precond, zp, x, S->r0);
    memcpy(S->px2, S->px1, sizeof(double) * n);
    memcpy(S->px1, x, sizeof(double) * n);
    for (long i = 0; i < n; ++i)
        S->r0[i] = rhs[i] - S->r0[i];
    memcpy(S->r1, S->r0, sizeof(double) * n);
    memcpy(S->p1, S->r0, sizeof(double) * n);
    r1r0 = dot_mode(mode, S->r1, S->r0, n, S->tmp);
    int pIter = 0;
    for (;;) {
        orc_apply_A(g, precond, zp, S->p1, S->ap);
        a = r1r0 / dot_march(mode, g, S->ap, S->r0, n, S->tmp);
        for (long i = 0; i < n; ++i)
            S->s1[i] = S->r1[i] - a * S->ap[i];
        orc_apply_A(g, precond, zp, S->s1, S->as);
        w = dot_march(mode, g, S->as, S->s1, n, S->tmp) /
            dot_march(mode, g, S->as, S->as, n, S->tmp);
        for (long i = 0; i < n; ++i) {
            x[i] += a * S->p1[i] + w * S->s1[i];
            S->r1[i] = S->s1[i] - w * S->as[i];
        }
        if ((sqrt(dot_march(mode, g, S->r1, S->r1, n, S->tmp)) / norm0) < tol)
            break;
        if (pIter >= maxit) {
            /* NaN/Inf reset, bicgstabRe2.hpp:133-153 */
            long check = 0;
            for (long i = 0; i < n; ++i)
                if (isnan(x[i]) || isinf(x[i])) {
                    x[i] = S->px1[i];
                    ++check;
                }
            if (check > n * 0.1) {
                for (long i = 0; i < n; ++i)
                    x[i] = S->px2[i];
                S->reset_members += (int)n;
                ++S->resets;
            } else if (check != 0) {
                S->reset_members += (int)check;
                ++S->resets;
            }
            break;
        }
        r1r0l = r1r0;
        r1r0 = dot_march(mode, g, S->r1, S->r0, n, S->tmp);
        b = (a / w) * (r1r0 / r1r0l);
        if (S->timestep < 10 || fabs(r1r0l) > tol * restart_factor) {
            for (long i = 0; i < n; ++i)
                S->p1[i] = S->r1[i] + b * (S->p1[i] - w * S->ap[i]);
        } else {
            for (long i = 0; i < n; ++i) {
                S->r0[i] = S->r1[i];
                S->p1[i] = S->r1[i];
            }
            ++S->restarts;
        }
        ++pIter;
        ++S->iter_total;
    }
    if (res_out)
        *res_out = sqrt(dot_march(mode, g, S->r1, S->r1, n, S->tmp)) / norm0;
    return pIter;
}

/* Pressure_transform_X_result Transform.hpp:8-25 */
void orc_transform(const orc_grid *g, const double *x, double *p)
{
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k)
                p[IC(g, i, j, k)] = x[ICAL(g, i, j, k)];
}

/* ------------------------------------------------------------------------ */
/* SOR: coefficients and update of SorPipeLine_seq SorPipeLine.seq.hpp:31-123 */
/* redblack = 0: lexicographic (the reference order); 1: red (i+j+k even)    */
/* then black.  Convergence: max |dp| over a sweep (SOR_ERROR_MAXVAL).       */
/* ------------------------------------------------------------------------ */
int orc_sor(const orc_grid *g, int redblack, int zp, double dt, double omega, double tol,
            int itmax, const double *us, const double *vs, const double *ws, double *p,
            double *err_out)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const long n = (long)g->nxc * g->nyc * g->nzc;
    const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz, *Dxs = g->Dxs, *Dys = g->Dys,
                 *Dzs = g->Dzs;
    double *cf = (double *)malloc(sizeof(double) * 8 * n);
    const double rdt = 1.0 / dt;
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const long ii = ICAL(g, i, j, k) * 8;
                cf[ii + 0] = Dy[j] * Dz[k] / Dxs[i];
                cf[ii + 1] = Dy[j] * Dz[k] / Dxs[i - 1];
                cf[ii + 2] = Dx[i] * Dz[k] / Dys[j];
                cf[ii + 3] = Dx[i] * Dz[k] / Dys[j - 1];
                cf[ii + 4] = Dx[i] * Dy[j] / Dzs[k];
                cf[ii + 5] = Dx[i] * Dy[j] / Dzs[k - 1];
                const double coef = Dy[j] * Dz[k] / Dxs[i] + Dy[j] * Dz[k] / Dxs[i - 1] +
                                    Dx[i] * Dz[k] / Dys[j] + Dx[i] * Dz[k] / Dys[j - 1] +
                                    Dx[i] * Dy[j] / Dzs[k] + Dx[i] * Dy[j] / Dzs[k - 1];
                cf[ii + 7] = -1.0 / (coef);
                const long icel = IC(g, i, j, k);
                const double mChange =
                    (us[icel] - us[IC(g, i - 1, j, k)]) * Dy[j] * Dz[k] +
                    (vs[icel] - vs[IC(g, i, j - 1, k)]) * Dx[i] * Dz[k] +
                    (ws[icel] - ws[IC(g, i, j, k - 1)]) * Dx[i] * Dy[j];
                cf[ii + 6] = mChange * rdt;
            }
    const long sx = (long)ny * nz, sy = nz;
    const int nzc = g->nzc;
    int iters = 0;
    double residual;
    do {
        residual = 0.0;
        for (int colour = 0; colour < (redblack ? 2 : 1); ++colour)
            for (int i = GC; i < nx - GC; ++i)
                for (int j = GC; j < ny - GC; ++j)
                    for (int k = GC; k < nz - GC; ++k) {
                        if (redblack && ((i + j + k) & 1) != colour)
                            continue;
                        const long icel = IC(g, i, j, k);
                        const long ii8 = ICAL(g, i, j, k) * 8;
                        const long kp = (zp && k == nz - GC - 1) ? icel + 1 - nzc : icel + 1;
                        const long km = (zp && k == GC) ? icel - 1 + nzc : icel - 1;
                        const double pNEW =
                            (-p[icel + sx] * cf[ii8] - p[icel - sx] * cf[ii8 + 1] -
                             p[icel + sy] * cf[ii8 + 2] - p[icel - sy] * cf[ii8 + 3] -
                             p[kp] * cf[ii8 + 4] - p[km] * cf[ii8 + 5] + cf[ii8 + 6]) *
                            cf[ii8 + 7];
                        const double pChange = fabs(pNEW - p[icel]);
                        p[icel] += omega * (pNEW - p[icel]);
                        if (pChange > residual)
                            residual = pChange;
                    }
        ++iters;
    } while (residual > tol && iters < itmax);
    if (err_out)
        *err_out = residual;
    free(cf);
    return iters - 1;
}

/* ------------------------------------------------------------------------ */
/* Projection + direct forcing: update_UandF_omp updateUandF.hpp:90-180      */
/* ------------------------------------------------------------------------ */
void orc_update_uf(const orc_grid *g, double dt, int zp, int outflow_convective,
                   const double *p, const double *eta, const double *us, const double *vs,
                   const double *ws, double *u, double *v, double *w, double *f)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const long G = (long)nx * ny * nz;
    const double uS = 0.0, vS = 0.0, wS = 0.0;
    for (int i = GC; i < nx - GC - 1; ++i)
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const long ic = IC(g, i, j, k), icP = IC(g, i + 1, j, k);
                const double T2u = us[ic] - dt * (p[icP] - p[ic]) / g->Dxs[i];
                const double e = eta[ic] + (eta[icP] - eta[ic]) * ((g->Dxs[i] * 0.5) / g->Dx[i]);
                u[ic] = eta[ic] * uS + (1.0 - e) * T2u;
                f[ic + 0 * G] = (u[ic] - T2u) / dt;
            }
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC - 1; ++j)
            for (int k = GC; k < nz - GC; ++k) {
                const long ic = IC(g, i, j, k), icP = IC(g, i, j + 1, k);
                const double T2v = vs[ic] - dt * (p[icP] - p[ic]) / g->Dys[j];
                const double e = eta[ic] + (eta[icP] - eta[ic]) * ((g->Dys[j] * 0.5) / g->Dy[j]);
                v[ic] = eta[ic] * vS + (1.0 - e) * T2v;
                f[ic + 1 * G] = (v[ic] - T2v) / dt;
            }
    const int kend = zp ? nz - GC : nz - GC - 1;
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < kend; ++k) {
                const long ic = IC(g, i, j, k);
                /* periodic: the face k = nz-3 couples to the interior cell k = 2 */
                const long icP = (zp && k == nz - GC - 1) ? IC(g, i, j, GC) : IC(g, i, j, k + 1);
                const double T2w = ws[ic] - dt * (p[icP] - p[ic]) / g->Dzs[k];
                const double e = eta[ic] + (eta[icP] - eta[ic]) * ((g->Dzs[k] * 0.5) / g->Dz[k]);
                w[ic] = eta[ic] * wS + (1.0 - e) * T2w;
                f[ic + 2 * G] = (w[ic] - T2w) / dt;
            }
    if (outflow_convective) {
        /* the outlet face is advanced by the boundary condition, not projected:
         * carry u*_b (= u^n_b) into the new field */
        for (int j = GC; j < ny - GC; ++j)
            for (int k = GC; k < nz - GC; ++k)
                u[IC(g, nx - 3, j, k)] = us[IC(g, nx - 3, j, k)];
    }
}

/* ------------------------------------------------------------------------ */
/* Immersed-boundary mask: DFIB_CylinderZ setEta.hpp:136-204                 */
/* ------------------------------------------------------------------------ */
void orc_eta_cylinder_z(const orc_grid *g, double center_x, double center_y, double Radius,
                        int nSubGrids, double *eta)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double *sdx = (double *)malloc(sizeof(double) * (nSubGrids + 1));
    double *sdy = (double *)malloc(sizeof(double) * (nSubGrids + 1));
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j) {
            const long icel = (long)i * nz * ny + (long)j * nz;
            const double Diagonal = sqrt(g->Dx[i] * g->Dx[i] + g->Dy[j] * g->Dy[j]) * 0.5;
            const double Distance = sqrt(pow(((g->X[i] + g->X[i + 1]) * 0.5) - center_x, 2) +
                                         pow(((g->Y[j] + g->Y[j + 1]) * 0.5) - center_y, 2));
            if (fabs(Distance - Radius) < Diagonal) {
                long xi = 0;
                const double dyg = g->Dy[j] / (double)(nSubGrids);
                const double dxg = g->Dx[i] / (double)(nSubGrids);
                for (int ii = 0; ii < nSubGrids + 1; ii++) {
                    sdy[ii] = g->Y[j] + (double)ii * dyg;
                    sdx[ii] = g->X[i] + (double)ii * dxg;
                }
                for (int ii = 0; ii < nSubGrids; ii++)
                    for (int jj = 0; jj < nSubGrids; jj++) {
                        const double Sdist = sqrt(pow((sdx[ii] + sdx[ii + 1]) * 0.5 - center_x, 2) +
                                                  pow((sdy[jj] + sdy[jj + 1]) * 0.5 - center_y, 2));
                        if (Sdist <= Radius)
                            ++xi;
                    }
                eta[icel] = (double)xi * pow(nSubGrids, -2);
            } else if (Distance > Radius) {
                eta[icel] = 0.;
            } else {
                eta[icel] = 1.;
            }
        }
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k)
                eta[IC(g, i, j, k)] = eta[IC(g, i, j, 0)];
    free(sdx);
    free(sdy);
}

/* Smoothed mask for differentiation (no reference counterpart):
 * eta = H(R - dist) with H(phi) = 0 (phi < -eps), 1 (phi > eps),
 * 0.5*(1 + phi/eps + sin(pi*phi/eps)/pi) otherwise; same plane layout as above. */
void orc_eta_smoothed(const orc_grid *g, double cx, double cy, double R, double eps, double *eta)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double pi = 3.14159265358979323846;
    for (int i = GC; i < nx - GC; ++i)
        for (int j = GC; j < ny - GC; ++j) {
            const long icel = (long)i * nz * ny + (long)j * nz;
            const double xc = (g->X[i] + g->X[i + 1]) * 0.5 - cx;
            const double yc = (g->Y[j] + g->Y[j + 1]) * 0.5 - cy;
            const double phi = R - sqrt(xc * xc + yc * yc);
            double e;
            if (phi < -eps) e = 0.0;
            else if (phi > eps) e = 1.0;
            else e = 0.5 * (1.0 + phi / eps + sin(pi * phi / eps) / pi);
            eta[icel] = e;
        }
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k)
                eta[IC(g, i, j, k)] = eta[IC(g, i, j, 0)];
}

/* ------------------------------------------------------------------------ */
/* Drag / lift: CD_CL + virtualF_Int virtualForceIntergrator.hpp:67-111      */
/* ------------------------------------------------------------------------ */
void orc_cd_cl(const orc_grid *g, int mode, const double *f, double *cd, double *cl)
{
    const long G = (long)g->nx * g->ny * g->nz;
    const long n = (long)g->nxc * g->nyc * g->nzc;
    double *t = (double *)malloc(sizeof(double) * n);
    double sum[2];
    for (int d = 0; d < 2; ++d) {
        for (int i = GC; i < g->nx - GC; ++i)
            for (int j = GC; j < g->ny - GC; ++j)
                for (int k = GC; k < g->nz - GC; ++k)
                    t[ICAL(g, i, j, k)] =
                        f[IC(g, i, j, k) + G * d] * (g->Dx[i] * g->Dy[j] * g->Dz[k]);
        sum[d] = reduce(mode, t, n);
    }
    *cd = -2.0 * sum[0] / g->lz;
    *cl = -2.0 * sum[1] / g->lz;
    free(t);
}

/* get_Cfl cfl.hpp:3-79 (signed sum per cell, max over cells, times dt) */
double orc_cfl(const orc_grid *g, double dt, int zp, const double *u, const double *v,
               const double *w)
{
    double cmax = 0.0;
    int first = 1;
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k) {
                const long ic = IC(g, i, j, k);
                double c = 0.0;
                if (i < g->nx - GC - 1) c += u[ic] / g->Dxs[i];
                if (j < g->ny - GC - 1) c += v[ic] / g->Dys[j];
                if (k < g->nz - GC - 1 || zp) c += w[ic] / g->Dzs[k];
                if (first || c > cmax) cmax = c;
                first = 0;
            }
    return cmax * dt;
}

/* CheckSteadyStatebyMaxVal_omp CheckSteadystate.hpp:247-307 */
void orc_maxdiff(const orc_grid *g, const double *u0, const double *v0, const double *w0,
                 const double *u3, const double *v3, const double *w3, double out[3])
{
    double m[3] = {0.0, 0.0, 0.0};
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k) {
                const long ic = IC(g, i, j, k);
                const double du = u0[ic] - u3[ic], dv = v0[ic] - v3[ic], dw = w0[ic] - w3[ic];
                if (du > m[0]) m[0] = du;
                if (dv > m[1]) m[1] = dv;
                if (dw > m[2]) m[2] = dw;
            }
    out[0] = m[0];
    out[1] = m[1];
    out[2] = m[2];
}

/* The same maximum over the faces the projection computes this step (orc_update_uf's
 * ranges: u for i < nx-3, v for j < ny-3, w for k < nz-3, or every k when z is
 * periodic).  The reference's T3 keeps, on the faces its update skips, values no step
 * refreshes (the initial field: copyVel copies T3 into T0, never back), so those faces
 * are left out here. */
void orc_maxdiff_faces(const orc_grid *g, int zp, const double *u0, const double *v0,
                       const double *w0, const double *u3, const double *v3, const double *w3,
                       double out[3])
{
    double m[3] = {0.0, 0.0, 0.0};
    const int kend = zp ? g->nz - GC : g->nz - GC - 1;
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k) {
                const long ic = IC(g, i, j, k);
                const double du = u0[ic] - u3[ic], dv = v0[ic] - v3[ic], dw = w0[ic] - w3[ic];
                if (i < g->nx - GC - 1 && du > m[0]) m[0] = du;
                if (j < g->ny - GC - 1 && dv > m[1]) m[1] = dv;
                if (k < kend && dw > m[2]) m[2] = dw;
            }
    out[0] = m[0];
    out[1] = m[1];
    out[2] = m[2];
}

/* Seeded initial field (no reference counterpart; the reference starts from
 * u = 1, v = w = 0, RUN_CPU.cpp:66-76): u = u0 + amp*(2U - 1) on interior cells,
 * U from splitmix64(seed ^ icelCal). */
static inline uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_init_velocity(const orc_grid *g, double u0, double amp, unsigned long long seed,
                       double *u, double *v, double *w)
{
    const long G = (long)g->nx * g->ny * g->nz;
    for (long c = 0; c < G; ++c) {
        u[c] = u0;
        v[c] = 0.0;
        w[c] = 0.0;
    }
    for (int i = GC; i < g->nx - GC; ++i)
        for (int j = GC; j < g->ny - GC; ++j)
            for (int k = GC; k < g->nz - GC; ++k) {
                const uint64_t h = splitmix64((uint64_t)seed ^ (uint64_t)ICAL(g, i, j, k));
                const double U = (double)(h >> 11) * 0x1.0p-53;
                u[IC(g, i, j, k)] = u0 + amp * (2.0 * U - 1.0);
            }
}

/* ------------------------------------------------------------------------ */
/* Time loop: runSeqOmp RUN_CPU.cpp:66-109 (init) and :257-436 (steps)       */
/* ------------------------------------------------------------------------ */
struct orc_sim {
    orc_grid g;
    orc_bc bc;
    orc_params prm;
    double *f0[3], *f1[3], *f3[3], *p, *eta, *f, *x, *b;
    double *viseff, *smag_coef; /* Smagorinsky (prm.smag_cs > 0) */
    orc_bicg *bicg;
    int loop;
    double time;
};

orc_sim *orc_sim_new(const orc_grid *g, const orc_bc *bc, const orc_params *prm)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->g = *g;
    s->bc = *bc;
    s->prm = *prm;
    const long G = (long)g->nx * g->ny * g->nz;
    const long n = (long)g->nxc * g->nyc * g->nzc;
    for (int c = 0; c < 3; ++c) {
        s->f0[c] = (double *)calloc((size_t)G, sizeof(double));
        s->f1[c] = (double *)calloc((size_t)G, sizeof(double));
        s->f3[c] = (double *)calloc((size_t)G, sizeof(double));
    }
    s->p = (double *)calloc((size_t)G, sizeof(double));
    s->eta = (double *)calloc((size_t)G, sizeof(double));
    s->f = (double *)calloc((size_t)(3 * G), sizeof(double));
    s->x = (double *)calloc((size_t)n, sizeof(double));
    s->b = (double *)calloc((size_t)n, sizeof(double));
    s->bicg = orc_bicg_new(n);
    if (prm->smag_cs > 0.0) {
        s->viseff = (double *)calloc((size_t)n, sizeof(double));
        s->smag_coef = (double *)calloc((size_t)n, sizeof(double));
        orc_smagorinsky_coef(g, prm->smag_cs, s->smag_coef);
    }
    s->loop = 1;
    s->time = 0.0;
    return s;
}

void orc_sim_free(orc_sim *s)
{
    if (!s) return;
    for (int c = 0; c < 3; ++c) {
        free(s->f0[c]);
        free(s->f1[c]);
        free(s->f3[c]);
    }
    free(s->p); free(s->eta); free(s->f); free(s->x); free(s->b);
    free(s->viseff); free(s->smag_coef);
    orc_bicg_free(s->bicg);
    free(s);
}

double *orc_sim_field(orc_sim *s, int which)
{
    if (which >= 0 && which < 3) return s->f0[which];
    if (which >= 3 && which < 6) return s->f1[which - 3];
    if (which >= 6 && which < 9) return s->f3[which - 6];
    switch (which) {
    case 9: return s->p;
    case 10: return s->eta;
    case 11: return s->f;
    case 12: return s->x;
    case 13: return s->b;
    }
    return NULL;
}

void orc_sim_start(orc_sim *s)
{
    /* RUN_CPU.cpp:104-109: T1 = T3 = T0; BC_staggered_main(T0, p); BC_staggered_copy(T0 -> T1) */
    const long G = (long)s->g.nx * s->g.ny * s->g.nz;
    for (int c = 0; c < 3; ++c) {
        memcpy(s->f1[c], s->f0[c], sizeof(double) * G);
        memcpy(s->f3[c], s->f0[c], sizeof(double) * G);
    }
    orc_bc_main(&s->g, &s->bc, 0, s->prm.dt, s->f0[0], s->f0[1], s->f0[2], s->p);
    orc_bc_copy(&s->g, s->f0[0], s->f0[1], s->f0[2], s->f1[0], s->f1[1], s->f1[2]);
}

int orc_sim_step(orc_sim *s, double out[11])
{
    const orc_grid *g = &s->g;
    const orc_params *P = &s->prm;
    const int zp = s->bc.z_periodic;
    const long G = (long)g->nx * g->ny * g->nz;
    int iters = 0, solved = 0;
    double err = 0.0;
    /* 0. eddy viscosity from T0 (RUN_CPU.cpp:261-263) */
    if (s->viseff)
        orc_smagorinsky(g, s->smag_coef, s->f0[0], s->f0[1], s->f0[2], s->viseff);
    /* 1. u* (RUN_CPU.cpp:268) */
    orc_convdiff_les(g, P->dt, P->nu, s->viseff, zp, s->f0[0], s->f0[1], s->f0[2], s->f1[0],
                     s->f1[1], s->f1[2]);
    const double cfl = orc_cfl(g, P->dt, zp, s->f1[0], s->f1[1], s->f1[2]);
    /* 2. pressure (RUN_CPU.cpp:285-322) */
    if (s->loop > 1 || !P->skip_first_pressure) {
        solved = 1;
        if (P->solver == ORC_SOLVER_BICGSTAB) {
            orc_rhs(g, P->dt, P->precond, zp, s->f1[0], s->f1[1], s->f1[2], s->b);
            iters = orc_bicg_solve(s->bicg, g, P->precond, zp, P->reduction, P->tol, P->maxit,
                                   s->b, s->x, &err);
            orc_transform(g, s->x, s->p);
        } else {
            iters = orc_sor(g, P->solver == ORC_SOLVER_RBSOR, zp, P->dt, P->omega, P->tol,
                            P->maxit, s->f1[0], s->f1[1], s->f1[2], s->p, &err);
        }
    }
    /* 3. projection + forcing (RUN_CPU.cpp:369) */
    orc_update_uf(g, P->dt, zp, s->bc.outflow_convective, s->p, s->eta, s->f1[0], s->f1[1],
                  s->f1[2], s->f3[0], s->f3[1], s->f3[2], s->f);
    /* steady-state check max(T0 - T3) (RUN_CPU.cpp:380-384) */
    double dmax[3];
    orc_maxdiff_faces(g, zp, s->f0[0], s->f0[1], s->f0[2], s->f3[0], s->f3[1], s->f3[2], dmax);
    /* 4. T3 -> T0 (RUN_CPU.cpp:391) */
    if (P->swap_buffers) {
        for (int c = 0; c < 3; ++c) {
            double *t = s->f0[c];
            s->f0[c] = s->f3[c];
            s->f3[c] = t;
        }
    } else {
        for (int c = 0; c < 3; ++c)
            memcpy(s->f0[c], s->f3[c], sizeof(double) * G);
    }
    /* 5. ghosts (RUN_CPU.cpp:399-400) */
    orc_bc_main(g, &s->bc, 1, P->dt, s->f0[0], s->f0[1], s->f0[2], s->p);
    orc_bc_copy(g, s->f0[0], s->f0[1], s->f0[2], s->f1[0], s->f1[1], s->f1[2]);
    /* 6. drag/lift (RUN_CPU.cpp:406-407), reported at the time before the increment */
    double cd, cl;
    orc_cd_cl(g, P->reduction, s->f, &cd, &cl);
    if (out) {
        out[0] = cd;
        out[1] = cl;
        out[2] = s->time;
        out[3] = iters;
        out[4] = err;
        out[5] = cfl;
        out[6] = s->bicg->restarts;
        out[7] = solved;
        out[8] = dmax[0];
        out[9] = dmax[1];
        out[10] = dmax[2];
    }
    s->time += P->dt;
    s->loop++;
    return iters;
}
