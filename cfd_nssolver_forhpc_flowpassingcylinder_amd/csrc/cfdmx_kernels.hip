// cfdmx_kernels.hip -- gfx950 kernels of the time step.
//
// Every kernel restates the reference expression by expression (same operands,
// same association; the library is compiled with -ffp-contract=off and divisions
// by grid metrics use the exactly rounded divr), so results equal the CPU
// restatement bit for bit; sums use the canonical order of cfdmx_device.hpp.
#include "cfdmx_internal.hpp"

namespace cfdmx {

// ---------------------------------------------------------------------------
// QUICK face values, QuickUnitA / QuickUnitB (Convection_Difussion.hpp:20-90).
// f[0..4] = phi at I-2, I-1, I, I+1, I+2.  The differences over D (A) or Ds (B)
// are formed once and shared with the diffusion term, which uses the same
// quotients (Convection_Difussion.hpp:257-263,339-345,420-426).
// ---------------------------------------------------------------------------
struct QuickOut {
    double P, N;
    double dm1, d0, d1; // A: g0 - g-1, g1 - g0, g2 - g1   B: h-1 - h-2, h0 - h-1, h1 - h0
};

// a / b for the metric quotients: the correctly rounded division, or, when every spacing
// of the axis is a power of two (P2), the product with rb = 1/b, which is then exact and
// so the same value (both are the correctly rounded a * 2^-k)
template <bool P2>
__device__ __forceinline__ double dv(double a, double b, double rb)
{
    return P2 ? a * rb : divr(a, b, rb);
}

template <bool P2 = false>
__device__ __forceinline__ void quick_grads_A(const Axis &a, int I, const double f[5],
                                              double &gm1, double &g0, double &g1, double &g2)
{
    gm1 = dv<P2>(f[1] - f[0], a.D[I - 1], a.rD[I - 1]);
    g0 = dv<P2>(f[2] - f[1], a.D[I], a.rD[I]);
    g1 = dv<P2>(f[3] - f[2], a.D[I + 1], a.rD[I + 1]);
    g2 = dv<P2>(f[4] - f[3], a.D[I + 2], a.rD[I + 2]);
}

// face values from the four A quotients g(-1..2) = (f[q+1] - f[q]) / D[I+q-1]
__device__ __forceinline__ void quickA_g(const Axis &a, int I, double Up, double Un,
                                         const double f[5], double gm1, double g0, double g1,
                                         double g2, double &P, double &N)
{
    const double cp = (Up > 0.0 ? a.qa[0] : a.qa[1])[I];
    const double dp = Up > 0.0 ? (g1 - g0) : (g2 - g1);
    P = 0.5 * (f[2] + f[3]) - cp * dp;
    const double cn = (Un > 0.0 ? a.qa[2] : a.qa[3])[I];
    const double dn = Un > 0.0 ? (g0 - gm1) : (g1 - g0);
    N = 0.5 * (f[1] + f[2]) - cn * dn;
}

template <bool P2 = false>
__device__ __forceinline__ void quickA(const Axis &a, int I, double Up, double Un,
                                       const double f[5], double &P, double &N, double &g0,
                                       double &g1)
{
    double gm1, g2;
    quick_grads_A<P2>(a, I, f, gm1, g0, g1, g2);
    quickA_g(a, I, Up, Un, f, gm1, g0, g1, g2, P, N);
}

// face values from the four B quotients h(-2..1) = (f[q+1] - f[q]) / Ds[I+q-2]
__device__ __forceinline__ void quickB_h(const Axis &a, int I, double Up, double Un,
                                         const double f[5], double hm2, double hm1, double h0,
                                         double h1, double &P, double &N)
{
    const double cp = (Up > 0.0 ? a.qb[0] : a.qb[1])[I];
    const double dp = Up > 0.0 ? (h0 - hm1) : (h1 - h0);
    P = 0.5 * (f[2] + f[3]) - cp * dp;
    const double cn = (Un > 0.0 ? a.qb[2] : a.qb[3])[I];
    const double dn = Un > 0.0 ? (hm1 - hm2) : (h0 - hm1);
    N = 0.5 * (f[1] + f[2]) - cn * dn;
}

template <bool P2 = false>
__device__ __forceinline__ void quickB(const Axis &a, int I, double Up, double Un,
                                       const double f[5], double &P, double &N, double &hm1,
                                       double &h0)
{
    const double hm2 = dv<P2>(f[1] - f[0], a.Ds[I - 2], a.rDs[I - 2]);
    hm1 = dv<P2>(f[2] - f[1], a.Ds[I - 1], a.rDs[I - 1]);
    h0 = dv<P2>(f[3] - f[2], a.Ds[I], a.rDs[I]);
    const double h1 = dv<P2>(f[4] - f[3], a.Ds[I + 1], a.rDs[I + 1]);
    quickB_h(a, I, Up, Un, f, hm2, hm1, h0, h1, P, N);
}

__device__ __forceinline__ void load5(const double *phi, int c, int s, double f[5])
{
    f[0] = phi[c - 2 * s];
    f[1] = phi[c - s];
    f[2] = phi[c];
    f[3] = phi[c + s];
    f[4] = phi[c + 2 * s];
}

// ---------------------------------------------------------------------------
// ConvectionDifussion (Convection_Difussion.hpp:174-450): T0 -> T1 for every
// interior cell's u, v and w faces, skipping the last face in each component's own
// direction (:198-202,278-284,361-367) unless z is periodic.  Also the per-cell
// CFL sum of get_Cfl (cfl.hpp:24-63) as a block maximum.
//
// Plane-marching and LDS-tiled: a block owns a tile of TJ rows x TK cells of the (j, k)
// plane (TK the power of two >= nzc, 4 to 32; TJ = 256 / TK; one cell per thread)
// and marches over kChunk planes in x.  A 3-slot LDS ring holds u, v, w of the planes
// behind, at and ahead of the current one over the tile plus two rows and two cells of
// halo on each side (the y and z stencils and the cross terms); the thread's own column
// at two planes behind and ahead (the rest of the x stencil) stays in registers.  The
// next ring plane is loaded into registers a plane ahead, and the 1-D metric tables of
// the chunk, the tile's rows and its k range sit in LDS.  Every value is formed with the
// reference's expressions from the same operands, so the result is bit-identical to a
// cell-by-cell evaluation.
// ---------------------------------------------------------------------------
constexpr int kCdTables = 13; // D, Ds, rD, rDs, IT, qa[4], qb[4]

// copy of the 13 tables of one axis for ghosted indices [lo, lo + n) into LDS at dst
// (table t at dst + t*n), and the Axis that reads them with the global indices
__device__ __forceinline__ Axis cd_axis_lds(const Axis &a, int lo, int n, int lim, double *dst)
{
    const double *src[kCdTables] = {a.D,     a.Ds,    a.rD,    a.rDs,   a.IT,
                                    a.qa[0], a.qa[1], a.qa[2], a.qa[3], a.qb[0],
                                    a.qb[1], a.qb[2], a.qb[3]};
    for (int e = threadIdx.x; e < kCdTables * n; e += blockDim.x) {
        const int t = e / n, x = lo + e % n;
        dst[e] = (x >= 0 && x < lim) ? src[t][x] : 0.0;
    }
    Axis r;
    double *base[kCdTables];
    for (int t = 0; t < kCdTables; ++t) base[t] = dst + t * n - lo;
    r.D = base[0];
    r.Ds = base[1];
    r.rD = base[2];
    r.rDs = base[3];
    r.IT = base[4];
    for (int q = 0; q < 4; ++q) {
        r.qa[q] = base[5 + q];
        r.qb[q] = base[9 + q];
    }
    return r;
}

__host__ __device__ inline size_t cd_lds_bytes(const Geo &g)
{
    const size_t P = (size_t)(g.cd_tj + 4) * (g.cd_tk + 4);
    return (9 * P + kCdTables * ((size_t)(kChunk + 4) + (g.cd_tj + 4) + (g.cd_tk + 4))) *
           sizeof(double);
}

// ZP2: every z spacing (Dz and Dzs) is a power of two, as on the uniform grids whose nz
// divides lz by a power of two: the z quotients are then exact products (dv)
template <int NE, bool LES, bool ZP2>
__global__ void __launch_bounds__(kThreads)
k_convdiff(Geo g, const double *__restrict__ u, const double *__restrict__ v,
           const double *__restrict__ w, double *__restrict__ us, double *__restrict__ vs,
           double *__restrict__ ws, double mdt, double nudt, double dt,
           double *__restrict__ cfl_part, double nu, const double *__restrict__ viseff)
{
    const int b = logical_block(g);
    if (b >= g.cd_nb) return;
    extern __shared__ double lds[];
    const int TK = g.cd_tk, TJ = g.cd_tj, RW = TK + 4, P = (TJ + 4) * RW;
    const int kt = b % g.cd_nkt, jt = (b / g.cd_nkt) % g.cd_njt, ch = b / (g.cd_nkt * g.cd_njt);
    const int k0 = kt * TK, j0 = jt * TJ, i0 = ch * kChunk;
    const int I0 = i0 + kGC, len = min(kChunk, g.nxc - i0);
    double *ring = lds;                 // [slot][field][P]
    double *mt = lds + 9 * P;           // metric tables
    const Axis X = cd_axis_lds(g.ax[0], I0 - 2, kChunk + 4, g.nx, mt);
    const Axis Y = cd_axis_lds(g.ax[1], j0, TJ + 4, g.ny, mt + kCdTables * (kChunk + 4));
    const Axis Z = cd_axis_lds(g.ax[2], k0, TK + 4,
                               g.nz, mt + kCdTables * (kChunk + 4 + TJ + 4));
    // own cell
    const int tj = threadIdx.x / TK, tk = threadIdx.x - tj * TK;
    const int J = j0 + tj + kGC, K = k0 + tk + kGC;
    const bool own = (j0 + tj < g.nyc) && (k0 + tk < g.nzc);
    const int c = (tj + 2) * RW + tk + 2;
    // ring tiles: element e of a plane is (row e / RW, column e % RW) at ghosted
    // (J, K) = (j0 + row, k0 + column); NE = ceil(P / 256) elements per thread
    auto tile_index = [&](int I, int e) -> int {
        const int row = e / RW, col = e - row * RW;
        const int Jg = j0 + row, Kg = k0 + col;
        const int ok = (int)(e < P) & (int)(Jg < g.ny) & (int)(Kg < g.nz) & (int)(I < g.nx);
        return ((I * g.ny + Jg) * g.nz + Kg) & -ok;
    };
    double pu[NE], pv[NE], pw[NE];
    auto load_tile = [&](int I) {
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int n = tile_index(I, threadIdx.x + q * kThreads);
            pu[q] = u[n];
            pv[q] = v[n];
            pw[q] = w[n];
        }
    };
    auto put_tile = [&](int slot) {
        double *S = ring + slot * 3 * P;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = threadIdx.x + q * kThreads;
            if (e < P) {
                S[e] = pu[q];
                S[P + e] = pv[q];
                S[2 * P + e] = pw[q];
            }
        }
    };
    // own column at ghosted plane I (zero-masked index when out of the grid)
    auto own_index = [&](int I) -> int {
        const int ok = (int)own & (int)(I < g.nx);
        return ((I * g.ny + J) * g.nz + K) & -ok;
    };
    // prologue: ring planes I0-1, I0, I0+1 (slots 0, 1, 2), own column at I0-2 and I0+2
    for (int p = 0; p < 3; ++p) {
        load_tile(I0 - 1 + p);
        put_tile(p);
    }
    int nm2 = own_index(I0 - 2), np2 = own_index(I0 + 2);
    double um2 = u[nm2], vm2 = v[nm2], wm2 = w[nm2];
    double up2 = u[np2], vp2 = v[np2], wp2 = w[np2];
    load_tile(I0 + 2);
    __syncthreads();
    // x quotients of the QUICK stencils carried from plane to plane: A for u,
    // qa(I') = (u[I'] - u[I'-1]) / D[I'] at I' = I-1, I, I+1; B for v and w,
    // hb(I') = (f[I'+1] - f[I']) / Ds[I'] at I' = I-2, I-1, I
    double gu[3], gv[3], gw[3];
    {
        const double *S0 = ring, *S1 = ring + 3 * P, *S2 = ring + 6 * P;
        gu[0] = divr(S0[c] - um2, X.D[I0 - 1], X.rD[I0 - 1]);
        gu[1] = divr(S1[c] - S0[c], X.D[I0], X.rD[I0]);
        gu[2] = divr(S2[c] - S1[c], X.D[I0 + 1], X.rD[I0 + 1]);
        gv[0] = divr(S0[P + c] - vm2, X.Ds[I0 - 2], X.rDs[I0 - 2]);
        gv[1] = divr(S1[P + c] - S0[P + c], X.Ds[I0 - 1], X.rDs[I0 - 1]);
        gv[2] = divr(S2[P + c] - S1[P + c], X.Ds[I0], X.rDs[I0]);
        gw[0] = divr(S0[2 * P + c] - wm2, X.Ds[I0 - 2], X.rDs[I0 - 2]);
        gw[1] = divr(S1[2 * P + c] - S0[2 * P + c], X.Ds[I0 - 1], X.rDs[I0 - 1]);
        gw[2] = divr(S2[2 * P + c] - S1[2 * P + c], X.Ds[I0], X.rDs[I0]);
    }
    double cmax = -1.0e300;
    for (int s = 0; s < len; ++s) {
        const int I = I0 + s;
        const double *Sm = ring + (s % 3) * 3 * P, *Sc = ring + ((s + 1) % 3) * 3 * P,
                     *Sp = ring + ((s + 2) % 3) * 3 * P;
        const double *um = Sm, *vm = Sm + P, *wm = Sm + 2 * P;
        const double *uc = Sc, *vc = Sc + P, *wc = Sc + 2 * P;
        const double *upl = Sp, *vpl = Sp + P, *wpl = Sp + 2 * P;
        const double gu3 = divr(up2 - upl[c], X.D[I + 2], X.rD[I + 2]);
        const double gv3 = divr(vp2 - vpl[c], X.Ds[I + 1], X.rDs[I + 1]);
        const double gw3 = divr(wp2 - wpl[c], X.Ds[I + 1], X.rDs[I + 1]);
        if (own) {
            const int ic = gidx(g, I, J, K);
            double cfl = 0.0;
            // the diffusion's nu*dt; with the Smagorinsky model nu = 1/Re + Viseff of the
            // cell (Convection_Difussion.hpp:251-255)
            const double nudt_c =
                LES ? (nu + viseff[((I - kGC) * g.nyc + (J - kGC)) * g.nzc + (K - kGC)]) * dt
                    : nudt;
            // ---------------- u (:205-273) ----------------
            if (I < g.nx - 3) {
                const double IT = X.IT[I];
                const double Up = 0.5 * (upl[c] + uc[c]);
                const double Un = 0.5 * (um[c] + uc[c]);
                const double Vp = vc[c] + (vpl[c] - vc[c]) * IT;
                const double Vn = vc[c - RW] + (vpl[c - RW] - vc[c - RW]) * IT;
                const double Wp = wc[c] + (wpl[c] - wc[c]) * IT;
                const double Wn = wc[c - 1] + (wpl[c - 1] - wc[c - 1]) * IT;
                double uPx, uNx, uPy, uNy, uPz, uNz, ax0, ax1, ay0, ay1, az0, az1;
                const double fx[5] = {um2, um[c], uc[c], upl[c], up2};
                quickA_g(X, I, Up, Un, fx, gu[0], gu[1], gu[2], gu3, uPx, uNx);
                ax0 = gu[1];
                ax1 = gu[2];
                const double fy[5] = {uc[c - 2 * RW], uc[c - RW], uc[c], uc[c + RW], uc[c + 2 * RW]};
                quickB(Y, J, Vp, Vn, fy, uPy, uNy, ay0, ay1);
                const double fz[5] = {uc[c - 2], uc[c - 1], uc[c], uc[c + 1], uc[c + 2]};
                quickB<ZP2>(Z, K, Wp, Wn, fz, uPz, uNz, az0, az1);
                const double convection =
                    mdt * (divr(uPx * uPx - uNx * uNx, X.Ds[I], X.rDs[I]) +
                           divr(uPy * Vp - uNy * Vn, Y.D[J], Y.rD[J]) +
                           dv<ZP2>(uPz * Wp - uNz * Wn, Z.D[K], Z.rD[K]));
                // ((u[xp]-u)/Dx[i+1] - (u-u[xm])/Dx[i]) = ax1 - ax0, etc.
                const double difussion = nudt_c * (divr(ax1 - ax0, X.Ds[I], X.rDs[I]) +
                                                 divr(ay1 - ay0, Y.D[J], Y.rD[J]) +
                                                 dv<ZP2>(az1 - az0, Z.D[K], Z.rD[K]));
                const double r = uc[c] + convection + difussion;
                us[ic] = r;
                cfl += divr(r, X.Ds[I], X.rDs[I]);
            }
            // ---------------- v (:286-354) ----------------
            if (J < g.ny - 3) {
                const double IT = Y.IT[J];
                const double Vp = 0.5 * (vc[c + RW] + vc[c]);
                const double Vn = 0.5 * (vc[c - RW] + vc[c]);
                const double Up = uc[c] + (uc[c + RW] - uc[c]) * IT;
                const double Un = um[c] + (um[c + RW] - um[c]) * IT;
                const double Wp = wc[c] + (wc[c + RW] - wc[c]) * IT;
                const double Wn = wc[c - 1] + (wc[c - 1 + RW] - wc[c - 1]) * IT;
                double vPx, vNx, vPy, vNy, vPz, vNz, ax0, ax1, ay0, ay1, az0, az1;
                const double fx[5] = {vm2, vm[c], vc[c], vpl[c], vp2};
                quickB_h(X, I, Up, Un, fx, gv[0], gv[1], gv[2], gv3, vPx, vNx);
                ax0 = gv[1];
                ax1 = gv[2];
                const double fy[5] = {vc[c - 2 * RW], vc[c - RW], vc[c], vc[c + RW], vc[c + 2 * RW]};
                quickA(Y, J, Vp, Vn, fy, vPy, vNy, ay0, ay1);
                const double fz[5] = {vc[c - 2], vc[c - 1], vc[c], vc[c + 1], vc[c + 2]};
                quickB<ZP2>(Z, K, Wp, Wn, fz, vPz, vNz, az0, az1);
                const double convection =
                    mdt * (divr(Up * vPx - Un * vNx, X.D[I], X.rD[I]) +
                           divr(vPy * vPy - vNy * vNy, Y.Ds[J], Y.rDs[J]) +
                           dv<ZP2>(Wp * vPz - Wn * vNz, Z.D[K], Z.rD[K]));
                const double difussion = nudt_c * (divr(ax1 - ax0, X.D[I], X.rD[I]) +
                                                 divr(ay1 - ay0, Y.Ds[J], Y.rDs[J]) +
                                                 dv<ZP2>(az1 - az0, Z.D[K], Z.rD[K]));
                const double r = vc[c] + convection + difussion;
                vs[ic] = r;
                cfl += divr(r, Y.Ds[J], Y.rDs[J]);
            }
            // ---------------- w (:368-437) ----------------
            if (K < g.nz - 3 || g.zper) {
                const double IT = Z.IT[K];
                const double Wp = 0.5 * (wc[c + 1] + wc[c]);
                const double Wn = 0.5 * (wc[c - 1] + wc[c]);
                const double Up = uc[c] + (uc[c + 1] - uc[c]) * IT;
                const double Un = um[c] + (um[c + 1] - um[c]) * IT;
                const double Vp = vc[c] + (vc[c + 1] - vc[c]) * IT;
                const double Vn = vc[c - RW] + (vc[c - RW + 1] - vc[c - RW]) * IT;
                double wPx, wNx, wPy, wNy, wPz, wNz, ax0, ax1, ay0, ay1, az0, az1;
                const double fx[5] = {wm2, wm[c], wc[c], wpl[c], wp2};
                quickB_h(X, I, Up, Un, fx, gw[0], gw[1], gw[2], gw3, wPx, wNx);
                ax0 = gw[1];
                ax1 = gw[2];
                const double fy[5] = {wc[c - 2 * RW], wc[c - RW], wc[c], wc[c + RW], wc[c + 2 * RW]};
                quickB(Y, J, Vp, Vn, fy, wPy, wNy, ay0, ay1);
                const double fz[5] = {wc[c - 2], wc[c - 1], wc[c], wc[c + 1], wc[c + 2]};
                quickA<ZP2>(Z, K, Wp, Wn, fz, wPz, wNz, az0, az1);
                const double convection =
                    mdt * (divr(Up * wPx - Un * wNx, X.D[I], X.rD[I]) +
                           divr(Vp * wPy - Vn * wNy, Y.D[J], Y.rD[J]) +
                           dv<ZP2>(wPz * wPz - wNz * wNz, Z.Ds[K], Z.rDs[K]));
                const double difussion = nudt_c * (divr(ax1 - ax0, X.D[I], X.rD[I]) +
                                                 divr(ay1 - ay0, Y.D[J], Y.rD[J]) +
                                                 dv<ZP2>(az1 - az0, Z.Ds[K], Z.rDs[K]));
                const double r = wc[c] + convection + difussion;
                ws[ic] = r;
                cfl += dv<ZP2>(r, Z.Ds[K], Z.rDs[K]);
            }
            cmax = cfl > cmax ? cfl : cmax;
            if (g.zper) {
                // u* z-ghosts are images of the interior (face k = 1 == nz-3)
                const int nzc = g.nzc;
                int img = -1;
                if (K == 2) img = ic + nzc;
                else if (K == 3) img = ic + nzc;
                else if (K == g.nz - 3) img = ic - nzc;
                else if (K == g.nz - 4) img = ic - nzc;
                if (img >= 0) {
                    us[img] = us[ic];
                    vs[img] = vs[ic];
                    ws[img] = ws[ic];
                }
            }
        }
        // shift the own column (plane I-1 becomes the one two behind) and the quotients
        gu[0] = gu[1], gu[1] = gu[2], gu[2] = gu3;
        gv[0] = gv[1], gv[1] = gv[2], gv[2] = gv3;
        gw[0] = gw[1], gw[1] = gw[2], gw[2] = gw3;
        um2 = um[c];
        vm2 = vm[c];
        wm2 = wm[c];
        __syncthreads();
        // plane I+2 into the slot of plane I-1; loads for the next step
        put_tile(s % 3);
        np2 = own_index(I + 3);
        up2 = u[np2];
        vp2 = v[np2];
        wp2 = w[np2];
        load_tile(I + 3);
        __syncthreads();
    }
    __shared__ double red[4];
    const double wmx = wave_max(cmax);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wmx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = red[q] > m ? red[q] : m;
        cfl_part[b] = m * dt;
    }
}

// ---------------------------------------------------------------------------
// SmagorinskyModel (SmagorinskyModel.hpp:30-79, GradientGauss :6-27): per interior cell,
// central differences 0.5/Dm * (phi[m] - phi[p]) of u over (Dxs, Dy, Dz), v over
// (Dx, Dys, Dz) and w over (Dx, Dy, Dzs); temp = |2 ux^2 + vy^2 + wz^2 + (uy + vx)^2 +
// (uz + wx)^2 + (vz + wy)^2| (the factor 2 on the first term only, as written there);
// Viseff = coef * sqrt(temp) with coef = (Cs delta)^2, delta = pow(Dx Dy Dz, 1/3),
// tabulated on the host.  One thread per cell, icelCal order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_smagorinsky(Geo g, const double *__restrict__ u, const double *__restrict__ v,
              const double *__restrict__ w, const double *__restrict__ coef,
              double *__restrict__ viseff)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N) return;
    int i, j, k;
    decode(g, n, i, j, k);
    const int I = i + kGC, J = j + kGC, K = k + kGC;
    const int c = gidx(g, I, J, K), sx = g.sx, sy = g.sy;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const double ux = 0.5 / X.Ds[I] * (u[c - sx] - u[c + sx]);
    const double uy = 0.5 / Y.D[J] * (u[c - sy] - u[c + sy]);
    const double uz = 0.5 / Z.D[K] * (u[c - 1] - u[c + 1]);
    const double vx = 0.5 / X.D[I] * (v[c - sx] - v[c + sx]);
    const double vy = 0.5 / Y.Ds[J] * (v[c - sy] - v[c + sy]);
    const double vz = 0.5 / Z.D[K] * (v[c - 1] - v[c + 1]);
    const double wx = 0.5 / X.D[I] * (w[c - sx] - w[c + sx]);
    const double wy = 0.5 / Y.D[J] * (w[c - sy] - w[c + sy]);
    const double wz = 0.5 / Z.Ds[K] * (w[c - 1] - w[c + 1]);
    const double a = uy + vx, b = uz + wx, d = vz + wy;
    const double temp = fabs(2.0 * (ux * ux) + vy * vy + wz * wz + a * a + b * b + d * d);
    viseff[n] = coef[n] * sqrt(temp);
}

// ---------------------------------------------------------------------------
// Plot3D q arrays (OutputPLOT3D_Qfile, qfile.hpp:63-103): p, cell-centred u, v, w
// (0.5 * (face below + face above) in FP64, rounded to float) and eta, each an interior
// array with i fastest.  The fields have k fastest, so a 32 x 32 (i, k) tile of one j
// row goes through LDS: reads coalesce along k, writes along i.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_plot3d_q(Geo g, const double *__restrict__ u, const double *__restrict__ v,
           const double *__restrict__ w, const double *__restrict__ p,
           const double *__restrict__ eta, float *__restrict__ out)
{
    __shared__ float tile[32][33];
    const int i0 = blockIdx.x * 32, k0 = blockIdx.y * 32, j = blockIdx.z;
    const size_t N = (size_t)g.N;
    for (int q = 0; q < 5; ++q) {
        for (int ii = threadIdx.y; ii < 32; ii += 8) {
            const int i = i0 + ii, k = k0 + threadIdx.x;
            if (i < g.nxc && k < g.nzc) {
                const int ic = gidx(g, i + kGC, j + kGC, k + kGC);
                double val;
                if (q == 0) val = p[ic];
                else if (q == 1) val = 0.5 * (u[ic - g.sx] + u[ic]);
                else if (q == 2) val = 0.5 * (v[ic - g.sy] + v[ic]);
                else if (q == 3) val = 0.5 * (w[ic - 1] + w[ic]);
                else val = eta[ic];
                tile[ii][threadIdx.x] = (float)val;
            }
        }
        __syncthreads();
        for (int kk = threadIdx.y; kk < 32; kk += 8) {
            const int i = i0 + threadIdx.x, k = k0 + kk;
            if (i < g.nxc && k < g.nzc)
                out[q * N + ((size_t)k * g.nyc + j) * g.nxc + i] = tile[threadIdx.x][kk];
        }
        __syncthreads();
    }
}

void launch_plot3d_q(const Geo &g, hipStream_t s, const double *u, const double *v,
                     const double *w, const double *p, const double *eta, float *out)
{
    hipLaunchKernelGGL(k_plot3d_q, dim3((g.nxc + 31) / 32, (g.nzc + 31) / 32, g.nyc),
                       dim3(32, 8), 0, s, g, u, v, w, p, eta, out);
}

// ---------------------------------------------------------------------------
// Boundary conditions, BC_staggered_main (Boundary_Condition.hpp:37-176) with the
// Num/Dir table, applied per face group in the reference's order (x, y, z faces),
// plus pressure ghosts.  With t1 != nullptr the same ghost values are written into
// T1, which is BC_staggered_copy (updateBC.hpp:6-110).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void put2(double *a, double *b, int c0, int c1, double val)
{
    a[c0] = val;
    a[c1] = val;
    if (b) {
        b[c0] = val;
        b[c1] = val;
    }
}

__global__ void k_bc_x(Geo g, BCParams bc, int advance, double *u, double *v, double *w,
                       double *p, double *u1, double *v1, double *w1)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.cyz) return;
    const int J = m / g.nzc + kGC, K = m % g.nzc + kGC;
    const int nx = g.nx;
    const int c0 = gidx(g, 0, J, K), c1 = gidx(g, 1, J, K), c2 = gidx(g, 2, J, K);
    // face 0 (:72-80)
    const double d0 = bc.inflow ? bc.inflow[J] : bc.dir[0][0];
    put2(u, u1, c1, c0, d0 + bc.num[0][0] * u[c2]);
    put2(v, v1, c1, c0, bc.dir[0][1] * 2. + (2.0 * bc.num[0][1] - 1.0) * v[c2]);
    put2(w, w1, c1, c0, bc.dir[0][2] * 2. + (2.0 * bc.num[0][2] - 1.0) * w[c2]);
    // face 1 (:84-94); convective outflow advances u on the outlet face
    const int e4 = gidx(g, nx - 4, J, K), e3 = gidx(g, nx - 3, J, K),
              e2 = gidx(g, nx - 2, J, K), e1 = gidx(g, nx - 1, J, K);
    double ub;
    if (bc.convective && advance) {
        const double uo = u[e3];
        ub = uo - bc.conv_coef * (uo - u[e4]);
    } else {
        ub = bc.dir[1][0] + bc.num[1][0] * u[e4];
    }
    u[e1] = ub;
    u[e2] = ub;
    u[e3] = ub;
    if (u1) {
        u1[e1] = ub;
        u1[e2] = ub;
        u1[e3] = ub;
    }
    put2(v, v1, e2, e1, bc.dir[1][1] * 2. + (2.0 * bc.num[1][1] - 1.0) * v[e3]);
    put2(w, w1, e2, e1, bc.dir[1][2] * 2. + (2.0 * bc.num[1][2] - 1.0) * w[e3]);
    // pressure (:149-155)
    const double pw = p[c2];
    p[c1] = pw;
    p[c0] = pw;
    const double pe = p[e3];
    p[e2] = pe;
    p[e1] = pe;
}

// outlet flux correction (no reference counterpart): canonical sums over the
// interior (j,k) of u_in*Dy*Dz, u_out*Dy*Dz and Dy*Dz
__global__ void __launch_bounds__(kThreads)
k_flux_partial(Geo g, const double *u, double *part, int nb)
{
    const int L = blockIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < kCells; ++c) {
        const int m = L * kPerBlock + c * kThreads + threadIdx.x;
        if (m >= g.cyz) break;
        const int J = m / g.nzc + kGC, K = m % g.nzc + kGC;
        const double a = g.ax[1].D[J] * g.ax[2].D[K];
        v[0] += u[gidx(g, 1, J, K)] * a;
        v[1] += u[gidx(g, g.nx - 3, J, K)] * a;
        v[2] += a;
    }
    block_sums<3>(v, part, nb, L);
}

// every block folds the few partials itself (the same order, so the same correction)
// and corrects kThreads cells of the outlet plane
__global__ void __launch_bounds__(kThreads)
k_flux_apply(Geo g, const double *part, int nb, double *u, double *u1)
{
    double t[3];
    fold_partials<3>(part, nb, nb, t);
    const double corr = (t[0] - t[1]) / t[2];
    const int m = blockIdx.x * kThreads + threadIdx.x;
    if (m < g.cyz) {
        const int J = m / g.nzc + kGC, K = m % g.nzc + kGC;
        const int e3 = gidx(g, g.nx - 3, J, K);
        const double ub = u[e3] + corr;
        u[e3] = ub;
        u[e3 + g.sx] = ub;
        u[e3 + 2 * g.sx] = ub;
        if (u1) {
            u1[e3] = ub;
            u1[e3 + g.sx] = ub;
            u1[e3 + 2 * g.sx] = ub;
        }
    }
}

__global__ void k_bc_y(Geo g, BCParams bc, double *u, double *v, double *w, double *p,
                       double *u1, double *v1, double *w1)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.nxc * g.nzc) return;
    const int I = m / g.nzc + kGC, K = m % g.nzc + kGC;
    const int ny = g.ny;
    const int c0 = gidx(g, I, 0, K), c1 = gidx(g, I, 1, K), c2 = gidx(g, I, 2, K);
    // face 2 (:97-105)
    put2(u, u1, c1, c0, bc.dir[2][0] * 2. + (2.0 * bc.num[2][0] - 1.0) * u[c2]);
    put2(v, v1, c1, c0, bc.dir[2][1] + bc.num[2][1] * v[c2]);
    put2(w, w1, c1, c0, bc.dir[2][2] * 2. + (2.0 * bc.num[2][2] - 1.0) * w[c2]);
    // face 3 (:107-117)
    const int e4 = gidx(g, I, ny - 4, K), e3 = gidx(g, I, ny - 3, K),
              e2 = gidx(g, I, ny - 2, K), e1 = gidx(g, I, ny - 1, K);
    const double vb = bc.dir[3][1] + bc.num[3][1] * v[e4];
    v[e1] = vb;
    v[e2] = vb;
    v[e3] = vb;
    if (v1) {
        v1[e1] = vb;
        v1[e2] = vb;
        v1[e3] = vb;
    }
    put2(u, u1, e2, e1, bc.dir[3][0] * 2. + (2.0 * bc.num[3][0] - 1.0) * u[e3]);
    put2(w, w1, e2, e1, bc.dir[3][2] * 2. + (2.0 * bc.num[3][2] - 1.0) * w[e3]);
    // pressure (:158-164)
    const double ps = p[c2];
    p[c1] = ps;
    p[c0] = ps;
    const double pn = p[e3];
    p[e2] = pn;
    p[e1] = pn;
}

__global__ void k_bc_z(Geo g, BCParams bc, double *u, double *v, double *w, double *p,
                       double *u1, double *v1, double *w1)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int nz = g.nz, nzc = g.nzc;
    if (g.zper) {
        // periodic: ghost k <- k + nzc (k < 2) or k - nzc (k >= nz-2) for u, v, w and p
        // in every (i,j) column (the face k = nz-3 reads x/y ghost columns); T1 receives
        // the interior columns, as BC_staggered_copy does
        if (m >= g.nx * g.ny) return;
        const int I = m / g.ny, J = m % g.ny;
        const int b = gidx(g, I, J, 0);
        const bool inner = I >= kGC && I < g.nx - kGC && J >= kGC && J < g.ny - kGC;
        double *f[4] = {u, v, w, p};
        double *f1[4] = {u1, v1, w1, nullptr};
        for (int q = 0; q < 4; ++q) {
            const double a1 = f[q][b + 1 + nzc], a0 = f[q][b + nzc];
            const double a2 = f[q][b + nz - 2 - nzc], a3 = f[q][b + nz - 1 - nzc];
            f[q][b + 1] = a1;
            f[q][b + 0] = a0;
            f[q][b + nz - 2] = a2;
            f[q][b + nz - 1] = a3;
            if (f1[q] && inner) {
                f1[q][b + 1] = a1;
                f1[q][b + 0] = a0;
                f1[q][b + nz - 2] = a2;
                f1[q][b + nz - 1] = a3;
                if (q == 2) f1[q][b + nz - 3] = f[q][b + nz - 3];
            }
        }
        return;
    }
    if (m >= g.nxc * g.nyc) return;
    const int I = m / g.nyc + kGC, J = m % g.nyc + kGC;
    const int b = gidx(g, I, J, 0);
    // face 4 (:121-129)
    put2(u, u1, b + 1, b, bc.dir[4][0] * 2. + (2.0 * bc.num[4][0] - 1.0) * u[b + 2]);
    put2(v, v1, b + 1, b, bc.dir[4][1] * 2. + (2.0 * bc.num[4][1] - 1.0) * v[b + 2]);
    put2(w, w1, b + 1, b, bc.dir[4][2] + bc.num[4][2] * w[b + 2]);
    // face 5 (:133-143)
    const double wb = bc.dir[5][2] + bc.num[5][2] * w[b + nz - 4];
    w[b + nz - 1] = wb;
    w[b + nz - 2] = wb;
    w[b + nz - 3] = wb;
    if (w1) {
        w1[b + nz - 1] = wb;
        w1[b + nz - 2] = wb;
        w1[b + nz - 3] = wb;
    }
    put2(u, u1, b + nz - 2, b + nz - 1,
         bc.dir[5][0] * 2. + (2.0 * bc.num[5][0] - 1.0) * u[b + nz - 3]);
    put2(v, v1, b + nz - 2, b + nz - 1,
         bc.dir[5][1] * 2. + (2.0 * bc.num[5][1] - 1.0) * v[b + nz - 3]);
    // pressure (:166-172)
    const double pb = p[b + 2];
    p[b + 1] = pb;
    p[b + 0] = pb;
    const double pt = p[b + nz - 3];
    p[b + nz - 2] = pt;
    p[b + nz - 1] = pt;
}

// BC_staggered_copy (updateBC.hpp:6-110) on its own: ghost entries T0 -> T1
__global__ void k_bc_copy(Geo g, const double *u0, const double *v0, const double *w0,
                          double *u1, double *v1, double *w1)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (m < g.cyz) {
        const int J = m / g.nzc + kGC, K = m % g.nzc + kGC;
        const int cs[5] = {0, 1, nx - 3, nx - 2, nx - 1};
        for (int q = 0; q < 5; ++q) {
            const int c = gidx(g, cs[q], J, K);
            u1[c] = u0[c];
            if (q != 2) {
                v1[c] = v0[c];
                w1[c] = w0[c];
            }
        }
        return;
    }
    m -= g.cyz;
    if (m < g.nxc * g.nzc) {
        const int I = m / g.nzc + kGC, K = m % g.nzc + kGC;
        const int cs[5] = {0, 1, ny - 3, ny - 2, ny - 1};
        for (int q = 0; q < 5; ++q) {
            const int c = gidx(g, I, cs[q], K);
            v1[c] = v0[c];
            if (q != 2) {
                u1[c] = u0[c];
                w1[c] = w0[c];
            }
        }
        return;
    }
    m -= g.nxc * g.nzc;
    if (m < g.nxc * g.nyc) {
        const int I = m / g.nyc + kGC, J = m % g.nyc + kGC;
        const int cs[5] = {0, 1, nz - 3, nz - 2, nz - 1};
        for (int q = 0; q < 5; ++q) {
            const int c = gidx(g, I, J, cs[q]);
            w1[c] = w0[c];
            if (q != 2) {
                u1[c] = u0[c];
                v1[c] = v0[c];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pressure operator rows.  SPE (SPE_sparseMatrix.hpp:42-106): slots -x, -y, -z,
// diag, +z, +y, +x with -y holding -yp(j) and +y holding -ym(j) as in the
// reference, boundary faces folded into the diagonal.  JACOBI: the row-scaled CSR
// matrix of createPressureMatrixCSR (structureGrid.hpp:557-654), entries
// (-coef)*jp with jp = 1/diag.  Zero boundary slots are skipped, which equals
// adding 0*x for finite x.
// ---------------------------------------------------------------------------
struct RowCoef {
    double v[7];
    bool has[7];
    int off[7];
};

template <int PRE>
__device__ __forceinline__ void row_coef(const Geo &g, int i, int j, int k, RowCoef &rc)
{
    const double xm = g.pa[0].m[i], xp = g.pa[0].p[i];
    const double ym = g.pa[1].m[j], yp = g.pa[1].p[j];
    const double zm = g.pa[2].m[k], zp = g.pa[2].p[k];
    const bool kf = (k == 0), kl = (k == g.nzc - 1);
    rc.has[0] = i != 0;
    rc.has[1] = j != 0;
    rc.has[2] = !kf || g.zper;
    rc.has[3] = true;
    rc.has[4] = !kl || g.zper;
    rc.has[5] = j != g.nyc - 1;
    rc.has[6] = i != g.nxc - 1;
    rc.off[0] = -g.cyz;
    rc.off[1] = -g.nzc;
    rc.off[2] = kf ? g.nzc - 1 : -1;
    rc.off[3] = 0;
    rc.off[4] = kl ? -(g.nzc - 1) : 1;
    rc.off[5] = g.nzc;
    rc.off[6] = g.cyz;
    if (PRE == 0) {
        double ijk = xm + ym + zm + xp + yp + zp;
        if (i == 0) ijk -= xm;
        if (i == g.nxc - 1) ijk -= xp;
        if (j == 0) ijk -= ym;
        if (j == g.nyc - 1) ijk -= yp;
        if (!g.zper) {
            if (kf) ijk -= zm;
            if (kl) ijk -= zp;
        }
        rc.v[0] = -xm;
        rc.v[1] = -yp;
        rc.v[2] = -zm;
        rc.v[3] = ijk;
        rc.v[4] = -zp;
        rc.v[5] = -ym;
        rc.v[6] = -xp;
    } else {
        double aa = 0.0, bb = 0.0, cc = 0.0;
        if (i == 0) aa = -xm;
        if (i == g.nxc - 1) aa = -xp;
        if (j == 0) bb = -ym;
        if (j == g.nyc - 1) bb = -yp;
        if (!g.zper) {
            if (kf) cc = -zm;
            if (kl) cc = -zp;
        }
        const double ijk = xm + ym + zm + xp + yp + zp + aa + bb + cc;
        const double jp = 1.0 / ijk;
        rc.v[0] = -xm * jp;
        rc.v[1] = -ym * jp;
        rc.v[2] = -zm * jp;
        rc.v[3] = ijk * jp;
        rc.v[4] = -zp * jp;
        rc.v[5] = -yp * jp;
        rc.v[6] = -xp * jp;
    }
}

// The seven values are fetched unconditionally (a missing neighbour reads the cell
// itself) and a missing term adds +0.0: the sum starts at +0.0 and so is never -0.0,
// which makes +0.0 an exact identity -- the result equals skipping the term.  Without a
// branch around each load, all seven are in flight together.
template <class F>
__device__ __forceinline__ double row_apply(const RowCoef &rc, int n, F &&val)
{
    double x[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) x[q] = val(rc.has[q] ? n + rc.off[q] : n);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) s += rc.has[q] ? rc.v[q] * x[q] : 0.0;
    return s;
}

template <int PRE>
__global__ void __launch_bounds__(kThreads)
k_apply(Geo g, const double *__restrict__ x, double *__restrict__ r)
{
    const int L = logical_block(g);
    if (L >= g.nblocks) return;
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        RowCoef rc;
        row_coef<PRE>(g, i, j, k, rc);
        r[n] = row_apply(rc, n, [&](int m) { return x[m]; });
    }
}

// The transpose, r = A^T y (no reference counterpart): what the adjoint of the pressure
// solve needs where A is not symmetric (the SPE operator on the SML grid, with its
// stretched widths and y-slot swap; the row-scaled CSR operator).  Every coupling is
// mutual in position: row n has a neighbour in slot q exactly when that row holds n in
// slot 6 - q, also across a periodic z wrap.  So column n takes, in a row's slot order,
// slot 6 - q of the row in its slot q, from +0.0, skipping the absent slots.  Each
// neighbour row's coefficients are formed by row_coef, as in k_apply.
template <int PRE>
__global__ void __launch_bounds__(kThreads)
k_apply_t(Geo g, const double *__restrict__ y, double *__restrict__ r)
{
    const int L = logical_block(g);
    if (L >= g.nblocks) return;
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        RowCoef me;
        row_coef<PRE>(g, i, j, k, me);
        const int km = k == 0 ? g.nzc - 1 : k - 1, kp = k == g.nzc - 1 ? 0 : k + 1;
        const int ni[7] = {i - 1, i, i, i, i, i, i + 1};
        const int nj[7] = {j, j - 1, j, j, j, j + 1, j};
        const int nk[7] = {k, k, km, k, kp, k, k};
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            if (!me.has[q]) continue;
            double v;
            if (q == 3) {
                v = me.v[3];
            } else {
                RowCoef nb;
                row_coef<PRE>(g, ni[q], nj[q], nk[q], nb);
                v = nb.v[6 - q];
            }
            s += v * y[n + me.off[q]];
        }
        r[n] = s;
    }
}

// createBMatrix_omp (b_matrix.hpp:47-93): b = -(1/dt) div u*, times jp on the CSR path
template <int PRE>
__global__ void __launch_bounds__(kThreads)
k_rhs(Geo g, const double *__restrict__ us, const double *__restrict__ vs,
      const double *__restrict__ ws, double ddt, double *__restrict__ b)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        const int I = i + kGC, J = j + kGC, K = k + kGC;
        const int ic = gidx(g, I, J, K);
        const double div = divr(us[ic] - us[ic - g.sx], X.D[I], X.rD[I]) +
                           divr(vs[ic] - vs[ic - g.sy], Y.D[J], Y.rD[J]) +
                           divr(ws[ic] - ws[ic - 1], Z.D[K], Z.rD[K]);
        if (PRE == 1) {
            // jp = 1/diag of createPressureMatrixCSR (structureGrid.hpp:590-603)
            const double xm = g.pa[0].m[i], xp = g.pa[0].p[i];
            const double ym = g.pa[1].m[j], yp = g.pa[1].p[j];
            const double zm = g.pa[2].m[k], zp = g.pa[2].p[k];
            double aa = 0.0, bb = 0.0, cc = 0.0;
            if (i == 0) aa = -xm;
            if (i == g.nxc - 1) aa = -xp;
            if (j == 0) bb = -ym;
            if (j == g.nyc - 1) bb = -yp;
            if (!g.zper) {
                if (k == 0) cc = -zm;
                if (k == g.nzc - 1) cc = -zp;
            }
            const double jp = 1.0 / (xm + ym + zm + xp + yp + zp + aa + bb + cc);
            b[n] = -jp * ddt * div;
        } else {
            b[n] = -ddt * div;
        }
    }
}

// The same right-hand side with four consecutive cells of a z-row per thread (nzc a
// multiple of 4, so a thread's cells share i and j, and the ghosted index of the first
// is even): u* and v* at the cells and at their -x / -y neighbours and b move as 16-byte
// vectors, and w* at the cells plus the one below.  Each cell's expression is k_rhs's.
__device__ __forceinline__ void load4(const double *p, double (&v)[4])
{
    const double2 a = reinterpret_cast<const double2 *>(p)[0];
    const double2 b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y;
}

template <int PRE>
__global__ void __launch_bounds__(kThreads)
k_rhs4(Geo g, const double *__restrict__ us, const double *__restrict__ vs,
       const double *__restrict__ ws, double ddt, double *__restrict__ b)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    const int n0 = L * kPerBlock + 4 * threadIdx.x;
    if (n0 >= g.N) return;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    int i, j, k0;
    decode(g, n0, i, j, k0);
    const int I = i + kGC, J = j + kGC;
    const int ic = gidx(g, I, J, k0 + kGC);
    double uc[4], uw[4], vc[4], vs4[4], wc[4];
    load4(us + ic, uc);
    load4(us + ic - g.sx, uw);
    load4(vs + ic, vc);
    load4(vs + ic - g.sy, vs4);
    load4(ws + ic, wc);
    const double wb = ws[ic - 1];
    double out[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = k0 + q, K = k + kGC;
        const double div = divr(uc[q] - uw[q], X.D[I], X.rD[I]) +
                           divr(vc[q] - vs4[q], Y.D[J], Y.rD[J]) +
                           divr(wc[q] - (q > 0 ? wc[q - 1] : wb), Z.D[K], Z.rD[K]);
        if (PRE == 1) {
            const double xm = g.pa[0].m[i], xp = g.pa[0].p[i];
            const double ym = g.pa[1].m[j], yp = g.pa[1].p[j];
            const double zm = g.pa[2].m[k], zp = g.pa[2].p[k];
            double aa = 0.0, bb = 0.0, cc = 0.0;
            if (i == 0) aa = -xm;
            if (i == g.nxc - 1) aa = -xp;
            if (j == 0) bb = -ym;
            if (j == g.nyc - 1) bb = -yp;
            if (!g.zper) {
                if (k == 0) cc = -zm;
                if (k == g.nzc - 1) cc = -zp;
            }
            const double jp = 1.0 / (xm + ym + zm + xp + yp + zp + aa + bb + cc);
            out[q] = -jp * ddt * div;
        } else {
            out[q] = -ddt * div;
        }
    }
    reinterpret_cast<double2 *>(b + n0)[0] = make_double2(out[0], out[1]);
    reinterpret_cast<double2 *>(b + n0)[1] = make_double2(out[2], out[3]);
}

// ---------------------------------------------------------------------------
// BiCGSTAB, bicgstabRe2::solve (bicgstabRe2.hpp:82-180), as three fused phases per
// iteration with the scalars kept on the device:
//   A: p = r + beta*(p - omega*Ap) (or p = r, and r0 = r on a restart),
//      Ap = A p, partial (Ap, r0)                                   -> alpha
//   B: s = r - alpha*Ap, As = A s, partials (As, s), (As, As)       -> omega
//   C: x += alpha*p + omega*s, r = s - omega*As, partials (r,r), (r,r0)
//      -> convergence test, beta, restart decision
// Each phase starts by folding the previous phase's per-block partials (the same
// fixed order in every block, so every block holds identical scalars); block 0 stores
// them in KState for the later phases.  k_bicg_check ends a batch of iterations.
// In A and B the new vector (p or s) of the block's own 1024 cells is formed once
// into LDS; stencil neighbours outside that range are formed from global memory
// with the same expression, so every value is bit-identical wherever it is formed.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_bicg_init(Geo g, int pre, const double *__restrict__ b, const double *__restrict__ x,
            double *__restrict__ r0, double *__restrict__ r, double *__restrict__ px1,
            double *__restrict__ part, KState *)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    double acc[2] = {0.0, 0.0};
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        RowCoef rc;
        if (pre) row_coef<1>(g, i, j, k, rc);
        else row_coef<0>(g, i, j, k, rc);
        const double ax = row_apply(rc, n, [&](int m) { return x[m]; });
        const double bn = b[n];
        const double rr = bn - ax;
        r0[n] = rr;
        r[n] = rr;
        px1[n] = x[n];
        acc[0] += bn * bn;
        acc[1] += rr * rr;
    }
    block_sums<2>(acc, part, g.nblocks, L);
}

// The same initialisation with four consecutive cells of a z-row per thread (nzc a
// multiple of 4): x at the cells and at their x and y neighbours, b and the three outputs
// move as 16-byte vectors, and the z neighbours are the cells beside each other plus one
// value at either end.  Each row's terms and order are k_bicg_init's.  The partial sums
// keep k_bicg_init's order, which the solver's stopping test depends on: each cell's
// (b, b) and (r, r) terms go through LDS and every thread adds the cells that k_bicg_init
// gives it (c * kThreads + thread, c = 0..3) in that order.
__global__ void __launch_bounds__(kThreads)
k_bicg_init4(Geo g, int pre, const double *__restrict__ b, const double *__restrict__ x,
             double *__restrict__ r0, double *__restrict__ r, double *__restrict__ px1,
             double *__restrict__ part, KState *)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    __shared__ double sbb[kPerBlock], srr[kPerBlock];
    const int n0 = L * kPerBlock + 4 * threadIdx.x;
    if (n0 < g.N) {
        int i, j, k0;
        decode(g, n0, i, j, k0);
        double xc[4], xxm[4], xxp[4], xym[4], xyp[4], bv[4];
        const bool hxm = i != 0, hxp = i != g.nxc - 1, hym = j != 0, hyp = j != g.nyc - 1;
        load4(x + n0, xc);
        load4(x + (hxm ? n0 - g.cyz : n0), xxm);
        load4(x + (hxp ? n0 + g.cyz : n0), xxp);
        load4(x + (hym ? n0 - g.nzc : n0), xym);
        load4(x + (hyp ? n0 + g.nzc : n0), xyp);
        load4(b + n0, bv);
        const double zlo = x[k0 > 0 ? n0 - 1 : (g.zper ? n0 + g.nzc - 1 : n0)];
        const double zhi = x[k0 + 3 < g.nzc - 1 ? n0 + 4 : (g.zper ? n0 + 4 - g.nzc : n0 + 3)];
        double rv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            RowCoef rc;
            if (pre) row_coef<1>(g, i, j, k0 + q, rc);
            else row_coef<0>(g, i, j, k0 + q, rc);
            const double xv[7] = {xxm[q], xym[q], q > 0 ? xc[q - 1] : zlo, xc[q],
                                  q < 3 ? xc[q + 1] : zhi, xyp[q], xxp[q]};
            double ax = 0.0;
#pragma unroll
            for (int t = 0; t < 7; ++t) ax += rc.has[t] ? rc.v[t] * xv[t] : 0.0;
            rv[q] = bv[q] - ax;
            sbb[4 * threadIdx.x + q] = bv[q] * bv[q];
            srr[4 * threadIdx.x + q] = rv[q] * rv[q];
        }
        reinterpret_cast<double2 *>(r0 + n0)[0] = make_double2(rv[0], rv[1]);
        reinterpret_cast<double2 *>(r0 + n0)[1] = make_double2(rv[2], rv[3]);
        reinterpret_cast<double2 *>(r + n0)[0] = make_double2(rv[0], rv[1]);
        reinterpret_cast<double2 *>(r + n0)[1] = make_double2(rv[2], rv[3]);
        reinterpret_cast<double2 *>(px1 + n0)[0] = make_double2(xc[0], xc[1]);
        reinterpret_cast<double2 *>(px1 + n0)[1] = make_double2(xc[2], xc[3]);
    }
    __syncthreads();
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int c = 0; c < kCells; ++c) {
        const int t = c * kThreads + threadIdx.x;
        if (L * kPerBlock + t < g.N) {
            acc[0] += sbb[t];
            acc[1] += srr[t];
        }
    }
    block_sums<2>(acc, part, g.nblocks, L);
}

// Plane-marching stencil phases.  A block owns the plane-index tile
// q in [q0, q0+1024) (q = j*nzc + k) for the planes i in [i0, i0+8); even chunks march
// down (i = i0+7 .. i0), odd chunks up, so that two neighbouring chunks read their
// shared halo planes at the same time (the second read hits L2).  An LDS ring holds
// the new vector (p for A, s for B) on the planes behind, at and ahead of the current
// one over the tile plus one z-row (nzc cells) on each side, so every stencil
// neighbour is read from LDS.  Each ring value is formed once per block with the
// reference's expression; cells of neighbouring tiles/chunks are formed again there,
// identically.  Phase A forms p = r + beta*q from q = p - omega*Ap, which phase C of
// the previous iteration stores in place of p (the same two roundings as
// r + beta*(p - omega*Ap), bicgstabRe2.hpp:161-163).
// T is the storage and arithmetic type of the Krylov vectors (double; float for the
// inner solve of the mixed-precision mode); sums are always accumulated in double.
template <typename T>
struct MarchIO {
    const T *r, *q, *ap; // inputs (A: r, q; B: r, Ap)
    T *r0;               // A: restart target
    T *out_p, *out_ap;   // A outputs (p only where it is stored: see k_march_C)
    T *out_as;           // B output
    double *part;
};

// planes of march block `chunk`: first plane and direction
__device__ __forceinline__ void march_span(const Geo &g, int chunk, int &first, int &dir, int &len)
{
    const int i0 = chunk * kChunk, i1 = min(i0 + kChunk, g.nxc);
    len = i1 - i0;
    dir = (chunk & 1) ? 1 : -1;
    first = dir > 0 ? i0 : i1 - 1;
}

// LDS of a march block: x coefficients of the chunk's planes, y coefficients of the
// tile's rows, z coefficients (double), then two plane slots (T)
__host__ __device__ inline int march_rows(int nzc) { return kTile / nzc + 2; }

template <int PRE>
__device__ __forceinline__ void march_row(const Geo &g, int i, int j, int k, double xm, double xp,
                                          double ym, double yp, double zm, double zp,
                                          double v[7])
{
    if (PRE == 0) {
        double ijk = xm + ym + zm + xp + yp + zp;
        if (i == 0) ijk -= xm;
        if (i == g.nxc - 1) ijk -= xp;
        if (j == 0) ijk -= ym;
        if (j == g.nyc - 1) ijk -= yp;
        if (!g.zper) {
            if (k == 0) ijk -= zm;
            if (k == g.nzc - 1) ijk -= zp;
        }
        v[0] = -xm;
        v[1] = -yp;
        v[2] = -zm;
        v[3] = ijk;
        v[4] = -zp;
        v[5] = -ym;
        v[6] = -xp;
    } else {
        double aa = 0.0, bb = 0.0, cc = 0.0;
        if (i == 0) aa = -xm;
        if (i == g.nxc - 1) aa = -xp;
        if (j == 0) bb = -ym;
        if (j == g.nyc - 1) bb = -yp;
        if (!g.zper) {
            if (k == 0) cc = -zm;
            if (k == g.nzc - 1) cc = -zp;
        }
        const double ijk = xm + ym + zm + xp + yp + zp + aa + bb + cc;
        const double jp = 1.0 / ijk;
        v[0] = -xm * jp;
        v[1] = -ym * jp;
        v[2] = -zm * jp;
        v[3] = ijk * jp;
        v[4] = -zp * jp;
        v[5] = -yp * jp;
        v[6] = -xp * jp;
    }
}

// Start of iteration m in phase A: the end of iteration m-1 (bicgstabRe2.hpp:129-174)
// from phase C's partials (r,r), (r,r0).  Returns false when the solve has ended.
__device__ __forceinline__ bool bicg_iteration_start(const Geo &g, KState *st, int m,
                                                     const double *partC, bool writer,
                                                     int &mode, double &beta)
{
    if (m == 0) {
        mode = 2; // p = r0 (bicgstabRe2.hpp:101-102)
        beta = 0.0;
        return true;
    }
    double t[2];
    fold_partials<2>(partC, g.nmarch, g.nmarch, t);
    const double res = sqrt(t[0]) / st->norm0; // L2Norm(r1) / norm0
    const int it = m - 1;                      // pIter of the iteration that just ended
    if (res < st->tol || it >= st->maxit) {
        if (writer) {
            st->res = res;
            st->iter = it;
            st->done = res < st->tol ? 1 : 2;
        }
        return false;
    }
    const double rho_l = st->rho[(m - 1) & 1];
    const double rho = t[1];
    beta = (st->alpha / st->omega) * (rho / rho_l);
    mode = (st->timestep < 10 || fabs(rho_l) > st->tol * st->restart_factor) ? 0 : 1;
    if (writer) {
        st->rho[m & 1] = rho;
        st->res = res;
        st->iter = m;
        st->beta = beta;
        st->mode = mode;
        if (mode == 1) st->restarts += 1;
    }
    return true;
}

// v when ok is 1, +0 when ok is 0, by masking the bits (a select here becomes a branch
// around the LDS store, which costs the plane loop its overlap of loads and stores)
__device__ __forceinline__ double keep_if(double v, int ok)
{
    return __longlong_as_double(__double_as_longlong(v) & -(long long)ok);
}
__device__ __forceinline__ float keep_if(float v, int ok)
{
    return __int_as_float(__float_as_int(v) & -ok);
}

// A march block keeps the new vector (p for A, s for B) of its own cells on the planes
// behind, at and ahead of the current one in registers (the x neighbours), and writes
// the current plane -- its own cells plus one z-row of halo on each side of the tile --
// to one of two LDS slots for the y and z neighbours.  Two slots need one barrier per
// plane: the slot written at step s was last read at step s-2.  The raw inputs of the
// plane two ahead (own cells) and of the next plane's halo are loaded one step before
// they are formed.  Halo elements per thread: NH = ceil(2 nzc / 256).
// NH = 1: at most 128 VGPRs, so that four blocks fit each CU at once.
//
// NT: the accesses read or written once per iteration (phase A's dot operand and its
// p store, phase B's As store, every phase C load and its x store) take the streaming
// (nontemporal) hint.  The launcher sets it when the iteration's seven vectors exceed
// the 256 MiB memory-side cache: there the hint leaves more cache to the stencil
// inputs, which are reread as halos.  The iteration measured 209 -> 188 us at
// 512x256x64 (470 MB) and 157 -> 149 us at 384x256x64 (352 MB), but 103 -> 108 us at
// 256x256x64 (235 MB) and 60 -> 66 us at 256x128x32, where the vectors fit the cache
// and the hint gives up their reuse.  The values are the same either way.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T v, T *p)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int PHASE, int PRE, int NH, bool FULL, bool NT, typename T>
__global__ void __launch_bounds__(kThreads, NH == 1 ? 4 : 1)
k_march(Geo g, MarchIO<T> io, KState *st, int m, const double *part_in)
{
    if (st->done) return;
    int b = blockIdx.x;
    if (g.xcd_remap) b = (b & 7) * ((gridDim.x + 7) >> 3) + (b >> 3);
    if (b >= g.nmarch) return;
    extern __shared__ double lds[];
    constexpr int C = kMarchCells;
    const int tile = b % g.ntile, chunk = b / g.ntile;
    int ib, d, len;
    march_span(g, chunk, ib, d, len);
    const int i0 = chunk * kChunk;
    const int q0 = tile * kTile, qn = min(kTile, g.cyz - q0);
    const int H = g.nzc, WP = kTile + 2 * H;
    const int RY = march_rows(g.nzc);
    double *cx = lds, *cy = cx + 2 * kChunk, *cz = cy + 2 * RY;
    T *slots = reinterpret_cast<T *>(cz + 2 * g.nzc);
    const bool writer = (b == 0 && threadIdx.x == 0);
    int mode = 0;
    double beta = 0.0, alpha = 0.0;
    T betaT = (T)0, alphaT = (T)0;
    // 1-D operator coefficients: the chunk's planes, the tile's rows j in
    // [jlo, jlo + rows), all k
    const int jlo = (int)fdiv((unsigned)q0, g.dz);
    const int rows = (int)fdiv((unsigned)(q0 + qn - 1), g.dz) - jlo + 1;
    for (int t = threadIdx.x; t < rows; t += kThreads) {
        cy[t] = g.pa[1].m[jlo + t];
        cy[RY + t] = g.pa[1].p[jlo + t];
    }
    for (int t = threadIdx.x; t < g.nzc; t += kThreads) {
        cz[t] = g.pa[2].m[t];
        cz[g.nzc + t] = g.pa[2].p[t];
    }
    if (threadIdx.x < len) {
        cx[threadIdx.x] = g.pa[0].m[i0 + threadIdx.x];
        cx[kChunk + threadIdx.x] = g.pa[0].p[i0 + threadIdx.x];
    }
    // Loads are branch-free: a lane whose cell is outside the grid, or whose plane is
    // not wanted, reads element 0, and the value is masked to zero when formed.  With
    // no load under a branch the compiler waits for exactly the values it needs.  q is
    // loaded whatever the mode (the first iteration and a restart ignore it), so that
    // loads can be issued before the scalars are known.
    const T *src2 = PHASE == 0 ? io.q : io.ap;
    auto plane_ok = [&](int i, bool want) -> int {
        return (int)want & (int)(i >= 0) & (int)(i < g.nxc);
    };
    auto own_ok = [&](int c) -> int { return FULL ? 1 : (int)(c * kThreads + threadIdx.x < qn); };
    // halo element h of this thread: slot index t (rows above, then below the tile)
    auto halo_t = [&](int h) {
        const int x = threadIdx.x + h * kThreads;
        return x < H ? x : qn + x;
    };
    auto halo_ok = [&](int h) -> int {
        const int x = threadIdx.x + h * kThreads, q = q0 - H + halo_t(h);
        return (int)(x < 2 * H) & (int)(q >= 0) & (int)(q < g.cyz);
    };
    auto load_own = [&](int i, bool want, T (&a)[C], T (&c2)[C]) {
        const int pk = plane_ok(i, want);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int n = (i * g.cyz + q0 + c * kThreads + (int)threadIdx.x) & -(pk & own_ok(c));
            a[c] = io.r[n];
            c2[c] = src2[n];
        }
    };
    auto load_halo = [&](int i, bool want, T (&a)[NH], T (&c2)[NH]) {
        const int pk = plane_ok(i, want);
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const int n = (i * g.cyz + q0 - H + halo_t(h)) & -(pk & halo_ok(h));
            a[h] = io.r[n];
            c2[h] = src2[n];
        }
    };
    // the new vector from the raw inputs (PHASE 0: p = r + beta*q, or r; PHASE 1:
    // s = r - alpha*Ap), zero outside the grid
    auto form = [&](T a, T c2, int ok) -> T {
        T v;
        if (PHASE == 0) v = mode == 0 ? a + betaT * c2 : a;
        else v = a - alphaT * c2;
        return keep_if(v, ok);
    };
    // the second operand of phase A's dot: r0, or r on a restart (read once: NT)
    auto load_r0 = [&](int i, bool want, T (&o)[C]) {
        const T *src = mode == 1 ? io.r : io.r0;
        const int pk = (int)want;
#pragma unroll
        for (int c = 0; c < C; ++c)
            o[c] = ld_stream<NT>(
                src + ((i * g.cyz + q0 + c * kThreads + (int)threadIdx.x) & -(pk & own_ok(c))));
    };
    int jc[C], kc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int q = q0 + c * kThreads + threadIdx.x;
        const int j = (int)fdiv((unsigned)q, g.dz);
        jc[c] = j - jlo;
        kc[c] = q - j * g.nzc;
    }
    // own cells of the planes behind (ob), at (oc) and ahead (oa) of the current one
    T ob[C], oc[C], oa[C];
    T ra[C], rc[C];   // raw inputs of the plane two ahead (own cells)
    T ha[NH], hc[NH]; // raw inputs of the current plane's halo
    T r0c[C];         // phase A: dot operand of the current plane
    {
        T pa[C], pc[C], na[C], nc[C];
        load_own(ib - d, true, ra, rc);
        load_own(ib, true, pa, pc);
        load_own(ib + d, true, na, nc);
        load_halo(ib, true, ha, hc);
        // the scalars, from the previous phase's partials, while the first planes load
        if (PHASE == 0) {
            if (!bicg_iteration_start(g, st, m, part_in, writer, mode, beta)) return;
        } else {
            double t[1];
            fold_partials<1>(part_in, g.nmarch, g.nmarch, t);
            alpha = st->rho[m & 1] / t[0]; // a = r1r0 / dot(ap, r0)
            if (writer) st->alpha = alpha;
        }
        betaT = (T)beta;
        alphaT = (T)alpha;
        const int okb = plane_ok(ib - d, true), okc = plane_ok(ib, true),
                  oka = plane_ok(ib + d, true);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int ow = own_ok(c);
            ob[c] = form(ra[c], rc[c], okb & ow);
            oc[c] = form(pa[c], pc[c], okc & ow);
            oa[c] = form(na[c], nc[c], oka & ow);
        }
    }
    load_own(ib + 2 * d, len > 1, ra, rc);
    double acc[2] = {0.0, 0.0};
    for (int s = 0; s < len; ++s) {
        const int i = ib + s * d;
        T *S = slots + (s & 1) * WP;
        // the current plane into its slot: own cells, then the halo
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (FULL || c * kThreads + (int)threadIdx.x < qn) S[H + c * kThreads + threadIdx.x] = oc[c];
        {
            const int okp = plane_ok(i, true);
#pragma unroll
            for (int h = 0; h < NH; ++h)
                if (threadIdx.x + h * kThreads < 2 * H)
                    S[halo_t(h)] = form(ha[h], hc[h], okp & halo_ok(h));
        }
        // next loads: the halo of the next plane and phase A's dot operand of this
        // one (the own cells of the plane three ahead follow this plane's compute)
        load_halo(i + d, s + 1 < len, ha, hc);
        if (PHASE == 0) load_r0(i, true, r0c);
        __syncthreads();
        const double xm = cx[i - i0], xp = cx[kChunk + i - i0];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int tq = c * kThreads + threadIdx.x;
            if (!FULL && tq >= qn) break;
            const int jj = jc[c], k = kc[c], j = jlo + jj;
            double v[7];
            march_row<PRE>(g, i, j, k, xm, xp, cy[jj], cy[RY + jj], cz[k], cz[g.nzc + k], v);
            // All neighbour reads are unconditional (the halo rows and the zeroed
            // planes outside the grid are in place), and a missing neighbour adds +0.0
            // in place of its term.  The running sum starts at +0.0 and so is never
            // -0.0, which makes +0.0 an exact identity: the result equals skipping the
            // term, as the reference row does.  The 7-point sum accumulates in double
            // for every T: in float the Laplacian's cancellation on smooth vectors
            // stalls the Krylov process.
            const int tc = tq + H;
            const double n0 = (double)(d > 0 ? ob[c] : oa[c]);
            const double n1 = (double)S[tc - g.nzc];
            const double n2 = (double)S[k == 0 ? tc + g.nzc - 1 : tc - 1];
            const double n3 = (double)oc[c];
            const double n4 = (double)S[k == g.nzc - 1 ? tc - (g.nzc - 1) : tc + 1];
            const double n5 = (double)S[tc + g.nzc];
            const double n6 = (double)(d > 0 ? oa[c] : ob[c]);
            double acc7 = 0.0;
            acc7 += i != 0 ? v[0] * n0 : 0.0;
            acc7 += j != 0 ? v[1] * n1 : 0.0;
            acc7 += (k != 0 || g.zper) ? v[2] * n2 : 0.0;
            acc7 += v[3] * n3;
            acc7 += (k != g.nzc - 1 || g.zper) ? v[4] * n4 : 0.0;
            acc7 += j != g.nyc - 1 ? v[5] * n5 : 0.0;
            acc7 += i != g.nxc - 1 ? v[6] * n6 : 0.0;
            const T sum = (T)acc7;
            const int n = i * g.cyz + q0 + tq;
            const T own = oc[c];
            const double sd = (double)sum;
            if (PHASE == 0) {
                // p is stored only on grids whose vectors stay in the caches (NT off);
                // with the streaming hints phase C forms it again from r and q, which
                // are still in place (8 bytes per cell less).  Ap (read by phase B)
                // keeps the default policy, which measured faster
                if (!NT) io.out_p[n] = own;
                io.out_ap[n] = sum;
                const T r0v = r0c[c];
                if (mode == 1) io.r0[n] = r0v;
                acc[0] += sd * (double)r0v;
            } else {
                st_stream<NT>(sum, io.out_as + n);
                acc[0] += sd * (double)own;
                acc[1] += sd * sd;
            }
            // one cell's neighbour reads at a time in flight: keeps the register count
            // within four blocks per CU
            __builtin_amdgcn_sched_barrier(0);
        }
        // shift the register queue: the plane two ahead becomes the one ahead
        {
            const int ok2 = plane_ok(i + 2 * d, s + 1 < len);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                ob[c] = oc[c];
                oc[c] = oa[c];
                oa[c] = form(ra[c], rc[c], ok2 & own_ok(c));
            }
        }
        load_own(i + 3 * d, s + 2 < len, ra, rc);
    }
    if (PHASE == 0) {
        double a1s[1] = {acc[0]};
        block_sums<1>(a1s, io.part, g.nmarch, b);
    } else {
        block_sums<2>(acc, io.part, g.nmarch, b);
    }
}

// phase C in the march mapping (elementwise): x, r update (bicgstabRe2.hpp:123-127),
// the partials of (r,r), (r,r0), and q = p - omega*Ap for the next phase A.  With the
// streaming hints (NT: the vectors exceed the memory-side cache) p is formed again here
// from r and q, with phase A's expression and the mode and beta it left in the state
// (p = r + beta q, or p = r on the first iteration and on a restart), so it is the same
// value phase A used, and A stores Ap only: 8 bytes per cell less (at 512x256x64 the
// iteration took 185 -> 175 us).  On smaller grids phase A stores p (pin) and C reads it:
// there the recomputing form measured slower (the C launch 18 -> 25 us at 256x128x32),
// the stored p being still in the caches.
template <bool NT, typename T>
__global__ void __launch_bounds__(kThreads)
k_march_C(Geo g, T *__restrict__ r, const T *__restrict__ ap, T *__restrict__ q,
          const T *__restrict__ pin, const T *__restrict__ as, const T *__restrict__ r0,
          T *__restrict__ x, const double *part_in, double *__restrict__ part_out, KState *st)
{
    if (st->done) return;
    int b = blockIdx.x;
    if (g.xcd_remap) b = (b & 7) * ((gridDim.x + 7) >> 3) + (b >> 3);
    if (b >= g.nmarch) return;
    double t[2];
    fold_partials<2>(part_in, g.nmarch, g.nmarch, t);
    const double omega = t[0] / t[1]; // w = dot(as, s) / dot(as, as)
    if (b == 0 && threadIdx.x == 0) st->omega = omega;
    const T alpha = (T)st->alpha, om = (T)omega;
    const int mode = st->mode;
    const T betaT = (T)st->beta;
    const int tile = b % g.ntile, chunk = b / g.ntile;
    int ib, d, len;
    march_span(g, chunk, ib, d, len);
    const int q0 = tile * kTile, qn = min(kTile, g.cyz - q0);
    double acc[2] = {0.0, 0.0};
    for (int s = 0; s < len; ++s) {
        const int i = ib + s * d;
#pragma unroll
        for (int c = 0; c < kMarchCells; ++c) {
            const int tq = c * kThreads + threadIdx.x;
            if (tq < qn) {
                const int n = i * g.cyz + q0 + tq;
                // every operand is read once (NT: streaming loads, see k_march).  x is
                // next read by the next phase C, so it is stored with the hint too; r and
                // q, which the next phase A reads at once, keep plain stores (hinted,
                // they measured slower)
                const T apn = ld_stream<NT>(ap + n);
                const T rold = ld_stream<NT>(r + n);
                T pn;
                if (NT) {
                    const T qv = ld_stream<NT>(q + n);
                    pn = mode == 0 ? rold + betaT * qv : rold;
                } else {
                    pn = pin[n];
                }
                const T sn = rold - alpha * apn;
                st_stream<NT>(ld_stream<NT>(x + n) + (alpha * pn + om * sn), x + n);
                const T rn = sn - om * ld_stream<NT>(as + n);
                r[n] = rn;
                q[n] = pn - om * apn;
                const double rd = (double)rn;
                acc[0] += rd * rd;
                acc[1] += rd * (double)ld_stream<NT>(r0 + n);
            }
        }
    }
    block_sums<2>(acc, part_out, g.nmarch, b);
}

// ---- mixed precision (no reference counterpart): FP64 iterative refinement around
// an inner BiCGSTAB in T (float; double for the same algorithm in FP64).  The outer
// residual r = b - A x is formed in FP64 on the same operator rows as k_bicg_init; the
// inner solve A e = r starts from e = 0 with r0 = p = T(r).  Its first x/r update writes
// e as 0 + its increment, so e needs no pass of its own here.  Partials: (b,b), (r,r) in
// FP64 and (T(r), T(r)).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_mixed_resid(Geo g, int pre, const double *__restrict__ b, const double *__restrict__ x,
              T *__restrict__ r32, T *__restrict__ r0_32, double *__restrict__ part)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    double acc[3] = {0.0, 0.0, 0.0};
    auto cell = [&](int n) {
        int i, j, k;
        decode(g, n, i, j, k);
        RowCoef rc;
        if (pre) row_coef<1>(g, i, j, k, rc);
        else row_coef<0>(g, i, j, k, rc);
        const double ax = row_apply(rc, n, [&](int m) { return x[m]; });
        const double bn = b[n];
        const double rr = bn - ax;
        const T rf = (T)rr;
        r32[n] = rf;
        r0_32[n] = rf;
        acc[0] += bn * bn;
        acc[1] += rr * rr;
        acc[2] += (double)rf * (double)rf;
    };
    // a full block has no bounds test, so its cells' loads overlap
    if ((L + 1) * kPerBlock <= g.N) {
#pragma unroll
        for (int c = 0; c < kCells; ++c) cell(L * kPerBlock + c * kThreads + threadIdx.x);
    } else {
        for (int c = 0; c < kCells; ++c) {
            const int n = L * kPerBlock + c * kThreads + threadIdx.x;
            if (n >= g.N) break;
            cell(n);
        }
    }
    block_sums<3>(acc, part, g.nblocks, L);
}

// The same residual with four consecutive cells of a z-row per thread (nzc a multiple of
// 4): x's centre, x and y neighbours, b and the three outputs move as vectors, and the
// z neighbours are the cells beside each other, so a thread issues 8 loads and 2 stores
// for 4 cells instead of 32 and 8.  Each row's terms and order are k_mixed_resid's; the
// partial sums add a thread's cells in another order.
template <typename T> struct Vec4T;
template <> struct Vec4T<float> {
    static __device__ __forceinline__ void store(float *p, const float (&v)[4])
    {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    }
};
template <> struct Vec4T<double> {
    static __device__ __forceinline__ void store(double *p, const double (&v)[4])
    {
        reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
        reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
    }
};

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_mixed_resid4(Geo g, int pre, const double *__restrict__ b, const double *__restrict__ x,
               T *__restrict__ r32, T *__restrict__ r0_32, double *__restrict__ part)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    double acc[3] = {0.0, 0.0, 0.0};
    const int n0 = L * kPerBlock + 4 * threadIdx.x;
    if (n0 < g.N) {
        int i, j, k0;
        decode(g, n0, i, j, k0);
        double xc[4], xxm[4], xxp[4], xym[4], xyp[4], bv[4];
        const bool hxm = i != 0, hxp = i != g.nxc - 1, hym = j != 0, hyp = j != g.nyc - 1;
        load4(x + n0, xc);
        load4(x + (hxm ? n0 - g.cyz : n0), xxm);
        load4(x + (hxp ? n0 + g.cyz : n0), xxp);
        load4(x + (hym ? n0 - g.nzc : n0), xym);
        load4(x + (hyp ? n0 + g.nzc : n0), xyp);
        load4(b + n0, bv);
        const double zlo = x[k0 > 0 ? n0 - 1 : (g.zper ? n0 + g.nzc - 1 : n0)];
        const double zhi = x[k0 + 3 < g.nzc - 1 ? n0 + 4 : (g.zper ? n0 + 4 - g.nzc : n0 + 3)];
        T rv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            RowCoef rc;
            if (pre) row_coef<1>(g, i, j, k0 + q, rc);
            else row_coef<0>(g, i, j, k0 + q, rc);
            const double xv[7] = {xxm[q], xym[q], q > 0 ? xc[q - 1] : zlo, xc[q],
                                  q < 3 ? xc[q + 1] : zhi, xyp[q], xxp[q]};
            double ax = 0.0;
#pragma unroll
            for (int t = 0; t < 7; ++t) ax += rc.has[t] ? rc.v[t] * xv[t] : 0.0;
            const double rr = bv[q] - ax;
            rv[q] = (T)rr;
            acc[0] += bv[q] * bv[q];
            acc[1] += rr * rr;
            acc[2] += (double)rv[q] * (double)rv[q];
        }
        Vec4T<T>::store(r32 + n0, rv);
        Vec4T<T>::store(r0_32 + n0, rv);
    }
    block_sums<3>(acc, part, g.nblocks, L);
}

// x += e (the correction, accumulated in FP64 by the inner solve), keeping x in xb so
// that a pass found to raise the residual can be dropped with x restored bit for bit
__global__ void k_mixed_update(int n, double *__restrict__ x, const double *__restrict__ e,
                               double *__restrict__ xb)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const double xo = x[q];
        xb[q] = xo;
        x[q] = xo + e[q];
    }
}

// the start of a multigrid pressure solve from the last two solutions: x <- 2 x - xp,
// xp <- x (the previous solution, kept for the next step)
__global__ void k_extrapolate(int n, double *__restrict__ x, double *__restrict__ xp)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const double xo = x[q];
        x[q] = 2.0 * xo - xp[q];
        xp[q] = xo;
    }
}

__global__ void __launch_bounds__(kThreads)
k_mixed_fold(const double *part, int nb, double *out)
{
    double t[3];
    fold_partials<3>(part, nb, nb, t);
    if (threadIdx.x == 0) {
        out[0] = t[0];
        out[1] = t[1];
        out[2] = t[2];
    }
}

// end of a batch: the convergence test of the last iteration m (bicgstabRe2.hpp:129-153)
__global__ void __launch_bounds__(kThreads)
k_bicg_check(Geo g, const double *partC, KState *st, int m)
{
    if (st->done) return;
    double t[2];
    fold_partials<2>(partC, g.nmarch, g.nmarch, t);
    if (threadIdx.x == 0) {
        const double res = sqrt(t[0]) / st->norm0;
        st->res = res;
        if (res < st->tol) {
            st->done = 1;
            st->iter = m;
        } else if (m >= st->maxit) {
            st->done = 2;
            st->iter = m;
        }
    }
}

// scalars after the initial residual (bicgstabRe2.hpp:88-105)
__global__ void __launch_bounds__(kThreads)
k_bicg_fin0(const double *part, int nb, KState *st)
{
    double t[2];
    fold_partials<2>(part, nb, nb, t);
    if (threadIdx.x == 0) {
        st->timestep += 1;
        st->norm0 = sqrt(t[0]); // L2Norm(rhs)
        st->rho[0] = t[1];      // dot(r1, r0)
        st->iter = 0;
        st->done = 0;
        st->mode = 2;
    }
}

// NaN/Inf reset after the iteration limit (bicgstabRe2.hpp:133-153)
__global__ void k_bicg_nanfix(int n, double *x, const double *px1, int *count)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double v = x[m];
    if (isnan(v) || isinf(v)) {
        x[m] = px1[m];
        atomicAdd(count, 1);
    }
}

__global__ void k_copy(int n, const double *a, double *b)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) b[m] = a[m];
}

// ---------------------------------------------------------------------------
// Red-black SOR with the coefficients and update of SorPipeLine_seq
// (SorPipeLine.seq.hpp:40-123); ghosts of p hold the last boundary values.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_sor_rhs(Geo g, const double *us, const double *vs, const double *ws, double rdt, double *cf6)
{
    const int L = logical_block(g);
    if (L >= g.nblocks) return;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        const int I = i + kGC, J = j + kGC, K = k + kGC;
        const int ic = gidx(g, I, J, K);
        const double mChange = (us[ic] - us[ic - g.sx]) * Y.D[J] * Z.D[K] +
                               (vs[ic] - vs[ic - g.sy]) * X.D[I] * Z.D[K] +
                               (ws[ic] - ws[ic - 1]) * X.D[I] * Y.D[J];
        cf6[n] = mChange * rdt;
    }
}

__global__ void __launch_bounds__(kThreads)
k_sor_half(Geo g, int colour, double omega, const double *__restrict__ cf6,
           double *__restrict__ p, double *__restrict__ part, const SorState *st)
{
    if (st->done) return;
    const int L = logical_block(g);
    if (L >= g.nblocks) return;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    double mx = 0.0;
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        const int I = i + kGC, J = j + kGC, K = k + kGC;
        if (((I + J + K) & 1) != colour) continue;
        const int ic = gidx(g, I, J, K);
        const double cf0 = divr(Y.D[J] * Z.D[K], X.Ds[I], X.rDs[I]);
        const double cf1 = divr(Y.D[J] * Z.D[K], X.Ds[I - 1], X.rDs[I - 1]);
        const double cf2 = divr(X.D[I] * Z.D[K], Y.Ds[J], Y.rDs[J]);
        const double cf3 = divr(X.D[I] * Z.D[K], Y.Ds[J - 1], Y.rDs[J - 1]);
        const double cf4 = divr(X.D[I] * Y.D[J], Z.Ds[K], Z.rDs[K]);
        const double cf5 = divr(X.D[I] * Y.D[J], Z.Ds[K - 1], Z.rDs[K - 1]);
        const double cf7 = -1.0 / (cf0 + cf1 + cf2 + cf3 + cf4 + cf5);
        const int kp = (g.zper && K == g.nz - 3) ? ic + 1 - g.nzc : ic + 1;
        const int km = (g.zper && K == kGC) ? ic - 1 + g.nzc : ic - 1;
        const double pc = p[ic];
        const double pNEW = (-p[ic + g.sx] * cf0 - p[ic - g.sx] * cf1 - p[ic + g.sy] * cf2 -
                             p[ic - g.sy] * cf3 - p[kp] * cf4 - p[km] * cf5 + cf6[n]) *
                            cf7;
        const double pChange = fabs(pNEW - pc);
        p[ic] = pc + omega * (pNEW - pc);
        mx = pChange > mx ? pChange : mx;
    }
    __shared__ double red[4];
    const double wm = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = red[q] > m ? red[q] : m;
        part[colour * g.nblocks + L] = m;
    }
}

__global__ void __launch_bounds__(kThreads)
k_sor_fin(const double *part, int nb, SorState *st)
{
    if (st->done) return;
    __shared__ double red[4];
    double mx = 0.0;
    for (int m = threadIdx.x; m < 2 * nb; m += kThreads) mx = part[m] > mx ? part[m] : mx;
    const double wm = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = red[q] > m ? red[q] : m;
        st->residual = m;
        st->iters += 1;
        if (!(m > st->tol && st->iters < st->itmax)) st->done = 1;
    }
}

// ---------------------------------------------------------------------------
// Red-black SOR on colour-split storage (even nzc).  Colour c's cells of interior row
// (i, j) are k = 2m + q with q = (i + j + c) & 1; they are stored at (i nyc + j) nzc/2
// + m of that colour's half (c * N/2 + ...).  Every neighbour of a cell has the other
// colour: its x and y neighbours are the other half's entries at the same m of rows
// (i +- 1, j) and (i, j +- 1), its z neighbours are entries m - 1 + q and m + q of its
// own row (wrapping with periodic z).  A half sweep reads the other colour and its own
// colour's p and right-hand side and writes its own colour: whole cache lines of only
// the values it uses (16 bytes per cell of the grid per half sweep).  Neighbours
// outside the interior are the ghost p, fixed during the solve, as in
// SorPipeLine_seq.  Coefficients and update are k_sor_half's, so results are the same.
// ---------------------------------------------------------------------------
struct RBIdx {
    int i, j, k, m, q, row, ic;
};

__device__ __forceinline__ RBIdx rb_decode(const Geo &g, const FastDiv &dh, int colour, int idx)
{
    RBIdx r;
    const int hz = g.nzc >> 1;
    r.row = (int)fdiv((unsigned)idx, dh);
    r.m = idx - r.row * hz;
    r.i = (int)fdiv((unsigned)r.row, g.dy);
    r.j = r.row - r.i * g.nyc;
    r.q = (r.i + r.j + colour) & 1;
    r.k = 2 * r.m + r.q;
    r.ic = gidx(g, r.i + kGC, r.j + kGC, r.k + kGC);
    return r;
}

// right-hand side cf6 (k_sor_rhs's) and p of both colours, from the natural order
__global__ void __launch_bounds__(kThreads)
k_sor_split(Geo g, FastDiv dh, const double *__restrict__ us, const double *__restrict__ vs,
            const double *__restrict__ ws, double rdt, const double *__restrict__ pg,
            double *__restrict__ cf6c, double *__restrict__ pc)
{
    const int half = g.N >> 1;
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= 2 * half) return;
    const int colour = t >= half ? 1 : 0, idx = t - colour * half;
    const RBIdx r = rb_decode(g, dh, colour, idx);
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const int I = r.i + kGC, J = r.j + kGC, K = r.k + kGC, ic = r.ic;
    const double mChange = (us[ic] - us[ic - g.sx]) * Y.D[J] * Z.D[K] +
                           (vs[ic] - vs[ic - g.sy]) * X.D[I] * Z.D[K] +
                           (ws[ic] - ws[ic - 1]) * X.D[I] * Y.D[J];
    cf6c[t] = mChange * rdt;
    pc[t] = pg[ic];
}

// p of both colours back into the natural order
__global__ void __launch_bounds__(kThreads)
k_sor_merge(Geo g, FastDiv dh, const double *__restrict__ pc, double *__restrict__ pg)
{
    const int half = g.N >> 1;
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= 2 * half) return;
    const int colour = t >= half ? 1 : 0, idx = t - colour * half;
    const RBIdx r = rb_decode(g, dh, colour, idx);
    pg[r.ic] = pc[t];
}

// one block updates kRBCells * 256 cells of its colour, the cells of a thread 256
// apart (each load instruction of a wave covers consecutive entries)
constexpr int kRBCells = 4;

// CONSTC: every metric table is one constant (the uniform grid), so the six face
// coefficients and -1/sum are the same for every cell; the host forms them once with the
// same expressions (sc), which gives the same values as forming them per cell.
template <bool CONSTC>
__global__ void __launch_bounds__(kThreads)
k_sor_rb(Geo g, FastDiv dh, int colour, double omega, SorCoef sc,
         const double *__restrict__ cf6c, double *__restrict__ pc,
         const double *__restrict__ pg, double *__restrict__ part, const SorState *st)
{
    if (st->done) return;
    const int half = g.N >> 1, hz = g.nzc >> 1;
    const int nb = (half + kRBCells * kThreads - 1) / (kRBCells * kThreads);
    const int L = dealt_block();
    if (L >= nb) return;
    const double *__restrict__ other = pc + (colour ^ 1) * half;
    double *__restrict__ own = pc + colour * half;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const int cyh = g.nyc * hz;
    double mx = 0.0;
#pragma unroll
    for (int c = 0; c < kRBCells; ++c) {
        const int idx = (L * kRBCells + c) * kThreads + threadIdx.x;
        if (idx < half) {
            const RBIdx r = rb_decode(g, dh, colour, idx);
            const int I = r.i + kGC, J = r.j + kGC, K = r.k + kGC, ic = r.ic;
            const int rb = r.row * hz;
            const double pE = r.i < g.nxc - 1 ? other[idx + cyh] : pg[ic + g.sx];
            const double pW = r.i > 0 ? other[idx - cyh] : pg[ic - g.sx];
            const double pN = r.j < g.nyc - 1 ? other[idx + hz] : pg[ic + g.sy];
            const double pS = r.j > 0 ? other[idx - hz] : pg[ic - g.sy];
            const double pT = r.k < g.nzc - 1 ? other[idx + r.q]
                                              : (g.zper ? other[rb] : pg[ic + 1]);
            const double pB = r.k > 0 ? other[idx - 1 + r.q]
                                      : (g.zper ? other[rb + hz - 1] : pg[ic - 1]);
            const double p0 = own[idx];
            const double rhs = cf6c[idx];
            double cf0, cf1, cf2, cf3, cf4, cf5, cf7;
            if (CONSTC) {
                cf0 = sc.c[0], cf1 = sc.c[1], cf2 = sc.c[2], cf3 = sc.c[3];
                cf4 = sc.c[4], cf5 = sc.c[5], cf7 = sc.c[6];
            } else {
                cf0 = divr(Y.D[J] * Z.D[K], X.Ds[I], X.rDs[I]);
                cf1 = divr(Y.D[J] * Z.D[K], X.Ds[I - 1], X.rDs[I - 1]);
                cf2 = divr(X.D[I] * Z.D[K], Y.Ds[J], Y.rDs[J]);
                cf3 = divr(X.D[I] * Z.D[K], Y.Ds[J - 1], Y.rDs[J - 1]);
                cf4 = divr(X.D[I] * Y.D[J], Z.Ds[K], Z.rDs[K]);
                cf5 = divr(X.D[I] * Y.D[J], Z.Ds[K - 1], Z.rDs[K - 1]);
                cf7 = -1.0 / (cf0 + cf1 + cf2 + cf3 + cf4 + cf5);
            }
            const double pNEW =
                (-pE * cf0 - pW * cf1 - pN * cf2 - pS * cf3 - pT * cf4 - pB * cf5 + rhs) * cf7;
            const double d = fabs(pNEW - p0);
            mx = d > mx ? d : mx;
            own[idx] = p0 + omega * (pNEW - p0);
        }
    }
    __shared__ double red[4];
    const double wm = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = red[q] > m ? red[q] : m;
        part[colour * nb + L] = m;
    }
}

// The same half sweep with two neighbouring cells of a row per thread (nzc a multiple of
// 4, so a row holds whole pairs): their own, right-hand-side and x and y neighbour values
// are read as one 16-byte vector each, and their z neighbours as the pair of the other
// colour beside them plus one value, which halves the load instructions.  Each cell's
// update is k_sor_rb's, term for term.
template <bool CONSTC>
__global__ void __launch_bounds__(kThreads)
k_sor_rb2(Geo g, FastDiv dh, int colour, double omega, SorCoef sc,
          const double *__restrict__ cf6c, double *__restrict__ pc,
          const double *__restrict__ pg, double *__restrict__ part, const SorState *st)
{
    if (st->done) return;
    const int half = g.N >> 1, hz = g.nzc >> 1;
    const int nb = (half + kRBCells * kThreads - 1) / (kRBCells * kThreads);
    const int L = dealt_block();
    if (L >= nb) return;
    const double *__restrict__ other = pc + (colour ^ 1) * half;
    double *__restrict__ own = pc + colour * half;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const int cyh = g.nyc * hz;
    double mx = 0.0;
#pragma unroll
    for (int c = 0; c < kRBCells / 2; ++c) {
        const int idx = 2 * ((L * (kRBCells / 2) + c) * kThreads + threadIdx.x);
        if (idx < half) {
            const RBIdx r = rb_decode(g, dh, colour, idx);
            const int rb = r.row * hz, q = r.q;
            const int icA = r.ic, icB = r.ic + 2; // the pair's cells are k and k + 2
            const double2 p0 = *reinterpret_cast<const double2 *>(own + idx);
            const double2 rhs = *reinterpret_cast<const double2 *>(cf6c + idx);
            double2 pE, pW, pN, pS;
            if (r.i < g.nxc - 1) pE = *reinterpret_cast<const double2 *>(other + idx + cyh);
            else pE = make_double2(pg[icA + g.sx], pg[icB + g.sx]);
            if (r.i > 0) pW = *reinterpret_cast<const double2 *>(other + idx - cyh);
            else pW = make_double2(pg[icA - g.sx], pg[icB - g.sx]);
            if (r.j < g.nyc - 1) pN = *reinterpret_cast<const double2 *>(other + idx + hz);
            else pN = make_double2(pg[icA + g.sy], pg[icB + g.sy]);
            if (r.j > 0) pS = *reinterpret_cast<const double2 *>(other + idx - hz);
            else pS = make_double2(pg[icA - g.sy], pg[icB - g.sy]);
            // z: the other colour's pair at this m, and the value before (q = 0) or after
            // (q = 1) it
            const double2 oz = *reinterpret_cast<const double2 *>(other + idx);
            double pBA, pTA, pBB, pTB;
            if (q == 0) {
                pBA = r.m > 0 ? other[idx - 1] : (g.zper ? other[rb + hz - 1] : pg[icA - 1]);
                pTA = oz.x;
                pBB = oz.x;
                pTB = oz.y;
            } else {
                pBA = oz.x;
                pTA = oz.y;
                pBB = oz.y;
                pTB = r.m + 1 < hz - 1 ? other[idx + 2] : (g.zper ? other[rb] : pg[icB + 1]);
            }
            const double pe[2] = {pE.x, pE.y}, pw[2] = {pW.x, pW.y}, pn[2] = {pN.x, pN.y},
                         ps[2] = {pS.x, pS.y}, pt[2] = {pTA, pTB}, pb[2] = {pBA, pBB},
                         po[2] = {p0.x, p0.y}, rh[2] = {rhs.x, rhs.y};
            double out[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                double cf0, cf1, cf2, cf3, cf4, cf5, cf7;
                if (CONSTC) {
                    cf0 = sc.c[0], cf1 = sc.c[1], cf2 = sc.c[2], cf3 = sc.c[3];
                    cf4 = sc.c[4], cf5 = sc.c[5], cf7 = sc.c[6];
                } else {
                    const int I = r.i + kGC, J = r.j + kGC, K = r.k + kGC + 2 * h;
                    cf0 = divr(Y.D[J] * Z.D[K], X.Ds[I], X.rDs[I]);
                    cf1 = divr(Y.D[J] * Z.D[K], X.Ds[I - 1], X.rDs[I - 1]);
                    cf2 = divr(X.D[I] * Z.D[K], Y.Ds[J], Y.rDs[J]);
                    cf3 = divr(X.D[I] * Z.D[K], Y.Ds[J - 1], Y.rDs[J - 1]);
                    cf4 = divr(X.D[I] * Y.D[J], Z.Ds[K], Z.rDs[K]);
                    cf5 = divr(X.D[I] * Y.D[J], Z.Ds[K - 1], Z.rDs[K - 1]);
                    cf7 = -1.0 / (cf0 + cf1 + cf2 + cf3 + cf4 + cf5);
                }
                const double pNEW = (-pe[h] * cf0 - pw[h] * cf1 - pn[h] * cf2 - ps[h] * cf3 -
                                     pt[h] * cf4 - pb[h] * cf5 + rh[h]) *
                                    cf7;
                const double d = fabs(pNEW - po[h]);
                mx = d > mx ? d : mx;
                out[h] = po[h] + omega * (pNEW - po[h]);
            }
            *reinterpret_cast<double2 *>(own + idx) = make_double2(out[0], out[1]);
        }
    }
    __shared__ double red[4];
    const double wm = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = red[q] > m ? red[q] : m;
        part[colour * nb + L] = m;
    }
}

// end of a colour-split sweep: max |dp| over the 2 nb partials, with eight loads in
// flight per thread
__global__ void __launch_bounds__(kThreads)
k_sor_rb_fin(const double *part, int nb2, SorState *st)
{
    if (st->done) return;
    __shared__ double red[4];
    double mx = 0.0;
    for (int m0 = threadIdx.x; m0 < nb2; m0 += 8 * kThreads) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int m = m0 + u * kThreads;
            v[u] = m < nb2 ? part[m] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) mx = v[u] > mx ? v[u] : mx;
    }
    const double wm = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = red[q] > m ? red[q] : m;
        st->residual = m;
        st->iters += 1;
        if (!(m > st->tol && st->iters < st->itmax)) st->done = 1;
    }
}

// ---------------------------------------------------------------------------
// update_UandF_omp (updateUandF.hpp:90-180) fused with Pressure_transform_X_result
// (Transform.hpp:8-25) and the drag/lift partial sums of virtualF_Int
// (virtualForceIntergrator.hpp:79-86).  FROM_X: pressure read from X_result.
// ---------------------------------------------------------------------------
// One cell's three faces (updateUandF.hpp:90-180 with Transform.hpp's pressure): the
// corrected velocity T2 = U* - dt dp/dn, the masked value (1 - eta_face) T2 (the solid
// velocity is 0) and the forcing (new - T2) / dt.  pc, pe, pn, pt: pressure at the cell and
// its +x, +y, +z neighbours; ec..et: the mask likewise; us, vs, ws: U* at the cell.
struct ProjIn {
    double pc, pe, pn, pt, ec, ee, en, et, us, vs, ws;
};
struct ProjFaces {
    double un, vn, wn, fx, fy, fz;
};
__device__ __forceinline__ ProjFaces project_faces(const Geo &g, int I, int J, int K, double dt,
                                                   double rdt, const ProjIn &a)
{
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const double uS = 0.0, vS = 0.0, wS = 0.0;
    ProjFaces o;
    // x face
    const double T2u = a.us - divr(dt * (a.pe - a.pc), X.Ds[I], X.rDs[I]);
    const double eu = a.ec + (a.ee - a.ec) * X.IT[I];
    o.un = a.ec * uS + (1.0 - eu) * T2u;
    o.fx = divr(o.un - T2u, dt, rdt);
    // y face
    const double T2v = a.vs - divr(dt * (a.pn - a.pc), Y.Ds[J], Y.rDs[J]);
    const double ev = a.ec + (a.en - a.ec) * Y.IT[J];
    o.vn = a.ec * vS + (1.0 - ev) * T2v;
    o.fy = divr(o.vn - T2v, dt, rdt);
    // z face
    const double T2w = a.ws - divr(dt * (a.pt - a.pc), Z.Ds[K], Z.rDs[K]);
    const double ew = a.ec + (a.et - a.ec) * Z.IT[K];
    o.wn = a.ec * wS + (1.0 - ew) * T2w;
    o.fz = divr(o.wn - T2w, dt, rdt);
    return o;
}

// DIFF: also the steady-state check of CheckSteadyStatebyMaxVal_omp
// (CheckSteadystate.hpp:247-307), max(T0 - T3) per component starting at 0, over the
// faces written here, from the previous field u0, v0, w0: block maxima into dpart
template <int FROM_X, int DIFF>
__global__ void __launch_bounds__(kThreads)
k_project(Geo g, double dt, double rdt, int convective, const double *__restrict__ xres,
          double *__restrict__ p, const double *__restrict__ eta,
          const double *__restrict__ us, const double *__restrict__ vs,
          const double *__restrict__ ws, double *__restrict__ u, double *__restrict__ v,
          double *__restrict__ w, double *__restrict__ f, double *__restrict__ part,
          const double *__restrict__ u0, const double *__restrict__ v0,
          const double *__restrict__ w0, double *__restrict__ dpart)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    double acc[2] = {0.0, 0.0};
    double dm[3] = {0.0, 0.0, 0.0};
    // Every load is unconditional: where a face is skipped (the last interior x or y
    // plane, the last z plane without periodicity) its neighbour index falls back to the
    // cell itself and the face's values are computed and dropped.  Only the stores and
    // the partial sums depend on the face tests, and each stored value is the same
    // expression as the reference's, so the result is unchanged; the cell's loads are
    // then all in flight together instead of one face after another.
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        const int I = i + kGC, J = j + kGC, K = k + kGC;
        const int ic = gidx(g, I, J, K);
        const bool ox = I < g.nx - 3, oy = J < g.ny - 3, oz = K < g.nz - 3 || g.zper;
        const bool wrap = (K == g.nz - 3);
        const int icE = ox ? ic + g.sx : ic, icN = oy ? ic + g.sy : ic;
        const int icP = oz ? (wrap ? ic - (g.nzc - 1) : ic + 1) : ic;
        const double pc = FROM_X ? xres[n] : p[ic];
        const double pe = FROM_X ? xres[ox ? n + g.cyz : n] : p[icE];
        const double pn = FROM_X ? xres[oy ? n + g.nzc : n] : p[icN];
        const double pt = FROM_X ? xres[oz ? (wrap ? n - (g.nzc - 1) : n + 1) : n] : p[icP];
        const ProjFaces o = project_faces(g, I, J, K, dt, rdt,
                                          {pc, pe, pn, pt, eta[ic], eta[icE], eta[icN],
                                           eta[icP], us[ic], vs[ic], ws[ic]});
        if (FROM_X) p[ic] = pc;
        if (ox) {
            u[ic] = o.un;
            f[ic] = o.fx;
        } else if (convective) {
            u[ic] = us[ic];
        }
        if (oy) {
            v[ic] = o.vn;
            f[ic + g.G] = o.fy;
        }
        if (oz) {
            w[ic] = o.wn;
            f[ic + 2 * g.G] = o.fz;
        }
        // CD_CL sums every interior entry of f: a face not written here keeps its value
        const double fx = ox ? o.fx : f[ic], fy = oy ? o.fy : f[ic + g.G];
        const double vol = X.D[I] * Y.D[J] * Z.D[K];
        acc[0] += fx * vol;
        acc[1] += fy * vol;
        if (DIFF) {
            const double du = u0[ic] - o.un, dv = v0[ic] - o.vn, dw = w0[ic] - o.wn;
            if (ox && du > dm[0]) dm[0] = du;
            if (oy && dv > dm[1]) dm[1] = dv;
            if (oz && dw > dm[2]) dm[2] = dw;
        }
    }
    block_sums<2>(acc, part, g.nblocks, L);
    if (DIFF) block_maxes<3>(dm, dpart, g.nblocks, L);
}

// The same projection with four consecutive cells of a z-row per thread (nzc a multiple
// of 4: a thread's cells share i and j, and the first one's interior and ghosted indices
// are even).  The pressure, mask and velocity values of the cells and of their +x / +y
// neighbours, and every store, move as 16-byte vectors; the +z neighbours are the cells
// beside each other plus one value.  Each cell's values come from project_faces, as in
// k_project.  The drag partials keep k_project's order: the per-cell terms go through LDS
// and each thread adds the cells k_project gives it (c * kThreads + thread), so the
// partial sums, and Cd and Cl, are the same bits.
// GW > 0 also stores the z ghosts (k = 0, 1 and nz-2, nz-1) of the rows it writes, so
// that each row goes out whole: HBM takes rows with unwritten gaps between them ~1.4x
// slower.  The ghosts of u, v, w and p get values the boundary conditions of the same
// step replace (k_bc_z rewrites them, and nothing reads them before); GW == 2 also
// stores f's ghosts, with the zeros the caller knows they hold.
template <int FROM_X, int DIFF, int GW>
__global__ void __launch_bounds__(kThreads)
k_project4(Geo g, double dt, double rdt, int convective, const double *__restrict__ xres,
           double *__restrict__ p, const double *__restrict__ eta,
           const double *__restrict__ us, const double *__restrict__ vs,
           const double *__restrict__ ws, double *__restrict__ u, double *__restrict__ v,
           double *__restrict__ w, double *__restrict__ f, double *__restrict__ part,
           const double *__restrict__ u0, const double *__restrict__ v0,
           const double *__restrict__ w0, double *__restrict__ dpart)
{
    const int L = dealt_block();
    if (L >= g.nblocks) return;
    __shared__ double sfx[kPerBlock], sfy[kPerBlock];
    const Axis &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    double dm[3] = {0.0, 0.0, 0.0};
    const int n0 = L * kPerBlock + 4 * threadIdx.x;
    if (n0 < g.N) {
        int i, j, k0;
        decode(g, n0, i, j, k0);
        const int I = i + kGC, J = j + kGC, K0 = k0 + kGC;
        const int ic = gidx(g, I, J, K0);
        const bool ox = I < g.nx - 3, oy = J < g.ny - 3;
        // the last cell's +z face: inside the row, wrapped (periodic z) or skipped
        const bool last = K0 + 3 == g.nz - 3;
        const bool oz3 = !last || g.zper;
        const int dz3 = last ? (g.zper ? 3 - (g.nzc - 1) : 3) : 4;
        const int icE = ox ? ic + g.sx : ic, icN = oy ? ic + g.sy : ic;
        double pc[4], pe[4], pn[4], pt3;
        if (FROM_X) {
            load4(xres + n0, pc);
            load4(xres + (ox ? n0 + g.cyz : n0), pe);
            load4(xres + (oy ? n0 + g.nzc : n0), pn);
            pt3 = xres[n0 + dz3];
        } else {
            load4(p + ic, pc);
            load4(p + icE, pe);
            load4(p + icN, pn);
            pt3 = p[ic + dz3];
        }
        double ec[4], ee[4], en[4], uc[4], vc[4], wc[4];
        load4(eta + ic, ec);
        load4(eta + icE, ee);
        load4(eta + icN, en);
        const double et3 = eta[ic + dz3];
        load4(us + ic, uc);
        load4(vs + ic, vc);
        load4(ws + ic, wc);
        // CD_CL sums every interior entry of f: on the last x (y) plane, whose x (y) faces
        // are not written here, the sums take the values f holds
        double fxo[4] = {0.0, 0.0, 0.0, 0.0}, fyo[4] = {0.0, 0.0, 0.0, 0.0};
        if (!ox) load4(f + ic, fxo);
        if (!oy) load4(f + g.G + ic, fyo);
        double un[4], vn[4], wn[4], fx[4], fy[4], fz[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const ProjFaces o = project_faces(
                g, I, J, K0 + q, dt, rdt,
                {pc[q], pe[q], pn[q], q < 3 ? pc[q + 1] : pt3, ec[q], ee[q], en[q],
                 q < 3 ? ec[q + 1] : et3, uc[q], vc[q], wc[q]});
            un[q] = o.un, vn[q] = o.vn, wn[q] = o.wn;
            fx[q] = o.fx, fy[q] = o.fy, fz[q] = o.fz;
            const double vol = X.D[I] * Y.D[J] * Z.D[K0 + q];
            sfx[4 * threadIdx.x + q] = (ox ? o.fx : fxo[q]) * vol;
            sfy[4 * threadIdx.x + q] = (oy ? o.fy : fyo[q]) * vol;
        }
        if (FROM_X) Vec4T<double>::store(p + ic, pc);
        if (ox) {
            Vec4T<double>::store(u + ic, un);
            Vec4T<double>::store(f + ic, fx);
        } else if (convective) {
            Vec4T<double>::store(u + ic, uc);
        }
        if (oy) {
            Vec4T<double>::store(v + ic, vn);
            Vec4T<double>::store(f + g.G + ic, fy);
        }
        if (oz3) {
            Vec4T<double>::store(w + ic, wn);
            Vec4T<double>::store(f + 2 * g.G + ic, fz);
        } else {
            reinterpret_cast<double2 *>(w + ic)[0] = make_double2(wn[0], wn[1]);
            w[ic + 2] = wn[2];
            reinterpret_cast<double2 *>(f + 2 * g.G + ic)[0] = make_double2(fz[0], fz[1]);
            f[2 * g.G + ic + 2] = fz[2];
        }
        if (GW) {
            // the two ghosts before the row's first cell and after its last
            for (int side = 0; side < 2; ++side) {
                if (side == 0 ? K0 != kGC : !last) continue;
                const int o = side == 0 ? ic - 2 : ic + 4, e = side == 0 ? 0 : 3;
                auto put = [o](double *a, double val) {
                    *reinterpret_cast<double2 *>(a + o) = make_double2(val, val);
                };
                if (FROM_X) put(p, pc[e]);
                if (ox) put(u, un[e]);
                else if (convective) put(u, uc[e]);
                if (oy) put(v, vn[e]);
                put(w, wn[e]);
                if (GW == 2) {
                    if (ox) put(f, 0.0);
                    if (oy) put(f + g.G, 0.0);
                    put(f + 2 * g.G, 0.0);
                }
            }
        }
        if (DIFF) {
            double a[4], b[4], c[4];
            load4(u0 + ic, a);
            load4(v0 + ic, b);
            load4(w0 + ic, c);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double du = a[q] - un[q], dv = b[q] - vn[q], dw = c[q] - wn[q];
                if (ox && du > dm[0]) dm[0] = du;
                if (oy && dv > dm[1]) dm[1] = dv;
                if ((q < 3 || oz3) && dw > dm[2]) dm[2] = dw;
            }
        }
    }
    __syncthreads();
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int c = 0; c < kCells; ++c) {
        const int t = c * kThreads + threadIdx.x;
        if (L * kPerBlock + t < g.N) {
            acc[0] += sfx[t];
            acc[1] += sfy[t];
        }
    }
    block_sums<2>(acc, part, g.nblocks, L);
    if (DIFF) block_maxes<3>(dm, dpart, g.nblocks, L);
}

// the three maxima of the steady-state check from the block maxima
__global__ void __launch_bounds__(kThreads) k_max3_fin(const double *part, int nb, double *out)
{
    __shared__ double red[3][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < 3; ++c) {
        double m = 0.0;
        for (int q = threadIdx.x; q < nb; q += kThreads) m = part[c * nb + q] > m ? part[c * nb + q] : m;
        const double w = wave_max(m);
        if (lane == 0) red[c][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) {
            double m = red[c][0];
            for (int q = 1; q < 4; ++q) m = red[c][q] > m ? red[c][q] : m;
            out[c] = m;
        }
}

__global__ void __launch_bounds__(kThreads)
k_cdcl_fin(const double *part, int nb, double lz, double *out)
{
    double t[2];
    fold_partials<2>(part, nb, nb, t);
    if (threadIdx.x == 0) {
        out[0] = -2.0 * t[0] / lz;
        out[1] = -2.0 * t[1] / lz;
    }
}

// drag/lift from a stored force field (CD_CL on its own)
__global__ void __launch_bounds__(kThreads)
k_force_partial(Geo g, const double *f, double *part)
{
    const int L = logical_block(g);
    if (L >= g.nblocks) return;
    double acc[2] = {0.0, 0.0};
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        const int I = i + kGC, J = j + kGC, K = k + kGC;
        const int ic = gidx(g, I, J, K);
        const double vol = g.ax[0].D[I] * g.ax[1].D[J] * g.ax[2].D[K];
        acc[0] += f[ic] * vol;
        acc[1] += f[ic + g.G] * vol;
    }
    block_sums<2>(acc, part, g.nblocks, L);
}

__global__ void __launch_bounds__(kThreads)
k_max_fin(const double *part, int nb, double *out)
{
    __shared__ double red[4];
    double mx = -1.0e300;
    for (int m = threadIdx.x; m < nb; m += kThreads) mx = part[m] > mx ? part[m] : mx;
    const double wm = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = red[q] > m ? red[q] : m;
        *out = m;
    }
}

// ---------------------------------------------------------------------------
// DFIB_CylinderZ (setEta.hpp:136-204): area fraction of each interior column from
// nsub x nsub sub-cell midpoints, then the k = 0 plane copied to k = 1..nz-2.
// ---------------------------------------------------------------------------
__global__ void k_eta_cylz(Geo g, const double *Xn, const double *Yn, double cx, double cy,
                           double R, int nsub, double inv_nsub2, double *eta)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.nxc * g.nyc) return;
    const int I = m / g.nyc + kGC, J = m % g.nyc + kGC;
    const double Dx = g.ax[0].D[I], Dy = g.ax[1].D[J];
    const double Diagonal = sqrt(Dx * Dx + Dy * Dy) * 0.5;
    const double ax = ((Xn[I] + Xn[I + 1]) * 0.5) - cx;
    const double ay = ((Yn[J] + Yn[J + 1]) * 0.5) - cy;
    const double Distance = sqrt(ax * ax + ay * ay);
    double e;
    if (fabs(Distance - R) < Diagonal) {
        const double dyg = Dy / (double)nsub;
        const double dxg = Dx / (double)nsub;
        long xi = 0;
        for (int ii = 0; ii < nsub; ++ii) {
            const double sx0 = Xn[I] + (double)ii * dxg, sx1 = Xn[I] + (double)(ii + 1) * dxg;
            const double dx = (sx0 + sx1) * 0.5 - cx;
            for (int jj = 0; jj < nsub; ++jj) {
                const double sy0 = Yn[J] + (double)jj * dyg,
                             sy1 = Yn[J] + (double)(jj + 1) * dyg;
                const double dy = (sy0 + sy1) * 0.5 - cy;
                if (sqrt(dx * dx + dy * dy) <= R) ++xi;
            }
        }
        e = (double)xi * inv_nsub2;
    } else if (Distance > R) {
        e = 0.;
    } else {
        e = 1.;
    }
    eta[(I * g.ny + J) * g.nz] = e;
}

__global__ void k_eta_copyk(Geo g, double *eta)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (g.nx - 1) * (g.ny - 1)) return;
    const int I = m / (g.ny - 1), J = m % (g.ny - 1);
    const int b = (I * g.ny + J) * g.nz;
    const double e = eta[b];
    for (int k = 1; k < g.nz - 1; ++k) eta[b + k] = e;
}

// seeded initial field (no reference counterpart)
__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_init_flow(Geo g, double u0, double amp, unsigned long long seed, double *u,
                            double *v, double *w, double *p)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.G) return;
    const int K = c % g.nz, t = c / g.nz, J = t % g.ny, I = t / g.ny;
    double uu = u0;
    if (I >= kGC && I < g.nx - kGC && J >= kGC && J < g.ny - kGC && K >= kGC && K < g.nz - kGC) {
        const uint64_t n = (uint64_t)(((I - kGC) * g.nyc + (J - kGC)) * g.nzc + (K - kGC));
        const uint64_t h = splitmix64((uint64_t)seed ^ n);
        const double U = (double)(h >> 11) * 0x1.0p-53;
        uu = u0 + amp * (2.0 * U - 1.0);
    }
    u[c] = uu;
    v[c] = 0.0;
    w[c] = 0.0;
    p[c] = 0.0;
}

__global__ void k_fill(int n, double val, double *a)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) a[m] = val;
}

// Pressure_transform_X_result (Transform.hpp:8-25)
__global__ void __launch_bounds__(kThreads)
k_transform(Geo g, const double *__restrict__ x, double *__restrict__ p)
{
    const int L = logical_block(g);
    if (L >= g.nblocks) return;
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n >= g.N) break;
        int i, j, k;
        decode(g, n, i, j, k);
        p[gidx(g, i + kGC, j + kGC, k + kGC)] = x[n];
    }
}

__global__ void k_set_int(int *a, int val) { *a = val; }

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int grid_blocks(const Geo &g)
{
    return g.xcd_remap ? ((g.nblocks + 7) / 8) * 8 : g.nblocks;
}
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

void launch_smagorinsky(const Geo &g, hipStream_t s, const double *u, const double *v,
                        const double *w, const double *coef, double *viseff)
{
    hipLaunchKernelGGL(k_smagorinsky, dim3((g.N + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       s, g, u, v, w, coef, viseff);
}

void launch_convdiff(const Geo &g, hipStream_t s, const double *u, const double *v,
                     const double *w, double *us, double *vs, double *ws, double dt,
                     double nu, double *cfl_part, const double *viseff)
{
    const int nb = g.xcd_remap ? ((g.cd_nb + 7) / 8) * 8 : g.cd_nb;
    const int P = (g.cd_tj + 4) * (g.cd_tk + 4);
    const size_t lds = cd_lds_bytes(g);
#define CFDMX_CD(NE, LES, ZP2)                                                                \
    hipLaunchKernelGGL((k_convdiff<NE, LES, ZP2>), dim3(nb), dim3(kThreads), lds, s, g, u, v, w, \
                       us, vs, ws, -dt, nu * dt, dt, cfl_part, nu, viseff)
    const bool les = viseff != nullptr, z2 = g.zpow2 != 0;
    if (P <= 2 * kThreads) {
        if (les) {
            if (z2) CFDMX_CD(2, true, true);
            else CFDMX_CD(2, true, false);
        } else {
            if (z2) CFDMX_CD(2, false, true);
            else CFDMX_CD(2, false, false);
        }
    } else {
        if (les) {
            if (z2) CFDMX_CD(3, true, true);
            else CFDMX_CD(3, true, false);
        } else {
            if (z2) CFDMX_CD(3, false, true);
            else CFDMX_CD(3, false, false);
        }
    }
#undef CFDMX_CD
}

void launch_bc(const Geo &g, hipStream_t s, const BCParams &bc, int advance, double *u,
               double *v, double *w, double *p, double *u1, double *v1, double *w1,
               double *flux_part)
{
    hipLaunchKernelGGL(k_bc_x, dim3(cdiv(g.cyz, 256)), dim3(256), 0, s, g, bc, advance, u, v, w,
                       p, u1, v1, w1);
    if (bc.flux_correction) {
        const int nb = cdiv(g.cyz, kPerBlock);
        hipLaunchKernelGGL(k_flux_partial, dim3(nb), dim3(kThreads), 0, s, g, u, flux_part, nb);
        hipLaunchKernelGGL(k_flux_apply, dim3(cdiv(g.cyz, kThreads)), dim3(kThreads), 0, s, g,
                           flux_part, nb, u, u1);
    }
    hipLaunchKernelGGL(k_bc_y, dim3(cdiv(g.nxc * g.nzc, 256)), dim3(256), 0, s, g, bc, u, v, w,
                       p, u1, v1, w1);
    const int nzcols = g.zper ? g.nx * g.ny : g.nxc * g.nyc;
    hipLaunchKernelGGL(k_bc_z, dim3(cdiv(nzcols, 256)), dim3(256), 0, s, g, bc, u, v, w, p, u1,
                       v1, w1);
}

void launch_bc_copy(const Geo &g, hipStream_t s, const double *u0, const double *v0,
                    const double *w0, double *u1, double *v1, double *w1)
{
    const int tot = g.cyz + g.nxc * g.nzc + g.nxc * g.nyc;
    hipLaunchKernelGGL(k_bc_copy, dim3(cdiv(tot, 256)), dim3(256), 0, s, g, u0, v0, w0, u1, v1,
                       w1);
}

void launch_rhs(const Geo &g, hipStream_t s, int pre, const double *us, const double *vs,
                const double *ws, double dt, double *b)
{
    const double ddt = 1.0 / dt;
    static_assert(kPerBlock == 4 * kThreads, "k_rhs4 covers a block's cells four per thread");
    if (g.nzc % 4 == 0) {
        if (pre)
            hipLaunchKernelGGL(k_rhs4<1>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, us, vs,
                               ws, ddt, b);
        else
            hipLaunchKernelGGL(k_rhs4<0>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, us, vs,
                               ws, ddt, b);
    } else if (pre)
        hipLaunchKernelGGL(k_rhs<1>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, us, vs, ws,
                           ddt, b);
    else
        hipLaunchKernelGGL(k_rhs<0>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, us, vs, ws,
                           ddt, b);
}

void launch_apply(const Geo &g, hipStream_t s, int pre, const double *x, double *r)
{
    if (pre)
        hipLaunchKernelGGL(k_apply<1>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, x, r);
    else
        hipLaunchKernelGGL(k_apply<0>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, x, r);
}

void launch_apply_t(const Geo &g, hipStream_t s, int pre, const double *y, double *r)
{
    if (pre)
        hipLaunchKernelGGL(k_apply_t<1>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, y, r);
    else
        hipLaunchKernelGGL(k_apply_t<0>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, y, r);
}

void launch_bicg_init(const Geo &g, hipStream_t s, int pre, const double *b, const double *x,
                      double *r0, double *r, double *px1, double *part, KState *st)
{
    if (g.nzc % 4 == 0)
        hipLaunchKernelGGL(k_bicg_init4, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, pre, b,
                           x, r0, r, px1, part, st);
    else
        hipLaunchKernelGGL(k_bicg_init, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, pre, b,
                           x, r0, r, px1, part, st);
}

static inline int march_blocks(const Geo &g)
{
    return g.xcd_remap ? ((g.nmarch + 7) / 8) * 8 : g.nmarch;
}
template <typename T>
static inline size_t march_lds(const Geo &g)
{
    return ((size_t)2 * kChunk + 2 * march_rows(g.nzc) + 2 * g.nzc) * sizeof(double) +
           (size_t)2 * (kTile + 2 * g.nzc) * sizeof(T);
}

// halo cells per thread: ceil(2 nzc / 256); 1 covers nzc <= 128, 5 the largest
// supported nzc (600)
// the streaming hint (NT in k_march) when the iteration's seven vectors r, q, r0, p,
// Ap, As, x exceed the 256 MiB memory-side cache
template <typename T>
static bool march_streaming(const Geo &g)
{
    return (size_t)7 * (size_t)g.N * sizeof(T) > ((size_t)256 << 20);
}
template <int PHASE, int NH, bool FULL, bool NT, typename T>
static void launch_march_nt(const Geo &g, hipStream_t s, int pre, const MarchIO<T> &io,
                            KState *st, int m, const double *part_in)
{
    if (pre)
        hipLaunchKernelGGL((k_march<PHASE, 1, NH, FULL, NT, T>), dim3(march_blocks(g)),
                           dim3(kThreads), march_lds<T>(g), s, g, io, st, m, part_in);
    else
        hipLaunchKernelGGL((k_march<PHASE, 0, NH, FULL, NT, T>), dim3(march_blocks(g)),
                           dim3(kThreads), march_lds<T>(g), s, g, io, st, m, part_in);
}
template <int PHASE, int NH, bool FULL, typename T>
static void launch_march_full(const Geo &g, hipStream_t s, int pre, const MarchIO<T> &io,
                              KState *st, int m, const double *part_in)
{
    if (march_streaming<T>(g))
        launch_march_nt<PHASE, NH, FULL, true>(g, s, pre, io, st, m, part_in);
    else
        launch_march_nt<PHASE, NH, FULL, false>(g, s, pre, io, st, m, part_in);
}
template <int PHASE, int NH, typename T>
static void launch_march_nh(const Geo &g, hipStream_t s, int pre, const MarchIO<T> &io,
                            KState *st, int m, const double *part_in)
{
    if (g.cyz % kTile == 0)
        launch_march_full<PHASE, NH, true>(g, s, pre, io, st, m, part_in);
    else
        launch_march_full<PHASE, NH, false>(g, s, pre, io, st, m, part_in);
}
template <int PHASE, typename T>
static void launch_march(const Geo &g, hipStream_t s, int pre, const MarchIO<T> &io, KState *st,
                         int m, const double *part_in)
{
    if (2 * g.nzc <= kThreads)
        launch_march_nh<PHASE, 1>(g, s, pre, io, st, m, part_in);
    else
        launch_march_nh<PHASE, 5>(g, s, pre, io, st, m, part_in);
}
template <typename T>
void launch_bicg_A(const Geo &g, hipStream_t s, int pre, BicgBuf<T> &B, double *const part[3],
                   KState *st, int m)
{
    MarchIO<T> io = {};
    io.r = B.r;
    io.q = B.q;
    io.r0 = B.r0;
    io.out_p = B.p;
    io.out_ap = B.ap;
    io.part = part[0];
    launch_march<0>(g, s, pre, io, st, m, part[2]);
}
template <typename T>
void launch_bicg_B(const Geo &g, hipStream_t s, int pre, BicgBuf<T> &B, double *const part[3],
                   KState *st, int m)
{
    MarchIO<T> io = {};
    io.r = B.r;
    io.ap = B.ap;
    io.out_as = B.as;
    io.part = part[1];
    launch_march<1>(g, s, pre, io, st, m, part[0]);
}
template <typename T>
void launch_bicg_C(const Geo &g, hipStream_t s, BicgBuf<T> &B, T *x, double *const part[3],
                   KState *st, int m)
{
    (void)m;
    if (march_streaming<T>(g))
        hipLaunchKernelGGL((k_march_C<true, T>), dim3(march_blocks(g)), dim3(kThreads), 0, s, g,
                           B.r, B.ap, B.q, B.p, B.as, B.r0, x, (const double *)part[1],
                           part[2], st);
    else
        hipLaunchKernelGGL((k_march_C<false, T>), dim3(march_blocks(g)), dim3(kThreads), 0, s,
                           g, B.r, B.ap, B.q, B.p, B.as, B.r0, x,
                           (const double *)part[1], part[2], st);
}
#define CFDMX_BICG_INSTANCES(T)                                                               \
    template void launch_bicg_A<T>(const Geo &, hipStream_t, int, BicgBuf<T> &,               \
                                   double *const[3], KState *, int);                          \
    template void launch_bicg_B<T>(const Geo &, hipStream_t, int, BicgBuf<T> &,               \
                                   double *const[3], KState *, int);                          \
    template void launch_bicg_C<T>(const Geo &, hipStream_t, BicgBuf<T> &, T *,               \
                                   double *const[3], KState *, int);
CFDMX_BICG_INSTANCES(double)
#undef CFDMX_BICG_INSTANCES

void launch_bicg_check(const Geo &g, hipStream_t s, double *const part[3], KState *st, int m)
{
    hipLaunchKernelGGL(k_bicg_check, dim3(1), dim3(kThreads), 0, s, g, (const double *)part[2],
                       st, m);
}

template <typename T>
void launch_mixed_resid(const Geo &g, hipStream_t s, int pre, const double *b, const double *x,
                        T *r, T *r0, double *part)
{
    if (g.nzc % 4 == 0 && kPerBlock == 4 * kThreads)
        hipLaunchKernelGGL(k_mixed_resid4<T>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, pre,
                           b, x, r, r0, part);
    else
        hipLaunchKernelGGL(k_mixed_resid<T>, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, pre,
                           b, x, r, r0, part);
}
void launch_mixed_update(const Geo &g, hipStream_t s, double *x, const double *e, double *xb)
{
    hipLaunchKernelGGL(k_mixed_update, dim3(2048), dim3(kThreads), 0, s, g.N, x, e, xb);
}
template void launch_mixed_resid<float>(const Geo &, hipStream_t, int, const double *,
                                        const double *, float *, float *, double *);
template void launch_mixed_resid<double>(const Geo &, hipStream_t, int, const double *,
                                         const double *, double *, double *, double *);
void launch_extrapolate(const Geo &g, hipStream_t s, double *x, double *xp)
{
    hipLaunchKernelGGL(k_extrapolate, dim3(2048), dim3(kThreads), 0, s, g.N, x, xp);
}

void launch_mixed_fold(const Geo &g, hipStream_t s, const double *part, double *out)
{
    hipLaunchKernelGGL(k_mixed_fold, dim3(1), dim3(kThreads), 0, s, part, g.nblocks, out);
}
void launch_bicg_fin0(const Geo &g, hipStream_t s, double *part, KState *st)
{
    hipLaunchKernelGGL(k_bicg_fin0, dim3(1), dim3(kThreads), 0, s, part, g.nblocks, st);
}

void launch_nanfix(hipStream_t s, int n, double *x, const double *px1, int *count)
{
    hipLaunchKernelGGL(k_bicg_nanfix, dim3(cdiv(n, 256)), dim3(256), 0, s, n, x, px1, count);
}

void launch_copy(hipStream_t s, int n, const double *a, double *b)
{
    hipLaunchKernelGGL(k_copy, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b);
}

void launch_fill(hipStream_t s, int n, double val, double *a)
{
    hipLaunchKernelGGL(k_fill, dim3(cdiv(n, 256)), dim3(256), 0, s, n, val, a);
}

void launch_sor_rhs(const Geo &g, hipStream_t s, const double *us, const double *vs,
                    const double *ws, double dt, double *cf6)
{
    hipLaunchKernelGGL(k_sor_rhs, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, us, vs, ws,
                       1.0 / dt, cf6);
}

void launch_sor_half(const Geo &g, hipStream_t s, int colour, double omega, const double *cf6,
                     double *p, double *part, SorState *st)
{
    hipLaunchKernelGGL(k_sor_half, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, colour, omega,
                       cf6, p, part, st);
}

// colour-split red-black SOR (even nzc): cf6c and pc hold N doubles each (red half,
// then black); part 2 * rb_blocks(g) doubles
int sor_rb_blocks(const Geo &g) { return cdiv(g.N >> 1, kRBCells * kThreads); }

void launch_sor_split(const Geo &g, hipStream_t s, const double *us, const double *vs,
                      const double *ws, double dt, const double *p, double *cf6c, double *pc)
{
    const FastDiv dh = make_fastdiv((unsigned)(g.nzc >> 1));
    hipLaunchKernelGGL(k_sor_split, dim3(cdiv(g.N, kThreads)), dim3(kThreads), 0, s, g, dh, us,
                       vs, ws, 1.0 / dt, p, cf6c, pc);
}

void launch_sor_merge(const Geo &g, hipStream_t s, const double *pc, double *p)
{
    const FastDiv dh = make_fastdiv((unsigned)(g.nzc >> 1));
    hipLaunchKernelGGL(k_sor_merge, dim3(cdiv(g.N, kThreads)), dim3(kThreads), 0, s, g, dh, pc,
                       p);
}

void launch_sor_rb(const Geo &g, hipStream_t s, int colour, double omega, const SorCoef &sc,
                   const double *cf6c, double *pc, const double *p, double *part, SorState *st)
{
    const FastDiv dh = make_fastdiv((unsigned)(g.nzc >> 1));
    const int nb = sor_rb_blocks(g);
    const int grid = g.xcd_remap ? ((nb + 7) / 8) * 8 : nb;
    const double *rhs = cf6c + colour * (g.N >> 1);
    if (g.nzc % 4 == 0) { // whole pairs per row: two cells per thread-step
        if (sc.constant)
            hipLaunchKernelGGL(k_sor_rb2<true>, dim3(grid), dim3(kThreads), 0, s, g, dh, colour,
                               omega, sc, rhs, pc, p, part, st);
        else
            hipLaunchKernelGGL(k_sor_rb2<false>, dim3(grid), dim3(kThreads), 0, s, g, dh, colour,
                               omega, sc, rhs, pc, p, part, st);
    } else if (sc.constant) {
        hipLaunchKernelGGL(k_sor_rb<true>, dim3(grid), dim3(kThreads), 0, s, g, dh, colour,
                           omega, sc, rhs, pc, p, part, st);
    } else {
        hipLaunchKernelGGL(k_sor_rb<false>, dim3(grid), dim3(kThreads), 0, s, g, dh, colour,
                           omega, sc, rhs, pc, p, part, st);
    }
}

void launch_sor_rb_fin(const Geo &g, hipStream_t s, double *part, SorState *st)
{
    hipLaunchKernelGGL(k_sor_rb_fin, dim3(1), dim3(kThreads), 0, s, part, 2 * sor_rb_blocks(g),
                       st);
}

void launch_sor_fin(const Geo &g, hipStream_t s, double *part, SorState *st)
{
    hipLaunchKernelGGL(k_sor_fin, dim3(1), dim3(kThreads), 0, s, part, g.nblocks, st);
}

void launch_project(const Geo &g, hipStream_t s, int from_x, double dt, int convective,
                    const double *xres, double *p, const double *eta, const double *us,
                    const double *vs, const double *ws, double *u, double *v, double *w,
                    double *f, double *part, const double *const *prev, double *dpart,
                    int ghosts)
{
    const double rdt = 1.0 / dt;
    const dim3 grid(grid_blocks(g)), block(kThreads);
    const double *u0 = prev ? prev[0] : nullptr, *v0 = prev ? prev[1] : nullptr,
                 *w0 = prev ? prev[2] : nullptr;
#define CFDMX_PROJECT(KERNEL, ...)                                                            \
    hipLaunchKernelGGL((KERNEL<__VA_ARGS__>), grid, block, 0, s, g, dt, rdt, convective,      \
                       xres, p, eta, us, vs, ws, u, v, w, f, part, u0, v0, w0, dpart)
#define CFDMX_PROJECT4(FX, DF)                                                                \
    do {                                                                                      \
        if (ghosts == 2) CFDMX_PROJECT(k_project4, FX, DF, 2);                                \
        else if (ghosts == 1) CFDMX_PROJECT(k_project4, FX, DF, 1);                           \
        else CFDMX_PROJECT(k_project4, FX, DF, 0);                                            \
    } while (0)
    const int df = dpart ? 1 : 0;
    if (g.nzc % 4 == 0) { // four cells per thread (k_project4 covers a block four per thread)
        if (from_x && df) CFDMX_PROJECT4(1, 1);
        else if (from_x) CFDMX_PROJECT4(1, 0);
        else if (df) CFDMX_PROJECT4(0, 1);
        else CFDMX_PROJECT4(0, 0);
    } else {
        if (from_x && df) CFDMX_PROJECT(k_project, 1, 1);
        else if (from_x) CFDMX_PROJECT(k_project, 1, 0);
        else if (df) CFDMX_PROJECT(k_project, 0, 1);
        else CFDMX_PROJECT(k_project, 0, 0);
    }
#undef CFDMX_PROJECT4
#undef CFDMX_PROJECT
}

void launch_max3_fin(const Geo &g, hipStream_t s, const double *dpart, double *out)
{
    hipLaunchKernelGGL(k_max3_fin, dim3(1), dim3(kThreads), 0, s, dpart, g.nblocks, out);
}

void launch_cdcl_fin(const Geo &g, hipStream_t s, const double *part, double lz, double *out)
{
    hipLaunchKernelGGL(k_cdcl_fin, dim3(1), dim3(kThreads), 0, s, part, g.nblocks, lz, out);
}

void launch_force_partial(const Geo &g, hipStream_t s, const double *f, double *part)
{
    hipLaunchKernelGGL(k_force_partial, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, f, part);
}

void launch_max_fin(const Geo &g, hipStream_t s, const double *part, double *out)
{
    hipLaunchKernelGGL(k_max_fin, dim3(1), dim3(kThreads), 0, s, part, g.cd_nb, out);
}

void launch_eta_cylz(const Geo &g, hipStream_t s, const double *Xn, const double *Yn, double cx,
                     double cy, double R, int nsub, double inv_nsub2, double *eta)
{
    hipLaunchKernelGGL(k_eta_cylz, dim3(cdiv(g.nxc * g.nyc, 64)), dim3(64), 0, s, g, Xn, Yn, cx,
                       cy, R, nsub, inv_nsub2, eta);
    hipLaunchKernelGGL(k_eta_copyk, dim3(cdiv((g.nx - 1) * (g.ny - 1), 256)), dim3(256), 0, s, g,
                       eta);
}

void launch_init_flow(const Geo &g, hipStream_t s, double u0, double amp,
                      unsigned long long seed, double *u, double *v, double *w, double *p)
{
    hipLaunchKernelGGL(k_init_flow, dim3(cdiv(g.G, 256)), dim3(256), 0, s, g, u0, amp, seed, u,
                       v, w, p);
}

void launch_transform(const Geo &g, hipStream_t s, const double *x, double *p)
{
    hipLaunchKernelGGL(k_transform, dim3(grid_blocks(g)), dim3(kThreads), 0, s, g, x, p);
}

void launch_set_int(hipStream_t s, int *a, int val)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, a, val);
}

} // namespace cfdmx
