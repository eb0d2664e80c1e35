// cfdmx_adjoint.hip -- transposed (adjoint) kernels of the time step for gradients of
// the time-averaged drag coefficient (no reference counterpart: the reference solver
// has no adjoint).  Every transpose is written in gather form (each thread owns the
// entries it writes), so results are deterministic without atomics.
//
// Forward map of one step (cfdmx_ctx.hip, step_once):
//   U* = QUICK conv-diff of U on the computed faces, copies of U on the entries the
//        boundary copy sets, periodic z-images        (Convection_Difussion.hpp:174-450)
//   b = -(1/dt) D U*;  x = A^+ b (BiCGSTAB, taken as an exact solve)
//   T2 = U* - dt G x;  U' = (1 - eta_f) T2;  f = -eta_f T2 / dt   (updateUandF.hpp:90-180)
//   Cd = -2/lz sum f_x dV;  U' <- BC(U')                (Boundary_Condition.hpp:37-176)
// The conv-diff transpose stores, per computed face, the 21 partial derivatives of
// the update (the upwind branch of each QUICK face value frozen at the forward
// state), then gathers them transposed.
#include "cfdmx_internal.hpp"

namespace cfdmx {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ void split(const Geo &g, int e, int &I, int &J, int &K)
{
    I = e / g.sx;
    const int r = e - I * g.sx;
    J = r / g.nz;
    K = r - J * g.nz;
}

// ---------------------------------------------------------------------------
// boundary-condition transposes, in the reverse of the forward face order
// ---------------------------------------------------------------------------
__device__ __forceinline__ void adj_copy(double *l, int dst, int src)
{
    const double t = l[dst];
    l[dst] = 0.0;
    l[src] += t;
}

// transpose of l[g1] = l[g2] = c + w * l[src]; returns the adjoint of the value
__device__ __forceinline__ double adj_ghost2(double *l, int g1, int g2, int src, double w)
{
    const double t = l[g1] + l[g2];
    l[g1] = 0.0;
    l[g2] = 0.0;
    l[src] += w * t;
    return t;
}

// periodic z-images (four ghosts per column), all columns or the interior ones
__global__ void k_adj_periodic(Geo g, int all_columns, double *l0, double *l1, double *l2)
{
    const int lo = all_columns ? 0 : kGC;
    const int ncx = g.nx - 2 * lo, ncy = g.ny - 2 * lo;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncx * ncy) return;
    const int i = t / ncy + lo, j = t % ncy + lo;
    const int c = gidx(g, i, j, 0), nz = g.nz, nzc = g.nzc;
    double *L[3] = {l2, l1, l0};
    for (int q = 0; q < 3; ++q) {
        double *l = L[q];
        adj_copy(l, c + nz - 1, c + nz - 1 - nzc);
        adj_copy(l, c + nz - 2, c + nz - 2 - nzc);
        adj_copy(l, c + 0, c + 0 + nzc);
        adj_copy(l, c + 1, c + 1 + nzc);
    }
}

// faces 5 then 4 (walls in z), one thread per interior (i, j)
__global__ void k_adj_bc_z(Geo g, BCParams bc, double *lu, double *lv, double *lw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxc * g.nyc) return;
    const int i = t / g.nyc + kGC, j = t % g.nyc + kGC;
    const int c = gidx(g, i, j, 0), nz = g.nz;
    adj_ghost2(lv, c + nz - 1, c + nz - 2, c + nz - 3, 2.0 * bc.num[5][1] - 1.0);
    adj_ghost2(lu, c + nz - 1, c + nz - 2, c + nz - 3, 2.0 * bc.num[5][0] - 1.0);
    {
        const double s = lw[c + nz - 1] + lw[c + nz - 2] + lw[c + nz - 3];
        lw[c + nz - 1] = lw[c + nz - 2] = lw[c + nz - 3] = 0.0;
        lw[c + nz - 4] += bc.num[5][2] * s;
    }
    adj_ghost2(lw, c + 0, c + 1, c + 2, bc.num[4][2]);
    adj_ghost2(lv, c + 0, c + 1, c + 2, 2.0 * bc.num[4][1] - 1.0);
    adj_ghost2(lu, c + 0, c + 1, c + 2, 2.0 * bc.num[4][0] - 1.0);
}

// faces 3 then 2, one thread per interior (i, k)
__global__ void k_adj_bc_y(Geo g, BCParams bc, double *lu, double *lv, double *lw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxc * g.nzc) return;
    const int i = t / g.nzc + kGC, k = t % g.nzc + kGC;
    const int ny = g.ny, sy = g.sy;
    const int c = gidx(g, i, 0, k);
    adj_ghost2(lw, c + (ny - 1) * sy, c + (ny - 2) * sy, c + (ny - 3) * sy,
               2.0 * bc.num[3][2] - 1.0);
    adj_ghost2(lu, c + (ny - 1) * sy, c + (ny - 2) * sy, c + (ny - 3) * sy,
               2.0 * bc.num[3][0] - 1.0);
    {
        const double s = lv[c + (ny - 1) * sy] + lv[c + (ny - 2) * sy] + lv[c + (ny - 3) * sy];
        lv[c + (ny - 1) * sy] = lv[c + (ny - 2) * sy] = lv[c + (ny - 3) * sy] = 0.0;
        lv[c + (ny - 4) * sy] += bc.num[3][1] * s;
    }
    adj_ghost2(lw, c, c + sy, c + 2 * sy, 2.0 * bc.num[2][2] - 1.0);
    adj_ghost2(lv, c, c + sy, c + 2 * sy, bc.num[2][1]);
    adj_ghost2(lu, c, c + sy, c + 2 * sy, 2.0 * bc.num[2][0] - 1.0);
}

// flux correction u_b <- u_b + (q_in - q_out) / area: per outlet cell the adjoint of
// u_b (kept in l[nx-3]) and the partial sums of the correction's adjoint
__global__ void __launch_bounds__(kThreads)
k_adj_flux_t(Geo g, double *lu, double *part, int nb)
{
    const int L = blockIdx.x;
    double v[1] = {0.0};
    for (int c = 0; c < kCells; ++c) {
        const int m = L * kPerBlock + c * kThreads + threadIdx.x;
        if (m >= g.cyz) break;
        const int J = m / g.nzc + kGC, K = m % g.nzc + kGC;
        const int e3 = gidx(g, g.nx - 3, J, K);
        const double t = lu[e3] + lu[e3 + g.sx] + lu[e3 + 2 * g.sx];
        lu[e3 + g.sx] = 0.0;
        lu[e3 + 2 * g.sx] = 0.0;
        lu[e3] = t;
        v[0] += t;
    }
    block_sums<1>(v, part, nb, L);
}

__global__ void __launch_bounds__(kThreads)
k_adj_flux_apply(Geo g, const double *part, int nb, double rarea, double *lu)
{
    double t[1];
    fold_partials<1>(part, nb, nb, t);
    const double lc = t[0] * rarea;
    for (int m = blockIdx.x * kThreads + threadIdx.x; m < g.cyz; m += gridDim.x * kThreads) {
        const int J = m / g.nzc + kGC, K = m % g.nzc + kGC;
        const double a = g.ax[1].D[J] * g.ax[2].D[K];
        lu[gidx(g, 1, J, K)] += lc * a;
        lu[gidx(g, g.nx - 3, J, K)] -= lc * a;
    }
}

// faces 1 then 0, one thread per interior (j, k); tin receives the adjoint of the
// inflow value of face 0 per (j, k)
__global__ void k_adj_bc_x(Geo g, BCParams bc, int advance, double *lu, double *lv, double *lw,
                           double *tin)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.cyz) return;
    const int J = t / g.nzc + kGC, K = t % g.nzc + kGC;
    const int sx = g.sx;
    const int e1 = gidx(g, g.nx - 1, J, K), e2 = e1 - sx, e3 = e1 - 2 * sx, e4 = e1 - 3 * sx;
    adj_ghost2(lw, e1, e2, e3, 2.0 * bc.num[1][2] - 1.0);
    adj_ghost2(lv, e1, e2, e3, 2.0 * bc.num[1][1] - 1.0);
    {
        const double s = lu[e1] + lu[e2] + lu[e3];
        lu[e1] = lu[e2] = lu[e3] = 0.0;
        if (bc.convective && advance) {
            lu[e3] += (1.0 - bc.conv_coef) * s;
            lu[e4] += bc.conv_coef * s;
        } else {
            lu[e4] += bc.num[1][0] * s;
        }
    }
    const int f0 = gidx(g, 0, J, K);
    adj_ghost2(lw, f0, f0 + sx, f0 + 2 * sx, 2.0 * bc.num[0][2] - 1.0);
    adj_ghost2(lv, f0, f0 + sx, f0 + 2 * sx, 2.0 * bc.num[0][1] - 1.0);
    tin[t] = adj_ghost2(lu, f0, f0 + sx, f0 + 2 * sx, bc.num[0][0]);
}

// inflow adjoint per ghosted j: the sum over k of the face-0 values (in k order)
__global__ void k_adj_inflow(Geo g, const double *tin, double *lin)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.nyc) return;
    double s = 0.0;
    for (int k = 0; k < g.nzc; ++k) s += tin[j * g.nzc + k];
    lin[j + kGC] += s;
}

// ---------------------------------------------------------------------------
// projection and IB forcing transposes
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool computed_face(const Geo &g, int d, int I, int J, int K)
{
    const bool in_i = I >= kGC && I < g.nx - kGC - (d == 0 ? 1 : 0);
    const bool in_j = J >= kGC && J < g.ny - kGC - (d == 1 ? 1 : 0);
    const bool in_k = K >= kGC && K < g.nz - kGC - ((d == 2 && !g.zper) ? 1 : 0);
    return in_i && in_j && in_k;
}

// per ghosted entry and component: the adjoints of T2 (written as the adjoint of U*)
// and of eta_f on the projected faces, zero elsewhere; the carried outlet face
__global__ void k_adj_project_faces(Geo g, double dt, int use_x, int convective, double wf,
                                    const double *x, const double *p, const double *eta,
                                    const double *us, const double *vs, const double *ws,
                                    const double *l3u, const double *l3v, const double *l3w,
                                    double *lsu, double *lsv, double *lsw, double *leu,
                                    double *lev, double *lew)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.G) return;
    int I, J, K;
    split(g, e, I, J, K);
    const double *S[3] = {us, vs, ws};
    const double *L3[3] = {l3u, l3v, l3w};
    double *LS[3] = {lsu, lsv, lsw};
    double *LE[3] = {leu, lev, lew};
    auto pv = [&](int i, int j, int k, int ee) -> double {
        return use_x ? x[((i - kGC) * g.nyc + (j - kGC)) * g.nzc + (k - kGC)] : p[ee];
    };
    for (int d = 0; d < 3; ++d) {
        double lt = 0.0, le = 0.0;
        if (computed_face(g, d, I, J, K)) {
            int ip = I, jp = J, kp = K;
            if (d == 0) ip = I + 1;
            if (d == 1) jp = J + 1;
            if (d == 2) kp = (g.zper && K == g.nz - kGC - 1) ? kGC : K + 1;
            const int Id = d == 0 ? I : d == 1 ? J : K;
            const int eP = gidx(g, ip, jp, kp);
            const double T2 = S[d][e] - dt * (pv(ip, jp, kp, eP) - pv(I, J, K, e)) / g.ax[d].Ds[Id];
            const double IT = g.ax[d].IT[Id];
            const double ef = eta[e] + (eta[eP] - eta[e]) * IT;
            const double lf = d == 0 ? wf * g.ax[0].D[I] * g.ax[1].D[J] * g.ax[2].D[K] : 0.0;
            lt = (1.0 - ef) * L3[d][e] - ef / dt * lf;
            le = -T2 * L3[d][e] - T2 / dt * lf;
        } else if (d == 0 && convective && I == g.nx - 3 && J >= kGC && J < g.ny - kGC &&
                   K >= kGC && K < g.nz - kGC) {
            lt = L3[0][e]; // the outlet face carries U*
        }
        LS[d][e] = lt;
        LE[d][e] = le;
    }
}

// per interior cell: the pressure adjoint lx (gathered from the six faces) and the
// eta adjoint, accumulated over steps
__global__ void k_adj_project_cells(Geo g, double dt, int use_x, const double *lsu,
                                    const double *lsv, const double *lsw, const double *leu,
                                    const double *lev, const double *lew, double *lx,
                                    double *leta)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N) return;
    int i, j, k;
    decode(g, n, i, j, k);
    const int I = i + kGC, J = j + kGC, K = k + kGC;
    const int e = gidx(g, I, J, K);
    const double *LS[3] = {lsu, lsv, lsw};
    const double *LE[3] = {leu, lev, lew};
    const int idx[3] = {I, J, K};
    const int str[3] = {g.sx, g.sy, 1};
    double sx = 0.0, se = 0.0;
    for (int d = 0; d < 3; ++d) {
        const int Id = idx[d];
        // the face on the + side of the cell
        if (computed_face(g, d, I, J, K)) {
            sx += dt / g.ax[d].Ds[Id] * LS[d][e];
            se += (1.0 - g.ax[d].IT[Id]) * LE[d][e];
        }
        // the face on the - side (periodic z: the face nz-3 for the cell k = 2)
        int em = -1, Im = Id - 1;
        if (Id - 1 >= kGC) {
            em = e - str[d];
        } else if (d == 2 && g.zper) {
            Im = g.nz - kGC - 1;
            em = e + (Im - Id);
        }
        if (em >= 0) {
            sx -= dt / g.ax[d].Ds[Im] * LS[d][em];
            se += g.ax[d].IT[Im] * LE[d][em];
        }
    }
    lx[n] = use_x ? sx : 0.0;
    leta[e] += se;
}

// transpose of the right-hand side b = -(1/dt) D U* (SPE path), added to the U*
// adjoint: per ghosted entry, the two cells whose divergence reads it
__global__ void k_adj_rhs(Geo g, double rdt, const double *lb, double *lsu, double *lsv,
                          double *lsw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.G) return;
    int I, J, K;
    split(g, e, I, J, K);
    double *LS[3] = {lsu, lsv, lsw};
    const int idx[3] = {I, J, K};
    const int lim[3] = {g.nx, g.ny, g.nz};
    auto interior = [&](int a, int b, int c) {
        return a >= kGC && a < g.nx - kGC && b >= kGC && b < g.ny - kGC && c >= kGC &&
               c < g.nz - kGC;
    };
    auto row = [&](int a, int b, int c) {
        return ((a - kGC) * g.nyc + (b - kGC)) * g.nzc + (c - kGC);
    };
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        if (interior(I, J, K)) s -= rdt * lb[row(I, J, K)] / g.ax[d].D[idx[d]];
        int I1 = I, J1 = J, K1 = K;
        if (d == 0) I1 += 1;
        if (d == 1) J1 += 1;
        if (d == 2) K1 += 1;
        if (idx[d] + 1 < lim[d] && interior(I1, J1, K1))
            s += rdt * lb[row(I1, J1, K1)] / g.ax[d].D[idx[d] + 1];
        LS[d][e] += s;
    }
}

// mean removal of the pressure adjoint (the minimum-norm solution of the singular
// system)
__global__ void __launch_bounds__(kThreads)
k_adj_sum_partial(Geo g, const double *a, double *part)
{
    const int L = logical_block(g);
    if (L >= g.nblocks) return;
    double v[1] = {0.0};
    for (int c = 0; c < kCells; ++c) {
        const int n = L * kPerBlock + c * kThreads + threadIdx.x;
        if (n < g.N) v[0] += a[n];
    }
    block_sums<1>(v, part, g.nblocks, L);
}

// one block: the mean into part[nblocks]
__global__ void __launch_bounds__(kThreads) k_adj_mean_fold(Geo g, double *part)
{
    double t[1];
    fold_partials<1>(part, g.nblocks, g.nblocks, t);
    if (threadIdx.x == 0) part[g.nblocks] = t[0] / (double)g.N;
}

__global__ void __launch_bounds__(kThreads)
k_adj_sub_mean(Geo g, const double *part, double *a)
{
    const double mean = part[g.nblocks];
    for (int n = blockIdx.x * kThreads + threadIdx.x; n < g.N; n += gridDim.x * kThreads)
        a[n] -= mean;
}

// ---------------------------------------------------------------------------
// conv-diff transpose
// ---------------------------------------------------------------------------
constexpr int kJSlots = 21;

// QUICK face-value weights over [f(+2), f(+1), f(0), f(-1), f(-2)] (QuickUnitA/B,
// Convection_Difussion.hpp:20-90) for the branch the transport velocity selects
__device__ void quick_weights(bool unitA, double Up, double Un, const double *Di,
                              const double *Ds, int I, double wp[5], double wn[5])
{
    for (int m = 0; m < 5; ++m) wp[m] = wn[m] = 0.0;
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

// Row slots of the update of component d at a face: 0 centre; 1..4 f at +1,+2,-1,-2
// along d; 5..8 along b1 = (d+1)%3; 9..12 along b2 = (d+2)%3; 13..16 the transport
// component b1 at {0, +s_d, -s_b1, -s_b1+s_d}; 17..20 b2 likewise.
__device__ __forceinline__ int own_slot(int base, int m) // m in {+1, +2, -1, -2}
{
    return base + (m == 1 ? 0 : m == 2 ? 1 : m == -1 ? 2 : 3);
}

__global__ void k_adj_cd_rows(Geo g, double dt, double nu, const double *u, const double *v,
                              const double *w, double *Jr)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.G) return;
    int idx[3];
    split(g, e, idx[0], idx[1], idx[2]);
    const double *U[3] = {u, v, w};
    const int str[3] = {g.sx, g.sy, 1};
    const size_t G = (size_t)g.G;
    for (int d = 0; d < 3; ++d) {
        double row[kJSlots];
        for (int q = 0; q < kJSlots; ++q) row[q] = 0.0;
        if (computed_face(g, d, idx[0], idx[1], idx[2])) {
            const double *f = U[d];
            const int Id = idx[d], sd = str[d];
            const Axis &Ad = g.ax[d];
            const double IT = Ad.IT[Id];
            auto addf = [&](int base, int m, double c) {
                if (m == 0) row[0] += c;
                else row[own_slot(base, m)] += c;
            };
            row[0] += 1.0; // identity
            {
                const double Up = 0.5 * (f[e + sd] + f[e]);
                const double Un = 0.5 * (f[e - sd] + f[e]);
                double wp[5], wn[5];
                quick_weights(true, Up, Un, Ad.D, Ad.Ds, Id, wp, wn);
                double fP = 0.0, fN = 0.0;
                for (int m = 0; m < 5; ++m) {
                    fP += wp[m] * f[e + (2 - m) * sd];
                    fN += wn[m] * f[e + (2 - m) * sd];
                }
                const double gP = -dt * 2.0 * fP / Ad.Ds[Id];
                const double gN = dt * 2.0 * fN / Ad.Ds[Id];
                for (int m = 0; m < 5; ++m) addf(1, 2 - m, wp[m] * gP + wn[m] * gN);
                const double cp = nu * dt / (Ad.D[Id + 1] * Ad.Ds[Id]);
                const double cm = nu * dt / (Ad.D[Id] * Ad.Ds[Id]);
                addf(1, 1, cp);
                addf(1, -1, cm);
                row[0] -= cp + cm;
            }
            for (int q = 1; q <= 2; ++q) {
                const int b = (d + q) % 3;
                const int Ib = idx[b], sb = str[b];
                const Axis &Ab = g.ax[b];
                const double *B = U[b];
                const int base = q == 1 ? 5 : 9, tb = q == 1 ? 13 : 17;
                const double Bp = B[e] + (B[e + sd] - B[e]) * IT;
                const double Bn = B[e - sb] + (B[e - sb + sd] - B[e - sb]) * IT;
                double wp[5], wn[5];
                quick_weights(false, Bp, Bn, Ab.D, Ab.Ds, Ib, wp, wn);
                double fP = 0.0, fN = 0.0;
                for (int m = 0; m < 5; ++m) {
                    fP += wp[m] * f[e + (2 - m) * sb];
                    fN += wn[m] * f[e + (2 - m) * sb];
                }
                const double gP = -dt * Bp / Ab.D[Ib];
                const double gN = dt * Bn / Ab.D[Ib];
                for (int m = 0; m < 5; ++m) addf(base, 2 - m, wp[m] * gP + wn[m] * gN);
                const double gBp = -dt * fP / Ab.D[Ib];
                const double gBn = dt * fN / Ab.D[Ib];
                row[tb + 0] += (1.0 - IT) * gBp;
                row[tb + 1] += IT * gBp;
                row[tb + 2] += (1.0 - IT) * gBn;
                row[tb + 3] += IT * gBn;
                const double cp = nu * dt / (Ab.Ds[Ib] * Ab.D[Ib]);
                const double cm = nu * dt / (Ab.Ds[Ib - 1] * Ab.D[Ib]);
                addf(base, 1, cp);
                addf(base, -1, cm);
                row[0] -= cp + cm;
            }
        }
        for (int q = 0; q < kJSlots; ++q) Jr[((size_t)d * kJSlots + q) * G + e] = row[q];
    }
}

// entries of U* that are copies of U (the boundary copy), less those the update and
// the periodic images overwrite
__device__ __forceinline__ bool copied(const Geo &g, int c, int I, int J, int K)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool iI = I >= kGC && I < nx - kGC, iJ = J >= kGC && J < ny - kGC,
               iK = K >= kGC && K < nz - kGC;
    if (g.zper && iI && iJ) {
        if (K < kGC || K >= nz - kGC) return false;
        if (c == 2 && K == nz - 3) return false;
    }
    bool in = false;
    if (iJ && iK) {
        if (I <= 1 || I >= nx - 2) in = true;
        if (I == nx - 3 && c == 0) in = true;
    }
    if (iI && iK) {
        if (J <= 1) in = true;
        if (J >= ny - 2 && c != 1) in = true;
        if (J >= ny - 3 && c == 1) in = true;
    }
    if (iI && iJ) {
        if (K <= 1) in = true;
        if (K >= nz - 2 && c != 2) in = true;
        if (K >= nz - 3 && c == 2) in = true;
    }
    return in;
}

// lU_c[e] = sum over the updates that read U_c[e] of (row entry) * (adjoint of U*),
// plus the copy
__global__ void k_adj_cd_gather(Geo g, const double *Jr, const double *lsu, const double *lsv,
                                const double *lsw, double *lu, double *lv, double *lw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.G) return;
    int I, J, K;
    split(g, e, I, J, K);
    const double *LS[3] = {lsu, lsv, lsw};
    double *LU[3] = {lu, lv, lw};
    const int str[3] = {g.sx, g.sy, 1};
    const size_t G = (size_t)g.G;
    auto term = [&](int d, int slot, int off) -> double {
        const int o = e - off;
        if (o < 0 || o >= g.G) return 0.0;
        return LS[d][o] * Jr[((size_t)d * kJSlots + slot) * G + o];
    };
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        // the update of component c reads c at 13 offsets
        s += term(c, 0, 0);
        const int b1 = (c + 1) % 3, b2 = (c + 2) % 3;
        const int axes[3] = {c, b1, b2}, bases[3] = {1, 5, 9};
        for (int a = 0; a < 3; ++a) {
            const int sa = str[axes[a]];
            s += term(c, own_slot(bases[a], 1), sa);
            s += term(c, own_slot(bases[a], 2), 2 * sa);
            s += term(c, own_slot(bases[a], -1), -sa);
            s += term(c, own_slot(bases[a], -2), -2 * sa);
        }
        // the updates of the other components read c as a transport velocity
        for (int d = 0; d < 3; ++d) {
            if (d == c) continue;
            const int tb = (c == (d + 1) % 3) ? 13 : 17;
            const int sd = str[d], sc = str[c];
            s += term(d, tb + 0, 0);
            s += term(d, tb + 1, sd);
            s += term(d, tb + 2, -sc);
            s += term(d, tb + 3, -sc + sd);
        }
        if (copied(g, c, I, J, K)) s += LS[c][e];
        LU[c][e] = s;
    }
}

// ---------------------------------------------------------------------------
// gradient assembly
// ---------------------------------------------------------------------------
// d eta / d (cx, cy, R) of the smoothed mask times the eta adjoint summed over k,
// partial sums over the (i, j) plane
__global__ void __launch_bounds__(kThreads)
k_adj_mask_partial(Geo g, const double *Xn, const double *Yn, double cx, double cy, double R,
                   double eps, const double *leta, double *part, int nb)
{
    const int L = blockIdx.x;
    const double pi = 3.14159265358979323846;
    double v[3] = {0.0, 0.0, 0.0};
    const int nplane = g.nxc * g.nyc;
    for (int c = 0; c < kCells; ++c) {
        const int m = L * kPerBlock + c * kThreads + threadIdx.x;
        if (m >= nplane) break;
        const int I = m / g.nyc + kGC, J = m % g.nyc + kGC;
        const double xc = (Xn[I] + Xn[I + 1]) * 0.5 - cx;
        const double yc = (Yn[J] + Yn[J + 1]) * 0.5 - cy;
        const double dist = sqrt(xc * xc + yc * yc);
        const double phi = R - dist;
        if (phi < -eps || phi > eps || dist == 0.0) continue;
        double le = 0.0;
        const int c0 = gidx(g, I, J, 0);
        for (int k = 0; k < g.nz; ++k) le += leta[c0 + k];
        const double h = 0.5 * (1.0 / eps + cos(pi * phi / eps) / eps);
        v[0] += le * h * xc / dist;
        v[1] += le * h * yc / dist;
        v[2] += le * h;
    }
    block_sums<3>(v, part, nb, L);
}

__global__ void __launch_bounds__(kThreads)
k_adj_fold3(const double *part, int nb, double *out)
{
    double t[3];
    fold_partials<3>(part, nb, nb, t);
    if (threadIdx.x == 0) {
        out[0] = t[0];
        out[1] = t[1];
        out[2] = t[2];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_adj_bc(const Geo &g, hipStream_t s, const BCParams &bc, int advance, double *lu,
                   double *lv, double *lw, double *tin, double *lin, double *part, double rarea)
{
    if (g.zper) {
        const int ncol = g.nx * g.ny;
        hipLaunchKernelGGL(k_adj_periodic, dim3(cdiv(ncol, 256)), dim3(256), 0, s, g, 1, lu, lv,
                           lw);
    } else {
        hipLaunchKernelGGL(k_adj_bc_z, dim3(cdiv(g.nxc * g.nyc, 256)), dim3(256), 0, s, g, bc, lu,
                           lv, lw);
    }
    hipLaunchKernelGGL(k_adj_bc_y, dim3(cdiv(g.nxc * g.nzc, 256)), dim3(256), 0, s, g, bc, lu, lv,
                       lw);
    if (bc.flux_correction) {
        const int nb = cdiv(g.cyz, kPerBlock);
        hipLaunchKernelGGL(k_adj_flux_t, dim3(nb), dim3(kThreads), 0, s, g, lu, part, nb);
        hipLaunchKernelGGL(k_adj_flux_apply, dim3(cdiv(g.cyz, kThreads)), dim3(kThreads), 0, s, g,
                           (const double *)part, nb, rarea, lu);
    }
    hipLaunchKernelGGL(k_adj_bc_x, dim3(cdiv(g.cyz, 256)), dim3(256), 0, s, g, bc, advance, lu,
                       lv, lw, tin);
    if (bc.inflow)
        hipLaunchKernelGGL(k_adj_inflow, dim3(cdiv(g.nyc, 64)), dim3(64), 0, s, g,
                           (const double *)tin, lin);
}

void launch_adj_project(const Geo &g, hipStream_t s, double dt, int use_x, int convective,
                        double wf, const double *x, const double *p, const double *eta,
                        double *const us[3], double *const l3[3], double *const ls[3],
                        double *const le[3], double *lx, double *leta)
{
    hipLaunchKernelGGL(k_adj_project_faces, dim3(cdiv(g.G, 256)), dim3(256), 0, s, g, dt, use_x,
                       convective, wf, x, p, eta, us[0], us[1], us[2], l3[0], l3[1], l3[2],
                       ls[0], ls[1], ls[2], le[0], le[1], le[2]);
    hipLaunchKernelGGL(k_adj_project_cells, dim3(cdiv(g.N, 256)), dim3(256), 0, s, g, dt, use_x,
                       ls[0], ls[1], ls[2], le[0], le[1], le[2], lx, leta);
}

void launch_adj_rhs(const Geo &g, hipStream_t s, double dt, const double *lb, double *const ls[3])
{
    hipLaunchKernelGGL(k_adj_rhs, dim3(cdiv(g.G, 256)), dim3(256), 0, s, g, 1.0 / dt, lb, ls[0],
                       ls[1], ls[2]);
}

void launch_adj_remove_mean(const Geo &g, hipStream_t s, double *a, double *part)
{
    const int nb = g.xcd_remap ? ((g.nblocks + 7) / 8) * 8 : g.nblocks;
    hipLaunchKernelGGL(k_adj_sum_partial, dim3(nb), dim3(kThreads), 0, s, g, (const double *)a,
                       part);
    hipLaunchKernelGGL(k_adj_mean_fold, dim3(1), dim3(kThreads), 0, s, g, part);
    hipLaunchKernelGGL(k_adj_sub_mean, dim3(cdiv(g.N, kThreads)), dim3(kThreads), 0, s, g,
                       (const double *)part, a);
}

void launch_adj_periodic_interior(const Geo &g, hipStream_t s, double *const l[3])
{
    hipLaunchKernelGGL(k_adj_periodic, dim3(cdiv(g.nxc * g.nyc, 256)), dim3(256), 0, s, g, 0,
                       l[0], l[1], l[2]);
}

void launch_adj_convdiff(const Geo &g, hipStream_t s, double dt, double nu,
                         const double *const U[3], double *Jr, double *const ls[3],
                         double *const lu[3])
{
    hipLaunchKernelGGL(k_adj_cd_rows, dim3(cdiv(g.G, 128)), dim3(128), 0, s, g, dt, nu, U[0],
                       U[1], U[2], Jr);
    hipLaunchKernelGGL(k_adj_cd_gather, dim3(cdiv(g.G, 256)), dim3(256), 0, s, g,
                       (const double *)Jr, ls[0], ls[1], ls[2], lu[0], lu[1], lu[2]);
}

void launch_adj_mask(const Geo &g, hipStream_t s, const double *Xn, const double *Yn, double cx,
                     double cy, double R, double eps, const double *leta, double *part,
                     double *out)
{
    const int nb = cdiv(g.nxc * g.nyc, kPerBlock);
    hipLaunchKernelGGL(k_adj_mask_partial, dim3(nb), dim3(kThreads), 0, s, g, Xn, Yn, cx, cy, R,
                       eps, leta, part, nb);
    hipLaunchKernelGGL(k_adj_fold3, dim3(1), dim3(kThreads), 0, s, (const double *)part, nb, out);
}

} // namespace cfdmx
