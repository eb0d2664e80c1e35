// cfdmx_mg.hip -- FP32 geometric multigrid preconditioner and the FP32 preconditioned
// BiCGSTAB of the mixed-precision pressure solve (no reference counterpart).
//
// The SPE operator on the uniform grid (SPE_sparseMatrix.hpp:42-106) is the 7-point
// Laplacian (A u)_c = sum over existing neighbours of c_d (u_c - u_nb), c_d = 1/h_d^2,
// with the Neumann faces folded into the diagonal and, with periodic z, the z
// neighbours wrapped.  Each coarser level halves the directions whose spacing is within
// a factor 2 of the smallest (z first on the flat cells of the cylinder grids), down to
// at most 512 cells or no even size left; its operator is the same Laplacian on the
// coarse spacings.  Cycle: V(2,2) with red-black Gauss-Seidel, restriction = mean of the
// children's residuals, prolongation = the parent's value; the coarsest level is solved
// by one workgroup (mean removed, Gauss-Seidel sweeps in LDS).
// Vectors are float; every inner product accumulates in double.
#include "cfdmx_internal.hpp"

namespace cfdmx {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ void mg_decode(const MGLevel &L, int n, int &i, int &j, int &k)
{
    k = n % L.nz;
    const int t = n / L.nz;
    j = t % L.ny;
    i = t / L.ny;
}

// sum of c_d * u_nb over the existing neighbours, and the diagonal
__device__ __forceinline__ void mg_row(const MGLevel &L, const float *u, int n, int i, int j,
                                      int k, float &s, float &d)
{
    s = 0.0f;
    d = 0.0f;
    const int sx = L.ny * L.nz, sy = L.nz;
    if (i > 0) { s += L.cx * u[n - sx]; d += L.cx; }
    if (i < L.nx - 1) { s += L.cx * u[n + sx]; d += L.cx; }
    if (j > 0) { s += L.cy * u[n - sy]; d += L.cy; }
    if (j < L.ny - 1) { s += L.cy * u[n + sy]; d += L.cy; }
    if (L.nz > 1) {
        if (L.zper) {
            s += L.cz * u[k > 0 ? n - 1 : n + L.nz - 1];
            s += L.cz * u[k < L.nz - 1 ? n + 1 : n - (L.nz - 1)];
            d += 2.0f * L.cz;
        } else {
            if (k > 0) { s += L.cz * u[n - 1]; d += L.cz; }
            if (k < L.nz - 1) { s += L.cz * u[n + 1]; d += L.cz; }
        }
    }
}

// red-black Gauss-Seidel half sweep; zero_first: u starts at zero (the other colour
// is written 0)
__global__ void k_mg_smooth(MGLevel L, int colour, int zero_first)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= L.n) return;
    int i, j, k;
    mg_decode(L, n, i, j, k);
    if (((i + j + k) & 1) != colour) {
        if (zero_first) L.u[n] = 0.0f;
        return;
    }
    float s, d;
    if (zero_first) {
        float dd = 0.0f;
        if (i > 0) dd += L.cx;
        if (i < L.nx - 1) dd += L.cx;
        if (j > 0) dd += L.cy;
        if (j < L.ny - 1) dd += L.cy;
        if (L.nz > 1) dd += L.zper ? 2.0f * L.cz : (k > 0 ? L.cz : 0.0f) + (k < L.nz - 1 ? L.cz : 0.0f);
        L.u[n] = L.f[n] / dd;
        return;
    }
    mg_row(L, L.u, n, i, j, k, s, d);
    L.u[n] = (L.f[n] + s) / d;
}

// coarse right-hand side: mean over the children of the fine residual f - A u
__global__ void k_mg_restrict(MGLevel F, MGLevel C)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= C.n) return;
    int I, J, K;
    mg_decode(C, m, I, J, K);
    float acc = 0.0f;
    for (int a = 0; a < C.fx; ++a)
        for (int b = 0; b < C.fy; ++b)
            for (int c = 0; c < C.fz; ++c) {
                const int i = I * C.fx + a, j = J * C.fy + b, k = K * C.fz + c;
                const int n = (i * F.ny + j) * F.nz + k;
                float s, d;
                mg_row(F, F.u, n, i, j, k, s, d);
                acc += F.f[n] - (d * F.u[n] - s);
            }
    C.f[m] = acc / (float)(C.fx * C.fy * C.fz);
}

__global__ void k_mg_prolong(MGLevel F, MGLevel C)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= F.n) return;
    int i, j, k;
    mg_decode(F, n, i, j, k);
    F.u[n] += C.u[((i / C.fx) * C.ny + j / C.fy) * C.nz + k / C.fz];
}

// coarsest level in one workgroup: f minus its mean, then Gauss-Seidel sweeps in LDS
constexpr int kMGCoarseMax = 512;
__global__ void __launch_bounds__(kThreads) k_mg_coarse(MGLevel L, int sweeps)
{
    __shared__ float u[kMGCoarseMax], f[kMGCoarseMax];
    __shared__ double red[4];
    double part = 0.0;
    for (int n = threadIdx.x; n < L.n; n += kThreads) {
        f[n] = L.f[n];
        u[n] = 0.0f;
        part += (double)L.f[n];
    }
    const double w = wave_fold(part);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    const float mean = (float)(((red[0] + red[1]) + (red[2] + red[3])) / (double)L.n);
    for (int n = threadIdx.x; n < L.n; n += kThreads) f[n] -= mean;
    __syncthreads();
    MGLevel S = L;
    S.u = u;
    for (int it = 0; it < sweeps; ++it)
        for (int colour = 0; colour < 2; ++colour) {
            for (int n = threadIdx.x; n < L.n; n += kThreads) {
                int i, j, k;
                mg_decode(L, n, i, j, k);
                if (((i + j + k) & 1) != colour) continue;
                float s, d;
                mg_row(S, u, n, i, j, k, s, d);
                u[n] = (f[n] + s) / d;
            }
            __syncthreads();
        }
    for (int n = threadIdx.x; n < L.n; n += kThreads) L.u[n] = u[n];
}

// ---- FP32 preconditioned BiCGSTAB pieces (linear cell order, canonical block sums) ----
// out = A in on level 0 (float, double accumulation of the 7 terms); partials of
// (out, w) and, with NS = 2, (out, out)
template <int NS>
__global__ void __launch_bounds__(kThreads)
k_pb_spmv(Geo g, MGLevel L, const float *__restrict__ in, float *__restrict__ out,
          const float *__restrict__ w, double *part, const PBState *st)
{
    if (st->done) return;
    const int B = logical_block(g);
    if (B >= g.nblocks) return;
    double acc[NS];
    for (int q = 0; q < NS; ++q) acc[q] = 0.0;
    for (int c = 0; c < kCells; ++c) {
        const int n = B * kPerBlock + c * kThreads + threadIdx.x;
        if (n < g.N) {
            int i, j, k;
            mg_decode(L, n, i, j, k);
            const int sx = L.ny * L.nz, sy = L.nz;
            double s = 0.0, d = 0.0;
            const double cx = L.cx, cy = L.cy, cz = L.cz;
            if (i > 0) { s += cx * in[n - sx]; d += cx; }
            if (i < L.nx - 1) { s += cx * in[n + sx]; d += cx; }
            if (j > 0) { s += cy * in[n - sy]; d += cy; }
            if (j < L.ny - 1) { s += cy * in[n + sy]; d += cy; }
            if (L.zper) {
                s += cz * in[k > 0 ? n - 1 : n + L.nz - 1];
                s += cz * in[k < L.nz - 1 ? n + 1 : n - (L.nz - 1)];
                d += 2.0 * cz;
            } else {
                if (k > 0) { s += cz * in[n - 1]; d += cz; }
                if (k < L.nz - 1) { s += cz * in[n + 1]; d += cz; }
            }
            const float o = (float)(d * (double)in[n] - s);
            out[n] = o;
            acc[0] += (double)o * (double)w[n];
            if (NS == 2) acc[NS - 1] += (double)o * (double)o;
        }
    }
    block_sums<NS>(acc, part, g.nblocks, B);
}

// p = r (first iteration) or r + beta (p - omega v)
__global__ void k_pb_update_p(Geo g, int first, const float *r, const float *v, float *p,
                              const PBState *st)
{
    if (st->done) return;
    const float beta = (float)st->beta, om = (float)st->omega;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < g.N; n += gridDim.x * blockDim.x)
        p[n] = first ? r[n] : r[n] + beta * (p[n] - om * v[n]);
}

// alpha = rho / (v, r0) (one workgroup)
__global__ void __launch_bounds__(kThreads) k_pb_alpha(Geo g, const double *part, PBState *st)
{
    if (st->done) return;
    double t[1];
    fold_partials<1>(part, g.nblocks, g.nblocks, t);
    if (threadIdx.x == 0) st->alpha = st->rho / t[0];
}

// omega = (t, s) / (t, t) (one workgroup)
__global__ void __launch_bounds__(kThreads) k_pb_omega(Geo g, const double *part, PBState *st)
{
    if (st->done) return;
    double t[2];
    fold_partials<2>(part, g.nblocks, g.nblocks, t);
    if (threadIdx.x == 0) st->omega = t[0] / t[1];
}

// s = r - alpha v
__global__ void k_pb_update_s(Geo g, const float *r, const float *v, float *s, const PBState *st)
{
    if (st->done) return;
    const float a = (float)st->alpha;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < g.N; n += gridDim.x * blockDim.x)
        s[n] = r[n] - a * v[n];
}

// e += alpha p^ + omega s^; r = s - omega t; partials (r, r), (r, r0)
__global__ void __launch_bounds__(kThreads)
k_pb_update_xr(Geo g, const float *ph, const float *sh, const float *s, const float *t,
               const float *r0, float *e, float *r, double *part_out, const PBState *st)
{
    if (st->done) return;
    const int B = logical_block(g);
    if (B >= g.nblocks) return;
    const float a = (float)st->alpha, om = (float)st->omega;
    double acc[2] = {0.0, 0.0};
    for (int c = 0; c < kCells; ++c) {
        const int n = B * kPerBlock + c * kThreads + threadIdx.x;
        if (n < g.N) {
            e[n] += a * ph[n] + om * sh[n];
            const float rn = s[n] - om * t[n];
            r[n] = rn;
            acc[0] += (double)rn * (double)rn;
            acc[1] += (double)rn * (double)r0[n];
        }
    }
    block_sums<2>(acc, part_out, g.nblocks, B);
}

// end of an iteration: residual test, rho and beta for the next one
__global__ void __launch_bounds__(kThreads) k_pb_end(Geo g, const double *part, PBState *st)
{
    if (st->done) return;
    double t[2];
    fold_partials<2>(part, g.nblocks, g.nblocks, t);
    if (threadIdx.x == 0) {
        const double res = sqrt(t[0]) / st->norm0;
        st->res = res;
        st->iter += 1;
        const double rho = t[1];
        st->beta = (rho / st->rho) * (st->alpha / st->omega);
        st->rho = rho;
        if (res < st->tol) st->done = 1;
        else if (st->iter >= st->maxit || !(res == res)) st->done = 2;
    }
}

// ---- launchers ----
void launch_mg_vcycle(hipStream_t s, const MGLevel *L, int nlev, int l)
{
    const MGLevel &F = L[l];
    if (l == nlev - 1) {
        hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(kThreads), 0, s, F, 60);
        return;
    }
    const int nb = cdiv(F.n, 256);
    hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(256), 0, s, F, 0, 1);
    hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(256), 0, s, F, 1, 0);
    hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(256), 0, s, F, 0, 0);
    hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(256), 0, s, F, 1, 0);
    const MGLevel &C = L[l + 1];
    hipLaunchKernelGGL(k_mg_restrict, dim3(cdiv(C.n, 256)), dim3(256), 0, s, F, C);
    launch_mg_vcycle(s, L, nlev, l + 1);
    hipLaunchKernelGGL(k_mg_prolong, dim3(nb), dim3(256), 0, s, F, C);
    for (int q = 0; q < 2; ++q) {
        hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(256), 0, s, F, 0, 0);
        hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(256), 0, s, F, 1, 0);
    }
}

static inline int lin_blocks(const Geo &g)
{
    return g.xcd_remap ? ((g.nblocks + 7) / 8) * 8 : g.nblocks;
}

void launch_pb_spmv(const Geo &g, hipStream_t s, const MGLevel &L0, int ns, const float *in,
                    float *out, const float *w, double *part, const PBState *st)
{
    if (ns == 1)
        hipLaunchKernelGGL(k_pb_spmv<1>, dim3(lin_blocks(g)), dim3(kThreads), 0, s, g, L0, in,
                           out, w, part, st);
    else
        hipLaunchKernelGGL(k_pb_spmv<2>, dim3(lin_blocks(g)), dim3(kThreads), 0, s, g, L0, in,
                           out, w, part, st);
}

void launch_pb_update_p(const Geo &g, hipStream_t s, int first, const float *r, const float *v,
                        float *p, const PBState *st)
{
    hipLaunchKernelGGL(k_pb_update_p, dim3(2048), dim3(256), 0, s, g, first, r, v, p, st);
}

void launch_pb_update_s(const Geo &g, hipStream_t s, const double *part, const float *r,
                        const float *v, float *sv, PBState *st)
{
    hipLaunchKernelGGL(k_pb_alpha, dim3(1), dim3(kThreads), 0, s, g, part, st);
    hipLaunchKernelGGL(k_pb_update_s, dim3(2048), dim3(256), 0, s, g, r, v, sv,
                       (const PBState *)st);
}

void launch_pb_update_xr(const Geo &g, hipStream_t s, const double *part_in, const float *ph,
                         const float *sh, const float *sv, const float *t, const float *r0,
                         float *e, float *r, double *part_out, PBState *st)
{
    hipLaunchKernelGGL(k_pb_omega, dim3(1), dim3(kThreads), 0, s, g, part_in, st);
    hipLaunchKernelGGL(k_pb_update_xr, dim3(lin_blocks(g)), dim3(kThreads), 0, s, g, ph, sh, sv,
                       t, r0, e, r, part_out, (const PBState *)st);
}

void launch_pb_end(const Geo &g, hipStream_t s, const double *part, PBState *st)
{
    hipLaunchKernelGGL(k_pb_end, dim3(1), dim3(kThreads), 0, s, g, part, st);
}

} // namespace cfdmx
