// cfdmx_device.hpp -- device-side building blocks shared by the gfx950 kernels.
//
// Geometry: ghosted fields use (I*ny + J)*nz + K with two ghost layers
// (structureGrid.hpp:34,153-156); interior cells are enumerated by the linear
// index n = (i*nyc + j)*nzc + k, i = I-2 etc. (structureGrid.hpp:176-181), which is
// also the layout of the pressure unknowns.
//
// Canonical work shapes: kernels that produce a sum use blocks of 256 threads.
//  linear: block L covers n in [1024L, 1024L+1024), thread t takes n = 1024L + 256c + t
//          for c = 0..3;
//  march:  block b = chunk*ntile + tile covers planes [8*chunk, 8*chunk+8) of the
//          plane-index tile [1024*tile, 1024*tile+1024), thread t takes, plane by plane,
//          q = 1024*tile + 256c + t for c = 0..3.
// Each thread accumulates its cells in that order, each 64-lane wave folds with
// a[l] += a[l+s] for s = 32..1, and the block combines its waves as
// (w0 + w1) + (w2 + w3).  The per-block partials are folded the same way by one
// block (thread t sums partials t, t+256, ...).  The result is independent of
// scheduling and is reproduced exactly by the CPU oracle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfdmx {

constexpr int kThreads = 256;
constexpr int kCells = 4;
constexpr int kPerBlock = kThreads * kCells;
constexpr int kGC = 2;
// plane-marching stencil kernels: a block owns kTile consecutive cells of the (j,k)
// plane index q = j*nzc + k and marches over kChunk planes i; thread t takes
// q = q0 + 256c + t, c = 0..kMarchCells-1, on every plane, and accumulates plane by
// plane
constexpr int kTile = 1024;
constexpr int kChunk = 8;
constexpr int kMarchCells = kTile / kThreads;

// n / d for n * d < 2^40, with m = ceil(2^40 / d)
struct FastDiv {
    unsigned d;
    unsigned long long m;
};
__host__ __device__ inline FastDiv make_fastdiv(unsigned d)
{
    FastDiv f;
    f.d = d;
    f.m = ((1ull << 40) + d - 1) / d;
    return f;
}
__device__ __forceinline__ unsigned fdiv(unsigned n, const FastDiv &f)
{
    return (unsigned)(((unsigned long long)n * f.m) >> 40);
}

// Correctly rounded a / b for b > 0 from rb = RN(1/b): two residual corrections
// (q += (a - b q) rb, with the residual exact under fma), which reproduce IEEE
// division for finite operands.  The residual is formed as -(b q - a): the same value
// (round-to-nearest is odd-symmetric), and with a = -0 it is -0, so that the
// corrections keep q = -0 and a zero numerator keeps its sign like a / b (every
// divisor here is a grid spacing or the time step, so b > 0).
__device__ __forceinline__ double divr(double a, double b, double rb)
{
    double q = a * rb;
    double t = __builtin_fma(q, b, -a);
    q = __builtin_fma(-t, rb, q);
    t = __builtin_fma(q, b, -a);
    q = __builtin_fma(-t, rb, q);
    return q;
}

// Per-axis metric tables (ghosted length), all computed on the host with the
// reference's own expressions.
struct Axis {
    const double *D, *Ds;   // grid::Dx, Dxs (or y, z)
    const double *rD, *rDs; // RN(1/D), RN(1/Ds)
    const double *IT;       // x_INTRPL: (Ds*0.5)/D, structureGrid.hpp:369-371
    // QUICK coefficients (Convection_Difussion.hpp:20-90):
    //  A: p1 = .125*D[I+1]*D[I+1]/Ds[I], p2 = .125*D[I+1]*D[I+1]/Ds[I+1],
    //     n1 = .125*D[I]*D[I]/Ds[I-1],   n2 = .125*D[I]*D[I]/Ds[I]
    //  B: p1 = .125*Ds[I]*Ds[I]/D[I],    p2 = .125*Ds[I]*Ds[I]/D[I+1],
    //     n1 = .125*Ds[I-1]*Ds[I-1]/D[I-1], n2 = .125*Ds[I-1]*Ds[I-1]/D[I]
    const double *qa[4], *qb[4];
};

// Pressure-operator coefficients on interior indices (SPE_sparseMatrix.hpp:52-58):
// m[i] = 1/D[i+2]/Ds[i+1], p[i] = 1/D[i+2]/Ds[i+2].
struct PAxis {
    const double *m, *p;
};

struct Geo {
    int nx, ny, nz;    // ghosted
    int nxc, nyc, nzc; // interior
    int N, G;
    int sx, sy;        // ghosted strides: ny*nz, nz
    int cyz;           // nyc*nzc
    int zper;          // periodic z
    int nblocks;       // ceil(N / 1024)
    int xcd_remap;
    int zpow2;         // every z spacing (Dz, Dzs) is a power of two
    FastDiv dz, dy;    // by nzc, by nyc
    int ntile, nchunk, nmarch; // plane tiles, plane chunks, march blocks
    // conv-diff tiles: cd_tk cells of k by cd_tj rows of j (cd_tk * cd_tj = 256), tiles
    // along k and j, blocks (tiles x plane chunks); the CFL partials are per block
    int cd_tk, cd_tj, cd_nkt, cd_njt, cd_nb;
    Axis ax[3];
    PAxis pa[3];
};

// logical block index; with remapping, consecutive logical blocks share an XCD
// (the hardware deals workgroups to the 8 XCDs round-robin)
__device__ __forceinline__ int logical_block(const Geo &g)
{
    const int b = blockIdx.x;
    if (!g.xcd_remap) return b;
    const int per = (gridDim.x + 7) >> 3; // launch pads gridDim to a multiple of 8
    return (b & 7) * per + (b >> 3);
}

// block index as dealt: the hardware gives consecutive workgroups to the 8 XCDs in turn,
// so neighbouring blocks run on different XCDs at once.  The cell-wise kernels (right-hand
// side, residual, projection) and the red-black half sweep measured faster this way at
// 512x256x64 (the half sweep 25.1 against 31.5 us per launch, the others by 2-7%); the
// plane marches and conv-diff, which reuse halo planes through an XCD's L2, keep
// logical_block.
__device__ __forceinline__ int dealt_block() { return blockIdx.x; }

__device__ __forceinline__ void decode(const Geo &g, int n, int &i, int &j, int &k)
{
    const unsigned t = fdiv((unsigned)n, g.dz);
    k = n - (int)t * g.nzc;
    i = (int)fdiv(t, g.dy);
    j = (int)t - i * g.nyc;
}

__device__ __forceinline__ int gidx(const Geo &g, int I, int J, int K)
{
    return (I * g.ny + J) * g.nz + K;
}

// canonical in-block fold; returns the block sum in thread 0
__device__ __forceinline__ double wave_fold(double v)
{
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1)
        v = v + __shfl_down(v, s, 64);
    return v;
}

template <int NS>
__device__ __forceinline__ void block_sums(double (&v)[NS], double *out_partials, int stride,
                                           int L)
{
    __shared__ double red[NS][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double w = wave_fold(v[s]);
        if (lane == 0) red[s][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            out_partials[s * stride + L] = (red[s][0] + red[s][1]) + (red[s][2] + red[s][3]);
    }
}

// second level: one block folds nb partials (stride apart) into NS totals
template <int NS>
__device__ __forceinline__ void fold_partials(const double *partials, int stride, int nb,
                                              double (&tot)[NS])
{
    __shared__ double red[NS][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        // same order as one add per partial; loads are issued 8 at a time and
        // out-of-range slots add +0.0, an identity here (the sum starts at +0.0
        // and so never holds -0.0)
        double acc = 0.0;
        for (int m0 = threadIdx.x; m0 < nb; m0 += 8 * kThreads) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int m = m0 + u * kThreads;
                v[u] = m < nb ? partials[s * stride + m] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        const double w = wave_fold(acc);
        if (lane == 0) red[s][wave] = w;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s)
        tot[s] = (red[s][0] + red[s][1]) + (red[s][2] + red[s][3]);
    __syncthreads();
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        const double o = __shfl_down(v, s, 64);
        v = o > v ? o : v;
    }
    return v;
}

// NS block maxima (starting from the values given) into out_partials[s * stride + L];
// a maximum is exact in any order, so no fixed order is kept
template <int NS>
__device__ __forceinline__ void block_maxes(double (&v)[NS], double *out_partials, int stride,
                                            int L)
{
    __shared__ double red[NS][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double w = wave_max(v[s]);
        if (lane == 0) red[s][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double m = red[s][0];
            for (int q = 1; q < 4; ++q) m = red[s][q] > m ? red[s][q] : m;
            out_partials[s * stride + L] = m;
        }
    }
}

} // namespace cfdmx
