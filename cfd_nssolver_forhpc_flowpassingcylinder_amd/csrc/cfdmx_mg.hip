// cfdmx_mg.hip -- geometric multigrid preconditioner and the preconditioned BiCGSTAB of
// the multigrid pressure solves (no reference counterpart), in T = float (the
// mixed-precision inner solve) or T = double (the same algorithm in FP64).
//
// The SPE operator on the uniform grid (SPE_sparseMatrix.hpp:42-106) is the 7-point
// Laplacian (A u)_c = sum over existing neighbours of c_d (u_c - u_nb), c_d = 1/h_d^2,
// with the Neumann faces folded into the diagonal and, with periodic z, the z
// neighbours wrapped.  Each coarser level halves the directions whose spacing is within
// a factor 2 of the smallest (z first on the flat cells of the cylinder grids), down to
// at most 512 cells or no even size left; its operator is the same Laplacian on the
// coarse spacings.  Cycle: V(2,2) with red-black Gauss-Seidel, restriction = mean of the
// children's residuals, prolongation = the parent's value; the coarsest level is solved
// exactly by fast diagonalisation (the 1-D eigenbases of its three axes, computed on the
// host).  Every inner product accumulates in double.
//
// A level with even nz stores u and f by colour (red cells, then black cells), and its
// kernels take one thread per z-pair (cells 2kk and 2kk+1 of an (i, j) row), so a half
// sweep runs one thread per cell of its colour and touches only the values it uses:
//   first sweep from u = 0 (red u = f/d, then black from the red values) -- one kernel;
//     on level 0 it also stores the natural-order right-hand side by colour;
//   half sweeps; on level 0 the last one writes the result in natural order;
//   restriction (one thread per coarse cell);
//   prolongation fused with the next red half sweep: the red cell reads its black
//   neighbours corrected on the fly; the black cells are left uncorrected, since the
//   black half sweep that follows overwrites them without reading them.
// Levels with odd nz are stored in natural order and use one thread per cell.  The
// arithmetic of every cell update (terms, weights, order) is the same in either layout.
#include "cfdmx_internal.hpp"

namespace cfdmx {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

template <typename T>
__device__ __forceinline__ void mg_coords(const MGLevelT<T> &L, int n, int &i, int &j, int &k)
{
    const unsigned t = fdiv((unsigned)n, L.dz);
    k = n - (int)t * L.nz;
    i = (int)fdiv(t, L.dy);
    j = (int)t - i * L.ny;
}

// z-pair m of a level with even nz: row (i, j) and the pair's first cell 2kk
template <typename T>
__device__ __forceinline__ void mg_pair(const MGLevelT<T> &L, int m, int &i, int &j, int &k0,
                                        int &row)
{
    const int half = L.nz >> 1;
    row = (int)fdiv((unsigned)m, L.dzh);
    k0 = 2 * (m - row * half);
    i = (int)fdiv((unsigned)row, L.dy);
    j = row - i * L.ny;
}

template <typename T>
__device__ __forceinline__ T mg_diag(const MGLevelT<T> &L, int i, int j, int k)
{
    T d = (T)0;
    if (i > 0) d += L.cx;
    if (i < L.nx - 1) d += L.cx;
    if (j > 0) d += L.cy;
    if (j < L.ny - 1) d += L.cy;
    if (L.nz > 1) {
        if (L.zper) {
            d += (T)2 * L.cz;
        } else {
            if (k > 0) d += L.cz;
            if (k < L.nz - 1) d += L.cz;
        }
    }
    return d;
}

// sum of c_d * v(nb) over the existing neighbours nb of cell n, in the order
// -x, +x, -y, +y, -z, +z; v(m, a, b, c) gives the value of cell m = (a, b, c).  A
// missing neighbour is taken as the cell itself with weight 0 (adding +0 changes no
// sum), so that every value is fetched without a branch and the loads overlap.
template <typename T, class V>
__device__ __forceinline__ T mg_nbsum(const MGLevelT<T> &L, int n, int i, int j, int k, V &&v)
{
    const int sx = L.ny * L.nz, sy = L.nz;
    const bool xm = i > 0, xp = i < L.nx - 1, ym = j > 0, yp = j < L.ny - 1;
    T s = (T)0;
    s += (xm ? L.cx : (T)0) * v(xm ? n - sx : n, xm ? i - 1 : i, j, k);
    s += (xp ? L.cx : (T)0) * v(xp ? n + sx : n, xp ? i + 1 : i, j, k);
    s += (ym ? L.cy : (T)0) * v(ym ? n - sy : n, i, ym ? j - 1 : j, k);
    s += (yp ? L.cy : (T)0) * v(yp ? n + sy : n, i, yp ? j + 1 : j, k);
    if (L.nz > 1) {
        const bool zm = L.zper || k > 0, zp = L.zper || k < L.nz - 1;
        const int km = k > 0 ? k - 1 : (zm ? L.nz - 1 : k);
        const int kp = k < L.nz - 1 ? k + 1 : (zp ? 0 : k);
        s += (zm ? L.cz : (T)0) * v(n + (km - k), i, j, km);
        s += (zp ? L.cz : (T)0) * v(n + (kp - k), i, j, kp);
    }
    return s;
}

// storage index of cell (i, j, k) of a level (natural or by colour, see MGLevelT)
template <typename T>
__device__ __forceinline__ int mg_at(const MGLevelT<T> &L, int i, int j, int k)
{
    const int row = i * L.ny + j;
    if (!L.split) return row * L.nz + k;
    return ((i + j + k) & 1) * (L.n >> 1) + row * (L.nz >> 1) + (k >> 1);
}

// mg_nbsum by coordinates: v(a, b, c) gives the value of cell (a, b, c); the same
// terms, weights and order, and the cell itself with weight 0 for a missing neighbour
template <typename T, class V>
__device__ __forceinline__ T mg_nbsum_c(const MGLevelT<T> &L, int i, int j, int k, V &&v)
{
    const bool xm = i > 0, xp = i < L.nx - 1, ym = j > 0, yp = j < L.ny - 1;
    T s = (T)0;
    s += (xm ? L.cx : (T)0) * v(xm ? i - 1 : i, j, k);
    s += (xp ? L.cx : (T)0) * v(xp ? i + 1 : i, j, k);
    s += (ym ? L.cy : (T)0) * v(i, ym ? j - 1 : j, k);
    s += (yp ? L.cy : (T)0) * v(i, yp ? j + 1 : j, k);
    if (L.nz > 1) {
        const bool zm = L.zper || k > 0, zp = L.zper || k < L.nz - 1;
        const int km = k > 0 ? k - 1 : (zm ? L.nz - 1 : k);
        const int kp = k < L.nz - 1 ? k + 1 : (zp ? 0 : k);
        s += (zm ? L.cz : (T)0) * v(i, j, km);
        s += (zp ? L.cz : (T)0) * v(i, j, kp);
    }
    return s;
}

// ---- levels with even nz: stored by colour, one thread per z-pair ----
// Thread m is the z-pair (2kk, 2kk+1) of row (i, j); its cell of colour c is
// k = 2kk + ((i + j + c) & 1), at c*n/2 + m.  The neighbours of that cell are the
// other colour's cells at the same kk of the rows (i +- 1, j) and (i, j +- 1), and at
// kk - 1 + p and kk + p of its own row (p = (i + j + c) & 1; periodic z wraps kk).  A
// half sweep so reads the other colour's half and its own colour's f, and writes its
// own colour's half: whole cache lines of the values it uses.

// half sweep over colour c: the sum of mg_nbsum (same terms and order; a missing
// neighbour reads the cell itself with weight 0)
template <typename T>
__device__ __forceinline__ T mg_split_update(const MGLevelT<T> &L, const T *u, int m, int row,
                                             int i, int j, int kk, int c, int &own, int &k)
{
    const int half = L.nz >> 1, hn = L.n >> 1;
    const int pc = (i + j + c) & 1;
    k = 2 * kk + pc;
    own = c * hn + m;
    const int ob = (1 - c) * hn;
    const int sxh = L.ny * half;
    const bool xm = i > 0, xp = i < L.nx - 1, ym = j > 0, yp = j < L.ny - 1;
    T s = (T)0;
    s += (xm ? L.cx : (T)0) * u[xm ? ob + m - sxh : own];
    s += (xp ? L.cx : (T)0) * u[xp ? ob + m + sxh : own];
    s += (ym ? L.cy : (T)0) * u[ym ? ob + m - half : own];
    s += (yp ? L.cy : (T)0) * u[yp ? ob + m + half : own];
    // nz is even, so nz > 1
    const bool zm = L.zper || k > 0, zp = L.zper || k < L.nz - 1;
    const int kkm = (k > 0) ? kk - 1 + pc : half - 1;  // k - 1, or nz - 1 wrapped
    const int kkp = (k < L.nz - 1) ? kk + pc : 0;      // k + 1, or 0 wrapped
    s += (zm ? L.cz : (T)0) * u[zm ? ob + row * half + kkm : own];
    s += (zp ? L.cz : (T)0) * u[zp ? ob + row * half + kkp : own];
    return s;
}

// the half sweep of colour c at z-pair m (the kernels below and the tail kernel)
template <typename T>
__device__ __forceinline__ void mg_smooth_at(const MGLevelT<T> &L, int m, int colour)
{
    int i, j, k0, row;
    mg_pair(L, m, i, j, k0, row);
    int own, k;
    const T s = mg_split_update(L, L.u, m, row, i, j, k0 >> 1, colour, own, k);
    L.u[own] = (L.f[own] + s) / mg_diag(L, i, j, k);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth(MGLevelT<T> L, int colour)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (L.n >> 1)) return;
    mg_smooth_at(L, m, colour);
}

// The same half sweep with two neighbouring cells of a colour per thread (FP64 levels
// with nz a multiple of 4, so a row holds whole pairs and a pair's first index is even):
// their f, the x and y neighbours of both and the other colour's pair beside them move as
// one 16-byte vector each, and the z neighbours take one more value.  Each cell's terms,
// weights and order are mg_split_update's, and an absent neighbour reads the cell itself
// with weight 0, as there, so the result is the same.
template <typename T> struct Pair;
template <> struct Pair<double> {
    using V = double2;
};

template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth2(MGLevelT<T> L, int colour)
{
    using V = typename Pair<T>::V;
    const int m = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    const int half = L.nz >> 1, hn = L.n >> 1;
    if (m >= hn) return;
    int i, j, k0, row;
    mg_pair(L, m, i, j, k0, row);
    const int kk = k0 >> 1; // even: the pair is (kk, kk + 1)
    const int pc = (i + j + colour) & 1;
    const int own = colour * hn + m, ob = (1 - colour) * hn;
    const int sxh = L.ny * half;
    const bool xm = i > 0, xp = i < L.nx - 1, ym = j > 0, yp = j < L.ny - 1;
    const T *u = L.u;
    const V vxm = *reinterpret_cast<const V *>(u + (xm ? ob + m - sxh : own));
    const V vxp = *reinterpret_cast<const V *>(u + (xp ? ob + m + sxh : own));
    const V vym = *reinterpret_cast<const V *>(u + (ym ? ob + m - half : own));
    const V vyp = *reinterpret_cast<const V *>(u + (yp ? ob + m + half : own));
    const V oz = *reinterpret_cast<const V *>(u + ob + m);
    const V fv = *reinterpret_cast<const V *>(L.f + own);
    // z neighbours (cells A = kk and B = kk + 1 at k = 2 kk + pc and k + 2): with pc = 0,
    // A's -z is the other colour at kk - 1 (wrapped to half - 1 at k = 0) and the rest are
    // the pair beside; with pc = 1, B's +z is the other colour at kk + 2 (wrapped to 0 at
    // k = nz - 1)
    const int kA = 2 * kk + pc, kB = kA + 2;
    T zmA, zpA, zmB, zpB;
    bool hzmA = true, hzpB = true;
    if (pc == 0) {
        hzmA = L.zper || kA > 0;
        zmA = u[hzmA ? (kA > 0 ? ob + m - 1 : ob + row * half + half - 1) : own];
        zpA = oz.x;
        zmB = oz.x;
        zpB = oz.y;
    } else {
        zmA = oz.x;
        zpA = oz.y;
        zmB = oz.y;
        hzpB = L.zper || kB < L.nz - 1;
        zpB = u[hzpB ? (kB < L.nz - 1 ? ob + m + 2 : ob + row * half) : own + 1];
    }
    const T wx = L.cx, wy = L.cy, wz = L.cz, z0 = (T)0;
    T sA = (T)0, sB = (T)0;
    sA += (xm ? wx : z0) * vxm.x;
    sB += (xm ? wx : z0) * vxm.y;
    sA += (xp ? wx : z0) * vxp.x;
    sB += (xp ? wx : z0) * vxp.y;
    sA += (ym ? wy : z0) * vym.x;
    sB += (ym ? wy : z0) * vym.y;
    sA += (yp ? wy : z0) * vyp.x;
    sB += (yp ? wy : z0) * vyp.y;
    sA += (hzmA ? wz : z0) * zmA;
    sB += wz * zmB;
    sA += wz * zpA;
    sB += (hzpB ? wz : z0) * zpB;
    V out;
    out.x = (fv.x + sA) / mg_diag(L, i, j, kA);
    out.y = (fv.y + sB) / mg_diag(L, i, j, kB);
    *reinterpret_cast<V *>(L.u + own) = out;
}

// the last half sweep of level 0 (black), writing the result of both colours to the
// natural-order output vector
template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth_out(MGLevelT<T> L)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (L.n >> 1)) return;
    int i, j, k0, row;
    mg_pair(L, m, i, j, k0, row);
    int own, k;
    const T s = mg_split_update(L, L.u, m, row, i, j, k0 >> 1, 1, own, k);
    const T ub = (L.f[own] + s) / mg_diag(L, i, j, k);
    const int kr = k0 + ((i + j) & 1);
    L.uout[row * L.nz + k] = ub;
    L.uout[row * L.nz + kr] = L.u[m];
}

// first sweep from u = 0: red u = f/d, then black u = (f + sum c (f/d)(red nb)) / d.
// On level 0 f comes in natural order (fin) and is stored by colour here.
template <typename T>
__device__ __forceinline__ void mg_smooth_first_at(const MGLevelT<T> &L, int m)
{
    int i, j, k0, row;
    mg_pair(L, m, i, j, k0, row);
    const int hn = L.n >> 1;
    const int kr = k0 + ((i + j) & 1), kb = k0 + ((i + j + 1) & 1);
    T fr, fb;
    if (L.fin) {
        fr = L.fin[row * L.nz + kr];
        fb = L.fin[row * L.nz + kb];
        L.f[m] = fr;
        L.f[hn + m] = fb;
    } else {
        fr = L.f[m];
        fb = L.f[hn + m];
    }
    L.u[m] = fr / mg_diag(L, i, j, kr);
    const T *fin = L.fin, *f = L.f;
    const T s = mg_nbsum_c(L, i, j, kb, [&](int a, int b, int c) {
        const T fv = fin ? fin[(a * L.ny + b) * L.nz + c] : f[mg_at(L, a, b, c)];
        return fv / mg_diag(L, a, b, c);
    });
    L.u[hn + m] = (fb + s) / mg_diag(L, i, j, kb);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth_first(MGLevelT<T> L)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (L.n >> 1)) return;
    mg_smooth_first_at(L, m);
}

// prolongation fused with the red half sweep (see the file comment)
template <typename T>
__device__ __forceinline__ void mg_prolong_smooth_at(const MGLevelT<T> &F, const MGLevelT<T> &C,
                                                     int m)
{
    int i, j, k0, row;
    mg_pair(F, m, i, j, k0, row);
    const int k = k0 + ((i + j) & 1);
    const int shx = C.fx >> 1, shy = C.fy >> 1, shz = C.fz >> 1;
    const T *u = F.u, *e = C.u;
    const T s = mg_nbsum_c(F, i, j, k, [&](int a, int b, int c) {
        return u[mg_at(F, a, b, c)] + e[mg_at(C, a >> shx, b >> shy, c >> shz)];
    });
    F.u[m] = (F.f[m] + s) / mg_diag(F, i, j, k);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong_smooth(MGLevelT<T> F, MGLevelT<T> C)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (F.n >> 1)) return;
    mg_prolong_smooth_at(F, C, m);
}

// ---- z-only coarsening (fx = fy = 1, fz = 2) of a level stored by colour: the
// children of coarse cell (i, j, kk) are fine z-pair m, so these run one thread per
// z-pair with the fast neighbour indexing ----

// coarse right-hand side: the mean of the two children's residuals f - A u (the same
// expression and order as k_mg_restrict)
template <typename T>
__global__ void __launch_bounds__(256) k_mg_restrict_z(MGLevelT<T> F, MGLevelT<T> C)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (F.n >> 1)) return;
    int i, j, k0, row;
    mg_pair(F, m, i, j, k0, row);
    const int kk = k0 >> 1;
    T r[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        int own, k;
        const T s = mg_split_update(F, F.u, m, row, i, j, kk, c, own, k);
        r[c] = F.f[own] - (mg_diag(F, i, j, k) * F.u[own] - s);
    }
    // children in z order: cell 2kk has colour (i + j) & 1
    const bool even_red = ((i + j) & 1) == 0;
    T acc = (T)0;
    acc += even_red ? r[0] : r[1];
    acc += even_red ? r[1] : r[0];
    C.f[mg_at(C, i, j, kk)] = acc / (T)2;
}

// prolongation fused with the red half sweep: each black neighbour is corrected by its
// parent's value, the parent of fine cell (a, b, c) being coarse cell (a, b, c/2)
template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong_smooth_z(MGLevelT<T> F, MGLevelT<T> C)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (F.n >> 1)) return;
    int i, j, k0, row;
    mg_pair(F, m, i, j, k0, row);
    const int kk = k0 >> 1;
    const int half = F.nz >> 1, hn = F.n >> 1, sxh = F.ny * half;
    const int pr = (i + j) & 1;
    const int k = k0 + pr;
    const T *u = F.u, *e = C.u;
    // fine black index (other half) and coarse index of each neighbour, in the order of
    // mg_nbsum; a missing neighbour is the red cell itself, weight 0
    const bool xm = i > 0, xp = i < F.nx - 1, ym = j > 0, yp = j < F.ny - 1;
    const bool zm = F.zper || k > 0, zp = F.zper || k < F.nz - 1;
    const int kzm = k > 0 ? k - 1 : F.nz - 1, kzp = k < F.nz - 1 ? k + 1 : 0;
    const int kkm = (k > 0) ? kk - 1 + pr : half - 1;
    const int kkp = (k < F.nz - 1) ? kk + pr : 0;
    // branch-free: the indices are selected, then both values are loaded unconditionally
    auto val = [&](bool ok, int fine, int a, int b, int c) {
        const int fi = ok ? hn + fine : m;
        const int ci = mg_at(C, ok ? a : i, ok ? b : j, (ok ? c : k) >> 1);
        return u[fi] + e[ci];
    };
    T s = (T)0;
    s += (xm ? F.cx : (T)0) * val(xm, m - sxh, i - 1, j, k);
    s += (xp ? F.cx : (T)0) * val(xp, m + sxh, i + 1, j, k);
    s += (ym ? F.cy : (T)0) * val(ym, m - half, i, j - 1, k);
    s += (yp ? F.cy : (T)0) * val(yp, m + half, i, j + 1, k);
    s += (zm ? F.cz : (T)0) * val(zm, row * half + kkm, i, j, kzm);
    s += (zp ? F.cz : (T)0) * val(zp, row * half + kkp, i, j, kzp);
    F.u[m] = (F.f[m] + s) / mg_diag(F, i, j, k);
}

// ---- levels with odd nz: one thread per cell ----

// half sweep; zero_first: u starts at zero (the other colour is written 0).  With
// periodic z and odd nz > 1 the planes k = 0 and k = nz-1 hold the same colour and are
// neighbours, so a half sweep there runs in two launches, planes = 1 (k < nz-1) and then
// planes = 2 (k = nz-1): the first reads the old k = nz-1 values, the second the new
// k = 0 values, and no cell is read while another thread writes it.  planes = 0: all.
template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth_cell(MGLevelT<T> L, int colour, int zero_first,
                                                        int planes)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= L.n) return;
    int i, j, k;
    mg_coords(L, n, i, j, k);
    if (((i + j + k) & 1) != colour) {
        if (zero_first) L.u[n] = (T)0;
        return;
    }
    if ((planes == 1 && k == L.nz - 1) || (planes == 2 && k != L.nz - 1)) return;
    const T *u = L.u;
    const T s = zero_first ? (T)0 : mg_nbsum(L, n, i, j, k, [&](int q, int, int, int) { return u[q]; });
    L.u[n] = (L.f[n] + s) / mg_diag(L, i, j, k);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong(MGLevelT<T> F, MGLevelT<T> C)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= F.n) return;
    int i, j, k;
    mg_coords(F, n, i, j, k);
    F.u[n] += C.u[mg_at(C, i >> (C.fx >> 1), j >> (C.fy >> 1), k >> (C.fz >> 1))];
}

// coarse right-hand side: mean over the children of the fine residual f - A u
template <typename T>
__device__ __forceinline__ void mg_restrict_at(const MGLevelT<T> &F, const MGLevelT<T> &C, int m)
{
    int I, J, K;
    mg_coords(C, m, I, J, K);
    const T *u = F.u;
    T acc = (T)0;
    // at most 2 x 2 x 2 children in the order (a, b, c), unrolled and branch-free so that
    // their loads are in flight together: a child that does not exist evaluates child
    // (0, 0, 0) and adds +0, an identity (the sum starts at +0 and so never holds -0)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q >> 2, b = (q >> 1) & 1, c = q & 1;
        const bool ok = a < C.fx && b < C.fy && c < C.fz;
        const int i = I * C.fx + (ok ? a : 0), j = J * C.fy + (ok ? b : 0),
                  k = K * C.fz + (ok ? c : 0);
        const int n = mg_at(F, i, j, k);
        const T s = mg_nbsum_c(F, i, j, k, [&](int x, int y, int z) {
            return u[mg_at(F, x, y, z)];
        });
        const T r = F.f[n] - (mg_diag(F, i, j, k) * u[n] - s);
        acc += ok ? r : (T)0;
    }
    C.f[mg_at(C, I, J, K)] = acc / (T)(C.fx * C.fy * C.fz);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_restrict(MGLevelT<T> F, MGLevelT<T> C)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= C.n) return;
    mg_restrict_at(F, C, m);
}

// coarsest level: the exact solve by fast diagonalisation.  The level's operator is the
// Kronecker sum of three 1-D chains with orthonormal eigenvectors Qx, Qy, Qz, so
// u = Q diag(1 / lambda) Q^T f with Q = Qx (x) Qy (x) Qz and lambda = lx + ly + lz; the
// constant mode (lambda = 0) is dropped, which returns the mean-free solution of A u = f
// for the mean-free part of f.  One workgroup: the six 1-D transforms run through LDS,
// a barrier apart.
constexpr int kMGCoarseThreads = 1024;
template <typename T>
__device__ __forceinline__ void mg_coarse_solve(const MGLevelT<T> &L, T *a, T *b)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const T *Qx = L.fd, *Qy = Qx + nx * nx, *Qz = Qy + ny * ny;
    const T *lx = Qz + nz * nz, *ly = lx + nx, *lz = ly + ny;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int n = tid; n < L.n; n += nt) a[n] = L.f[n];
    __syncthreads();
    // forward: coefficients along x, then y, then z (divided by the eigenvalue)
    for (int n = tid; n < L.n; n += nt) {
        int m, j, k;
        mg_coords(L, n, m, j, k);
        T acc = (T)0;
        for (int i = 0; i < nx; ++i) acc += Qx[m * nx + i] * a[(i * ny + j) * nz + k];
        b[n] = acc;
    }
    __syncthreads();
    for (int n = tid; n < L.n; n += nt) {
        int mx, m, k;
        mg_coords(L, n, mx, m, k);
        T acc = (T)0;
        for (int j = 0; j < ny; ++j) acc += Qy[m * ny + j] * b[(mx * ny + j) * nz + k];
        a[n] = acc;
    }
    __syncthreads();
    for (int n = tid; n < L.n; n += nt) {
        int mx, my, m;
        mg_coords(L, n, mx, my, m);
        T acc = (T)0;
        for (int k = 0; k < nz; ++k) acc += Qz[m * nz + k] * a[(mx * ny + my) * nz + k];
        const T lam = lx[mx] + ly[my] + lz[m];
        b[n] = (mx == 0 && my == 0 && m == 0) ? (T)0 : acc / lam;
    }
    __syncthreads();
    // inverse: back along z, y, x
    for (int n = tid; n < L.n; n += nt) {
        int mx, my, k;
        mg_coords(L, n, mx, my, k);
        T acc = (T)0;
        for (int m = 0; m < nz; ++m) acc += Qz[m * nz + k] * b[(mx * ny + my) * nz + m];
        a[n] = acc;
    }
    __syncthreads();
    for (int n = tid; n < L.n; n += nt) {
        int mx, j, k;
        mg_coords(L, n, mx, j, k);
        T acc = (T)0;
        for (int m = 0; m < ny; ++m) acc += Qy[m * ny + j] * a[(mx * ny + m) * nz + k];
        b[n] = acc;
    }
    __syncthreads();
    for (int n = tid; n < L.n; n += nt) {
        int i, j, k;
        mg_coords(L, n, i, j, k);
        T acc = (T)0;
        for (int m = 0; m < nx; ++m) acc += Qx[m * nx + i] * b[(m * ny + j) * nz + k];
        L.u[n] = acc;
    }
}

template <typename T>
__global__ void __launch_bounds__(kMGCoarseThreads) k_mg_coarse(MGLevelT<T> L)
{
    extern __shared__ double coarse_lds[];
    T *a = reinterpret_cast<T *>(coarse_lds);
    mg_coarse_solve(L, a, a + L.n);
}

// The two coarsest levels in one workgroup: the level F above the coarsest (stored by
// colour) from its first sweep to its last, with the restriction to the coarsest level
// C, C's exact solve and the prolongation between.  These levels hold a few thousand
// cells, where a launch per half sweep costs more than the sweep.  Here F's u (and its
// f, where both fit in 64 KiB of LDS with C's two work arrays) stay in LDS from the
// first step to the last, the steps a barrier apart, and only F's final u goes back to
// global memory for the prolongation to the finer level.  Every cell is updated by the
// same device function as in the per-step kernels (the restriction and prolongation by
// their general forms, which give the same values as the z-only ones), so the result
// is the same bits.
constexpr size_t kMGTailLds = 65536;
template <typename T>
__host__ __device__ inline bool mg_tail_f_in_lds(int fn, int cn)
{
    return (2 * (size_t)fn + 2 * (size_t)cn) * sizeof(T) <= kMGTailLds;
}

template <typename T, bool F_LDS>
__global__ void __launch_bounds__(kMGCoarseThreads) k_mg_tail(MGLevelT<T> F, MGLevelT<T> C)
{
    extern __shared__ double coarse_lds[];
    const int tid = threadIdx.x, nt = blockDim.x, hn = F.n >> 1;
    T *su = reinterpret_cast<T *>(coarse_lds);
    T *sf = su + F.n;
    T *ca = F_LDS ? sf + F.n : sf, *cb = ca + C.n;
    MGLevelT<T> FL = F, CL = C;
    FL.u = su;
    if (F_LDS) {
        FL.f = sf;
        for (int n = tid; n < F.n; n += nt) sf[n] = F.f[n];
        __syncthreads();
    }
    // C's right-hand side is built in the first work array and its result left there
    CL.f = ca;
    CL.u = ca;
    for (int m = tid; m < hn; m += nt) mg_smooth_first_at(FL, m);
    __syncthreads();
    for (int m = tid; m < hn; m += nt) mg_smooth_at(FL, m, 0);
    __syncthreads();
    for (int m = tid; m < hn; m += nt) mg_smooth_at(FL, m, 1);
    __syncthreads();
    for (int m = tid; m < C.n; m += nt) mg_restrict_at(FL, CL, m);
    __syncthreads();
    mg_coarse_solve(CL, ca, cb);
    __syncthreads();
    for (int m = tid; m < hn; m += nt) mg_prolong_smooth_at(FL, CL, m);
    __syncthreads();
    for (int m = tid; m < hn; m += nt) mg_smooth_at(FL, m, 1);
    __syncthreads();
    for (int m = tid; m < hn; m += nt) mg_smooth_at(FL, m, 0);
    __syncthreads();
    for (int m = tid; m < hn; m += nt) mg_smooth_at(FL, m, 1);
    __syncthreads();
    for (int n = tid; n < F.n; n += nt) F.u[n] = su[n];
}

// ---- preconditioned BiCGSTAB pieces (linear cell order, canonical block sums) ----
// out = A in on level 0 (double accumulation of the 7 terms); partials of (out, w) and,
// with NS = 2, (out, out)
template <int NS, typename T>
__global__ void __launch_bounds__(kThreads)
k_pb_spmv(Geo g, MGLevelT<T> L, const T *__restrict__ in, T *__restrict__ out,
          const T *__restrict__ w, double *part, const PBState *st)
{
    if (st->done) return;
    const int B = logical_block(g);
    if (B >= g.nblocks) return;
    double acc[NS];
    for (int q = 0; q < NS; ++q) acc[q] = 0.0;
    auto cell = [&](int n) {
        {
            int i, j, k;
            mg_coords(L, n, i, j, k);
            // branch-free: a missing neighbour reads the cell itself with weight 0
            const int sx = L.ny * L.nz, sy = L.nz;
            const bool xm = i > 0, xp = i < L.nx - 1, ym = j > 0, yp = j < L.ny - 1;
            const bool zm = L.zper || k > 0, zp = L.zper || k < L.nz - 1;
            const double cx = L.cx, cy = L.cy, cz = L.cz;
            const double wxm = xm ? cx : 0.0, wxp = xp ? cx : 0.0, wym = ym ? cy : 0.0,
                         wyp = yp ? cy : 0.0, wzm = zm ? cz : 0.0, wzp = zp ? cz : 0.0;
            const int km = k > 0 ? n - 1 : (L.zper ? n + L.nz - 1 : n);
            const int kp = k < L.nz - 1 ? n + 1 : (L.zper ? n - (L.nz - 1) : n);
            double s = 0.0;
            s += wxm * (double)in[xm ? n - sx : n];
            s += wxp * (double)in[xp ? n + sx : n];
            s += wym * (double)in[ym ? n - sy : n];
            s += wyp * (double)in[yp ? n + sy : n];
            s += wzm * (double)in[km];
            s += wzp * (double)in[kp];
            const double d = wxm + wxp + wym + wyp + wzm + wzp;
            const T o = (T)(d * (double)in[n] - s);
            out[n] = o;
            acc[0] += (double)o * (double)w[n];
            if (NS == 2) acc[NS - 1] += (double)o * (double)o;
        }
    };
    // a full block has no bounds test, so its cells' loads overlap
    if ((B + 1) * kPerBlock <= g.N) {
#pragma unroll
        for (int c = 0; c < kCells; ++c) cell(B * kPerBlock + c * kThreads + threadIdx.x);
    } else {
        for (int c = 0; c < kCells; ++c) {
            const int n = B * kPerBlock + c * kThreads + threadIdx.x;
            if (n < g.N) cell(n);
        }
    }
    block_sums<NS>(acc, part, g.nblocks, B);
}

// The same product with four consecutive cells of a z-row per thread (nz a multiple of
// 4): their centres, x and y neighbours are read as one vector each and their z
// neighbours are each other, so a thread issues 7 loads for 4 cells instead of 28.
// Each cell's terms, weights and order are k_pb_spmv's, so out is the same; the dot
// partials add a thread's cells in another order.
template <typename T> struct Vec4;
template <> struct Vec4<float> {
    float v[4];
    __device__ __forceinline__ void load(const float *p)
    {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
    }
    __device__ __forceinline__ void store(float *p) const
    {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    }
};
template <> struct Vec4<double> {
    double v[4];
    __device__ __forceinline__ void load(const double *p)
    {
        const double2 a = reinterpret_cast<const double2 *>(p)[0];
        const double2 b = reinterpret_cast<const double2 *>(p)[1];
        v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y;
    }
    __device__ __forceinline__ void store(double *p) const
    {
        reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
        reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
    }
};

template <int NS, typename T>
__global__ void __launch_bounds__(kThreads)
k_pb_spmv4(Geo g, MGLevelT<T> L, const T *__restrict__ in, T *__restrict__ out,
           const T *__restrict__ w, double *part, const PBState *st)
{
    if (st->done) return;
    const int B = logical_block(g);
    if (B >= g.nblocks) return;
    double acc[NS];
    for (int q = 0; q < NS; ++q) acc[q] = 0.0;
    const int n0 = B * kPerBlock + 4 * threadIdx.x;
    if (n0 < g.N) {
        int i, j, k0;
        mg_coords(L, n0, i, j, k0);
        const int sx = L.ny * L.nz, sy = L.nz;
        const bool xm = i > 0, xp = i < L.nx - 1, ym = j > 0, yp = j < L.ny - 1;
        const double cx = L.cx, cy = L.cy, cz = L.cz;
        const double wxm = xm ? cx : 0.0, wxp = xp ? cx : 0.0, wym = ym ? cy : 0.0,
                     wyp = yp ? cy : 0.0;
        Vec4<T> c, vxm, vxp, vym, vyp, vw;
        c.load(in + n0);
        vxm.load(in + (xm ? n0 - sx : n0));
        vxp.load(in + (xp ? n0 + sx : n0));
        vym.load(in + (ym ? n0 - sy : n0));
        vyp.load(in + (yp ? n0 + sy : n0));
        vw.load(w + n0);
        const int klo = k0 > 0 ? n0 - 1 : (L.zper ? n0 + L.nz - 1 : n0);
        const int khi = k0 + 3 < L.nz - 1 ? n0 + 4 : (L.zper ? n0 + 4 - L.nz : n0 + 3);
        const T zlo = in[klo], zhi = in[khi];
        Vec4<T> o;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + q;
            const bool zm = L.zper || k > 0, zp = L.zper || k < L.nz - 1;
            const double wzm = zm ? cz : 0.0, wzp = zp ? cz : 0.0;
            const T zmv = q > 0 ? c.v[q - 1] : zlo, zpv = q < 3 ? c.v[q + 1] : zhi;
            double s = 0.0;
            s += wxm * (double)vxm.v[q];
            s += wxp * (double)vxp.v[q];
            s += wym * (double)vym.v[q];
            s += wyp * (double)vyp.v[q];
            s += wzm * (double)zmv;
            s += wzp * (double)zpv;
            const double d = wxm + wxp + wym + wyp + wzm + wzp;
            o.v[q] = (T)(d * (double)c.v[q] - s);
            acc[0] += (double)o.v[q] * (double)vw.v[q];
            if (NS == 2) acc[NS - 1] += (double)o.v[q] * (double)o.v[q];
        }
        o.store(out + n0);
    }
    block_sums<NS>(acc, part, g.nblocks, B);
}

// p = r (first iteration) or r + beta (p - omega v)
template <typename T>
__global__ void k_pb_update_p(Geo g, int first, const T *r, const T *v, T *p, const PBState *st)
{
    if (st->done) return;
    const T beta = (T)st->beta, om = (T)st->omega;
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
template <typename T>
__global__ void k_pb_update_s(Geo g, const T *r, const T *v, T *s, const PBState *st)
{
    if (st->done) return;
    const T a = (T)st->alpha;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < g.N; n += gridDim.x * blockDim.x)
        s[n] = r[n] - a * v[n];
}

// e += alpha p^ + omega s^; r = s - omega t; partials (r, r), (r, r0).  The correction e
// and its increments are FP64 in both modes: in FP32 its own rounding (6e-8 of e),
// multiplied by the operator's condition number (~1e5 here), would leave a true residual
// near 1e-2 of the pass's start, whatever the inner tolerance; in FP64 the pass's floor is
// the FP32 recursion's drift from the true residual, ~1e-6.  FIRST (the solve's first
// iteration, e = 0): e is written as +0.0 plus the increment without being read, which is
// the sum the zeroed vector would give, -0.0 turning to +0.0 included.
template <typename T, bool FIRST>
__global__ void __launch_bounds__(kThreads)
k_pb_update_xr(Geo g, const T *ph, const T *sh, const T *s, const T *t, const T *r0, double *e,
               T *r, double *part_out, const PBState *st)
{
    if (st->done) return;
    const int B = logical_block(g);
    if (B >= g.nblocks) return;
    const T om = (T)st->omega;
    const double ad = st->alpha, omd = st->omega;
    double acc[2] = {0.0, 0.0};
    for (int c = 0; c < kCells; ++c) {
        const int n = B * kPerBlock + c * kThreads + threadIdx.x;
        if (n < g.N) {
            const double inc = ad * (double)ph[n] + omd * (double)sh[n];
            if constexpr (FIRST) e[n] = 0.0 + inc;
            else e[n] += inc;
            const T rn = s[n] - om * t[n];
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
template <typename T>
void launch_mg_vcycle(hipStream_t s, const MGLevelT<T> *L, int nlev, int l)
{
    const MGLevelT<T> &F = L[l];
    if (l == nlev - 1) {
        hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(kMGCoarseThreads),
                           2 * (size_t)F.n * sizeof(T), s, F);
        return;
    }
    const MGLevelT<T> &C = L[l + 1];
    const size_t tail_lds =
        ((mg_tail_f_in_lds<T>(F.n, C.n) ? 2 : 1) * (size_t)F.n + 2 * (size_t)C.n) * sizeof(T);
    if (l > 0 && l == nlev - 2 && (F.nz & 1) == 0 && F.n <= kMGTailCells &&
        tail_lds <= kMGTailLds) {
        if (mg_tail_f_in_lds<T>(F.n, C.n))
            hipLaunchKernelGGL((k_mg_tail<T, true>), dim3(1), dim3(kMGCoarseThreads), tail_lds, s,
                               F, C);
        else
            hipLaunchKernelGGL((k_mg_tail<T, false>), dim3(1), dim3(kMGCoarseThreads), tail_lds, s,
                               F, C);
        return;
    }
    if ((F.nz & 1) == 0) {
        const int nb = cdiv(F.n >> 1, 256);
        // a plain half sweep: in FP64 two cells of a colour per thread where rows hold
        // whole pairs (measured 4% faster per V-cycle at 512x256x64); in FP32 one cell per
        // thread (pairs measured 3% slower there, four cells level with one)
        auto smooth = [&](int colour) {
            if constexpr (std::is_same<T, double>::value) {
                if ((F.nz & 3) == 0) {
                    hipLaunchKernelGGL(k_mg_smooth2<T>, dim3(cdiv(F.n >> 2, 256)), dim3(256), 0,
                                       s, F, colour);
                    return;
                }
            }
            hipLaunchKernelGGL(k_mg_smooth<T>, dim3(nb), dim3(256), 0, s, F, colour);
        };
        hipLaunchKernelGGL(k_mg_smooth_first<T>, dim3(nb), dim3(256), 0, s, F);
        smooth(0);
        smooth(1);
        const bool zonly = C.fx == 1 && C.fy == 1 && C.fz == 2;
        if (zonly)
            hipLaunchKernelGGL(k_mg_restrict_z<T>, dim3(nb), dim3(256), 0, s, F, C);
        else
            hipLaunchKernelGGL(k_mg_restrict<T>, dim3(cdiv(C.n, 256)), dim3(256), 0, s, F, C);
        launch_mg_vcycle(s, L, nlev, l + 1);
        if (zonly)
            hipLaunchKernelGGL(k_mg_prolong_smooth_z<T>, dim3(nb), dim3(256), 0, s, F, C);
        else
            hipLaunchKernelGGL(k_mg_prolong_smooth<T>, dim3(nb), dim3(256), 0, s, F, C);
        smooth(1);
        smooth(0);
        if (F.uout)
            hipLaunchKernelGGL(k_mg_smooth_out<T>, dim3(nb), dim3(256), 0, s, F);
        else
            smooth(1);
        return;
    }
    const int nb = cdiv(F.n, 256);
    // a half sweep that reads neighbours; two launches where the periodic z wrap joins
    // two cells of one colour (see k_mg_smooth_cell)
    const bool wrap_same = F.zper && F.nz > 1;
    auto half_sweep = [&](int colour) {
        if (wrap_same) {
            hipLaunchKernelGGL(k_mg_smooth_cell<T>, dim3(nb), dim3(256), 0, s, F, colour, 0, 1);
            hipLaunchKernelGGL(k_mg_smooth_cell<T>, dim3(nb), dim3(256), 0, s, F, colour, 0, 2);
        } else {
            hipLaunchKernelGGL(k_mg_smooth_cell<T>, dim3(nb), dim3(256), 0, s, F, colour, 0, 0);
        }
    };
    hipLaunchKernelGGL(k_mg_smooth_cell<T>, dim3(nb), dim3(256), 0, s, F, 0, 1, 0);
    half_sweep(1);
    half_sweep(0);
    half_sweep(1);
    hipLaunchKernelGGL(k_mg_restrict<T>, dim3(cdiv(C.n, 256)), dim3(256), 0, s, F, C);
    launch_mg_vcycle(s, L, nlev, l + 1);
    hipLaunchKernelGGL(k_mg_prolong<T>, dim3(nb), dim3(256), 0, s, F, C);
    for (int q = 0; q < 2; ++q) {
        half_sweep(0);
        half_sweep(1);
    }
}

static inline int lin_blocks(const Geo &g)
{
    return g.xcd_remap ? ((g.nblocks + 7) / 8) * 8 : g.nblocks;
}

template <typename T>
void launch_pb_spmv(const Geo &g, hipStream_t s, const MGLevelT<T> &L0, int ns, const T *in,
                    T *out, const T *w, double *part, const PBState *st)
{
    // four cells of a z-row per thread when rows hold whole fours (kPerBlock = 4 * 256,
    // so each block still covers the same 1024 cells)
    const bool by4 = (L0.nz % 4) == 0 && kPerBlock == 4 * kThreads;
    if (ns == 1) {
        if (by4)
            hipLaunchKernelGGL((k_pb_spmv4<1, T>), dim3(lin_blocks(g)), dim3(kThreads), 0, s, g,
                               L0, in, out, w, part, st);
        else
            hipLaunchKernelGGL((k_pb_spmv<1, T>), dim3(lin_blocks(g)), dim3(kThreads), 0, s, g,
                               L0, in, out, w, part, st);
    } else {
        if (by4)
            hipLaunchKernelGGL((k_pb_spmv4<2, T>), dim3(lin_blocks(g)), dim3(kThreads), 0, s, g,
                               L0, in, out, w, part, st);
        else
            hipLaunchKernelGGL((k_pb_spmv<2, T>), dim3(lin_blocks(g)), dim3(kThreads), 0, s, g,
                               L0, in, out, w, part, st);
    }
}

template <typename T>
void launch_pb_update_p(const Geo &g, hipStream_t s, int first, const T *r, const T *v, T *p,
                        const PBState *st)
{
    hipLaunchKernelGGL(k_pb_update_p<T>, dim3(2048), dim3(256), 0, s, g, first, r, v, p, st);
}

template <typename T>
void launch_pb_update_s(const Geo &g, hipStream_t s, const double *part, const T *r, const T *v,
                        T *sv, PBState *st)
{
    hipLaunchKernelGGL(k_pb_alpha, dim3(1), dim3(kThreads), 0, s, g, part, st);
    hipLaunchKernelGGL(k_pb_update_s<T>, dim3(2048), dim3(256), 0, s, g, r, v, sv,
                       (const PBState *)st);
}

template <typename T>
void launch_pb_update_xr(const Geo &g, hipStream_t s, int first, const double *part_in,
                         const T *ph, const T *sh, const T *sv, const T *t, const T *r0,
                         double *e, T *r, double *part_out, PBState *st)
{
    hipLaunchKernelGGL(k_pb_omega, dim3(1), dim3(kThreads), 0, s, g, part_in, st);
    if (first)
        hipLaunchKernelGGL((k_pb_update_xr<T, true>), dim3(lin_blocks(g)), dim3(kThreads), 0, s,
                           g, ph, sh, sv, t, r0, e, r, part_out, (const PBState *)st);
    else
        hipLaunchKernelGGL((k_pb_update_xr<T, false>), dim3(lin_blocks(g)), dim3(kThreads), 0, s,
                           g, ph, sh, sv, t, r0, e, r, part_out, (const PBState *)st);
}

void launch_pb_end(const Geo &g, hipStream_t s, const double *part, PBState *st)
{
    hipLaunchKernelGGL(k_pb_end, dim3(1), dim3(kThreads), 0, s, g, part, st);
}

#define CFDMX_MG_INSTANCES(T)                                                                 \
    template void launch_mg_vcycle<T>(hipStream_t, const MGLevelT<T> *, int, int);            \
    template void launch_pb_spmv<T>(const Geo &, hipStream_t, const MGLevelT<T> &, int,       \
                                    const T *, T *, const T *, double *, const PBState *);    \
    template void launch_pb_update_p<T>(const Geo &, hipStream_t, int, const T *, const T *,  \
                                        T *, const PBState *);                                \
    template void launch_pb_update_s<T>(const Geo &, hipStream_t, const double *, const T *,  \
                                        const T *, T *, PBState *);                           \
    template void launch_pb_update_xr<T>(const Geo &, hipStream_t, int, const double *,       \
                                         const T *, const T *, const T *, const T *,          \
                                         const T *, double *, T *, double *, PBState *);
CFDMX_MG_INSTANCES(float)
CFDMX_MG_INSTANCES(double)
#undef CFDMX_MG_INSTANCES

} // namespace cfdmx
