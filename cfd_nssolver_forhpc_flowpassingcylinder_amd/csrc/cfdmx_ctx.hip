// cfdmx_ctx.hip -- context, grid generation and the C-ABI of include/cfdmx.h.
#include "cfdmx.h"
#include "cfdmx_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

using namespace cfdmx;

struct cfdmx_ctx {
    cfdmx_config cfg;
    Geo g;
    hipStream_t stream = nullptr;
    std::string err;
    // host metrics (grid::X ... Dzs)
    std::vector<double> X, Y, Z, Dx, Dy, Dz, Dxs, Dys, Dzs;
    // device memory
    std::vector<void *> allocs;
    double *metric = nullptr;
    double *Xn = nullptr, *Yn = nullptr;
    double *t0[3] = {}, *t1[3] = {}, *t3[3] = {};
    double *p = nullptr, *eta = nullptr, *f = nullptr;
    // every z ghost of f (k = 0, 1, nz-2, nz-1: entries nothing computes) holds +0.0, as the
    // reset leaves them; an upload or a checkpoint load tests the values it brings.  The
    // time step's projection then stores them too, so that f's rows go out whole
    int f_ghosts_zero = 0;
    double *x = nullptr, *b = nullptr, *px1 = nullptr, *px2 = nullptr;
    BicgBuffers B = {};
    // multigrid-preconditioned solve (mixed precision: float; FP64 multigrid: double):
    // vectors of the preconditioned inner BiCGSTAB (r, r0, p, v, s, t, p^, s^) and the
    // correction e, the multigrid levels of the type in use, the inner-solve state, the
    // three folded sums of the FP64 residual
    void *pb[8] = {};
    void *ecorr = nullptr;
    double *xback = nullptr; // x before the last refinement pass, to drop a rising pass
    std::vector<MGLevelT<float>> mg32;
    std::vector<MGLevelT<double>> mg64;
    PBState *pbs = nullptr, *pbs_h = nullptr;
    // V-cycles captured as HIP graphs, one per element type and level-0 binding
    struct VGraph {
        hipGraphExec_t exec;
        size_t tsize;
        const void *rhs;
        void *out;
    };
    std::vector<VGraph> vgraphs;
    double *pb_part = nullptr;
    double *mix = nullptr, *mix_h = nullptr;
    int mix_outer = 0;
    // Plot3D q output: staged float arrays on the device, pinned host copy filled on a
    // second stream, the file written when the next output or a flush comes
    struct P3D {
        float *dev = nullptr, *host = nullptr;
        hipStream_t side = nullptr;
        hipEvent_t ready = nullptr;
        bool pending = false;
        std::string path;
        float reyn = 0.0f, time = 0.0f;
    } p3d;
    // smoothed immersed-boundary mask (cfdmx_set_eta_smoothed), the adjoint's parameters
    int mask_smoothed = 0;
    double mask_cx = 0.0, mask_cy = 0.0, mask_r = 0.0, mask_eps = 0.0;
    // adjoint of the drag: tape, checkpoints and adjoint fields (allocated on first use)
    struct Adj {
        int K = 0, ncp = 0;
        std::vector<void *> mem;
        double *tapeU = nullptr, *tapeP = nullptr, *tapeX = nullptr;
        std::vector<int> solved;
        double *cpU = nullptr, *cpP = nullptr, *cpX = nullptr, *cpPx1 = nullptr,
               *cpPx2 = nullptr;
        KState *cpK = nullptr;
        std::vector<int> cp_loop, cp_resets;
        std::vector<double> cp_time;
        double *lu[3] = {}, *ls[3] = {}, *le[3] = {}, *us[3] = {};
        double *lx = nullptr, *mu = nullptr, *mpx1 = nullptr, *mpx2 = nullptr;
        // the pressure solution (X_result) of every forward step before the taped
        // segment, so that the reverse pass recomputes those steps without their solves;
        // null where it did not fit
        double *allX = nullptr;
        size_t allX_n = 0;
        KState *kadj = nullptr, *kadj_h = nullptr;
        double *leta = nullptr, *Jr = nullptr, *lin = nullptr, *tin = nullptr;
        double *apart = nullptr, *out = nullptr, *out_h = nullptr;
    } adj;
    double *part = nullptr, *cfl_part = nullptr, *flux_part = nullptr;
    double *diff_part = nullptr; // block maxima of the steady-state check (3 * nblocks)
    // Smagorinsky model (cfg.smagorinsky_cs > 0): eddy viscosity per interior cell and its
    // factor (Cs delta)^2
    double *viseff = nullptr, *smag_coef = nullptr;
    double *bpart[3] = {}; // per-phase BiCGSTAB partials (A, B, C)
    double *scal = nullptr; // [0] cfl, [1] cd, [2] cl, [3..5] max(T0 - T3) of u, v, w
    double *inflow = nullptr;
    KState *kst = nullptr;
    SorState *sst = nullptr;
    SorCoef sorc = {}; // red-black SOR coefficients of a grid with constant spacings
    // red-black SOR working set of a context configured for it (even nzc): p of both
    // colours, then the right-hand side of both colours, in one allocation
    double *sorw = nullptr;
    int *count = nullptr;
    // pinned host mirrors
    KState *kst_h = nullptr;
    SorState *sst_h = nullptr;
    double *scal_h = nullptr;
    int *count_h = nullptr;
    BCParams bc = {};
    int loop = 1;
    double time = 0.0;
    int resets = 0;
    int unconverged = 0; // pressure solves that ended with the residual not below tol
    int started = 0;
    // per-kernel HIP-event timing on the context's stream (cfdmx_profile)
    struct ProfRec {
        int id;
        hipEvent_t a, b;
    };
    // this context's device and that device's address of the file-scope scratch array
    // of cfdmx_cd_cl_noalloc
    int device = -1;
    double *scratch_dev = nullptr;
    int prof_on = 0;
    double prof_ms[CFDMX_K_COUNT] = {};
    long long prof_n[CFDMX_K_COUNT] = {};
    std::vector<hipEvent_t> prof_pool;
    std::vector<ProfRec> prof_pending;
};

static thread_local std::string g_create_error;

// ---------------------------------------------------------------------------
// Scratch memory of cfdmx_cd_cl_noalloc: the block partials of the drag/lift
// reduction (two sums per 1024-cell block) in a file-scope device array instead of
// context memory.  Every GPU holds its own copy for the life of the process; host code
// uses only the per-device address that hipGetSymbolAddress returns (kept in
// g_scratch at the first context creation on that device), never the array's name.
// Calls on one GPU share the array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int kScratchBlocks = 65536; // 64 Mi interior cells: 1024 x 512 x 128
static __device__ __align__(256) double scratch[2 * kScratchBlocks];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices] = {};

#define FAIL(code, ...)                                                                        \
    do {                                                                                       \
        char buf_[512];                                                                        \
        std::snprintf(buf_, sizeof buf_, __VA_ARGS__);                                         \
        ctx->err = buf_;                                                                       \
        return code;                                                                           \
    } while (0)

#define HIPCHK(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            FAIL(CFDMX_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,       \
                 __LINE__);                                                                    \
    } while (0)

#define LAUNCHCHK() HIPCHK(hipGetLastError())

static hipEvent_t prof_event(cfdmx_ctx *ctx)
{
    if (!ctx->prof_pool.empty()) {
        hipEvent_t e = ctx->prof_pool.back();
        ctx->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// HIP events around one launch when profiling is on
#define PROF(id, stmt)                                                                         \
    do {                                                                                       \
        hipEvent_t a_ = nullptr, b_ = nullptr;                                                 \
        if (ctx->prof_on && (a_ = prof_event(ctx)) != nullptr)                                 \
            (void)hipEventRecord(a_, ctx->stream);                                             \
        stmt;                                                                                  \
        if (a_) {                                                                              \
            b_ = prof_event(ctx);                                                              \
            if (b_) {                                                                          \
                (void)hipEventRecord(b_, ctx->stream);                                         \
                ctx->prof_pending.push_back({(id), a_, b_});                                   \
            } else {                                                                           \
                ctx->prof_pool.push_back(a_);                                                  \
            }                                                                                  \
        }                                                                                      \
    } while (0)

// accumulate finished event pairs (call after the stream has been synchronised)
static void prof_harvest(cfdmx_ctx *ctx)
{
    for (auto &r : ctx->prof_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            ctx->prof_ms[r.id] += ms;
            ctx->prof_n[r.id] += 1;
        }
        ctx->prof_pool.push_back(r.a);
        ctx->prof_pool.push_back(r.b);
    }
    ctx->prof_pending.clear();
}

// ---------------------------------------------------------------------------
// grid generation: grid::init_uniform (structureGrid.hpp:336-367),
// unit_SML / init_nonuniform_SML (:250-329), init_gC_dx (:433-486)
// ---------------------------------------------------------------------------
static void gc_dx(std::vector<double> &X, std::vector<double> &D, std::vector<double> &Ds, int n)
{
    X[1] = 2 * X[2] - X[3];
    X[0] = 2 * X[1] - X[2];
    X[n - 1] = 2 * X[n - 2] - X[n - 3];
    X[n] = 2 * X[n - 1] - X[n - 2];
    for (int i = kGC; i < n - kGC - 1; ++i) Ds[i] = (X[i + 2] - X[i]) / 2.0;
    for (int i = kGC; i < n - kGC - 1; ++i) D[i] = (X[i + 1] - X[i]);
    D[0] = D[1] = D[2];
    Ds[0] = Ds[1] = Ds[2];
    D[n - 1] = D[n - 2] = D[n - 3] = X[n - 2] - X[n - 3];
    Ds[n - 1] = Ds[n - 2] = Ds[n - 3] = Ds[n - 4];
}

static bool unit_sml(std::vector<double> &a, double lsml, int ns, int off_ns, double l, int n)
{
    const int ncal = n - 2 * kGC;
    const double dS = lsml / ns, d = (l - lsml) / (ncal - ns);
    a[kGC] = 0.0;
    if (ns != 0) {
        if ((n - ns) % 2 != 0) return false; // structureGrid.hpp:271
        int nlEnd0 = (ncal - ns) / 2;
        int nsEnd0 = nlEnd0 + ns;
        const int nlEnd1 = nsEnd0 + nlEnd0;
        nlEnd0 += off_ns;
        nsEnd0 += off_ns;
        if (nlEnd0 <= 0) return false; // structureGrid.hpp:287
        for (int i = 0; i < nlEnd0; i++) a[i + kGC + 1] = a[i + kGC] + d;
        for (int i = nlEnd0; i < nsEnd0; ++i) a[i + kGC + 1] = a[i + kGC] + dS;
        for (int i = nsEnd0; i < nlEnd1; i++) a[i + kGC + 1] = a[i + kGC] + d;
    } else {
        for (int i = kGC; i < n - kGC; i++) a[i + 1] = a[i] + l / ncal;
    }
    return true;
}

static int build_grid(cfdmx_ctx *ctx)
{
    const cfdmx_config &c = ctx->cfg;
    const int nx = c.nx + 4, ny = c.ny + 4, nz = c.nz + 4;
    ctx->X.assign(nx + 1, 0.0);
    ctx->Y.assign(ny + 1, 0.0);
    ctx->Z.assign(nz + 1, 0.0);
    ctx->Dx.assign(nx, 0.0);
    ctx->Dxs.assign(nx, 0.0);
    ctx->Dy.assign(ny, 0.0);
    ctx->Dys.assign(ny, 0.0);
    ctx->Dz.assign(nz, 0.0);
    ctx->Dzs.assign(nz, 0.0);
    if (c.grid_kind == CFDMX_GRID_UNIFORM) {
        const double dx = c.lx / (nx - 2.0 * kGC), dy = c.ly / (ny - 2.0 * kGC),
                     dz = c.lz / (nz - 2.0 * kGC);
        ctx->X[0] = 0.0 - 2.0 * dx;
        for (int i = 1; i < nx + 1; ++i) ctx->X[i] = ctx->X[i - 1] + dx;
        ctx->Y[0] = 0.0 - 2.0 * dy;
        for (int j = 1; j < ny + 1; ++j) ctx->Y[j] = ctx->Y[j - 1] + dy;
        ctx->Z[0] = 0.0 - 2.0 * dz;
        for (int k = 1; k < nz + 1; ++k) ctx->Z[k] = ctx->Z[k - 1] + dz;
        for (int i = 0; i < nx; ++i) ctx->Dx[i] = ctx->Dxs[i] = dx;
        for (int j = 0; j < ny; ++j) ctx->Dy[j] = ctx->Dys[j] = dy;
        for (int k = 0; k < nz; ++k) ctx->Dz[k] = ctx->Dzs[k] = dz;
        return CFDMX_OK;
    }
    if (c.grid_kind == CFDMX_GRID_SML) {
        if (!unit_sml(ctx->X, 0.3 * c.lx, 100, -4, c.lx, nx) ||
            !unit_sml(ctx->Y, 0.3 * c.ly, 100, 0, c.ly, ny) ||
            !unit_sml(ctx->Z, 0.5 * c.lz, 0, 0, c.lz, nz))
            FAIL(CFDMX_ERR_ARG, "SML grid: nx must be even and >= 110, ny even and >= 102 "
                                "(structureGrid.hpp:271,287)");
        gc_dx(ctx->X, ctx->Dx, ctx->Dxs, nx);
        gc_dx(ctx->Y, ctx->Dy, ctx->Dys, ny);
        gc_dx(ctx->Z, ctx->Dz, ctx->Dzs, nz);
        return CFDMX_OK;
    }
    FAIL(CFDMX_ERR_ARG, "unknown grid_kind %d", c.grid_kind);
}

template <class T> static int dalloc(cfdmx_ctx *ctx, T **ptr, size_t count)
{
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 8));
    ctx->allocs.push_back(p);
    *ptr = static_cast<T *>(p);
    return CFDMX_OK;
}

#define ALLOC(ptr, count)                                                                      \
    do {                                                                                       \
        int rc_ = dalloc(ctx, &(ptr), (size_t)(count));                                        \
        if (rc_) return rc_;                                                                   \
    } while (0)

// ---- multigrid-preconditioned solve: levels, coarse inverse, vectors ----
template <typename T> static std::vector<MGLevelT<T>> &mg_levels(cfdmx_ctx *ctx);
template <> std::vector<MGLevelT<float>> &mg_levels<float>(cfdmx_ctx *ctx) { return ctx->mg32; }
template <> std::vector<MGLevelT<double>> &mg_levels<double>(cfdmx_ctx *ctx) { return ctx->mg64; }

// Eigenpairs of the 1-D operator of one axis of a multigrid level: c (2 u_i - u_(i-1) -
// u_(i+1)) over n cells, with the end faces folded into the diagonal (Neumann:
// q_m(i) = cos(pi m (i + 1/2) / n), lambda_m = c (2 - 2 cos(pi m / n))) or wrapped
// (periodic: the real Fourier modes, lambda_m = c (2 - 2 cos(2 pi m / n))).  Q holds
// the orthonormal modes as rows (Q[m n + i]); a single cell has the one mode 1, with 0.
static void axis_eigen(int n, double c, bool periodic, std::vector<double> &Q,
                       std::vector<double> &lam)
{
    const double pi = 3.14159265358979323846;
    Q.assign((size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    if (n == 1) {
        Q[0] = 1.0;
        return;
    }
    if (!periodic) {
        for (int m = 0; m < n; ++m) {
            const double a = m == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
            for (int i = 0; i < n; ++i) Q[(size_t)m * n + i] = a * std::cos(pi * m * (i + 0.5) / n);
            lam[m] = c * (2.0 - 2.0 * std::cos(pi * m / n));
        }
        return;
    }
    int row = 0;
    auto put = [&](int m, bool sine, double a) {
        for (int i = 0; i < n; ++i) {
            const double t = 2.0 * pi * m * i / n;
            Q[(size_t)row * n + i] = a * (sine ? std::sin(t) : std::cos(t));
        }
        lam[row] = c * (2.0 - 2.0 * std::cos(2.0 * pi * m / n));
        ++row;
    };
    put(0, false, std::sqrt(1.0 / n));
    for (int m = 1; 2 * m < n; ++m) {
        put(m, false, std::sqrt(2.0 / n));
        put(m, true, std::sqrt(2.0 / n));
    }
    if (n % 2 == 0) put(n / 2, false, std::sqrt(1.0 / n));
}


// Levels of the multigrid preconditioner (level 0 is the SPE operator of the uniform
// grid), the inner-solve vectors and state, in arithmetic type T.
// right-hand side and result of one V-cycle on level 0 (natural-order vectors)
template <typename T> static void mg_bind_level0(MGLevelT<T> &L0, const T *rhs, T *out)
{
    if (L0.split) {
        L0.fin = rhs;
        L0.uout = out;
    } else {
        L0.f = const_cast<T *>(rhs);
        L0.u = out;
    }
}

// one V-cycle on rhs -> out (natural-order level-0 vectors).  With cfg.mg_graphs the
// cycle's launches are captured once per (type, rhs, out) into a HIP graph and replayed;
// the cycle has no memory sets or library calls, only kernels.
template <typename T> static int run_vcycle(cfdmx_ctx *ctx, const T *rhs, T *out)
{
    std::vector<MGLevelT<T>> &L = mg_levels<T>(ctx);
    const int nlev = (int)L.size();
    hipStream_t s = ctx->stream;
    mg_bind_level0(L[0], rhs, out);
    if (!ctx->cfg.mg_graphs) {
        launch_mg_vcycle<T>(s, L.data(), nlev, 0);
        LAUNCHCHK();
        return CFDMX_OK;
    }
    for (const auto &v : ctx->vgraphs)
        if (v.tsize == sizeof(T) && v.rhs == rhs && v.out == out) {
            HIPCHK(hipGraphLaunch(v.exec, s));
            return CFDMX_OK;
        }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_mg_vcycle<T>(s, L.data(), nlev, 0);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    if (le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        FAIL(CFDMX_ERR_HIP, "multigrid: capturing the V-cycle failed: %s",
             hipGetErrorString(le != hipSuccess ? le : ce));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess)
        FAIL(CFDMX_ERR_HIP, "multigrid: instantiating the V-cycle graph failed: %s",
             hipGetErrorString(ie));
    ctx->vgraphs.push_back({exec, sizeof(T), rhs, out});
    HIPCHK(hipGraphLaunch(exec, s));
    return CFDMX_OK;
}

template <typename T> static int setup_multigrid(cfdmx_ctx *ctx)
{
    const Geo &g = ctx->g;
    const size_t N = (size_t)g.N;
    for (int q = 0; q < 8; ++q) {
        T *z = nullptr;
        ALLOC(z, N);
        ctx->pb[q] = z;
    }
    {
        double *z = nullptr; // the correction is FP64 in both modes (k_pb_update_xr)
        ALLOC(z, N);
        ctx->ecorr = z;
        ALLOC(ctx->xback, N);
    }
    std::vector<MGLevelT<T>> &lev = mg_levels<T>(ctx);
    lev.clear();
    auto divs = [](MGLevelT<T> &L) {
        L.dz = make_fastdiv((unsigned)L.nz);
        L.dy = make_fastdiv((unsigned)L.ny);
        L.dzh = make_fastdiv((unsigned)(L.nz >= 2 ? L.nz / 2 : 1));
    };
    MGLevelT<T> L = {};
    L.nx = g.nxc;
    L.ny = g.nyc;
    L.nz = g.nzc;
    L.n = (int)N;
    double h[3] = {ctx->Dx[kGC], ctx->Dy[kGC], ctx->Dz[kGC]};
    double cd[3] = {1.0 / (ctx->Dx[kGC] * ctx->Dxs[kGC]), 1.0 / (ctx->Dy[kGC] * ctx->Dys[kGC]),
                    1.0 / (ctx->Dz[kGC] * ctx->Dzs[kGC])};
    L.cx = (T)cd[0];
    L.cy = (T)cd[1];
    L.cz = (T)cd[2];
    L.fx = L.fy = L.fz = 1;
    L.zper = g.zper;
    divs(L);
    // a level with even nz is stored by colour; level 0 then has u and f of its own,
    // filled from and returned to the natural-order Krylov vectors
    L.split = L.nz % 2 == 0;
    if (L.split) {
        ALLOC(L.u, N);
        ALLOC(L.f, N);
    }
    lev.push_back(L);
    for (;;) {
        const MGLevelT<T> &P = lev.back();
        if (P.n <= kMGCoarseCells) break;
        const double hmin = std::min(h[0], std::min(h[1], h[2]));
        const int fx = (P.nx % 2 == 0 && P.nx >= 4 && h[0] <= 2.0 * hmin) ? 2 : 1;
        const int fy = (P.ny % 2 == 0 && P.ny >= 4 && h[1] <= 2.0 * hmin) ? 2 : 1;
        const int fz = (P.nz % 2 == 0 && P.nz >= 2 && h[2] <= 2.0 * hmin) ? 2 : 1;
        if (fx * fy * fz == 1) break;
        MGLevelT<T> C = {};
        C.nx = P.nx / fx;
        C.ny = P.ny / fy;
        C.nz = P.nz / fz;
        C.n = C.nx * C.ny * C.nz;
        h[0] *= fx;
        h[1] *= fy;
        h[2] *= fz;
        for (int a = 0; a < 3; ++a) cd[a] = 1.0 / (h[a] * h[a]);
        C.cx = (T)cd[0];
        C.cy = (T)cd[1];
        C.cz = (T)cd[2];
        C.fx = fx;
        C.fy = fy;
        C.fz = fz;
        C.zper = g.zper;
        divs(C);
        C.split = C.nz % 2 == 0;
        ALLOC(C.u, (size_t)C.n);
        ALLOC(C.f, (size_t)C.n);
        lev.push_back(C);
    }
    MGLevelT<T> &B = lev.back();
    if (B.n > kMGCoarseCells || B.nx > kMGCoarseAxis || B.ny > kMGCoarseAxis ||
        B.nz > kMGCoarseAxis)
        FAIL(CFDMX_ERR_ARG, "multigrid: the hierarchy of %dx%dx%d stops at %dx%dx%d (the exact "
                            "coarse solve takes at most %d cells and %d per axis); use sizes "
                            "with more factors of two",
             g.nxc, g.nyc, g.nzc, B.nx, B.ny, B.nz, kMGCoarseCells, kMGCoarseAxis);
    // the coarsest level is solved in natural order (a single level is then bound to the
    // Krylov vectors directly); its operator is the Kronecker sum of three 1-D chains
    B.split = 0;
    {
        std::vector<double> Qx, Qy, Qz, lx, ly, lz;
        axis_eigen(B.nx, (double)B.cx, false, Qx, lx);
        axis_eigen(B.ny, (double)B.cy, false, Qy, ly);
        axis_eigen(B.nz, (double)B.cz, B.zper != 0, Qz, lz);
        std::vector<T> pack;
        for (const auto *v : {&Qx, &Qy, &Qz, &lx, &ly, &lz})
            for (double d : *v) pack.push_back((T)d);
        T *dm = nullptr;
        ALLOC(dm, pack.size());
        HIPCHK(hipMemcpy(dm, pack.data(), pack.size() * sizeof(T), hipMemcpyHostToDevice));
        B.fd = dm;
    }
    ALLOC(ctx->pbs, 1);
    ALLOC(ctx->pb_part, 2 * (size_t)g.nblocks + 64);
    ALLOC(ctx->mix, 4);
    HIPCHK(hipHostMalloc((void **)&ctx->pbs_h, sizeof(PBState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&ctx->mix_h, 4 * sizeof(double), hipHostMallocDefault));
    return CFDMX_OK;
}

// metric tables of one axis (cfdmx_device.hpp:Axis) appended to buf
static void axis_tables(const std::vector<double> &D, const std::vector<double> &Ds,
                        std::vector<double> &buf, size_t off[13])
{
    const int n = (int)D.size();
    std::vector<double> t(n);
    auto push = [&](int slot, auto fn) {
        off[slot] = buf.size();
        for (int I = 0; I < n; ++I) buf.push_back(fn(I));
    };
    push(0, [&](int I) { return D[I]; });
    push(1, [&](int I) { return Ds[I]; });
    push(2, [&](int I) { return 1.0 / D[I]; });
    push(3, [&](int I) { return 1.0 / Ds[I]; });
    push(4, [&](int I) { return (Ds[I] * 0.5) / D[I]; });
    auto ok = [&](int I) { return I >= 1 && I + 1 < n; };
    push(5, [&](int I) { return ok(I) ? 0.125 * D[I + 1] * D[I + 1] / Ds[I] : 0.0; });
    push(6, [&](int I) { return ok(I) ? 0.125 * D[I + 1] * D[I + 1] / Ds[I + 1] : 0.0; });
    push(7, [&](int I) { return ok(I) ? 0.125 * D[I] * D[I] / Ds[I - 1] : 0.0; });
    push(8, [&](int I) { return ok(I) ? 0.125 * D[I] * D[I] / Ds[I] : 0.0; });
    push(9, [&](int I) { return ok(I) ? 0.125 * Ds[I] * Ds[I] / D[I] : 0.0; });
    push(10, [&](int I) { return ok(I) ? 0.125 * Ds[I] * Ds[I] / D[I + 1] : 0.0; });
    push(11, [&](int I) { return ok(I) ? 0.125 * Ds[I - 1] * Ds[I - 1] / D[I - 1] : 0.0; });
    push(12, [&](int I) { return ok(I) ? 0.125 * Ds[I - 1] * Ds[I - 1] / D[I] : 0.0; });
}

static int reset_run_state(cfdmx_ctx *ctx);

// SorPipeLine_seq's coefficients (SorPipeLine.seq.hpp:40-51), formed once when every
// spacing table is a single value (the uniform grid): then each cell's six quotients
// and -1/(their sum) are these values, evaluated with the same operations and order
static void sor_constant_coefficients(cfdmx_ctx *ctx)
{
    auto constant = [](const std::vector<double> &a) {
        for (double x : a)
            if (x != a[0]) return false;
        return true;
    };
    SorCoef &sc = ctx->sorc;
    sc.constant = constant(ctx->Dx) && constant(ctx->Dy) && constant(ctx->Dz) &&
                  constant(ctx->Dxs) && constant(ctx->Dys) && constant(ctx->Dzs);
    if (!sc.constant) return;
    const double Dx = ctx->Dx[0], Dy = ctx->Dy[0], Dz = ctx->Dz[0];
    const double Dxs = ctx->Dxs[0], Dys = ctx->Dys[0], Dzs = ctx->Dzs[0];
    sc.c[0] = Dy * Dz / Dxs;
    sc.c[1] = Dy * Dz / Dxs;
    sc.c[2] = Dx * Dz / Dys;
    sc.c[3] = Dx * Dz / Dys;
    sc.c[4] = Dx * Dy / Dzs;
    sc.c[5] = Dx * Dy / Dzs;
    sc.c[6] = -1.0 / (sc.c[0] + sc.c[1] + sc.c[2] + sc.c[3] + sc.c[4] + sc.c[5]);
}

static int setup_device(cfdmx_ctx *ctx)
{
    sor_constant_coefficients(ctx);
    cfdmx_config &c = ctx->cfg;
    Geo &g = ctx->g;
    g.nxc = c.nx;
    g.nyc = c.ny;
    g.nzc = c.nz;
    g.nx = c.nx + 4;
    g.ny = c.ny + 4;
    g.nz = c.nz + 4;
    g.N = c.nx * c.ny * c.nz;
    g.G = g.nx * g.ny * g.nz;
    g.sx = g.ny * g.nz;
    g.sy = g.nz;
    g.cyz = c.ny * c.nz;
    g.zper = c.z_periodic ? 1 : 0;
    g.nblocks = (g.N + kPerBlock - 1) / kPerBlock;
    g.xcd_remap = c.xcd_remap ? 1 : 0;
    {
        auto pow2 = [](double x) {
            int e = 0;
            return x > 0.0 && std::frexp(x, &e) == 0.5;
        };
        bool z2 = true;
        for (int k = 0; k < (int)ctx->Dz.size(); ++k)
            z2 = z2 && pow2(ctx->Dz[k]) && pow2(ctx->Dzs[k]);
        g.zpow2 = z2 ? 1 : 0;
    }
    g.dz = make_fastdiv((unsigned)c.nz);
    g.dy = make_fastdiv((unsigned)c.ny);
    g.ntile = (g.cyz + kTile - 1) / kTile;
    g.nchunk = (c.nx + kChunk - 1) / kChunk;
    g.nmarch = g.ntile * g.nchunk;
    g.cd_tk = 4;
    while (g.cd_tk < g.nzc && g.cd_tk < 32) g.cd_tk *= 2;
    g.cd_tj = kThreads / g.cd_tk;
    g.cd_nkt = (g.nzc + g.cd_tk - 1) / g.cd_tk;
    g.cd_njt = (g.nyc + g.cd_tj - 1) / g.cd_tj;
    g.cd_nb = g.cd_nkt * g.cd_njt * g.nchunk;

    // metric tables in one buffer
    std::vector<double> buf;
    size_t off[3][13];
    axis_tables(ctx->Dx, ctx->Dxs, buf, off[0]);
    axis_tables(ctx->Dy, ctx->Dys, buf, off[1]);
    axis_tables(ctx->Dz, ctx->Dzs, buf, off[2]);
    // pressure coefficients: xm(i) = 1/Dx[i+2]/Dxs[i+1], xp(i) = 1/Dx[i+2]/Dxs[i+2]
    size_t poff[3][2];
    const std::vector<double> *DD[3] = {&ctx->Dx, &ctx->Dy, &ctx->Dz};
    const std::vector<double> *DS[3] = {&ctx->Dxs, &ctx->Dys, &ctx->Dzs};
    const int nc[3] = {c.nx, c.ny, c.nz};
    for (int a = 0; a < 3; ++a) {
        poff[a][0] = buf.size();
        for (int i = 0; i < nc[a]; ++i)
            buf.push_back(1.0 / (*DD[a])[i + kGC] / (*DS[a])[i + kGC - 1]);
        poff[a][1] = buf.size();
        for (int i = 0; i < nc[a]; ++i)
            buf.push_back(1.0 / (*DD[a])[i + kGC] / (*DS[a])[i + kGC]);
    }
    const size_t xoff = buf.size();
    buf.insert(buf.end(), ctx->X.begin(), ctx->X.end());
    const size_t yoff = buf.size();
    buf.insert(buf.end(), ctx->Y.begin(), ctx->Y.end());
    ALLOC(ctx->metric, buf.size());
    HIPCHK(hipMemcpy(ctx->metric, buf.data(), buf.size() * sizeof(double),
                     hipMemcpyHostToDevice));
    const double *m = ctx->metric;
    for (int a = 0; a < 3; ++a) {
        Axis &A = g.ax[a];
        A.D = m + off[a][0];
        A.Ds = m + off[a][1];
        A.rD = m + off[a][2];
        A.rDs = m + off[a][3];
        A.IT = m + off[a][4];
        for (int q = 0; q < 4; ++q) {
            A.qa[q] = m + off[a][5 + q];
            A.qb[q] = m + off[a][9 + q];
        }
        g.pa[a].m = m + poff[a][0];
        g.pa[a].p = m + poff[a][1];
    }
    ctx->Xn = ctx->metric + xoff;
    ctx->Yn = ctx->metric + yoff;

    // fields
    const size_t G = g.G, N = g.N;
    for (int q = 0; q < 3; ++q) {
        ALLOC(ctx->t0[q], G);
        ALLOC(ctx->t1[q], G);
        ALLOC(ctx->t3[q], G);
    }
    ALLOC(ctx->p, G);
    ALLOC(ctx->eta, G);
    ALLOC(ctx->f, 3 * G);
    ALLOC(ctx->x, N);
    ALLOC(ctx->b, N);
    ALLOC(ctx->px1, N);
    ALLOC(ctx->px2, N);
    ALLOC(ctx->B.r0, N);
    ALLOC(ctx->B.r, N);
    ALLOC(ctx->B.as, N);
    ALLOC(ctx->B.q, N);
    ALLOC(ctx->B.p, N);
    ALLOC(ctx->B.ap, N);
    if (ctx->cfg.solver == CFDMX_SOLVER_RBSOR && g.nzc % 2 == 0) ALLOC(ctx->sorw, 2 * N);
    ALLOC(ctx->part, 4 * (size_t)(g.nblocks > g.nmarch ? g.nblocks : g.nmarch) + 64);
    ALLOC(ctx->cfl_part, (size_t)g.cd_nb + 8);
    ALLOC(ctx->diff_part, 3 * (size_t)g.nblocks + 8);
    if (ctx->cfg.smagorinsky_cs > 0.0) {
        // (Cs delta)^2 with delta = pow(Dx Dy Dz, 1/3) (SmagorinskyModel.hpp:74-76), on the
        // host with the C library's pow
        std::vector<double> coef(N);
        for (int i = 0; i < g.nxc; ++i)
            for (int j = 0; j < g.nyc; ++j)
                for (int k = 0; k < g.nzc; ++k) {
                    const double delta = std::pow(ctx->Dx[i + kGC] * ctx->Dy[j + kGC] *
                                                      ctx->Dz[k + kGC],
                                                  1.0 / 3.0);
                    const double cd = ctx->cfg.smagorinsky_cs * delta;
                    coef[((size_t)i * g.nyc + j) * g.nzc + k] = cd * cd;
                }
        ALLOC(ctx->smag_coef, N);
        ALLOC(ctx->viseff, N);
        HIPCHK(hipMemcpy(ctx->smag_coef, coef.data(), N * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(ctx->viseff, 0, N * sizeof(double)));
    }
    for (int q = 0; q < 3; ++q) ALLOC(ctx->bpart[q], 2 * (size_t)g.nmarch + 64);
    ALLOC(ctx->flux_part, 3 * ((size_t)g.cyz / kPerBlock + 2));
    ALLOC(ctx->scal, 16);
    ALLOC(ctx->inflow, (size_t)g.ny);
    ALLOC(ctx->kst, 1);
    ALLOC(ctx->sst, 1);
    ALLOC(ctx->count, 1);
    HIPCHK(hipHostMalloc((void **)&ctx->kst_h, sizeof(KState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&ctx->sst_h, sizeof(SorState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&ctx->scal_h, 16 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&ctx->count_h, sizeof(int), hipHostMallocDefault));
    if (c.precision == CFDMX_PRECISION_MIXED) {
        if (int rc = setup_multigrid<float>(ctx)) return rc;
    } else if (c.precision == CFDMX_PRECISION_FP64_MG) {
        if (int rc = setup_multigrid<double>(ctx)) return rc;
    }

    // reference initial state: u = 1, v = w = 0 in T0/T1/T3 (RUN_CPU.cpp:66-78),
    // p = 0, eta = 0, f = 0 (resize_variable), X_result = 0 (RUN_CPU.cpp:120)
    hipStream_t s = ctx->stream;
    for (int q = 0; q < 3; ++q) {
        const double val = q == 0 ? 1.0 : 0.0;
        launch_fill(s, (int)G, val, ctx->t0[q]);
        launch_fill(s, (int)G, val, ctx->t1[q]);
        launch_fill(s, (int)G, val, ctx->t3[q]);
    }
    launch_fill(s, (int)G, 0.0, ctx->p);
    launch_fill(s, (int)G, 0.0, ctx->eta);
    LAUNCHCHK();
    if (int rc = reset_run_state(ctx)) return rc;

    // boundary table
    BCParams &bc = ctx->bc;
    std::memcpy(bc.num, c.bc_num, sizeof bc.num);
    std::memcpy(bc.dir, c.bc_dir, sizeof bc.dir);
    bc.inflow = nullptr;
    bc.convective = c.outflow_convective;
    bc.conv_coef = c.convective_speed * c.dt / ctx->Dx[g.nx - 3];
    bc.flux_correction = c.flux_correction;
    return CFDMX_OK;
}

// whether every z ghost of the three components of a force field (3*G doubles) is +0.0
static bool force_zghosts_zero(const Geo &g, const double *f)
{
    const int ks[4] = {0, 1, g.nz - 2, g.nz - 1};
    for (int q = 0; q < 3; ++q)
        for (size_t row = 0; row < (size_t)g.nx * g.ny; ++row)
            for (int k : ks) {
                uint64_t bits;
                std::memcpy(&bits, f + (size_t)q * g.G + row * g.nz + k, sizeof bits);
                if (bits) return false;
            }
    return true;
}

// the pressure solve's state of a new run: X_result and its history, the Krylov vectors,
// the forcing and the solver states (a fresh bicgstabRe2, bicgstabRe2.hpp:17-52), the
// counters of the step report
static int reset_run_state(cfdmx_ctx *ctx)
{
    hipStream_t s = ctx->stream;
    const size_t G = (size_t)ctx->g.G, N = (size_t)ctx->g.N;
    launch_fill(s, (int)(3 * G), 0.0, ctx->f);
    ctx->f_ghosts_zero = 1;
    double *zs[] = {ctx->x, ctx->b, ctx->px1, ctx->px2, ctx->B.r0, ctx->B.r, ctx->B.as,
                    ctx->B.q, ctx->B.p, ctx->B.ap};
    for (double *z : zs) launch_fill(s, (int)N, 0.0, z);
    launch_fill(s, 16, 0.0, ctx->scal);
    LAUNCHCHK();
    KState ks = {};
    ks.tol = ctx->cfg.tol;
    ks.restart_factor = 1e-7; // pRestartFactor_ (bicgstabRe2.hpp:52)
    ks.maxit = ctx->cfg.maxit;
    ks.done = 1;
    *ctx->kst_h = ks;
    HIPCHK(hipMemcpyAsync(ctx->kst, ctx->kst_h, sizeof ks, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    ctx->resets = 0;
    ctx->unconverged = 0;
    return CFDMX_OK;
}

static int validate(cfdmx_ctx *ctx)
{
    const cfdmx_config &c = ctx->cfg;
    if (c.nx < 4 || c.ny < 4 || c.nz < 4)
        FAIL(CFDMX_ERR_ARG, "interior sizes must be >= 4 (got %d %d %d)", c.nx, c.ny, c.nz);
    const double G = (double)(c.nx + 4) * (c.ny + 4) * (c.nz + 4);
    if (3.0 * G >= 2147483647.0)
        FAIL(CFDMX_ERR_ARG, "grid too large for 32-bit indexing (3*G = %.0f)", 3.0 * G);
    const double N = (double)c.nx * c.ny * c.nz;
    // FastDiv: n / d is exact for n * d < 2^40, and n * ceil(2^40 / d) must fit 64 bits,
    // i.e. n < d * 2^24; the largest quotients taken are (i, j) row indices, so the
    // second bound is nx * ny < 2^24
    if (N * c.nz >= 1099511627776.0 || (N / c.nz) * c.ny >= 1099511627776.0 ||
        (double)c.nx * c.ny >= 16777216.0)
        FAIL(CFDMX_ERR_ARG, "grid too large for the index decoder");
    if (!(c.dt > 0.0) || !(c.re > 0.0)) FAIL(CFDMX_ERR_ARG, "dt and re must be positive");
    if (c.solver != CFDMX_SOLVER_BICGSTAB && c.solver != CFDMX_SOLVER_RBSOR)
        FAIL(CFDMX_ERR_ARG, "unknown solver %d", c.solver);
    if (c.precond != CFDMX_PRECOND_NONE && c.precond != CFDMX_PRECOND_JACOBI)
        FAIL(CFDMX_ERR_ARG, "unknown precond %d", c.precond);
    if (c.maxit < 0) FAIL(CFDMX_ERR_ARG, "maxit must be >= 0");
    if (!(c.smagorinsky_cs >= 0.0 && c.smagorinsky_cs < 1.0))
        FAIL(CFDMX_ERR_ARG, "smagorinsky_cs must be in [0, 1) (got %g)", c.smagorinsky_cs);
    if (c.steady_check != 0 && c.steady_check != 1)
        FAIL(CFDMX_ERR_ARG, "steady_check must be 0 or 1 (got %d)", c.steady_check);
    if (c.mg_extrapolate != 0 && c.mg_extrapolate != 1)
        FAIL(CFDMX_ERR_ARG, "mg_extrapolate must be 0 or 1 (got %d)", c.mg_extrapolate);
    if (c.precision != CFDMX_PRECISION_FP64 && c.precision != CFDMX_PRECISION_MIXED &&
        c.precision != CFDMX_PRECISION_FP64_MG)
        FAIL(CFDMX_ERR_ARG, "unknown precision %d", c.precision);
    if (c.precision != CFDMX_PRECISION_FP64) {
        if (c.solver != CFDMX_SOLVER_BICGSTAB)
            FAIL(CFDMX_ERR_ARG, "the multigrid pressure solve needs the BiCGSTAB solver");
        if (!(c.inner_tol >= 0.0 && c.inner_tol < 1.0))
            FAIL(CFDMX_ERR_ARG, "inner_tol must be in [0, 1) (0: the mode's default)");
        if (c.grid_kind != CFDMX_GRID_UNIFORM || c.precond != CFDMX_PRECOND_NONE)
            FAIL(CFDMX_ERR_ARG, "the multigrid pressure solve needs the uniform grid and the "
                                "SPE operator (precond none): its levels are uniform "
                                "Laplacians");
    }
    if (c.nz > 600) FAIL(CFDMX_ERR_ARG, "nz above 600 is not supported (LDS ring of the stencil)");
    if (c.z_periodic && c.solver == CFDMX_SOLVER_RBSOR && (c.nz % 2) != 0)
        FAIL(CFDMX_ERR_ARG, "red-black SOR with periodic z needs an even nz");
    if (c.flux_correction) {
        const bool walls = c.bc_num[2][1] == 0.0 && c.bc_dir[2][1] == 0.0 &&
                           c.bc_num[3][1] == 0.0 && c.bc_dir[3][1] == 0.0;
        const bool zclosed = c.z_periodic || (c.bc_num[4][2] == 0.0 && c.bc_dir[4][2] == 0.0 &&
                                              c.bc_num[5][2] == 0.0 && c.bc_dir[5][2] == 0.0);
        if (!walls || !zclosed)
            FAIL(CFDMX_ERR_ARG, "flux_correction needs zero normal velocity on the y faces "
                                "and periodic or zero-normal z faces");
    }
    return CFDMX_OK;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int cfdmx_config_default(cfdmx_config *cfg)
{
    if (!cfg) return CFDMX_ERR_ARG;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->nx = 128;
    cfg->ny = 64;
    cfg->nz = 4;
    cfg->lx = 20.0;
    cfg->ly = 10.0;
    cfg->lz = 0.5;
    cfg->grid_kind = CFDMX_GRID_UNIFORM;
    cfg->re = 100.0;
    cfg->dt = 0.002;
    // FLOWpassing.hpp:6-40
    const double num[6][3] = {{0, 0, 0}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}};
    const double dir[6][3] = {{1, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    std::memcpy(cfg->bc_num, num, sizeof num);
    std::memcpy(cfg->bc_dir, dir, sizeof dir);
    cfg->convective_speed = 1.0;
    cfg->solver = CFDMX_SOLVER_BICGSTAB;
    cfg->precond = CFDMX_PRECOND_NONE;
    cfg->tol = 1e-3;   // simpulationVariable::p_criteria (0_1_SimulationVariable.hpp:61)
    cfg->maxit = 5000; // bicgstabRe2::iters_max_ (bicgstabRe2.hpp:47)
    cfg->sor_omega = 1.8; // p_sor_omega (0_1_SimulationVariable.hpp:65)
    cfg->skip_first_pressure = 1;
    cfg->precision = CFDMX_PRECISION_FP64;
    cfg->inner_tol = 0.0; // the mode's own default (kInnerTolMixed, kInnerTolFp64)
    cfg->xcd_remap = 1;
    cfg->mg_graphs = 1;
    cfg->smagorinsky_cs = 0.0;
    cfg->steady_check = 0;
    cfg->mg_extrapolate = 1;
    return CFDMX_OK;
}

const char *cfdmx_last_error(const cfdmx_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

static int plot3d_flush_impl(cfdmx_ctx *ctx);

int cfdmx_destroy(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_OK;
    (void)plot3d_flush_impl(ctx);
    if (ctx->p3d.side) (void)hipStreamDestroy(ctx->p3d.side);
    if (ctx->p3d.ready) (void)hipEventDestroy(ctx->p3d.ready);
    if (ctx->p3d.host) (void)hipHostFree(ctx->p3d.host);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &v : ctx->vgraphs) (void)hipGraphExecDestroy(v.exec);
    for (void *p : ctx->allocs) (void)hipFree(p);
    prof_harvest(ctx);
    for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
    if (ctx->kst_h) (void)hipHostFree(ctx->kst_h);
    if (ctx->pbs_h) (void)hipHostFree(ctx->pbs_h);
    for (void *q : ctx->adj.mem) (void)hipFree(q);
    if (ctx->adj.allX) (void)hipFree(ctx->adj.allX);
    if (ctx->adj.kadj_h) (void)hipHostFree(ctx->adj.kadj_h);
    if (ctx->adj.out_h) (void)hipHostFree(ctx->adj.out_h);
    if (ctx->mix_h) (void)hipHostFree(ctx->mix_h);
    if (ctx->sst_h) (void)hipHostFree(ctx->sst_h);
    if (ctx->scal_h) (void)hipHostFree(ctx->scal_h);
    if (ctx->count_h) (void)hipHostFree(ctx->count_h);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CFDMX_OK;
}

// the scratch array's address on the current device, looked up once per device
static int scratch_setup(cfdmx_ctx *ctx)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        FAIL(CFDMX_ERR_HIP, "device %d outside the %d devices the library tracks", dev,
             kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *addr = nullptr;
        HIPCHK(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
        if (!addr) FAIL(CFDMX_ERR_HIP, "hipGetSymbolAddress returned no address for scratch");
        g_scratch[dev] = static_cast<double *>(addr);
    }
    ctx->device = dev;
    ctx->scratch_dev = g_scratch[dev];
    return CFDMX_OK;
}

int cfdmx_create(const cfdmx_config *cfg, cfdmx_ctx **out)
{
    if (!cfg || !out) {
        g_create_error = "cfdmx_create: null argument";
        return CFDMX_ERR_ARG;
    }
    *out = nullptr;
    cfdmx_ctx *ctx = new cfdmx_ctx();
    ctx->cfg = *cfg;
    int rc = validate(ctx);
    if (!rc) rc = build_grid(ctx);
    if (!rc) {
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            ctx->err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
            rc = CFDMX_ERR_HIP;
        }
    }
    if (!rc) rc = setup_device(ctx);
    if (!rc) rc = scratch_setup(ctx);
    if (rc) {
        g_create_error = ctx->err;
        cfdmx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return CFDMX_OK;
}

int cfdmx_sizes(const cfdmx_ctx *ctx, long long *G, long long *N)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (G) *G = ctx->g.G;
    if (N) *N = ctx->g.N;
    return CFDMX_OK;
}

int cfdmx_grid_metrics(const cfdmx_ctx *ctx, double *out)
{
    if (!ctx || !out) return CFDMX_ERR_ARG;
    size_t o = 0;
    for (const std::vector<double> *v : {&ctx->X, &ctx->Y, &ctx->Z, &ctx->Dx, &ctx->Dy, &ctx->Dz,
                                         &ctx->Dxs, &ctx->Dys, &ctx->Dzs})
        for (double d : *v) out[o++] = d;
    return CFDMX_OK;
}

static double **vel_set(cfdmx_ctx *ctx, int which)
{
    return which == 0 ? ctx->t0 : which == 1 ? ctx->t1 : which == 2 ? ctx->t3 : nullptr;
}

static int copy_h2d(cfdmx_ctx *ctx, double *dst, const double *src, size_t n)
{
    if (!src) FAIL(CFDMX_ERR_ARG, "null host pointer");
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFDMX_OK;
}

static int copy_d2h(cfdmx_ctx *ctx, double *dst, const double *src, size_t n)
{
    if (!dst) FAIL(CFDMX_ERR_ARG, "null host pointer");
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFDMX_OK;
}

int cfdmx_upload_velocity(cfdmx_ctx *ctx, int which, const double *u, const double *v,
                          const double *w)
{
    if (!ctx) return CFDMX_ERR_ARG;
    double **t = vel_set(ctx, which);
    if (!t) FAIL(CFDMX_ERR_ARG, "which must be 0 (T0), 1 (T1) or 2 (T3)");
    const double *h[3] = {u, v, w};
    for (int q = 0; q < 3; ++q) {
        int rc = copy_h2d(ctx, t[q], h[q], ctx->g.G);
        if (rc) return rc;
    }
    return CFDMX_OK;
}

int cfdmx_download_velocity(cfdmx_ctx *ctx, int which, double *u, double *v, double *w)
{
    if (!ctx) return CFDMX_ERR_ARG;
    double **t = vel_set(ctx, which);
    if (!t) FAIL(CFDMX_ERR_ARG, "which must be 0 (T0), 1 (T1) or 2 (T3)");
    double *h[3] = {u, v, w};
    for (int q = 0; q < 3; ++q) {
        int rc = copy_d2h(ctx, h[q], t[q], ctx->g.G);
        if (rc) return rc;
    }
    return CFDMX_OK;
}

int cfdmx_upload_pressure(cfdmx_ctx *ctx, const double *p)
{
    return ctx ? copy_h2d(ctx, ctx->p, p, ctx->g.G) : CFDMX_ERR_ARG;
}
int cfdmx_download_pressure(cfdmx_ctx *ctx, double *p)
{
    return ctx ? copy_d2h(ctx, p, ctx->p, ctx->g.G) : CFDMX_ERR_ARG;
}
int cfdmx_upload_xresult(cfdmx_ctx *ctx, const double *x)
{
    return ctx ? copy_h2d(ctx, ctx->x, x, ctx->g.N) : CFDMX_ERR_ARG;
}
int cfdmx_download_xresult(cfdmx_ctx *ctx, double *x)
{
    return ctx ? copy_d2h(ctx, x, ctx->x, ctx->g.N) : CFDMX_ERR_ARG;
}
int cfdmx_download_rhs(cfdmx_ctx *ctx, double *b)
{
    return ctx ? copy_d2h(ctx, b, ctx->b, ctx->g.N) : CFDMX_ERR_ARG;
}
int cfdmx_upload_rhs(cfdmx_ctx *ctx, const double *b)
{
    return ctx ? copy_h2d(ctx, ctx->b, b, ctx->g.N) : CFDMX_ERR_ARG;
}
int cfdmx_upload_eta(cfdmx_ctx *ctx, const double *eta)
{
    return ctx ? copy_h2d(ctx, ctx->eta, eta, ctx->g.G) : CFDMX_ERR_ARG;
}
int cfdmx_download_eta(cfdmx_ctx *ctx, double *eta)
{
    return ctx ? copy_d2h(ctx, eta, ctx->eta, ctx->g.G) : CFDMX_ERR_ARG;
}
int cfdmx_download_force(cfdmx_ctx *ctx, double *f)
{
    return ctx ? copy_d2h(ctx, f, ctx->f, 3 * (size_t)ctx->g.G) : CFDMX_ERR_ARG;
}

int cfdmx_upload_force(cfdmx_ctx *ctx, const double *f)
{
    if (!ctx || !f) return CFDMX_ERR_ARG;
    const int rc = copy_h2d(ctx, ctx->f, f, 3 * (size_t)ctx->g.G);
    if (!rc) ctx->f_ghosts_zero = force_zghosts_zero(ctx->g, f) ? 1 : 0;
    return rc;
}

int cfdmx_set_inflow(cfdmx_ctx *ctx, const double *profile)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (!profile) {
        ctx->bc.inflow = nullptr;
        return CFDMX_OK;
    }
    int rc = copy_h2d(ctx, ctx->inflow, profile, ctx->g.ny);
    if (rc) return rc;
    ctx->bc.inflow = ctx->inflow;
    return CFDMX_OK;
}

static int start_impl(cfdmx_ctx *ctx)
{
    hipStream_t s = ctx->stream;
    const Geo &g = ctx->g;
    for (int q = 0; q < 3; ++q) {
        HIPCHK(hipMemcpyAsync(ctx->t1[q], ctx->t0[q], g.G * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(ctx->t3[q], ctx->t0[q], g.G * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    }
    launch_bc(g, s, ctx->bc, 0, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->p, ctx->t1[0],
              ctx->t1[1], ctx->t1[2], ctx->flux_part);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(s));
    ctx->loop = 1;
    ctx->time = 0.0;
    ctx->started = 1;
    return CFDMX_OK;
}

int cfdmx_start(cfdmx_ctx *ctx) { return ctx ? start_impl(ctx) : CFDMX_ERR_ARG; }

int cfdmx_init_flow(cfdmx_ctx *ctx, double u0, double amp, unsigned long long seed)
{
    if (!ctx) return CFDMX_ERR_ARG;
    // a new run: the solver starts cold, as runSeqOmp's freshly built solver does
    if (int rc = reset_run_state(ctx)) return rc;
    launch_init_flow(ctx->g, ctx->stream, u0, amp, seed, ctx->t0[0], ctx->t0[1], ctx->t0[2],
                     ctx->p);
    LAUNCHCHK();
    return start_impl(ctx);
}

int cfdmx_set_eta_cylinder_z(cfdmx_ctx *ctx, double cx, double cy, double radius, int nsub)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (nsub <= 0) FAIL(CFDMX_ERR_ARG, "nsub must be positive");
    ctx->mask_smoothed = 0;
    launch_fill(ctx->stream, ctx->g.G, 0.0, ctx->eta); // RUN_CPU.cpp:228
    launch_eta_cylz(ctx->g, ctx->stream, ctx->Xn, ctx->Yn, cx, cy, radius, nsub,
                    std::pow(nsub, -2), ctx->eta);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFDMX_OK;
}

int cfdmx_set_eta_smoothed(cfdmx_ctx *ctx, double cx, double cy, double R, double eps)
{
    // smoothed Heaviside of R - distance on the k = 0 plane, copied along z like
    // DFIB_CylinderZ (setEta.hpp:200-203); evaluated on the host with libm
    if (!ctx) return CFDMX_ERR_ARG;
    if (!(eps > 0.0)) FAIL(CFDMX_ERR_ARG, "eps must be positive");
    const Geo &g = ctx->g;
    std::vector<double> eta((size_t)g.G, 0.0);
    const double pi = 3.14159265358979323846;
    for (int i = kGC; i < g.nx - kGC; ++i)
        for (int j = kGC; j < g.ny - kGC; ++j) {
            const double xc = (ctx->X[i] + ctx->X[i + 1]) * 0.5 - cx;
            const double yc = (ctx->Y[j] + ctx->Y[j + 1]) * 0.5 - cy;
            const double phi = R - std::sqrt(xc * xc + yc * yc);
            double e;
            if (phi < -eps) e = 0.0;
            else if (phi > eps) e = 1.0;
            else e = 0.5 * (1.0 + phi / eps + std::sin(pi * phi / eps) / pi);
            eta[((size_t)i * g.ny + j) * g.nz] = e;
        }
    for (int i = 0; i < g.nx - 1; ++i)
        for (int j = 0; j < g.ny - 1; ++j)
            for (int k = 1; k < g.nz - 1; ++k)
                eta[((size_t)i * g.ny + j) * g.nz + k] = eta[((size_t)i * g.ny + j) * g.nz];
    ctx->mask_smoothed = 1;
    ctx->mask_cx = cx;
    ctx->mask_cy = cy;
    ctx->mask_r = R;
    ctx->mask_eps = eps;
    return copy_h2d(ctx, ctx->eta, eta.data(), g.G);
}

int cfdmx_set_physics(cfdmx_ctx *ctx, double re, double dt)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (!(dt > 0.0) || !(re > 0.0)) FAIL(CFDMX_ERR_ARG, "dt and re must be positive");
    ctx->cfg.re = re;
    ctx->cfg.dt = dt;
    ctx->bc.conv_coef = ctx->cfg.convective_speed * dt / ctx->Dx[ctx->g.nx - 3];
    return CFDMX_OK;
}

int cfdmx_set_tolerance(cfdmx_ctx *ctx, double tol, int maxit)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (!(tol > 0.0) || maxit < 0) FAIL(CFDMX_ERR_ARG, "tol must be positive and maxit >= 0");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ctx->kst_h, ctx->kst, sizeof(KState), hipMemcpyDeviceToHost));
    ctx->kst_h->tol = tol;
    ctx->kst_h->maxit = maxit;
    HIPCHK(hipMemcpy(ctx->kst, ctx->kst_h, sizeof(KState), hipMemcpyHostToDevice));
    ctx->cfg.tol = tol;
    ctx->cfg.maxit = maxit;
    return CFDMX_OK;
}

int cfdmx_smagorinsky(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (!ctx->viseff)
        FAIL(CFDMX_ERR_STATE, "the Smagorinsky model is off (smagorinsky_cs = 0)");
    launch_smagorinsky(ctx->g, ctx->stream, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->smag_coef,
                       ctx->viseff);
    LAUNCHCHK();
    return CFDMX_OK;
}

int cfdmx_download_viseff(cfdmx_ctx *ctx, double *viseff)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (!ctx->viseff)
        FAIL(CFDMX_ERR_STATE, "the Smagorinsky model is off (smagorinsky_cs = 0)");
    return copy_d2h(ctx, viseff, ctx->viseff, ctx->g.N);
}

int cfdmx_convdiff(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    PROF(CFDMX_K_CONVDIFF,
         launch_convdiff(ctx->g, ctx->stream, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->t1[0],
                         ctx->t1[1], ctx->t1[2], ctx->cfg.dt, 1.0 / ctx->cfg.re, ctx->cfl_part,
                         ctx->viseff));
    launch_max_fin(ctx->g, ctx->stream, ctx->cfl_part, ctx->scal + 0);
    LAUNCHCHK();
    return CFDMX_OK;
}

int cfdmx_bc_main(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    launch_bc(ctx->g, ctx->stream, ctx->bc, 1, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->p,
              nullptr, nullptr, nullptr, ctx->flux_part);
    LAUNCHCHK();
    return CFDMX_OK;
}

int cfdmx_bc_copy(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    launch_bc_copy(ctx->g, ctx->stream, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->t1[0],
                   ctx->t1[1], ctx->t1[2]);
    LAUNCHCHK();
    return CFDMX_OK;
}

int cfdmx_pressure_rhs(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    PROF(CFDMX_K_RHS, launch_rhs(ctx->g, ctx->stream, ctx->cfg.precond, ctx->t1[0], ctx->t1[1],
                                 ctx->t1[2], ctx->cfg.dt, ctx->b));
    LAUNCHCHK();
    return CFDMX_OK;
}

// One FP64 BiCGSTAB solve A x = b (bicgstabRe2.hpp:82-180) on the given vectors and
// solver state; the Krylov work vectors and partials are shared scratch.
struct SolveTarget {
    const double *b;
    double *x;
    double **px1, **px2; // history copies (bicgstabRe2.hpp:94-95), swapped in place
    KState *kst, *kst_h;
};

static int solve_bicgstab_core(cfdmx_ctx *ctx, const SolveTarget &T, int *iters, double *res)
{
    hipStream_t s = ctx->stream;
    const Geo &g = ctx->g;
    const int pre = ctx->cfg.precond;
    std::swap(*T.px1, *T.px2); // px2 <- px1, then px1 <- x (bicgstabRe2.hpp:94-95)
    PROF(CFDMX_K_BICG_INIT, launch_bicg_init(g, s, pre, T.b, T.x, ctx->B.r0, ctx->B.r, *T.px1,
                                             ctx->part, T.kst));
    PROF(CFDMX_K_BICG_CHECK, launch_bicg_fin0(g, s, ctx->part, T.kst));
    LAUNCHCHK();
    int batch = 4, m = 0;
    for (;;) {
        for (int q = 0; q < batch; ++q, ++m) {
            PROF(CFDMX_K_BICG_A, launch_bicg_A(g, s, pre, ctx->B, ctx->bpart, T.kst, m));
            PROF(CFDMX_K_BICG_B, launch_bicg_B(g, s, pre, ctx->B, ctx->bpart, T.kst, m));
            PROF(CFDMX_K_BICG_C, launch_bicg_C(g, s, ctx->B, T.x, ctx->bpart, T.kst, m));
        }
        PROF(CFDMX_K_BICG_CHECK, launch_bicg_check(g, s, ctx->bpart, T.kst, m - 1));
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(T.kst_h, T.kst, sizeof(KState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        prof_harvest(ctx);
        if (T.kst_h->done) break;
        batch = batch < 64 ? batch * 2 : 64;
    }
    if (T.kst_h->done == 2) {
        // NaN/Inf reset at the iteration limit (bicgstabRe2.hpp:133-153)
        launch_set_int(s, ctx->count, 0);
        launch_nanfix(s, g.N, T.x, *T.px1, ctx->count);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(ctx->count_h, ctx->count, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const int check = *ctx->count_h;
        if (check > g.N * 0.1) {
            launch_copy(s, g.N, *T.px2, T.x);
            LAUNCHCHK();
            ++ctx->resets;
        } else if (check != 0) {
            ++ctx->resets;
        }
    }
    if (iters) *iters = T.kst_h->iter;
    if (res) *res = T.kst_h->res;
    return CFDMX_OK;
}

static int solve_bicgstab_impl(cfdmx_ctx *ctx, int *iters, double *res)
{
    SolveTarget T = {ctx->b, ctx->x, &ctx->px1, &ctx->px2, ctx->kst, ctx->kst_h};
    return solve_bicgstab_core(ctx, T, iters, res);
}

// Multigrid pressure solve (no reference counterpart): FP64 iterative refinement.  Each
// outer pass forms r = b - A x in FP64 and stops when ||r|| / ||b|| < tol; otherwise a
// BiCGSTAB preconditioned by a multigrid V-cycle (FP64 sums), in float for mixed
// precision or in double for the FP64 multigrid mode, solves A e = r to the relative
// tolerance max(inner_tol, 0.3 tol ||b|| / ||r||) and x += e in FP64.  The iteration
// limit bounds the inner iterations summed over the outer passes.  A pass that raises
// the FP64 residual (a stalled inner Krylov process: on an inconsistent system e can grow
// along the constant null vector until x + e has lost x's low bits) is dropped: x is
// restored from the copy the update kept, bit for bit, and the solve ends there.
static constexpr int kMaxOuter = 40;
// inner_tol = 0 takes the mode's default, the fastest measured at 512x256x64 (tol 1e-10,
// extrapolated starts): 1e-5 in both, where one pass usually reaches tol.  With the
// correction accumulated in FP64 the FP32 inner solve takes the pass as far as the FP64
// one does, in the same iterations (2.14 per step of the developed flow; 2.78 ms at 1e-5
// against 2.89 ms at 1e-2 in the mixed mode, 3.78 ms in the FP64 multigrid mode).
static constexpr double kInnerTolMixed = 1e-5, kInnerTolFp64 = 1e-5;

} // extern "C" (the multigrid solve below is templated on its arithmetic type)

// BiCGSTAB in T with the multigrid V-cycle as right preconditioner on A e = r:
// vectors r, r0, p, v, s, t, p^, s^ (pb[0..7]), solution e; returns the iterations
template <typename T>
static int solve_pb_inner(cfdmx_ctx *ctx, double rr, double tol, int maxit, int *iters,
                          double *res)
{
    hipStream_t s = ctx->stream;
    const Geo &g = ctx->g;
    T *const *pv = reinterpret_cast<T *const *>(ctx->pb);
    T *r = pv[0], *r0 = pv[1], *p = pv[2], *v = pv[3], *sv = pv[4], *t = pv[5], *ph = pv[6],
      *sh = pv[7];
    double *e = static_cast<double *>(ctx->ecorr);
    int rc = CFDMX_OK;
    PBState ps = {};
    ps.rho = rr; // (r, r0) with r0 = r
    ps.norm0 = std::sqrt(rr);
    ps.tol = tol;
    ps.maxit = maxit;
    ps.done = (ps.norm0 > 0.0 && maxit > 0) ? 0 : 1;
    *ctx->pbs_h = ps;
    HIPCHK(hipMemcpyAsync(ctx->pbs, ctx->pbs_h, sizeof ps, hipMemcpyHostToDevice, s));
    // e starts at 0: the first iteration's update writes it whole; a solve that ends
    // before its first iteration leaves e = 0 here
    if (ps.done) launch_fill(s, (int)g.N, 0.0, e);
    std::vector<MGLevelT<T>> &L = mg_levels<T>(ctx);
    for (int it = 0; !ctx->pbs_h->done; ++it) {
        PROF(CFDMX_K_PB, launch_pb_update_p<T>(g, s, it == 0, r, v, p, ctx->pbs));
        PROF(CFDMX_K_MG, rc = run_vcycle<T>(ctx, p, ph));
        if (rc) return rc;
        PROF(CFDMX_K_PB, launch_pb_spmv<T>(g, s, L[0], 1, ph, v, r0, ctx->part, ctx->pbs));
        PROF(CFDMX_K_PB, launch_pb_update_s<T>(g, s, ctx->part, r, v, sv, ctx->pbs));
        PROF(CFDMX_K_MG, rc = run_vcycle<T>(ctx, sv, sh));
        if (rc) return rc;
        PROF(CFDMX_K_PB, launch_pb_spmv<T>(g, s, L[0], 2, sh, t, sv, ctx->part, ctx->pbs));
        PROF(CFDMX_K_PB, launch_pb_update_xr<T>(g, s, it == 0, ctx->part, ph, sh, sv, t, r0, e,
                                                r, ctx->pb_part, ctx->pbs));
        launch_pb_end(g, s, ctx->pb_part, ctx->pbs);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(ctx->pbs_h, ctx->pbs, sizeof(PBState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        prof_harvest(ctx);
    }
    *iters = ctx->pbs_h->iter;
    *res = ctx->pbs_h->res;
    return CFDMX_OK;
}

template <typename T>
static int solve_refine(cfdmx_ctx *ctx, const double *b, double *x, int *iters, double *res)
{
    hipStream_t s = ctx->stream;
    const Geo &g = ctx->g;
    const cfdmx_config &c = ctx->cfg;
    T *const *pv = reinterpret_cast<T *const *>(ctx->pb);
    double *e = static_cast<double *>(ctx->ecorr);
    // the FP64 residual of x into pv[0], pv[1]; ||r|| / ||b||, and the inner solve's
    // start (T(r), T(r)) in rr
    auto residual = [&](double &rr, double &rel) -> int {
        PROF(CFDMX_K_BICG_INIT, launch_mixed_resid<T>(g, s, 0, b, x, pv[0], pv[1], ctx->part));
        launch_mixed_fold(g, s, ctx->part, ctx->mix);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(ctx->mix_h, ctx->mix, 3 * sizeof(double), hipMemcpyDeviceToHost,
                              s));
        HIPCHK(hipStreamSynchronize(s));
        prof_harvest(ctx);
        rel = std::sqrt(ctx->mix_h[1]) / std::sqrt(ctx->mix_h[0]);
        rr = ctx->mix_h[2];
        return CFDMX_OK;
    };
    int total = 0, outer = 0;
    double res64 = 0.0, rr = 0.0;
    int rc = residual(rr, res64);
    if (rc) return rc;
    const double itol = c.inner_tol > 0.0
                            ? c.inner_tol
                            : (std::is_same<T, float>::value ? kInnerTolMixed : kInnerTolFp64);
    for (;;) {
        if (!(res64 >= c.tol)) break; // converged (or NaN: nothing left to refine)
        if (total >= c.maxit || outer >= kMaxOuter) break;
        int it = 0;
        double ires = 0.0;
        rc = solve_pb_inner<T>(ctx, rr, std::fmax(itol, 0.3 * c.tol / res64), c.maxit - total,
                               &it, &ires);
        if (rc) return rc;
        total += it;
        ++outer;
        if (!std::isfinite(ires)) break; // diverged: keep x, report res64
        PROF(CFDMX_K_BICG_INIT, launch_mixed_update(g, s, x, e, ctx->xback));
        double rr_new = 0.0, res_new = 0.0;
        if ((rc = residual(rr_new, res_new))) return rc;
        if (!(res_new <= res64)) {
            // the pass raises the residual: x as it was, whose residual is res64
            launch_copy(s, (int)g.N, ctx->xback, x);
            LAUNCHCHK();
            break;
        }
        res64 = res_new;
        rr = rr_new;
    }
    ctx->mix_outer = outer;
    if (iters) *iters = total;
    if (res) *res = res64;
    return CFDMX_OK;
}

// the multigrid solve of A x = b (pressure, or the pressure adjoint) in the
// configured precision
static int solve_mg_target(cfdmx_ctx *ctx, const double *b, double *x, int *iters, double *res)
{
    return ctx->cfg.precision == CFDMX_PRECISION_MIXED
               ? solve_refine<float>(ctx, b, x, iters, res)
               : solve_refine<double>(ctx, b, x, iters, res);
}

// The time step's pressure solve in a multigrid mode.  With mg_extrapolate the start is
// 2 x_n - x_(n-1) from the two solutions before this one; px1 (in these modes only this
// history, and checkpointed) holds x_(n-1).  The solves before this step's follow from
// the step counter (the first step has none with skip_first_pressure), so a run resumed
// from a checkpoint or recomputed from an adjoint checkpoint starts the same way.
static int solve_mixed_impl(cfdmx_ctx *ctx, int *iters, double *res)
{
    if (ctx->cfg.mg_extrapolate) {
        const int solved_before = (ctx->loop - 1) - (ctx->cfg.skip_first_pressure ? 1 : 0);
        if (solved_before >= 2) launch_extrapolate(ctx->g, ctx->stream, ctx->x, ctx->px1);
        else launch_copy(ctx->stream, ctx->g.N, ctx->x, ctx->px1);
        LAUNCHCHK();
    }
    return solve_mg_target(ctx, ctx->b, ctx->x, iters, res);
}

extern "C" {

int cfdmx_solve_bicgstab(cfdmx_ctx *ctx, int *iters, double *residual)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (ctx->cfg.precision != CFDMX_PRECISION_FP64) return solve_mixed_impl(ctx, iters, residual);
    return solve_bicgstab_impl(ctx, iters, residual);
}

static int solve_rbsor_impl(cfdmx_ctx *ctx, int *iters, double *err)
{
    hipStream_t s = ctx->stream;
    const Geo &g = ctx->g;
    SorState ss = {};
    ss.tol = ctx->cfg.tol;
    ss.itmax = ctx->cfg.maxit;
    *ctx->sst_h = ss;
    HIPCHK(hipMemcpyAsync(ctx->sst, ctx->sst_h, sizeof ss, hipMemcpyHostToDevice, s));
    // even nzc: colour-split storage, p's natural order written back at the end; the
    // working set is the context's own when it is configured for red-black SOR, else the
    // Krylov vector r0 (p of both colours) and b (the right-hand side).  Odd nzc: the
    // natural-order half sweep
    const bool split = (g.nzc % 2) == 0;
    double *pc = ctx->sorw ? ctx->sorw : ctx->B.r0;
    double *rhs = ctx->sorw ? ctx->sorw + g.N : ctx->b;
    if (split)
        PROF(CFDMX_K_RHS, launch_sor_split(g, s, ctx->t1[0], ctx->t1[1], ctx->t1[2],
                                           ctx->cfg.dt, ctx->p, rhs, pc));
    else
        PROF(CFDMX_K_RHS, launch_sor_rhs(g, s, ctx->t1[0], ctx->t1[1], ctx->t1[2],
                                         ctx->cfg.dt, ctx->b));
    LAUNCHCHK();
    int batch = 4;
    for (;;) {
        for (int q = 0; q < batch; ++q) {
            for (int colour = 0; colour < 2; ++colour) {
                if (split)
                    PROF(CFDMX_K_SOR, launch_sor_rb(g, s, colour, ctx->cfg.sor_omega,
                                                    ctx->sorc, rhs, pc, ctx->p, ctx->part,
                                                    ctx->sst));
                else
                    PROF(CFDMX_K_SOR, launch_sor_half(g, s, colour, ctx->cfg.sor_omega,
                                                      ctx->b, ctx->p, ctx->part, ctx->sst));
            }
            if (split) launch_sor_rb_fin(g, s, ctx->part, ctx->sst);
            else launch_sor_fin(g, s, ctx->part, ctx->sst);
        }
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(ctx->sst_h, ctx->sst, sizeof ss, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        prof_harvest(ctx);
        if (ctx->sst_h->done) break;
        batch = batch < 64 ? batch * 2 : 64;
    }
    if (split) {
        launch_sor_merge(g, s, pc, ctx->p);
        LAUNCHCHK();
    }
    if (iters) *iters = ctx->sst_h->iters - 1; // SorPipeLine.seq.hpp:123
    if (err) *err = ctx->sst_h->residual;
    return CFDMX_OK;
}

int cfdmx_solve_rbsor(cfdmx_ctx *ctx, int *iters, double *err)
{
    if (!ctx) return CFDMX_ERR_ARG;
    return solve_rbsor_impl(ctx, iters, err);
}

int cfdmx_transform(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    launch_transform(ctx->g, ctx->stream, ctx->x, ctx->p);
    LAUNCHCHK();
    return CFDMX_OK;
}

int cfdmx_project_ib(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    launch_project(ctx->g, ctx->stream, 0, ctx->cfg.dt, ctx->cfg.outflow_convective, ctx->x,
                   ctx->p, ctx->eta, ctx->t1[0], ctx->t1[1], ctx->t1[2], ctx->t3[0],
                   ctx->t3[1], ctx->t3[2], ctx->f, ctx->part);
    launch_cdcl_fin(ctx->g, ctx->stream, ctx->part, ctx->cfg.lz, ctx->scal + 1);
    LAUNCHCHK();
    return CFDMX_OK;
}

int cfdmx_swap_t0_t3(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    for (int q = 0; q < 3; ++q) std::swap(ctx->t0[q], ctx->t3[q]);
    return CFDMX_OK;
}

static int read_scalars(cfdmx_ctx *ctx)
{
    HIPCHK(hipMemcpyAsync(ctx->scal_h, ctx->scal, 16 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_harvest(ctx);
    return CFDMX_OK;
}

int cfdmx_cd_cl(cfdmx_ctx *ctx, double *cd, double *cl)
{
    if (!ctx) return CFDMX_ERR_ARG;
    launch_force_partial(ctx->g, ctx->stream, ctx->f, ctx->part);
    launch_cdcl_fin(ctx->g, ctx->stream, ctx->part, ctx->cfg.lz, ctx->scal + 1);
    LAUNCHCHK();
    int rc = read_scalars(ctx);
    if (rc) return rc;
    if (cd) *cd = ctx->scal_h[1];
    if (cl) *cl = ctx->scal_h[2];
    return CFDMX_OK;
}

int cfdmx_cd_cl_noalloc(cfdmx_ctx *ctx, const double *f_dev, double *cdcl_dev)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (!f_dev || !cdcl_dev) FAIL(CFDMX_ERR_ARG, "cfdmx_cd_cl_noalloc: null device pointer");
    if (ctx->g.nblocks > kScratchBlocks)
        FAIL(CFDMX_ERR_ARG,
             "cfdmx_cd_cl_noalloc: %d interior cells exceed the scratch limit of %d "
             "(%d reduction blocks of 1024 cells)",
             ctx->g.N, kScratchBlocks * 1024, kScratchBlocks);
    if (!ctx->scratch_dev) FAIL(CFDMX_ERR_STATE, "cfdmx_cd_cl_noalloc: no scratch address");
    // every block writes its two partials, so nothing left by an earlier call is read
    launch_force_partial(ctx->g, ctx->stream, f_dev, ctx->scratch_dev);
    launch_cdcl_fin(ctx->g, ctx->stream, ctx->scratch_dev, ctx->cfg.lz, cdcl_dev);
    LAUNCHCHK();
    return CFDMX_OK;
}

int cfdmx_cfl(cfdmx_ctx *ctx, double *cfl)
{
    if (!ctx) return CFDMX_ERR_ARG;
    int rc = read_scalars(ctx);
    if (rc) return rc;
    if (cfl) *cfl = ctx->scal_h[0];
    return CFDMX_OK;
}

int cfdmx_apply_operator(cfdmx_ctx *ctx, const double *x, double *r)
{
    if (!ctx) return CFDMX_ERR_ARG;
    int rc = copy_h2d(ctx, ctx->B.as, x, ctx->g.N);
    if (rc) return rc;
    launch_apply(ctx->g, ctx->stream, ctx->cfg.precond, ctx->B.as, ctx->B.r);
    LAUNCHCHK();
    return copy_d2h(ctx, r, ctx->B.r, ctx->g.N);
}

int cfdmx_apply_operator_transpose(cfdmx_ctx *ctx, const double *y, double *r)
{
    if (!ctx) return CFDMX_ERR_ARG;
    int rc = copy_h2d(ctx, ctx->B.as, y, ctx->g.N);
    if (rc) return rc;
    launch_apply_t(ctx->g, ctx->stream, ctx->cfg.precond, ctx->B.as, ctx->B.r);
    LAUNCHCHK();
    return copy_d2h(ctx, r, ctx->B.r, ctx->g.N);
}

// one iteration of the runSeqOmp loop body (RUN_CPU.cpp:257-436).  x_given (the
// adjoint's recompute): the pressure solution this step is known to reach, copied into
// X_result in place of the right-hand side and the solve; every other stage is the same,
// so the step gives the same bits as the one that solved.
static int step_once(cfdmx_ctx *ctx, int *iters, double *err, int *solved,
                     const double *x_given = nullptr)
{
    hipStream_t s = ctx->stream;
    const Geo &g = ctx->g;
    const cfdmx_config &c = ctx->cfg;
    int rc = CFDMX_OK;
    // the eddy viscosity from T0 (RUN_CPU.cpp:261-263)
    if (ctx->viseff && (rc = cfdmx_smagorinsky(ctx))) return rc;
    rc = cfdmx_convdiff(ctx);
    if (rc) return rc;
    *iters = 0;
    *err = 0.0;
    *solved = 0;
    if (ctx->loop > 1 || !c.skip_first_pressure) {
        *solved = 1;
        if (x_given) {
            HIPCHK(hipMemcpyAsync(ctx->x, x_given, (size_t)g.N * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
        } else {
            if (c.solver == CFDMX_SOLVER_BICGSTAB) {
                rc = cfdmx_pressure_rhs(ctx);
                if (!rc)
                    rc = c.precision != CFDMX_PRECISION_FP64 ? solve_mixed_impl(ctx, iters, err)
                                                             : solve_bicgstab_impl(ctx, iters, err);
            } else {
                rc = solve_rbsor_impl(ctx, iters, err);
            }
            if (rc) return rc;
            if (!(*err < c.tol)) ++ctx->unconverged;
        }
    }
    const int from_x = (c.solver == CFDMX_SOLVER_BICGSTAB && *solved) ? 1 : 0;
    // the steady-state check reads T0 (u^n) beside the new field the projection writes
    const double *const prev[3] = {ctx->t0[0], ctx->t0[1], ctx->t0[2]};
    PROF(CFDMX_K_PROJECT,
         launch_project(g, s, from_x, c.dt, c.outflow_convective, ctx->x, ctx->p, ctx->eta,
                        ctx->t1[0], ctx->t1[1], ctx->t1[2], ctx->t3[0], ctx->t3[1], ctx->t3[2],
                        ctx->f, ctx->part, c.steady_check ? prev : nullptr,
                        c.steady_check ? ctx->diff_part : nullptr,
                        ctx->f_ghosts_zero ? 2 : 1)); // launch_bc below rewrites the z ghosts
    launch_cdcl_fin(g, s, ctx->part, c.lz, ctx->scal + 1);
    if (c.steady_check) launch_max3_fin(g, s, ctx->diff_part, ctx->scal + 3);
    for (int q = 0; q < 3; ++q) std::swap(ctx->t0[q], ctx->t3[q]);
    PROF(CFDMX_K_BC, launch_bc(g, s, ctx->bc, 1, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->p,
                               ctx->t1[0], ctx->t1[1], ctx->t1[2], ctx->flux_part));
    LAUNCHCHK();
    ctx->time += c.dt;
    ctx->loop += 1;
    return CFDMX_OK;
}

int cfdmx_step(cfdmx_ctx *ctx, int nsteps, cfdmx_step_info *info)
{
    if (!ctx) return CFDMX_ERR_ARG;
    if (!ctx->started) FAIL(CFDMX_ERR_STATE, "call cfdmx_init_flow or cfdmx_start first");
    int iters = 0, solved = 0;
    double err = 0.0, t_at = ctx->time;
    for (int q = 0; q < nsteps; ++q) {
        t_at = ctx->time;
        int rc = step_once(ctx, &iters, &err, &solved);
        if (rc) return rc;
    }
    if (info) {
        int rc = read_scalars(ctx);
        if (rc) return rc;
        info->cfl = ctx->scal_h[0];
        info->cd = ctx->scal_h[1];
        info->cl = ctx->scal_h[2];
        info->time = t_at;
        info->iters = iters;
        info->err = err;
        info->solved = solved;
        HIPCHK(hipMemcpyAsync(ctx->kst_h, ctx->kst, sizeof(KState), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        info->restarts = ctx->kst_h->restarts;
        info->resets = ctx->resets;
        for (int q = 0; q < 3; ++q)
            info->maxdiff[q] = ctx->cfg.steady_check ? ctx->scal_h[3 + q] : 0.0;
        info->unconverged = ctx->unconverged;
    }
    return CFDMX_OK;
}

int cfdmx_synchronize(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_harvest(ctx);
    return CFDMX_OK;
}

int cfdmx_profile(cfdmx_ctx *ctx, int enable)
{
    if (!ctx) return CFDMX_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_harvest(ctx);
    for (int q = 0; q < CFDMX_K_COUNT; ++q) {
        ctx->prof_ms[q] = 0.0;
        ctx->prof_n[q] = 0;
    }
    ctx->prof_on = enable ? 1 : 0;
    return CFDMX_OK;
}

int cfdmx_profile_read(cfdmx_ctx *ctx, int kernel, double *total_ms, long long *launches)
{
    if (!ctx || kernel < 0 || kernel >= CFDMX_K_COUNT) return CFDMX_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_harvest(ctx);
    if (total_ms) *total_ms = ctx->prof_ms[kernel];
    if (launches) *launches = ctx->prof_n[kernel];
    return CFDMX_OK;
}

// ---- Plot3D output (OutputPLOT3D_Qfile, qfile.hpp:6-108; OutputPlot3D_Xfile,
// xfileWrite.hpp:6-64) ----
static int plot3d_flush_impl(cfdmx_ctx *ctx)
{
    cfdmx_ctx::P3D &P = ctx->p3d;
    if (!P.pending) return CFDMX_OK;
    P.pending = false;
    HIPCHK(hipStreamSynchronize(P.side));
    FILE *fp = std::fopen(P.path.c_str(), "wb");
    if (!fp) FAIL(CFDMX_ERR_ARG, "plot3d: cannot open %s for writing", P.path.c_str());
    const int32_t hdr[4] = {1, ctx->g.nxc, ctx->g.nyc, ctx->g.nzc};
    const float fh[4] = {0.0f, 0.0f, P.reyn, P.time}; // mach, alpha, Re, time
    const size_t n = 5 * (size_t)ctx->g.N;
    const bool ok = std::fwrite(hdr, sizeof hdr, 1, fp) == 1 &&
                    std::fwrite(fh, sizeof fh, 1, fp) == 1 &&
                    std::fwrite(P.host, sizeof(float), n, fp) == n;
    if (std::fclose(fp) != 0 || !ok)
        FAIL(CFDMX_ERR_ARG, "plot3d: writing %s failed", P.path.c_str());
    return CFDMX_OK;
}

int cfdmx_plot3d_q(cfdmx_ctx *ctx, const char *path, double time)
{
    if (!ctx || !path) return CFDMX_ERR_ARG;
    int rc = plot3d_flush_impl(ctx);
    if (rc) return rc;
    cfdmx_ctx::P3D &P = ctx->p3d;
    const size_t n = 5 * (size_t)ctx->g.N;
    if (!P.dev) {
        ALLOC(P.dev, n);
        HIPCHK(hipHostMalloc((void **)&P.host, n * sizeof(float), hipHostMallocDefault));
        HIPCHK(hipStreamCreateWithFlags(&P.side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&P.ready, hipEventDisableTiming));
    }
    launch_plot3d_q(ctx->g, ctx->stream, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->p, ctx->eta,
                    P.dev);
    LAUNCHCHK();
    HIPCHK(hipEventRecord(P.ready, ctx->stream));
    HIPCHK(hipStreamWaitEvent(P.side, P.ready, 0));
    HIPCHK(hipMemcpyAsync(P.host, P.dev, n * sizeof(float), hipMemcpyDeviceToHost, P.side));
    P.path = path;
    P.reyn = (float)ctx->cfg.re;
    P.time = (float)time;
    P.pending = true;
    return CFDMX_OK;
}

int cfdmx_plot3d_flush(cfdmx_ctx *ctx)
{
    if (!ctx) return CFDMX_ERR_ARG;
    return plot3d_flush_impl(ctx);
}

// ---- FP64 checkpoint / resume (no reference counterpart: the reference's resume reads
// a float32 q file back, QfileRead in qfileRead.hpp:6-100, which loses precision and the
// solver state).  The file holds every array a later step reads, so that loading it into
// a context of the same configuration continues the run bit for bit.
namespace {
constexpr char kCkMagic[8] = {'C', 'F', 'D', 'M', 'X', 'C', 'K', '1'};
constexpr int32_t kCkVersion = 2;
struct CkHeader {
    char magic[8];
    int32_t version, nx, ny, nz, grid_kind, precision, solver, z_periodic;
    int64_t G, N;
    double lx, ly, lz;
    int32_t loop, resets, started, mask_smoothed, has_inflow, pad;
    double time, dt, re, mask_cx, mask_cy, mask_r, mask_eps;
    KState kst;
    SorState sst;
    // the configuration the saved steps ran with, beyond the grid: a load into a context
    // configured otherwise would not continue the run
    double bc_num[6][3], bc_dir[6][3];
    double convective_speed, smagorinsky_cs;
    int32_t precond, outflow_convective, flux_correction, unconverged;
};

// the first configuration entry that differs between a checkpoint and a context, or null
static const char *ck_mismatch(const CkHeader &h, const cfdmx_config &c)
{
    if (h.precision != c.precision) return "precision";
    if (h.solver != c.solver) return "solver";
    if (h.precond != c.precond) return "precond";
    if (std::memcmp(h.bc_num, c.bc_num, sizeof h.bc_num) != 0 ||
        std::memcmp(h.bc_dir, c.bc_dir, sizeof h.bc_dir) != 0)
        return "boundary-condition table (bc_num, bc_dir)";
    if (h.outflow_convective != c.outflow_convective) return "outflow_convective";
    if (h.outflow_convective && h.convective_speed != c.convective_speed)
        return "convective_speed";
    if (h.flux_correction != c.flux_correction) return "flux_correction";
    if (h.smagorinsky_cs != c.smagorinsky_cs) return "smagorinsky_cs";
    return nullptr;
}

// the checkpointed arrays and their lengths, in file order
static std::vector<std::pair<double *, size_t>> ck_arrays(cfdmx_ctx *ctx)
{
    const size_t G = (size_t)ctx->g.G, N = (size_t)ctx->g.N;
    std::vector<std::pair<double *, size_t>> a;
    for (int q = 0; q < 3; ++q) a.push_back({ctx->t0[q], G});
    for (int q = 0; q < 3; ++q) a.push_back({ctx->t1[q], G});
    for (int q = 0; q < 3; ++q) a.push_back({ctx->t3[q], G});
    a.push_back({ctx->p, G});
    a.push_back({ctx->eta, G});
    a.push_back({ctx->f, 3 * G});
    a.push_back({ctx->x, N});
    a.push_back({ctx->px1, N});
    a.push_back({ctx->px2, N});
    a.push_back({ctx->inflow, (size_t)ctx->g.ny});
    return a;
}
} // namespace

int cfdmx_checkpoint_save(cfdmx_ctx *ctx, const char *path)
{
    if (!ctx || !path) return CFDMX_ERR_ARG;
    if (!ctx->started) FAIL(CFDMX_ERR_STATE, "checkpoint: nothing to save before cfdmx_start");
    CkHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, kCkMagic, sizeof h.magic);
    const cfdmx_config &c = ctx->cfg;
    h.version = kCkVersion;
    h.nx = c.nx, h.ny = c.ny, h.nz = c.nz;
    h.grid_kind = c.grid_kind, h.precision = c.precision, h.solver = c.solver;
    h.z_periodic = c.z_periodic;
    h.G = ctx->g.G, h.N = ctx->g.N;
    h.lx = c.lx, h.ly = c.ly, h.lz = c.lz;
    h.loop = ctx->loop, h.resets = ctx->resets, h.started = ctx->started;
    h.mask_smoothed = ctx->mask_smoothed;
    h.has_inflow = ctx->bc.inflow != nullptr;
    h.time = ctx->time, h.dt = c.dt, h.re = c.re;
    h.mask_cx = ctx->mask_cx, h.mask_cy = ctx->mask_cy, h.mask_r = ctx->mask_r;
    h.mask_eps = ctx->mask_eps;
    std::memcpy(h.bc_num, c.bc_num, sizeof h.bc_num);
    std::memcpy(h.bc_dir, c.bc_dir, sizeof h.bc_dir);
    h.convective_speed = c.convective_speed;
    h.smagorinsky_cs = c.smagorinsky_cs;
    h.precond = c.precond;
    h.outflow_convective = c.outflow_convective;
    h.flux_correction = c.flux_correction;
    h.unconverged = ctx->unconverged;
    HIPCHK(hipMemcpyAsync(ctx->kst_h, ctx->kst, sizeof(KState), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->sst_h, ctx->sst, sizeof(SorState), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    h.kst = *ctx->kst_h;
    h.sst = *ctx->sst_h;
    const auto arrays = ck_arrays(ctx);
    size_t most = 0;
    for (const auto &a : arrays) most = std::max(most, a.second);
    std::vector<double> buf(most);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) FAIL(CFDMX_ERR_ARG, "checkpoint: cannot open %s for writing", path);
    bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1;
    for (const auto &a : arrays) {
        if (!ok) break;
        const int rc = copy_d2h(ctx, buf.data(), a.first, a.second);
        if (rc) {
            std::fclose(fp);
            return rc;
        }
        ok = std::fwrite(buf.data(), sizeof(double), a.second, fp) == a.second;
    }
    if (std::fclose(fp) != 0 || !ok) FAIL(CFDMX_ERR_ARG, "checkpoint: writing %s failed", path);
    return CFDMX_OK;
}

int cfdmx_checkpoint_load(cfdmx_ctx *ctx, const char *path)
{
    if (!ctx || !path) return CFDMX_ERR_ARG;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) FAIL(CFDMX_ERR_ARG, "checkpoint: cannot open %s", path);
    CkHeader h;
    if (std::fread(&h, sizeof h, 1, fp) != 1 ||
        std::memcmp(h.magic, kCkMagic, sizeof h.magic) != 0 || h.version != kCkVersion) {
        std::fclose(fp);
        FAIL(CFDMX_ERR_ARG, "checkpoint: %s is not a cfdmx checkpoint of version %d", path,
             kCkVersion);
    }
    const cfdmx_config &c = ctx->cfg;
    if (h.nx != c.nx || h.ny != c.ny || h.nz != c.nz || h.grid_kind != c.grid_kind ||
        h.lx != c.lx || h.ly != c.ly || h.lz != c.lz || h.z_periodic != c.z_periodic ||
        h.G != ctx->g.G || h.N != ctx->g.N) {
        std::fclose(fp);
        FAIL(CFDMX_ERR_ARG,
             "checkpoint: %s holds a %dx%dx%d grid (kind %d) of another size or domain than "
             "this context's %dx%dx%d (kind %d)",
             path, h.nx, h.ny, h.nz, h.grid_kind, c.nx, c.ny, c.nz, c.grid_kind);
    }
    if (const char *what = ck_mismatch(h, c)) {
        std::fclose(fp);
        FAIL(CFDMX_ERR_ARG,
             "checkpoint: %s was saved with another %s than this context's configuration",
             path, what);
    }
    const auto arrays = ck_arrays(ctx);
    size_t most = 0, total = sizeof h;
    for (const auto &a : arrays) {
        most = std::max(most, a.second);
        total += a.second * sizeof(double);
    }
    // a short file is refused before anything is overwritten
    if (std::fseek(fp, 0, SEEK_END) != 0 || std::ftell(fp) < (long)total ||
        std::fseek(fp, (long)sizeof h, SEEK_SET) != 0) {
        std::fclose(fp);
        FAIL(CFDMX_ERR_ARG, "checkpoint: %s is truncated", path);
    }
    std::vector<double> buf(most);
    for (const auto &a : arrays) {
        if (std::fread(buf.data(), sizeof(double), a.second, fp) != a.second) {
            std::fclose(fp);
            FAIL(CFDMX_ERR_ARG, "checkpoint: reading %s failed", path);
        }
        const int rc = copy_h2d(ctx, a.first, buf.data(), a.second);
        if (rc) {
            std::fclose(fp);
            return rc;
        }
        if (a.first == ctx->f) ctx->f_ghosts_zero = force_zghosts_zero(ctx->g, buf.data()) ? 1 : 0;
    }
    std::fclose(fp);
    *ctx->kst_h = h.kst;
    *ctx->sst_h = h.sst;
    HIPCHK(hipMemcpyAsync(ctx->kst, ctx->kst_h, sizeof(KState), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->sst, ctx->sst_h, sizeof(SorState), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->loop = h.loop;
    ctx->resets = h.resets;
    ctx->unconverged = h.unconverged;
    ctx->started = h.started;
    ctx->time = h.time;
    ctx->cfg.dt = h.dt;
    ctx->cfg.re = h.re;
    ctx->bc.conv_coef = ctx->cfg.convective_speed * h.dt / ctx->Dx[ctx->g.nx - 3];
    ctx->bc.inflow = h.has_inflow ? ctx->inflow : nullptr;
    ctx->mask_smoothed = h.mask_smoothed;
    ctx->mask_cx = h.mask_cx, ctx->mask_cy = h.mask_cy, ctx->mask_r = h.mask_r;
    ctx->mask_eps = h.mask_eps;
    return CFDMX_OK;
}

int cfdmx_plot3d_x(cfdmx_ctx *ctx, const char *path)
{
    if (!ctx || !path) return CFDMX_ERR_ARG;
    const int nxc = ctx->g.nxc, nyc = ctx->g.nyc, nzc = ctx->g.nzc;
    // cell centres (structureGrid.hpp:216-229): (X[i+gC] + X[i+gC+1]) / 2
    std::vector<float> xc(nxc), yc(nyc), zc(nzc);
    for (int i = 0; i < nxc; ++i) xc[i] = (float)((ctx->X[i + kGC] + ctx->X[i + kGC + 1]) / 2.0);
    for (int j = 0; j < nyc; ++j) yc[j] = (float)((ctx->Y[j + kGC] + ctx->Y[j + kGC + 1]) / 2.0);
    for (int k = 0; k < nzc; ++k) zc[k] = (float)((ctx->Z[k + kGC] + ctx->Z[k + kGC + 1]) / 2.0);
    const size_t n = (size_t)nxc * nyc * nzc;
    std::vector<float> buf(3 * n);
    size_t o = 0;
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < nzc; ++k)
            for (int j = 0; j < nyc; ++j)
                for (int i = 0; i < nxc; ++i) buf[o++] = a == 0 ? xc[i] : (a == 1 ? yc[j] : zc[k]);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) FAIL(CFDMX_ERR_ARG, "plot3d: cannot open %s for writing", path);
    const int32_t hdr[4] = {1, nxc, nyc, nzc};
    const bool ok = std::fwrite(hdr, sizeof hdr, 1, fp) == 1 &&
                    std::fwrite(buf.data(), sizeof(float), buf.size(), fp) == buf.size();
    if (std::fclose(fp) != 0 || !ok) FAIL(CFDMX_ERR_ARG, "plot3d: writing %s failed", path);
    return CFDMX_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Adjoint of the time-averaged drag coefficient (no reference counterpart)
// ---------------------------------------------------------------------------
// J = (1/n) sum of Cd over n steps from the current state, with U^0 = BC_start(U).
// The forward pass keeps a checkpoint (velocity, pressure, X_result, history vectors,
// solver state) every K steps; the reverse pass restores each checkpoint, recomputes
// its K steps onto a tape, and runs the transposed step kernels back through them.
// The pressure adjoint reuses the forward pressure solver (A is symmetric on the
// uniform grid): the FP64 BiCGSTAB with its own solution, history and state, or, in
// the multigrid precisions, the multigrid-preconditioned solve with FP64 refinement.  The context is returned to the state
// at entry (start boundary conditions applied).
static int adj_alloc(cfdmx_ctx *ctx, int K, int ncp)
{
    cfdmx_ctx::Adj &A = ctx->adj;
    if (A.K == K && A.ncp >= ncp && A.kadj) return CFDMX_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (void *q : A.mem) (void)hipFree(q);
    A.mem.clear();
    const size_t G = ctx->g.G, N = ctx->g.N;
    auto al = [&](double **p, size_t n) -> int {
        void *q = nullptr;
        HIPCHK(hipMalloc(&q, (n > 0 ? n : 1) * sizeof(double)));
        A.mem.push_back(q);
        *p = static_cast<double *>(q);
        return CFDMX_OK;
    };
    int rc = 0;
    rc |= al(&A.tapeU, (size_t)K * 3 * G);
    rc |= al(&A.tapeP, (size_t)K * G);
    rc |= al(&A.tapeX, (size_t)K * N);
    rc |= al(&A.cpU, (size_t)ncp * 3 * G);
    rc |= al(&A.cpP, (size_t)ncp * G);
    rc |= al(&A.cpX, (size_t)ncp * N);
    rc |= al(&A.cpPx1, (size_t)ncp * N);
    rc |= al(&A.cpPx2, (size_t)ncp * N);
    for (int q = 0; q < 3; ++q) {
        rc |= al(&A.lu[q], G);
        rc |= al(&A.ls[q], G);
        rc |= al(&A.le[q], G);
        rc |= al(&A.us[q], G);
    }
    rc |= al(&A.lx, N);
    rc |= al(&A.mu, N);
    rc |= al(&A.mpx1, N);
    rc |= al(&A.mpx2, N);
    rc |= al(&A.leta, G);
    rc |= al(&A.Jr, (size_t)63 * G);
    rc |= al(&A.lin, (size_t)ctx->g.ny);
    rc |= al(&A.tin, (size_t)ctx->g.cyz);
    rc |= al(&A.apart, 4 * ((size_t)ctx->g.nblocks + 64) + 3 * ((size_t)ctx->g.cyz / kPerBlock + 2));
    rc |= al(&A.out, 4);
    if (rc) return CFDMX_ERR_HIP;
    double *kp = nullptr, *kc = nullptr;
    rc |= al(&kp, (sizeof(KState) + 7) / 8);
    rc |= al(&kc, (size_t)ncp * ((sizeof(KState) + 7) / 8));
    if (rc) return CFDMX_ERR_HIP;
    A.kadj = reinterpret_cast<KState *>(kp);
    A.cpK = reinterpret_cast<KState *>(kc);
    if (!A.kadj_h) HIPCHK(hipHostMalloc((void **)&A.kadj_h, sizeof(KState), hipHostMallocDefault));
    if (!A.out_h) HIPCHK(hipHostMalloc((void **)&A.out_h, 4 * sizeof(double), hipHostMallocDefault));
    A.K = K;
    A.ncp = ncp;
    A.solved.assign((size_t)K, 0);
    A.cp_loop.assign((size_t)ncp, 0);
    A.cp_resets.assign((size_t)ncp, 0);
    A.cp_time.assign((size_t)ncp, 0.0);
    return CFDMX_OK;
}

static int d2d(cfdmx_ctx *ctx, double *dst, const double *src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return CFDMX_OK;
}

static int adj_save(cfdmx_ctx *ctx, int c)
{
    cfdmx_ctx::Adj &A = ctx->adj;
    const size_t G = ctx->g.G, N = ctx->g.N;
    int rc = 0;
    for (int q = 0; q < 3; ++q) rc |= d2d(ctx, A.cpU + ((size_t)c * 3 + q) * G, ctx->t0[q], G);
    rc |= d2d(ctx, A.cpP + (size_t)c * G, ctx->p, G);
    rc |= d2d(ctx, A.cpX + (size_t)c * N, ctx->x, N);
    rc |= d2d(ctx, A.cpPx1 + (size_t)c * N, ctx->px1, N);
    rc |= d2d(ctx, A.cpPx2 + (size_t)c * N, ctx->px2, N);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(A.cpK + c, ctx->kst, sizeof(KState), hipMemcpyDeviceToDevice,
                          ctx->stream));
    A.cp_loop[c] = ctx->loop;
    A.cp_time[c] = ctx->time;
    A.cp_resets[c] = ctx->resets;
    return CFDMX_OK;
}

static int adj_restore(cfdmx_ctx *ctx, int c)
{
    cfdmx_ctx::Adj &A = ctx->adj;
    const size_t G = ctx->g.G, N = ctx->g.N;
    int rc = 0;
    for (int q = 0; q < 3; ++q) rc |= d2d(ctx, ctx->t0[q], A.cpU + ((size_t)c * 3 + q) * G, G);
    rc |= d2d(ctx, ctx->p, A.cpP + (size_t)c * G, G);
    rc |= d2d(ctx, ctx->x, A.cpX + (size_t)c * N, N);
    rc |= d2d(ctx, ctx->px1, A.cpPx1 + (size_t)c * N, N);
    rc |= d2d(ctx, ctx->px2, A.cpPx2 + (size_t)c * N, N);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->kst, A.cpK + c, sizeof(KState), hipMemcpyDeviceToDevice,
                          ctx->stream));
    // the boundary copies into T1 that the step after the checkpoint reads
    launch_bc_copy(ctx->g, ctx->stream, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->t1[0],
                   ctx->t1[1], ctx->t1[2]);
    LAUNCHCHK();
    ctx->loop = A.cp_loop[c];
    ctx->time = A.cp_time[c];
    ctx->resets = A.cp_resets[c];
    return CFDMX_OK;
}

static int adj_check(cfdmx_ctx *ctx, int nsteps, int K)
{
    const cfdmx_config &c = ctx->cfg;
    if (nsteps < 1 || K < 1) FAIL(CFDMX_ERR_ARG, "nsteps and checkpoint_every must be >= 1");
    if (c.grid_kind != CFDMX_GRID_UNIFORM || c.precond != CFDMX_PRECOND_NONE ||
        c.solver != CFDMX_SOLVER_BICGSTAB)
        FAIL(CFDMX_ERR_ARG, "the drag adjoint needs the uniform grid, the SPE operator "
                            "(precond none) and BiCGSTAB");
    if (c.smagorinsky_cs > 0.0)
        FAIL(CFDMX_ERR_ARG, "the drag adjoint does not differentiate the Smagorinsky eddy "
                            "viscosity (smagorinsky_cs must be 0)");
    if (!ctx->started) FAIL(CFDMX_ERR_STATE, "call cfdmx_init_flow or cfdmx_start first");
    return CFDMX_OK;
}

// one time step onto tape slot `slot`: U^{n-1} and p before the step, X_result and the
// solved flag after it; x_given as in step_once
static int adj_tape_step(cfdmx_ctx *ctx, int slot, int *iters, const double *x_given = nullptr)
{
    cfdmx_ctx::Adj &A = ctx->adj;
    const size_t G = ctx->g.G, N = ctx->g.N;
    int rc = 0;
    for (int q = 0; q < 3; ++q)
        if ((rc = d2d(ctx, A.tapeU + ((size_t)slot * 3 + q) * G, ctx->t0[q], G))) return rc;
    if ((rc = d2d(ctx, A.tapeP + (size_t)slot * G, ctx->p, G))) return rc;
    int it = 0, solved = 0;
    double err = 0.0;
    if ((rc = step_once(ctx, &it, &err, &solved, x_given))) return rc;
    if ((rc = d2d(ctx, A.tapeX + (size_t)slot * N, ctx->x, N))) return rc;
    A.solved[slot] = solved;
    if (iters) *iters = it;
    return CFDMX_OK;
}

// the forward path of the objective: BC_start, then n steps; Cd summed in step order.
// Steps n >= tape_from (>= 0) go onto the tape at slot n - tape_from: the adjoint tapes
// its last segment here, which the reverse pass then need not recompute (every kernel
// is deterministic, so a recompute would give the same bits).  Checkpoints are kept
// every K steps before that, and with A.allX the pressure solution of every step before
// the tape.
static int adj_forward(cfdmx_ctx *ctx, int nsteps, int K, int keep_checkpoints, int tape_from,
                       double *J, int *iters_total)
{
    hipStream_t s = ctx->stream;
    launch_bc(ctx->g, s, ctx->bc, 0, ctx->t0[0], ctx->t0[1], ctx->t0[2], ctx->p, ctx->t1[0],
              ctx->t1[1], ctx->t1[2], ctx->flux_part);
    LAUNCHCHK();
    double sum = 0.0;
    int it_tot = 0;
    for (int n = 0; n < nsteps; ++n) {
        const bool taped = tape_from >= 0 && n >= tape_from;
        if (keep_checkpoints && n % K == 0 && !taped) {
            int rc = adj_save(ctx, n / K);
            if (rc) return rc;
        }
        int it = 0, solved = 0;
        double err = 0.0;
        int rc = taped ? adj_tape_step(ctx, n - tape_from, &it)
                       : step_once(ctx, &it, &err, &solved);
        if (!rc && keep_checkpoints && !taped && ctx->adj.allX)
            rc = d2d(ctx, ctx->adj.allX + (size_t)n * ctx->g.N, ctx->x, ctx->g.N);
        if (!rc) rc = read_scalars(ctx);
        if (rc) return rc;
        sum += ctx->scal_h[1];
        it_tot += it;
    }
    *J = sum / nsteps;
    if (iters_total) *iters_total = it_tot;
    return CFDMX_OK;
}

int cfdmx_drag_objective(cfdmx_ctx *ctx, int nsteps, double *J)
{
    if (!ctx || !J) return CFDMX_ERR_ARG;
    int rc = adj_check(ctx, nsteps, 1);
    if (!rc) rc = adj_alloc(ctx, 1, 1);
    if (!rc) rc = adj_save(ctx, 0);
    if (!rc) rc = adj_forward(ctx, nsteps, 1, 0, -1, J, nullptr);
    if (!rc) rc = adj_restore(ctx, 0);
    if (!rc) HIPCHK(hipStreamSynchronize(ctx->stream));
    return rc;
}

int cfdmx_adjoint_drag(cfdmx_ctx *ctx, int nsteps, int checkpoint_every, double *J,
                       double *grad_inflow, double *grad_mask, cfdmx_adjoint_info *info)
{
    if (!ctx || !J) return CFDMX_ERR_ARG;
    const int K = checkpoint_every;
    int rc = adj_check(ctx, nsteps, K);
    if (rc) return rc;
    if (grad_mask && !ctx->mask_smoothed)
        FAIL(CFDMX_ERR_STATE, "mask gradients need the smoothed mask (cfdmx_set_eta_smoothed)");
    const int ncp = (nsteps + K - 1) / K;
    rc = adj_alloc(ctx, K, ncp + 1);
    if (rc) return rc;
    cfdmx_ctx::Adj &A = ctx->adj;
    // the pressure solutions of the steps the reverse pass recomputes (all but the taped
    // last segment): N doubles per step, kept where the device has room, else the
    // recompute solves again (the same bits either way)
    {
        const size_t need = (size_t)(ncp - 1) * K * ctx->g.N;
        if (need > A.allX_n) {
            if (A.allX) (void)hipFree(A.allX);
            A.allX = nullptr;
            A.allX_n = 0;
            void *q = nullptr;
            if (need > 0 && hipMalloc(&q, need * sizeof(double)) == hipSuccess) {
                A.allX = static_cast<double *>(q);
                A.allX_n = need;
            } else {
                (void)hipGetLastError(); // an allocation failure is not an error here
            }
        }
    }
    hipStream_t s = ctx->stream;
    const Geo &g = ctx->g;
    const cfdmx_config &c = ctx->cfg;
    const size_t G = g.G, N = g.N;
    HIPCHK(hipStreamSynchronize(s));
    auto t0 = std::chrono::steady_clock::now();
    // entry state in the extra checkpoint slot ncp
    rc = adj_save(ctx, ncp);
    if (rc) return rc;
    int fwd_iters = 0;
    rc = adj_forward(ctx, nsteps, K, 1, (ncp - 1) * K, J, &fwd_iters);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();

    // reverse pass
    for (int q = 0; q < 3; ++q) launch_fill(s, (int)G, 0.0, A.lu[q]);
    launch_fill(s, (int)G, 0.0, A.leta);
    launch_fill(s, g.ny, 0.0, A.lin);
    launch_fill(s, (int)N, 0.0, A.mu);
    launch_fill(s, (int)N, 0.0, A.mpx1);
    launch_fill(s, (int)N, 0.0, A.mpx2);
    LAUNCHCHK();
    KState ks = {};
    ks.tol = c.tol;
    ks.restart_factor = 1e-7;
    ks.maxit = c.maxit;
    ks.done = 1;
    *A.kadj_h = ks;
    HIPCHK(hipMemcpyAsync(A.kadj, A.kadj_h, sizeof ks, hipMemcpyHostToDevice, s));
    double area = 0.0;
    for (int j = kGC; j < g.ny - kGC; ++j)
        for (int k = kGC; k < g.nz - kGC; ++k) area += ctx->Dy[j] * ctx->Dz[k];
    const double wf = (-2.0 / c.lz) / nsteps;
    const double nu = 1.0 / c.re;
    int adj_iters = 0;
    int mg_solves = 0; // multigrid pressure solves of the reverse pass so far
    for (int seg = ncp - 1; seg >= 0; --seg) {
        const int n0 = seg * K, n1 = std::min(n0 + K, nsteps);
        if (seg != ncp - 1) { // the last segment is on the tape from the forward pass
            rc = adj_restore(ctx, seg);
            if (rc) return rc;
            for (int n = n0; n < n1; ++n)
                if ((rc = adj_tape_step(ctx, n - n0, nullptr,
                                        A.allX ? A.allX + (size_t)n * N : nullptr)))
                    return rc;
        }
        for (int n = n1 - 1; n >= n0; --n) {
            const int slot = n - n0;
            double *U0[3];
            for (int q = 0; q < 3; ++q) U0[q] = A.tapeU + ((size_t)slot * 3 + q) * G;
            // boundary conditions at the end of step n
            launch_adj_bc(g, s, ctx->bc, 1, A.lu[0], A.lu[1], A.lu[2], A.tin, A.lin, A.apart,
                          1.0 / area);
            // U* of step n, recomputed from U^{n-1}
            for (int q = 0; q < 3; ++q)
                if ((rc = d2d(ctx, A.us[q], U0[q], G))) return rc;
            launch_convdiff(g, s, U0[0], U0[1], U0[2], A.us[0], A.us[1], A.us[2], c.dt, nu,
                            ctx->cfl_part);
            // projection and forcing
            launch_adj_project(g, s, c.dt, A.solved[slot], c.outflow_convective, wf,
                               A.tapeX + (size_t)slot * N, A.tapeP + (size_t)slot * G, ctx->eta,
                               A.us, A.lu, A.ls, A.le, A.lx, A.leta);
            LAUNCHCHK();
            if (A.solved[slot]) {
                // pressure: A mu = lx, mean-free, then the right-hand side's transpose
                int it = 0;
                double res = 0.0;
                if (c.precision == CFDMX_PRECISION_FP64) {
                    SolveTarget T = {A.lx, A.mu, &A.mpx1, &A.mpx2, A.kadj, A.kadj_h};
                    if ((rc = solve_bicgstab_core(ctx, T, &it, &res))) return rc;
                } else {
                    // the multigrid solve in the forward pass's precision; lx is made
                    // mean-free first (J does not depend on the mean of X_result, so
                    // its adjoint has zero mean up to rounding), so that the singular
                    // system is consistent and the FP64 refinement can reach tol
                    launch_adj_remove_mean(g, s, A.lx, A.apart);
                    // with mg_extrapolate, as in the forward solves: the start is
                    // 2 mu_k - mu_(k-1) from the two solutions before (mpx1 holds the
                    // older; the FP64 branch alone uses it otherwise)
                    if (c.mg_extrapolate) {
                        if (mg_solves >= 2) launch_extrapolate(g, s, A.mu, A.mpx1);
                        else launch_copy(s, (int)N, A.mu, A.mpx1);
                    }
                    LAUNCHCHK();
                    if ((rc = solve_mg_target(ctx, A.lx, A.mu, &it, &res))) return rc;
                    ++mg_solves;
                }
                adj_iters += it;
                launch_adj_remove_mean(g, s, A.mu, A.apart);
                launch_adj_rhs(g, s, c.dt, A.mu, A.ls);
            }
            if (g.zper) launch_adj_periodic_interior(g, s, A.ls);
            const double *Uc[3] = {U0[0], U0[1], U0[2]};
            launch_adj_convdiff(g, s, c.dt, nu, Uc, A.Jr, A.ls, A.lu);
            LAUNCHCHK();
        }
    }
    // U^0 = BC_start(U)
    launch_adj_bc(g, s, ctx->bc, 0, A.lu[0], A.lu[1], A.lu[2], A.tin, A.lin, A.apart,
                  1.0 / area);
    if (grad_mask)
        launch_adj_mask(g, s, ctx->Xn, ctx->Yn, ctx->mask_cx, ctx->mask_cy, ctx->mask_r,
                        ctx->mask_eps, A.leta, A.apart, A.out);
    LAUNCHCHK();
    if (grad_inflow) {
        if ((rc = copy_d2h(ctx, grad_inflow, A.lin, g.ny))) return rc;
        if (!ctx->bc.inflow)
            for (int j = 0; j < g.ny; ++j) grad_inflow[j] = 0.0;
    }
    if (grad_mask) {
        HIPCHK(hipMemcpyAsync(A.out_h, A.out, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int q = 0; q < 3; ++q) grad_mask[q] = A.out_h[q];
    }
    HIPCHK(hipStreamSynchronize(s));
    auto t2 = std::chrono::steady_clock::now();
    rc = adj_restore(ctx, ncp);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
    if (info) {
        info->forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        info->reverse_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        info->forward_iters = fwd_iters;
        info->adjoint_iters = adj_iters;
        info->nsteps = nsteps;
        info->checkpoints = ncp;
    }
    return CFDMX_OK;
}
