// ref_harness.cpp -- C entry points around the reference's own hot-path functions.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile directly against the
// reference headers under /root/reference/inc (nothing is copied); the output
// lands in oracle/_ref/.  It exists so that the C oracle (cfd_oracle.c) can be
// checked bit-for-bit against the reference's ConvectionDifussion,
// BC_staggered_main/copy, createBMatrix_omp, SPE_matrix, bicgstabRe2,
// Pressure_transform_X_result, SorPipeLine_seq, update_UandF_omp,
// DFIB_CylinderZ and virtualF_Int, and so that the reference can be timed as
// the CPU baseline.
//
// The reference fixes the grid size and domain as default member initialisers
// of class grid (structureGrid.hpp:20-40); grid has no constructor, so it is an
// aggregate and aggregate initialisation overrides those members.  The time
// loop below follows runSeqOmp (RUN_CPU.cpp:257-436) minus its file I/O.
#include "0_General.hpp"
#include "run/general.hpp"

#include <cstdio>
#include <memory>
#include <sstream>

namespace {

// bicgstabRe2::iters_max_ is private and fixed at 5000 (bicgstabRe2.hpp:47); the CPU
// baseline times a capped solve.  Explicit instantiation may name private members, so
// a pointer to the member is captured without touching the reference sources.
template <class Tag, typename Tag::type M> struct Expose {
    friend typename Tag::type member(Tag) { return M; }
};
struct SpeMaxit {
    typedef int solver::bicgstabRe2<mat::SPE_matrix<double>>::*type;
    friend type member(SpeMaxit);
};
struct CsrMaxit {
    typedef int solver::bicgstabRe2<mat::CSR_matrix<double>>::*type;
    friend type member(CsrMaxit);
};
template struct Expose<SpeMaxit, &solver::bicgstabRe2<mat::SPE_matrix<double>>::iters_max_>;
template struct Expose<CsrMaxit, &solver::bicgstabRe2<mat::CSR_matrix<double>>::iters_max_>;

struct RefCtx {
    std::unique_ptr<grid> g;
    simpulationVariable simu;
    divideLocal Lo;
    velocity T0, T1, T3;
    pressure t1;
    DfibArray Dfib;
    MxClass Mx;
    SORcoefficient Sor;
    shareMenory ShareM;
    int precond = 0;
    std::unique_ptr<solver::bicgstabRe2<mat::SPE_matrix<double>>> spe_solver;
    std::unique_ptr<solver::bicgstabRe2<mat::CSR_matrix<double>>> csr_solver;
    std::vector<double> metrics;
};

}  // namespace

extern "C" {

void *ref_new(int nxc, int nyc, int nzc, double lx, double ly, double lz, int sml, double Re,
              double dt, int precond, double tol, int threads)
{
    if (threads > 0) omp_set_num_threads(threads);
    auto *c = new RefCtx();
    c->g.reset(new grid{{"[x]", "[Y]", "[Z]"}, sml ? "non_uniform-SML" : "uniform", lx, ly, lz,
                        2, nxc, nyc, nzc});
    grid &gA = *c->g;
    c->simu.Locality = 0;
    c->simu.set_time_max(1.0);
    c->simu.set_dt(dt);
    c->simu.set_Re(Re);
    c->simu.p_criteria = tol;
    std::vector<int> grid_size{gA.nx, gA.ny, gA.nz};
    std::vector<int> dims{1, 1, 1};
    c->Lo.initTable(grid_size, dims);
    c->Lo.initLocal(0, 0, 0);
    // resize_variable (resize_and_simuSetting.hpp:5-32) without its stdout
    c->Dfib.eta.resize(gA.iceltot);
    c->Dfib.f.resize(gA.iceltot * 3);
    c->t1.p.resize(gA.iceltot);
    c->T0.resize(gA.iceltot);
    c->T1.resize(gA.iceltot);
    c->T3.resize(gA.iceltot);
    c->T0.iniU(1.0, 0.0, 0.0);
    c->T1.iniU(1.0, 0.0, 0.0);
    c->T3.iniU(1.0, 0.0, 0.0);
    c->t1.init_p(0.0);
    generateGride(c->simu, c->ShareM, c->Lo, gA);
    c->Mx.X_result.resize(gA.iceltotCal, 0.0);
    c->Mx.matB.resize(gA.iceltotCal, 0.0);
    c->precond = precond;
    if (precond) {
        c->Mx.matA_csr.set(gA.createPressureMatrixCSR());
        c->csr_solver.reset(new solver::bicgstabRe2<mat::CSR_matrix<double>>(c->Mx.matA_csr));
        c->csr_solver->setTolerance(tol);
    } else {
        c->Mx.matA_spe.resize(gA.nxCal, gA.nyCal, gA.nzCal);
        c->Mx.matA_spe.setupPressure(gA.gC, gA.Dx, gA.Dy, gA.Dz, gA.Dxs, gA.Dys, gA.Dzs);
        c->spe_solver.reset(new solver::bicgstabRe2<mat::SPE_matrix<double>>(c->Mx.matA_spe));
        c->spe_solver->setTolerance(tol);
    }
    std::fill(c->Dfib.eta.begin(), c->Dfib.eta.end(), 0.0);
    return c;
}

void ref_free(void *h) { delete static_cast<RefCtx *>(h); }

// which: 0..2 T0 u,v,w; 3..5 T1; 6..8 T3; 9 p; 10 eta; 11 f; 12 X_result; 13 matB
double *ref_field(void *h, int which)
{
    auto *c = static_cast<RefCtx *>(h);
    switch (which) {
    case 0: return c->T0.u.data();
    case 1: return c->T0.v.data();
    case 2: return c->T0.w.data();
    case 3: return c->T1.u.data();
    case 4: return c->T1.v.data();
    case 5: return c->T1.w.data();
    case 6: return c->T3.u.data();
    case 7: return c->T3.v.data();
    case 8: return c->T3.w.data();
    case 9: return c->t1.p.data();
    case 10: return c->Dfib.eta.data();
    case 11: return c->Dfib.f.data();
    case 12: return c->Mx.X_result.data();
    case 13: return c->Mx.matB.data();
    }
    return nullptr;
}

// metric arrays: X,Y,Z (n+1 each) then Dx,Dy,Dz,Dxs,Dys,Dzs (n each)
void ref_metrics(void *h, double *out)
{
    auto *c = static_cast<RefCtx *>(h);
    grid &g = *c->g;
    size_t o = 0;
    for (auto *v : {&g.X, &g.Y, &g.Z, &g.Dx, &g.Dy, &g.Dz, &g.Dxs, &g.Dys, &g.Dzs})
        for (double d : *v) out[o++] = d;
}

void ref_start(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    c->T1.u = c->T3.u = c->T0.u;  // RUN_CPU.cpp:104-109
    c->T1.v = c->T3.v = c->T0.v;
    c->T1.w = c->T3.w = c->T0.w;
    BC_staggered_main(c->Lo, c->T0, c->t1, *c->g);
    BC_staggered_copy(c->Lo, c->T0, c->T1, *c->g);
}

void ref_convdiff(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    ConvectionDifussion(c->simu, c->T0, c->T1, c->Lo, *c->g);
}

void ref_bc_main(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    BC_staggered_main(c->Lo, c->T0, c->t1, *c->g);
}

void ref_bc_copy(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    BC_staggered_copy(c->Lo, c->T0, c->T1, *c->g);
}

void ref_rhs(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    createBMatrix_omp(c->T1, c->Mx, c->simu, c->Lo, *c->g);
}

int ref_solve(void *h, double *err)
{
    auto *c = static_cast<RefCtx *>(h);
    std::pair<int, double> r = c->precond ? c->csr_solver->solve(c->Mx.matB, c->Mx.X_result)
                                          : c->spe_solver->solve(c->Mx.matB, c->Mx.X_result);
    if (err) *err = r.second;
    return r.first;
}

void ref_transform(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    Pressure_transform_X_result(c->t1, c->Mx, c->Lo, *c->g);
}

void ref_update(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    update_UandF_omp(c->Dfib, c->simu, c->t1, c->T1, c->T3, c->Lo, *c->g);
}

void ref_copyvel(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    copyVel(c->T3, c->T0);
}

void ref_cdcl(void *h, double *cd, double *cl)
{
    // CD_CL (virtualForceIntergrator.hpp:92-111) without its stdout/file output
    auto *c = static_cast<RefCtx *>(h);
    virtualF_Int(c->Dfib, c->Lo, *c->g, 0);
    virtualF_Int(c->Dfib, c->Lo, *c->g, 1);
    *cd = -2.0 * c->Dfib.ValSum[0] / c->g->lz;
    *cl = -2.0 * c->Dfib.ValSum[1] / c->g->lz;
}

void ref_eta(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    std::fill(c->Dfib.eta.begin(), c->Dfib.eta.end(), 0.0);
    DFIB_CylinderZ(c->Dfib, c->Lo, *c->g);
}

void ref_apply_A(void *h, const double *x, double *r)
{
    auto *c = static_cast<RefCtx *>(h);
    const size_t n = c->g->iceltotCal;
    std::vector<double> xv(x, x + n), rv(n);
    if (c->precond) c->Mx.matA_csr.multiply_omp(xv, rv);
    else c->Mx.matA_spe.multiply_omp(xv, rv);
    std::copy(rv.begin(), rv.end(), r);
}

int ref_sor_seq(void *h, double omega, double tol, int itmax, double *err)
{
    // SorPipeLine_seq builds its coefficient table only when simu.loop == 1
    // (SorPipeLine.seq.hpp:31-51); present loop == 1 so every call is well defined.
    auto *c = static_cast<RefCtx *>(h);
    const size_t saved = c->simu.loop;
    c->simu.loop = 1;
    c->simu.p_sor_omega = omega;
    c->simu.p_sor_iter_max = itmax;
    const double saved_tol = c->simu.p_criteria;
    c->simu.p_criteria = tol;
    std::streambuf *old = std::cout.rdbuf(nullptr);
    SorPipeLine_seq(c->ShareM, c->Sor, c->simu, c->T1, c->t1, c->Lo, *c->g);
    std::cout.rdbuf(old);
    c->simu.loop = saved;
    c->simu.p_criteria = saved_tol;
    if (err) *err = c->simu.error;
    return c->simu.iters;
}

// One iteration of the runSeqOmp loop body (RUN_CPU.cpp:257-436), I/O removed.
// out: cd, cl, time, iters, err, solved
void ref_step(void *h, int loop, double *out)
{
    auto *c = static_cast<RefCtx *>(h);
    grid &gA = *c->g;
    c->simu.loop = loop;
    ConvectionDifussion(c->simu, c->T0, c->T1, c->Lo, gA);
    int iters = 0;
    double err = 0.0;
    int solved = 0;
    if (c->simu.loop > 1) {
        createBMatrix_omp(c->T1, c->Mx, c->simu, c->Lo, gA);
        iters = ref_solve(h, &err);
        Pressure_transform_X_result(c->t1, c->Mx, c->Lo, gA);
        solved = 1;
    }
    update_UandF_omp(c->Dfib, c->simu, c->t1, c->T1, c->T3, c->Lo, gA);
    copyVel(c->T3, c->T0);
    BC_staggered_main(c->Lo, c->T0, c->t1, gA);
    BC_staggered_copy(c->Lo, c->T0, c->T1, gA);
    double cd = 0, cl = 0;
    ref_cdcl(h, &cd, &cl);
    if (out) {
        out[0] = cd;
        out[1] = cl;
        out[2] = c->simu.time;
        out[3] = iters;
        out[4] = err;
        out[5] = solved;
    }
    c->simu.time += c->simu.dt;  // finishloop (0_1_SimulationVariable.hpp:139-156)
}

// OutputPLOT3D_Qfile (qfile.hpp:6-108) writes mx_out/P3D<io>.q and OutputPlot3D_Xfile
// (xfileWrite.hpp:6-64) mx_out/P3D.x, relative to the working directory; both create
// ../mx_out when it is missing, so the caller provides mx_out/ and ../mx_out first.
int ref_plot3d_q(void *h, int io, double time)
{
    auto *c = static_cast<RefCtx *>(h);
    c->simu.currentFile = (size_t)io;
    c->simu.time = time;
    return OutputPLOT3D_Qfile(c->Dfib, c->simu, c->t1, c->T3, c->Lo, *c->g);
}

int ref_plot3d_x(void *h)
{
    auto *c = static_cast<RefCtx *>(h);
    return OutputPlot3D_Xfile(c->simu, *c->g);
}

// CheckSteadyStatebyMaxVal_omp (CheckSteadystate.hpp:247-307) on the harness's T0 and
// T3.  The reference prints the three maxima and returns nothing; its line is captured
// at 17 significant digits and read back, which restores the doubles exactly.
int ref_maxdiff(void *h, double *out)
{
    auto *c = static_cast<RefCtx *>(h);
    std::ostringstream buf;
    std::streambuf *old = std::cout.rdbuf(buf.rdbuf());
    const std::streamsize prec = std::cout.precision(17);
    CheckSteadyStatebyMaxVal_omp(c->simu, c->ShareM, c->T0, c->T3, c->Lo, *c->g);
    std::cout.precision(prec);
    std::cout.rdbuf(old);
    const std::string line = buf.str();
    const size_t eq = line.find('=');
    if (eq == std::string::npos) return -1;
    return std::sscanf(line.c_str() + eq + 1, " %lf , %lf , %lf", &out[0], &out[1], &out[2]) == 3
               ? 0
               : -1;
}

int ref_threads(void) { return omp_get_max_threads(); }

void ref_set_threads(int threads) { omp_set_num_threads(threads); }

void ref_set_maxit(void *h, int maxit)
{
    auto *c = static_cast<RefCtx *>(h);
    if (c->precond) (*c->csr_solver).*member(CsrMaxit()) = maxit;
    else (*c->spe_solver).*member(SpeMaxit()) = maxit;
}

void ref_set_tol(void *h, double tol)
{
    auto *c = static_cast<RefCtx *>(h);
    c->simu.p_criteria = tol;
    if (c->precond) c->csr_solver->setTolerance(tol);
    else c->spe_solver->setTolerance(tol);
}

}  // extern "C"
