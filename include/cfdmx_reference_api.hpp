// cfdmx_reference_api.hpp -- the reference's hot-path entry points, by their own names
// and signatures, over the C-ABI of cfdmx.h.
//
// For a maintainer of jimmy3769tw/CFD_NSsolver_forHPC_flowPassingCylinder: include this
// header after the reference's own headers (run/general.hpp) and qualify the calls in
// runSeqOmp (src/RUN_CPU.cpp:257-436) with cfdmx_ref::, or call cfdmx_ref::runSeqOmp
// instead of runSeqOmp to keep the whole time loop on the GPU.  The functions are
// templates over the reference's types (velocity, pressure, grid, simpulationVariable,
// divideLocal, MxClass, DfibArray, SORcoefficient, shareMenory), so this header needs
// none of the reference's headers itself.
//
// The stage functions move the host std::vectors they are given to the device and back
// on every call (the reference owns its fields on the host); runSeqOmp keeps every field
// resident and returns only the per-step report.  Errors are thrown as
// std::runtime_error / std::invalid_argument, as the reference does.
#pragma once

#include "cfdmx.h"

#include <cstdio>
#include <fstream>
#include <sys/stat.h>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cfdmx_ref {

struct Session {
    cfdmx_ctx *ctx = nullptr;
    long long G = 0, N = 0;
    ~Session()
    {
        if (ctx) cfdmx_destroy(ctx);
    }
};

inline Session &session()
{
    static Session s;
    return s;
}

// Device-context configuration used when the first call creates the context.  The
// defaults are the reference's: the FLOWpassing table (FLOWpassing.hpp:6-40), SPE
// operator, BiCGSTAB, tolerance simu.p_criteria, 5000 iterations.
inline cfdmx_config &config()
{
    static cfdmx_config c = [] {
        cfdmx_config x;
        cfdmx_config_default(&x);
        x.steady_check = 1; // CheckSteadyStatebyMaxVal_omp every step (RUN_CPU.cpp:380-384)
#ifdef TERBULENCE_SMAGORINSKY
        x.smagorinsky_cs = 0.18; // SmagorinskyModel.hpp:38
#endif
        return x;
    }();
    return c;
}

inline void check(int rc, const char *what)
{
    if (rc == CFDMX_OK) return;
    const std::string msg = std::string(what) + ": " + cfdmx_last_error(session().ctx);
    if (rc == CFDMX_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// the context for this process's grid (runSeqOmp has one grid, RUN_CPU.cpp:9-14);
// Re and dt come from config() until a call that carries the simulation variables
template <class Grid> cfdmx_ctx *context(const Grid &gA)
{
    Session &s = session();
    if (s.ctx) return s.ctx;
    cfdmx_config c = config();
    c.nx = gA.nxCal;
    c.ny = gA.nyCal;
    c.nz = gA.nzCal;
    c.lx = gA.lx;
    c.ly = gA.ly;
    c.lz = gA.lz;
    if (gA.Gridder == "uniform") c.grid_kind = CFDMX_GRID_UNIFORM;
    else if (gA.Gridder == "non_uniform-SML") c.grid_kind = CFDMX_GRID_SML;
    else throw std::runtime_error("grid ?"); // Generator.hpp:32
    check(cfdmx_create(&c, &s.ctx), "cfdmx_create");
    check(cfdmx_sizes(s.ctx, &s.G, &s.N), "cfdmx_sizes");
    return s.ctx;
}

template <class Grid, class Simu> cfdmx_ctx *context(const Grid &gA, const Simu &simu)
{
    cfdmx_ctx *c = context(gA);
    check(cfdmx_set_physics(c, simu.Re, simu.dt), "set Re, dt");
    return c;
}

inline cfdmx_ctx *ctx_or_throw()
{
    if (!session().ctx)
        throw std::runtime_error("cfdmx_ref: no device context yet (call a stage that takes "
                                 "the grid first)");
    return session().ctx;
}

template <class Vec> inline void fit(Vec &v, long long n)
{
    if ((long long)v.size() != n) v.resize((size_t)n);
}

template <class Vel> void put_vel(int which, Vel &v)
{
    const long long G = session().G;
    fit(v.u, G);
    fit(v.v, G);
    fit(v.w, G);
    check(cfdmx_upload_velocity(session().ctx, which, v.u.data(), v.v.data(), v.w.data()),
          "upload velocity");
}

template <class Vel> void get_vel(int which, Vel &v)
{
    check(cfdmx_download_velocity(session().ctx, which, v.u.data(), v.v.data(), v.w.data()),
          "download velocity");
}

// ConvectionDifussion, Convection_Difussion.hpp:174-180
template <class Simu, class Vel, class Lo, class Grid>
bool ConvectionDifussion(Simu &simu, Vel &oldVel, Vel &curtVel, const Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    put_vel(0, oldVel);
    put_vel(1, curtVel);
    check(cfdmx_convdiff(c), "ConvectionDifussion");
    get_vel(1, curtVel);
    return true;
}

// SmagorinskyModel, SmagorinskyModel.hpp:30-37 (called at RUN_CPU.cpp:261-263 when the
// reference is built with TERBULENCE_SMAGORINSKY): Viseff of TX from its velocity.  The
// context must have been created with config().smagorinsky_cs > 0 (the reference's
// constant is 0.18); the following ConvectionDifussion adds Viseff to 1/Re.
template <class Share, class Simu, class Vel, class Pre, class Lo, class Grid>
void SmagorinskyModel(Share &, Simu &simu, Vel &TX, Pre &, Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    put_vel(0, TX);
    check(cfdmx_smagorinsky(c), "SmagorinskyModel");
    fit(TX.Viseff, session().N);
    check(cfdmx_download_viseff(c, TX.Viseff.data()), "download Viseff");
}

// BC_staggered_main, Boundary_Condition.hpp:37-42
template <class Lo, class Vel, class Pre, class Grid>
bool BC_staggered_main(Lo &, Vel &vel, Pre &pre, Grid &gA)
{
    cfdmx_ctx *c = context(gA);
    put_vel(0, vel);
    fit(pre.p, session().G);
    check(cfdmx_upload_pressure(c, pre.p.data()), "upload pressure");
    check(cfdmx_bc_main(c), "BC_staggered_main");
    get_vel(0, vel);
    check(cfdmx_download_pressure(c, pre.p.data()), "download pressure");
    return true;
}

// BC_staggered_copy, updateBC.hpp:6-11
template <class Lo, class Vel, class Grid>
bool BC_staggered_copy(const Lo &, const Vel &oldVel, Vel &currtVel, Grid &gA)
{
    cfdmx_ctx *c = context(gA);
    Vel tmp = oldVel;
    put_vel(0, tmp);
    put_vel(1, currtVel);
    check(cfdmx_bc_copy(c), "BC_staggered_copy");
    get_vel(1, currtVel);
    return true;
}

// createBMatrix_omp, b_matrix.hpp:35-41
template <class Vel, class Mx, class Simu, class Lo, class Grid>
void createBMatrix_omp(Vel &T1, Mx &mx, Simu &simu, Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    put_vel(1, T1);
    check(cfdmx_pressure_rhs(c), "createBMatrix_omp");
    fit(mx.matB, session().N);
    check(cfdmx_download_rhs(c, mx.matB.data()), "download rhs");
}

// mat::SPE_matrix, SPE_sparseMatrix.hpp:22-108: the operator is built on the device
// from the context's grid; multiply_omp runs there.
template <class T> class SPE_matrix {
public:
    SPE_matrix() {}
    SPE_matrix(int nx, int ny, int nz) { resize(nx, ny, nz); }
    void resize(int nx, int ny, int nz) { ndim_ = nx * ny * nz; }
    template <class V>
    void setupPressure(int, const V &, const V &, const V &, const V &, const V &, const V &)
    {
    }
    void multiply_omp(std::vector<double> &x, std::vector<double> &r)
    {
        cfdmx_ctx *c = ctx_or_throw();
        fit(r, (long long)x.size());
        check(cfdmx_apply_operator(c, x.data(), r.data()), "SPE_matrix::multiply_omp");
    }
    int row() const { return ndim_; }

private:
    int ndim_ = 0;
};

// solver::bicgstabRe2, bicgstabRe2.hpp:10-42 (the history vectors and the restart state
// live on the device)
template <class matrixT> class bicgstabRe2 {
public:
    bicgstabRe2() {}
    explicit bicgstabRe2(matrixT &) {}
    void setTolerance(double tolerance)
    {
        tol_ = tolerance;
        if (session().ctx) check(cfdmx_set_tolerance(session().ctx, tol_, 5000), "setTolerance");
    }
    std::pair<int, double> solve(const std::vector<double> &rhs, std::vector<double> &x)
    {
        cfdmx_ctx *c = ctx_or_throw();
        if (tol_ > 0.0) check(cfdmx_set_tolerance(c, tol_, 5000), "setTolerance");
        check(cfdmx_upload_rhs(c, rhs.data()), "upload rhs");
        check(cfdmx_upload_xresult(c, x.data()), "upload x");
        int it = 0;
        double res = 0.0;
        check(cfdmx_solve_bicgstab(c, &it, &res), "bicgstabRe2::solve");
        check(cfdmx_download_xresult(c, x.data()), "download x");
        return std::make_pair(it, res);
    }

private:
    double tol_ = 0.0;
};

// Pressure_transform_X_result, Transform.hpp:8-13
template <class Pre, class Mx, class Lo, class Grid>
void Pressure_transform_X_result(Pre &t1, Mx &mx, Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA);
    fit(t1.p, session().G);
    check(cfdmx_upload_pressure(c, t1.p.data()), "upload pressure");
    check(cfdmx_upload_xresult(c, mx.X_result.data()), "upload x");
    check(cfdmx_transform(c), "Pressure_transform_X_result");
    check(cfdmx_download_pressure(c, t1.p.data()), "download pressure");
}

// SorPipeLine_omp, SorPipeLine.omp.hpp:4-12 (red-black ordering on the device)
template <class Share, class Sor, class Simu, class Vel, class Pre, class Lo, class Grid>
bool SorPipeLine_omp(Share &, Sor &, Simu &simu, Vel &T1, Pre &t1, Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    check(cfdmx_set_tolerance(c, simu.p_criteria, (int)simu.p_sor_iter_max), "tolerance");
    put_vel(1, T1);
    fit(t1.p, session().G);
    check(cfdmx_upload_pressure(c, t1.p.data()), "upload pressure");
    int it = 0;
    double err = 0.0;
    check(cfdmx_solve_rbsor(c, &it, &err), "SorPipeLine");
    check(cfdmx_download_pressure(c, t1.p.data()), "download pressure");
    simu.iters = it;
    simu.error = err;
    return true;
}

// update_UandF_omp, updateUandF.hpp:90-98
template <class Dfib, class Simu, class Pre, class Vel, class Lo, class Grid>
void update_UandF_omp(Dfib &dfib, const Simu &simu, const Pre &t1, const Vel &oldVel,
                      Vel &curtVel, const Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    const long long G = session().G;
    std::vector<double> p(t1.p.begin(), t1.p.end());
    fit(p, G);
    check(cfdmx_upload_pressure(c, p.data()), "upload pressure");
    fit(dfib.eta, G);
    check(cfdmx_upload_eta(c, dfib.eta.data()), "upload eta");
    Vel old = oldVel;
    put_vel(1, old);
    put_vel(2, curtVel);
    check(cfdmx_project_ib(c), "update_UandF_omp");
    get_vel(2, curtVel);
    fit(dfib.f, 3 * G);
    check(cfdmx_download_force(c, dfib.f.data()), "download f");
}

// DFIB_CylinderZ, setEta.hpp:136-140, with its fixed centre (6.5, ly/2), radius 0.5 and
// 100 x 100 sub-cells (setEta.hpp:145-152)
template <class Dfib, class Lo, class Grid> void DFIB_CylinderZ(Dfib &dfib, Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA);
    dfib.cylinderDimension = 1.0;
    dfib.cylinderCenter.assign({6.5, gA.ly / 2.0});
    check(cfdmx_set_eta_cylinder_z(c, 6.5, gA.ly / 2.0, 0.5, 100), "DFIB_CylinderZ");
    fit(dfib.eta, session().G);
    check(cfdmx_download_eta(c, dfib.eta.data()), "download eta");
}

// the stdout line and the Time_cDcL.dat row of CD_CL (virtualForceIntergrator.hpp:113-145):
// the file in Information/ of the working directory, its Tecplot header on the first step
template <class Simu> void cdcl_output(Simu &simu, double cd, double cl)
{
    std::cout << "[cD, cL] : " << cd << ", " << cl << std::endl;
    (void)::mkdir("Information", 0755);
    std::ofstream file("Information/Time_cDcL.dat", std::ios::out | std::ios::app);
    if (simu.loop == 1)
        file << "TITLE     = \"\"\nVARIABLES = \"simulation time\",\"C<sub>D</sub>\","
                "\"C<sub>L</sub>\"\nZONE T=\"" << simu.ZONE() << "\"";
    file << "\n" << simu.time << " " << cd << " " << cl;
}

// CD_CL, virtualForceIntergrator.hpp:92-148 (same stdout line and Time_cDcL.dat row)
template <class Simu, class Lo, class Dfib, class Grid>
bool CD_CL(Simu &simu, Lo &, Dfib &dfib, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    fit(dfib.f, 3 * session().G);
    check(cfdmx_upload_force(c, dfib.f.data()), "upload f");
    double cd = 0.0, cl = 0.0;
    check(cfdmx_cd_cl(c, &cd, &cl), "CD_CL");
    dfib.ValSum.assign({-cd * gA.lz / 2.0, -cl * gA.lz / 2.0, 0.0});
    cdcl_output(simu, cd, cl);
    return true;
}

// Plot3D files in mx_out/ (created when missing): the reference writes them relative to
// the working directory under these names (qfile.hpp:40-47, xfileWrite.hpp:37-39)
inline std::string plot3d_q_name(size_t io)
{
    std::string name = "mx_out/P3D";
    if (io < 10) name += "0";
    return name + std::to_string(io) + ".q";
}
inline void plot3d_dir() { (void)::mkdir("mx_out", 0755); }

// OutputPLOT3D_Qfile, qfile.hpp:6-12: p, cell-centred u, v, w and eta of the given fields
// to mx_out/P3D<nn>.q (the copy to host memory and the file write are asynchronous; see
// cfdmx_plot3d_q)
template <class Dfib, class Simu, class Pre, class Vel, class Lo, class Grid>
int OutputPLOT3D_Qfile(Dfib &dfib, Simu &simu, Pre &t1, Vel &vel, Lo &, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    const long long G = session().G;
    std::vector<double> p(t1.p.begin(), t1.p.end());
    fit(p, G);
    check(cfdmx_upload_pressure(c, p.data()), "upload pressure");
    fit(dfib.eta, G);
    check(cfdmx_upload_eta(c, dfib.eta.data()), "upload eta");
    Vel v = vel;
    put_vel(0, v);
    plot3d_dir();
    std::cout << "(Output Plaot3D):" << simu.currentFile << std::endl;
    check(cfdmx_plot3d_q(c, plot3d_q_name(simu.currentFile).c_str(), simu.get_time()),
          "OutputPLOT3D_Qfile");
    return 0;
}

// OutputPlot3D_Xfile, xfileWrite.hpp:6-9: the cell-centre coordinates to mx_out/P3D.x
template <class Simu, class Grid> int OutputPlot3D_Xfile(Simu &simu, Grid &gA)
{
    cfdmx_ctx *c = context(gA, simu);
    plot3d_dir();
    check(cfdmx_plot3d_x(c, "mx_out/P3D.x"), "OutputPlot3D_Xfile");
    return 0;
}

// runSeqOmp, RUN_CPU.cpp:9-440 / run/1_1_Sequential.hpp:5-10, with every field resident
// on the device: the reference initial state (u = 1, v = w = p = 0, a cold solver), its
// boundary conditions and mask, then the loop body per step; the per-step report goes to
// simu (iters, error), to stdout as in the reference (the solver line, the steady-state
// check's [DIFFENT U,V,W] line when config().steady_check is on, the [cD, cL] line) and
// to Information/Time_cDcL.dat, and the Plot3D files are written at its points: the grid
// and the initial field before the loop (RUN_CPU.cpp:84,241), then a q file whenever
// simu.get_writefile() says so (RUN_CPU.cpp:411-413), staged on the device while the
// next steps run.
template <class Grid, class Timer, class Simu>
bool runSeqOmp(Grid &gA, Timer &, Simu &simu, int, char **)
{
    cfdmx_ctx *c = context(gA, simu);
    plot3d_dir();
    check(cfdmx_plot3d_x(c, "mx_out/P3D.x"), "OutputPlot3D_Xfile");
    check(cfdmx_init_flow(c, 1.0, 0.0, 0ull), "initial field");
    if (simu.DfibMethod == "DFIB_Cylinder-Z")
        check(cfdmx_set_eta_cylinder_z(c, 6.5, gA.ly / 2.0, 0.5, 100), "DFIB_CylinderZ");
    else if (simu.DfibMethod != "OFF")
        throw std::invalid_argument("DFIB method??"); // RUN_CPU.cpp:236
    check(cfdmx_plot3d_q(c, plot3d_q_name(simu.currentFile).c_str(), simu.get_time()),
          "OutputPLOT3D_Qfile");
    for (simu.loop = 1; simu.get_finishloop(); simu.finishloop()) {
        cfdmx_step_info info;
        check(cfdmx_step(c, 1, &info), "time step");
        simu.iters = info.iters;
        simu.error = info.err;
        simu.printInfo();
        if (config().steady_check) // CheckSteadystate.hpp:302-305
            std::cout << "[DIFFENT U,V,W] = " << info.maxdiff[0] << ", " << info.maxdiff[1]
                      << ", " << info.maxdiff[2] << std::endl;
        cdcl_output(simu, info.cd, info.cl);
        if (simu.get_writefile())
            check(cfdmx_plot3d_q(c, plot3d_q_name(simu.currentFile).c_str(), simu.get_time()),
                  "OutputPLOT3D_Qfile");
    }
    check(cfdmx_plot3d_flush(c), "OutputPLOT3D_Qfile");
    return true;
}

} // namespace cfdmx_ref
