// End-to-end run of the C++ drop-in adapter (include/cfdmx_reference_api.hpp) against the
// reference's own time loop (TEST INFRASTRUCTURE ONLY).
//
// Built by oracle/Makefile (target dropin) against the reference headers under
// /root/reference/inc and libcfdmx.so; the binary lands in oracle/_ref/ like the other
// reference-derived checkers.  It runs, from the same initial state (u = 1, v = w = p = 0,
// the reference's DFIB_CylinderZ mask, its FLOWpassing boundary table):
//   1. the loop body of runSeqOmp (RUN_CPU.cpp:257-436) with the reference's functions,
//      on one OpenMP thread;
//   2. the same loop with every stage replaced by its cfdmx_ref:: drop-in (host vectors
//      moved to the device and back at each call);
//   3. cfdmx_ref::runSeqOmp (the whole loop on the device), which also writes the
//      reference's Plot3D files and Information/Time_cDcL.dat in the working directory.
// It prints one line per comparison and exits with status 1 when a bound is missed:
//   - step 1 (no pressure solve, RUN_CPU.cpp:303): 2 bit for bit against 1 for u, v, w;
//   - every step: 2 against 1 within 1e-10 of |u| (relative L2) for u, v, w and 1e-8 for
//     the mean-free pressure; BiCGSTAB sums in a fixed device order there and in index
//     order in the reference, so iterates differ by rounding and both solve to 1e-12;
//   - 3 against 2 bit for bit on the interior cells (the same kernels, fields resident or
//     not; copyVel is a copy in 2 and a buffer swap in 3, which differ only in ghost
//     corners no stage reads), and the same Cd, Cl rows in Time_cDcL.dat.
#include "0_General.hpp"
#include "run/general.hpp"

#include "cfdmx_reference_api.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <sys/stat.h>

namespace {

constexpr int kNx = 48, kNy = 24, kNz = 4, kSteps = 4;
constexpr double kRe = 100.0, kDt = 0.002, kTol = 1e-12;

struct State {
    velocity T0, T1, T3;
    pressure t1;
    DfibArray Dfib;
    MxClass Mx;
};

std::unique_ptr<grid> make_grid()
{
    // class grid is an aggregate (structureGrid.hpp:20-40): the size and domain are set
    // by aggregate initialisation, as in oracle/ref_harness.cpp
    return std::unique_ptr<grid>(
        new grid{{"[x]", "[Y]", "[Z]"}, "uniform", 20.0, 10.0, 0.5, 2, kNx, kNy, kNz});
}

void init_state(State &s, grid &gA)
{
    s.Dfib.eta.resize(gA.iceltot);
    s.Dfib.f.resize(gA.iceltot * 3);
    s.t1.p.resize(gA.iceltot);
    s.T0.resize(gA.iceltot);
    s.T1.resize(gA.iceltot);
    s.T3.resize(gA.iceltot);
    s.T0.iniU(1.0, 0.0, 0.0); // RUN_CPU.cpp:66-76
    s.T1.iniU(1.0, 0.0, 0.0);
    s.T3.iniU(1.0, 0.0, 0.0);
    s.t1.init_p(0.0);
    s.Mx.X_result.assign(gA.iceltotCal, 0.0);
    s.Mx.matB.assign(gA.iceltotCal, 0.0);
}

double rel_l2(const std::vector<double> &a, const std::vector<double> &b, double scale)
{
    double d = 0.0;
    for (size_t i = 0; i < a.size(); ++i) d += (a[i] - b[i]) * (a[i] - b[i]);
    return std::sqrt(d) / scale;
}

bool interior_equal(const std::vector<double> &a, const std::vector<double> &b, grid &gA)
{
    for (int i = gA.gC; i < gA.nx - gA.gC; ++i)
        for (int j = gA.gC; j < gA.ny - gA.gC; ++j)
            for (int k = gA.gC; k < gA.nz - gA.gC; ++k)
                if (a[gA.icel(i, j, k)] != b[gA.icel(i, j, k)]) return false;
    return true;
}

double norm(const std::vector<double> &a)
{
    double d = 0.0;
    for (double x : a) d += x * x;
    return std::sqrt(d);
}

std::vector<double> interior_mean_free(const std::vector<double> &p, grid &gA)
{
    std::vector<double> out;
    for (int i = gA.gC; i < gA.nx - gA.gC; ++i)
        for (int j = gA.gC; j < gA.ny - gA.gC; ++j)
            for (int k = gA.gC; k < gA.nz - gA.gC; ++k) out.push_back(p[gA.icel(i, j, k)]);
    double m = 0.0;
    for (double x : out) m += x;
    m /= (double)out.size();
    for (double &x : out) x -= m;
    return out;
}

int failures = 0;

void report(const char *what, int step, double value, double bound)
{
    const bool ok = value <= bound;
    std::printf("%s step %d: %.3e (bound %.1e) %s\n", what, step, value, bound,
                ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

} // namespace

int main()
{
    omp_set_num_threads(1);
    std::unique_ptr<grid> g = make_grid();
    grid &gA = *g;
    simpulationVariable simu;
    simu.Locality = 0;
    simu.set_time_max(kSteps * kDt);
    simu.set_dt(kDt);
    simu.set_Re(kRe);
    simu.p_criteria = kTol;
    simu.init_fileStep(1.0); // no q file inside the run
    simu.DfibMethod = "DFIB_Cylinder-Z";
    divideLocal Lo;
    std::vector<int> grid_size{gA.nx, gA.ny, gA.nz};
    std::vector<int> dims{1, 1, 1};
    Lo.initTable(grid_size, dims);
    Lo.initLocal(0, 0, 0);
    shareMenory ShareM;
    generateGride(simu, ShareM, Lo, gA);

    // 1. the reference
    State R;
    init_state(R, gA);
    R.Mx.matA_spe.resize(gA.nxCal, gA.nyCal, gA.nzCal);
    R.Mx.matA_spe.setupPressure(gA.gC, gA.Dx, gA.Dy, gA.Dz, gA.Dxs, gA.Dys, gA.Dzs);
    solver::bicgstabRe2<mat::SPE_matrix<double>> refSolver(R.Mx.matA_spe);
    refSolver.setTolerance(kTol);
    ::DFIB_CylinderZ(R.Dfib, Lo, gA);
    ::BC_staggered_main(Lo, R.T0, R.t1, gA);
    ::BC_staggered_copy(Lo, R.T0, R.T1, gA);

    // 2. the drop-in stages
    cfdmx_ref::config().tol = kTol;
    State D;
    init_state(D, gA);
    cfdmx_ref::SPE_matrix<double> A(gA.nxCal, gA.nyCal, gA.nzCal);
    A.setupPressure(gA.gC, gA.Dx, gA.Dy, gA.Dz, gA.Dxs, gA.Dys, gA.Dzs);
    cfdmx_ref::bicgstabRe2<cfdmx_ref::SPE_matrix<double>> pSolver(A);
    pSolver.setTolerance(kTol);
    cfdmx_ref::DFIB_CylinderZ(D.Dfib, Lo, gA);
    cfdmx_ref::BC_staggered_main(Lo, D.T0, D.t1, gA);
    cfdmx_ref::BC_staggered_copy(Lo, D.T0, D.T1, gA);

    std::vector<double> cd_d(kSteps + 1), cl_d(kSteps + 1);
    for (int step = 1; step <= kSteps; ++step) {
        simu.loop = step;
        simu.time = (step - 1) * kDt;
        // the reference (RUN_CPU.cpp:257-436, its file output left out)
        ::ConvectionDifussion(simu, R.T0, R.T1, Lo, gA);
        int it_r = 0, it_d = 0;
        if (step > 1) {
            ::createBMatrix_omp(R.T1, R.Mx, simu, Lo, gA);
            it_r = refSolver.solve(R.Mx.matB, R.Mx.X_result).first;
            ::Pressure_transform_X_result(R.t1, R.Mx, Lo, gA);
        }
        ::update_UandF_omp(R.Dfib, simu, R.t1, R.T1, R.T3, Lo, gA);
        ::copyVel(R.T3, R.T0);
        ::BC_staggered_main(Lo, R.T0, R.t1, gA);
        ::BC_staggered_copy(Lo, R.T0, R.T1, gA);
        virtualF_Int(R.Dfib, Lo, gA, 0);
        virtualF_Int(R.Dfib, Lo, gA, 1);
        const double cd_r = -2.0 * R.Dfib.ValSum[0] / gA.lz;
        const double cl_r = -2.0 * R.Dfib.ValSum[1] / gA.lz;
        // the drop-in, the same calls under cfdmx_ref::
        cfdmx_ref::ConvectionDifussion(simu, D.T0, D.T1, Lo, gA);
        if (step > 1) {
            cfdmx_ref::createBMatrix_omp(D.T1, D.Mx, simu, Lo, gA);
            std::tie(it_d, simu.error) = pSolver.solve(D.Mx.matB, D.Mx.X_result);
            cfdmx_ref::Pressure_transform_X_result(D.t1, D.Mx, Lo, gA);
        }
        cfdmx_ref::update_UandF_omp(D.Dfib, simu, D.t1, D.T1, D.T3, Lo, gA);
        ::copyVel(D.T3, D.T0);
        cfdmx_ref::BC_staggered_main(Lo, D.T0, D.t1, gA);
        cfdmx_ref::BC_staggered_copy(Lo, D.T0, D.T1, gA);
        std::streambuf *old = std::cout.rdbuf(nullptr); // CD_CL's stdout line
        (void)::mkdir("Information", 0755);
        cfdmx_ref::CD_CL(simu, Lo, D.Dfib, gA);
        std::cout.rdbuf(old);
        const double cd_dd = -2.0 * D.Dfib.ValSum[0] / gA.lz;
        const double cl_dd = -2.0 * D.Dfib.ValSum[1] / gA.lz;
        cd_d[step] = cd_dd;
        cl_d[step] = cl_dd;
        const double un = norm(R.T0.u);
        std::printf("step %d: iterations reference %d drop-in %d; Cd %.15g / %.15g\n", step,
                    it_r, it_d, cd_r, cd_dd);
        if (step == 1) {
            const bool same = R.T0.u == D.T0.u && R.T0.v == D.T0.v && R.T0.w == D.T0.w;
            std::printf("u, v, w bit for bit step 1: %s\n", same ? "ok" : "FAIL");
            if (!same) ++failures;
        }
        report("u", step, rel_l2(D.T0.u, R.T0.u, un), 1e-10);
        report("v", step, rel_l2(D.T0.v, R.T0.v, un), 1e-10);
        report("w", step, rel_l2(D.T0.w, R.T0.w, un), 1e-10);
        const std::vector<double> pr = interior_mean_free(R.t1.p, gA),
                                  pd = interior_mean_free(D.t1.p, gA);
        report("p", step, step > 1 ? rel_l2(pd, pr, norm(pr)) : 0.0, 1e-8);
        report("Cd", step, std::fabs(cd_dd - cd_r) / std::fabs(cd_r), 1e-12);
    }

    // 3. the whole loop on the device
    simu.loop = 1;
    simu.time = 0.0;
    simu.currentFile = 0;
    clockstruct timer;
    std::streambuf *old = std::cout.rdbuf(nullptr);
    const bool ran = cfdmx_ref::runSeqOmp(gA, timer, simu, 0, nullptr);
    std::cout.rdbuf(old);
    if (!ran) ++failures;
    cfdmx_ctx *c = cfdmx_ref::session().ctx;
    const long long G = cfdmx_ref::session().G;
    std::vector<double> u(G), v(G), w(G);
    cfdmx_download_velocity(c, 0, u.data(), v.data(), w.data());
    std::vector<double> p(G);
    cfdmx_download_pressure(c, p.data());
    const bool same = interior_equal(u, D.T0.u, gA) && interior_equal(v, D.T0.v, gA) &&
                      interior_equal(w, D.T0.w, gA) && interior_equal(p, D.t1.p, gA);
    std::printf("runSeqOmp interior u, v, w, p bit for bit against the stage loop: %s\n",
                same ? "ok" : "FAIL");
    if (!same) ++failures;
    // Time_cDcL.dat: the stage loop's CD_CL rows, then runSeqOmp's, the same text
    std::FILE *fp = std::fopen("Information/Time_cDcL.dat", "r");
    std::vector<std::string> rows;
    if (fp) {
        char line[512];
        while (std::fgets(line, sizeof line, fp)) {
            double t, cd, cl;
            // the reference writes each row after a newline and its header without one,
            // so an appended run's header follows the last row on the same line: the
            // three numbers are compared
            if (std::sscanf(line, "%lf %lf %lf", &t, &cd, &cl) == 3) {
                char buf[160];
                std::snprintf(buf, sizeof buf, "%.17g %.17g %.17g", t, cd, cl);
                rows.push_back(buf);
            }
        }
        std::fclose(fp);
    }
    bool rows_ok = rows.size() == 2 * (size_t)kSteps;
    for (int q = 0; rows_ok && q < kSteps; ++q) rows_ok = rows[q] == rows[q + kSteps];
    std::printf("Time_cDcL.dat: %zu rows, runSeqOmp's equal to the stage loop's: %s\n",
                rows.size(), rows_ok ? "ok" : "FAIL");
    for (size_t q = 0; q < rows.size(); ++q) std::printf("  row %zu: %s\n", q, rows[q].c_str());
    if (!rows_ok) ++failures;
    std::printf("%s\n", failures ? "DROPIN FAIL" : "DROPIN OK");
    return failures ? 1 : 0;
}
