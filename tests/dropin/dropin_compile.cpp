// Compile check of include/cfdmx_reference_api.hpp against the reference's own types:
// the drop-in adapter is instantiated with velocity, pressure, grid, simpulationVariable,
// divideLocal, MxClass, DfibArray, SORcoefficient and shareMenory exactly as runSeqOmp
// (src/RUN_CPU.cpp:257-436) passes them.  Built with -fsyntax-only (test infrastructure).
#include "0_General.hpp"
#include "run/general.hpp"

#include "cfdmx_reference_api.hpp"

// the Smagorinsky stage as the reference calls it under TERBULENCE_SMAGORINSKY
// (RUN_CPU.cpp:261-263)
void dropin_les(grid &gA, simpulationVariable &simu, velocity &T0, pressure &t1)
{
    shareMenory ShareM;
    divideLocal Lo;
    cfdmx_ref::SmagorinskyModel(ShareM, simu, T0, t1, Lo, gA);
}

bool dropin_loop(grid &gA, clockstruct &timer, simpulationVariable &simu, int argc, char **argv)
{
    DfibArray Dfib;
    pressure t1;
    velocity T0, T1, T3;
    SORcoefficient Sor;
    shareMenory ShareM;
    MxClass Mx;
    divideLocal Lo;
    cfdmx_ref::BC_staggered_main(Lo, T0, t1, gA);
    cfdmx_ref::BC_staggered_copy(Lo, T0, T1, gA);
    cfdmx_ref::SPE_matrix<double> A(gA.nxCal, gA.nyCal, gA.nzCal);
    A.setupPressure(gA.gC, gA.Dx, gA.Dy, gA.Dz, gA.Dxs, gA.Dys, gA.Dzs);
    cfdmx_ref::bicgstabRe2<cfdmx_ref::SPE_matrix<double>> pSolver(A);
    pSolver.setTolerance(simu.p_criteria);
    cfdmx_ref::DFIB_CylinderZ(Dfib, Lo, gA);
    cfdmx_ref::OutputPlot3D_Xfile(simu, gA);
    cfdmx_ref::OutputPLOT3D_Qfile(Dfib, simu, t1, T0, Lo, gA);
    for (simu.loop = 1; simu.get_finishloop(); simu.finishloop()) {
        cfdmx_ref::ConvectionDifussion(simu, T0, T1, Lo, gA);
        if (simu.loop > 1) {
            cfdmx_ref::createBMatrix_omp(T1, Mx, simu, Lo, gA);
            std::tie(simu.iters, simu.error) = pSolver.solve(Mx.matB, Mx.X_result);
            cfdmx_ref::Pressure_transform_X_result(t1, Mx, Lo, gA);
        }
        cfdmx_ref::update_UandF_omp(Dfib, simu, t1, T1, T3, Lo, gA);
        copyVel(T3, T0);
        cfdmx_ref::BC_staggered_main(Lo, T0, t1, gA);
        cfdmx_ref::BC_staggered_copy(Lo, T0, T1, gA);
        cfdmx_ref::CD_CL(simu, Lo, Dfib, gA);
        if (simu.get_writefile()) cfdmx_ref::OutputPLOT3D_Qfile(Dfib, simu, t1, T3, Lo, gA);
    }
    cfdmx_ref::SorPipeLine_omp(ShareM, Sor, simu, T1, t1, Lo, gA);
    std::vector<double> x(8), r(8);
    A.multiply_omp(x, r);
    return cfdmx_ref::runSeqOmp(gA, timer, simu, argc, argv);
}
