// dropin_wrap.cpp -- test harness: registers the GPU step solver in the
// reference's step-solver factory without changing the reference.
//
// The drop-in binary links the reference's own objects (its unmodified main
// loop included) with -Wl,--wrap=<symbol of
// abstractStepSolver::createStepSolver>, so main's call of the factory
// (Schemi.cpp:753) lands here: keys ending in "_GPU" or "_GPU_MIXED" build
// schemi::gpuStepSolverRK through createGpuStepSolver, every other key goes on
// to the reference's own factory (abstractStepSolver.cpp:56-146).  A
// maintainer adds the same two-line key test to createStepSolver itself.
#include <string>

#include "gpuStepSolverRK.hpp"

#ifndef CREATE_STEP_SOLVER_SYMBOL
#error "CREATE_STEP_SOLVER_SYMBOL (the factory's assembler name) must be defined"
#endif

using namespace schemi;

#define SG_FACTORY_ARGS                                                                         \
    const std::string &key, homogeneousPhase<cubicCell> &gasPhase_in,                           \
        const abstractLimiter &limiter_in, const abstractFlowSolver &fsolver_in,                \
        const bool &gravitationFlag_in, const vector &g_in,                                     \
        const boundaryConditionValue &boundaryConditionValueCalc_in, scalar &timeForTVD_in,     \
        scalar &timeForHancock_in, scalar &timeForFlowCalculation_in,                           \
        scalar &timeForTimeIntegration_in, const MPIHandler &parallelism_in,                    \
        const bool &diffusionFlag_in, const abstractMatrixSolver &msolver_in,                   \
        const abstractMatrixSolver &msolverEnthFl_in,                                           \
        const std::pair<scalar, scalar> &timestepCoeffs_in, scalar &timeForDiffusion_in,        \
        const std::vector<boundaryConditionType> &commonConditions_in,                          \
        const enthalpyFlow &enthalpyFlowFlag_in, const bool &linearFlag_in,                     \
        const volumeField<scalar> &minimalLengthScale_in, const timestep &sourceTimeFlag_in,    \
        const bool &molMassDiffusionFlag_in,                                                    \
        chemicalKinetics::abstractChemicalKinetics &chemKin_in,                                 \
        const bool &nonLinearityIteratonsFlag_in

#define SG_FACTORY_PASS                                                                         \
    key, gasPhase_in, limiter_in, fsolver_in, gravitationFlag_in, g_in,                         \
        boundaryConditionValueCalc_in, timeForTVD_in, timeForHancock_in,                        \
        timeForFlowCalculation_in, timeForTimeIntegration_in, parallelism_in, diffusionFlag_in, \
        msolver_in, msolverEnthFl_in, timestepCoeffs_in, timeForDiffusion_in,                   \
        commonConditions_in, enthalpyFlowFlag_in, linearFlag_in, minimalLengthScale_in,         \
        sourceTimeFlag_in, molMassDiffusionFlag_in, chemKin_in, nonLinearityIteratonsFlag_in

std::unique_ptr<abstractStepSolver> reference_factory(SG_FACTORY_ARGS)
    __asm__("__real_" CREATE_STEP_SOLVER_SYMBOL);
std::unique_ptr<abstractStepSolver> wrapped_factory(SG_FACTORY_ARGS)
    __asm__("__wrap_" CREATE_STEP_SOLVER_SYMBOL);

static bool gpu_key(const std::string &k)
{
    auto ends = [&](const std::string &s) {
        return k.size() >= s.size() && k.compare(k.size() - s.size(), s.size(), s) == 0;
    };
    return ends("_GPU") || ends("_GPU_MIXED");
}

std::unique_ptr<abstractStepSolver> wrapped_factory(SG_FACTORY_ARGS)
{
    if (gpu_key(key))
        return createGpuStepSolver(SG_FACTORY_PASS);
    return reference_factory(SG_FACTORY_PASS);
}
