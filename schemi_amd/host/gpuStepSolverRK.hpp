// gpuStepSolverRK.hpp -- the GPU step as a plugin of the reference's own
// step-solver interface.
//
// schemi::gpuStepSolverRK derives from the reference's
// schemi::abstractStepSolver (Schemi/headers/abstractStepSolver.hpp:26-106)
// and takes exactly the arguments abstractStepSolver::createStepSolver passes
// to its solvers (abstractStepSolver.hpp:78-101, abstractStepSolver.cpp:56-146),
// so a maintainer registers it in that factory under the keys
// "SecondOrderRK_GPU" and "SecondOrder_GPU" (createGpuStepSolver below) and
// the reference's main loop runs unchanged.  It reads the reference's objects
// directly: the mesh of the phase's fields (nCells, delta, zero,
// taskDimension), the mixture (mixtureIdeal: Mv, Cvv, Rv), the flux and
// limiter objects (their dynamic types), gravity, every field's boundary
// condition on each patch, and the diffusion settings (constant transport
// model, explicit / Jacobi / Gauss-Seidel solver, linear reconstruction).
//
// calculateStep() uploads the conserved fields of homogeneousPhase<cubicCell>
// (density[0..Ns], momentum, totalEnergy), advances them by one step of
// mesh::timestep() on the GPU and writes back every field the reference's
// step leaves consistent (Advection2dOrder.cpp:578-608): concentrations,
// densities, velocity, momentum, pressure, internal and total energy,
// temperature, and the Helmholtz energy and entropy through the phase's own
// mixture (Fv, Sv), plus mesh::timestepSourceRef() from the diffusion
// stage's source step for the main loop's "Courant-Source" time step
// (Schemi.cpp:828-833).  Errors are thrown as the reference's
// schemi::exception with its errors codes.
//
// Compiled against the reference's headers (-I Schemi/headers
// -I Schemi/enumerations -I Schemi/templates) and linked into the reference
// with gpuStepCore.cpp and libschemi_gpu.so.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "abstractStepSolver.hpp"
#include "gpuStepCore.hpp"

namespace schemi {

class gpuStepSolverRK final : public abstractStepSolver {
public:
    // stepSolver: SCHEMI_STEP_SECOND_ORDER_RK or SCHEMI_STEP_SECOND_ORDER;
    // precision: SCHEMI_PREC_FP64 or SCHEMI_PREC_MIXED; the rest as
    // secondOrderStepSolverRK (secondOrderStepSolverRK.hpp:17-38).
    gpuStepSolverRK(int stepSolver, int precision, homogeneousPhase<cubicCell> &gasPhase_in,
                    const abstractLimiter &limiter_in, const abstractFlowSolver &fsolver_in,
                    const bool &gravitationFlag_in, const vector &g_in,
                    const boundaryConditionValue &boundaryConditionValueCalc_in,
                    scalar &timeForTVD_in, scalar &timeForHancock_in,
                    scalar &timeForFlowCalculation_in, scalar &timeForTimeIntegration_in,
                    const MPIHandler &parallelism_in, const bool &diffusionFlag_in,
                    const abstractMatrixSolver &msolver_in,
                    const abstractMatrixSolver &msolverEnthFl_in,
                    const std::pair<scalar, scalar> &timestepCoeffs_in,
                    scalar &timeForDiffusion_in,
                    const std::vector<boundaryConditionType> &commonConditions_in,
                    const enthalpyFlow &enthalpyFlowFlag_in, const bool &linearFlag_in,
                    const volumeField<scalar> &minimalLengthScale_in,
                    const timestep &sourceTimeFlag_in, const bool &molMassDiffusionFlag_in,
                    chemicalKinetics::abstractChemicalKinetics &chemKin_in,
                    const bool &nonLinearityIteratonsFlag_in);

    gpuStepSolverRK(const gpuStepSolverRK &) = delete;
    auto &operator=(const gpuStepSolverRK &) = delete;
    ~gpuStepSolverRK() noexcept override;

    void calculateStep() override;

    // Courant step of the state after the last step, computed on the device
    // (Schemi.cpp:807-833 without the MPI reduction).
    double nextTimestep() const noexcept;

private:
    std::unique_ptr<schemi_gpu_host::stepCore> core;
};

// createStepSolver for the GPU keys: "SecondOrderRK_GPU" (SSP-RK2),
// "SecondOrder_GPU" (MUSCL-Hancock), each with the suffix "_MIXED" for the
// FP32-flux / FP64-accumulate path.  Other keys throw initialisationError.
std::unique_ptr<abstractStepSolver> createGpuStepSolver(
    const std::string &key, homogeneousPhase<cubicCell> &gasPhase_in,
    const abstractLimiter &limiter_in, const abstractFlowSolver &fsolver_in,
    const bool &gravitationFlag_in, const vector &g_in,
    const boundaryConditionValue &boundaryConditionValueCalc_in, scalar &timeForTVD_in,
    scalar &timeForHancock_in, scalar &timeForFlowCalculation_in,
    scalar &timeForTimeIntegration_in, const MPIHandler &parallelism_in,
    const bool &diffusionFlag_in, const abstractMatrixSolver &msolver_in,
    const abstractMatrixSolver &msolverEnthFl_in,
    const std::pair<scalar, scalar> &timestepCoeffs_in, scalar &timeForDiffusion_in,
    const std::vector<boundaryConditionType> &commonConditions_in,
    const enthalpyFlow &enthalpyFlowFlag_in, const bool &linearFlag_in,
    const volumeField<scalar> &minimalLengthScale_in, const timestep &sourceTimeFlag_in,
    const bool &molMassDiffusionFlag_in, chemicalKinetics::abstractChemicalKinetics &chemKin_in,
    const bool &nonLinearityIteratonsFlag_in);

}  // namespace schemi
