// gpuStepSolverRK.hpp -- C++ drop-in for the reference's SecondOrderRK step.
//
// Mirrors the reference's step-solver plugin interface
// (abstractStepSolver::calculateStep / createStepSolver,
//  Schemi/headers/abstractStepSolver.hpp:76-156,
//  Schemi/src/interfaces/abstractStepSolver.cpp:56-146) over the C ABI of
// include/schemi_gpu.h.  The solver holds a reference to the caller's cell
// fields (structure of arrays, reference cell order), like the reference's
// solver holds homogeneousPhase<cubicCell>&; calculateStep() advances them by
// one SSP-RK2 step on the GPU and leaves conserved and primitive fields
// consistent, so the caller's time loop (Courant step, output) is unchanged.
// Errors are thrown as schemi::exception with the reference's errors codes.
#pragma once

#include <array>
#include <exception>
#include <memory>
#include <string>
#include <valarray>
#include <vector>

#include "../../include/schemi_gpu.h"

namespace schemi {

// Schemi/enumerations/errorsEnum.hpp:15-30
enum class errors {
    noErrors, RiemannSolverError, negativeTemperatureError, tooBigOutputNumberError,
    systemError, boundaryConditionError, cubicEquationError, positivnessError, MPIError,
    fieldInitialisationError, meshGenerationError, initialisationError, NaNError
};
// Schemi/enumerations/flowSolverEnum.hpp:15-18
enum class flowSolver { HLL, HLLCF, HLLC, HLLCLM, KT, HLLC2p, Richtmyer };
// Schemi/enumerations/typeOfTVDLimiterEnum.hpp:15-32
enum class typeOfTVDLimiter {
    zero, minmod, vanLeer, linear, vanAlbada, HQUICK, vanLeer2, superbee, vanAlbada2,
    minmod2, Sweby, Koren, Koren2, UTCDFS, Schmidtmann
};
// Schemi/enumerations/boundaryConditionTypesEnum.hpp:15-34 (+ periodic)
enum class boundaryConditionType {
    blank, freeBoundary, slip, fixedValueCell, fixedValueSurface, innerSurface, calculated,
    calculatedTurbulentViscosity, calculatedTemperature, calculatedMassFraction,
    calculatedAverageMolarMass, calculatedNonidealityCorrectionPerDensity, calculatedCv,
    calculatedCvM, calculatedDensity, calculatedParallelBoundary, calculatedMolarFraction,
    periodic = 32
};

// Schemi/headers/exception.hpp:20-30
class exception : public std::exception {
    std::string msg;
    errors code;

public:
    exception(std::string m, errors c) : msg(std::move(m)), code(c) {}
    const char *what() const noexcept override { return msg.c_str(); }
    errors errCode() const noexcept { return code; }
};

// Cell fields of homogeneousPhase<cubicCell> on the hot path
// (Schemi/templates/bunchOfFields.hpp:28-45), one valarray per component.
struct phaseFields {
    std::vector<std::valarray<double>> concentration;  // c_0 (sum), c_1 .. c_Ns
    std::vector<std::valarray<double>> density;        // rho_0 (total), rho_1 .. rho_Ns
    std::array<std::valarray<double>, 3> velocity;
    std::array<std::valarray<double>, 3> momentum;
    std::valarray<double> pressure, internalEnergy, totalEnergy, temperature;

    phaseFields(std::size_t ncells, std::size_t ns);
};

// Deck settings the step reads (Schemi/main/Schemi.cpp:84-163).
struct stepSettings {
    schemi_mesh_desc mesh{};
    schemi_mix_desc mixture{};
    schemi_scheme_desc scheme{};
};

// abstractStepSolver (Schemi/headers/abstractStepSolver.hpp:76-156)
class abstractStepSolver {
public:
    virtual ~abstractStepSolver() noexcept = default;
    virtual void calculateStep() = 0;
};

// SecondOrderRK (SSP-RK2) or SecondOrder (MUSCL-Hancock) on the GPU, as
// settings.scheme.step_solver selects.  timestep is the caller's
// mesh::timestep() (mesh.hpp:151), read at every call.
class gpuStepSolverRK final : public abstractStepSolver {
public:
    gpuStepSolverRK(phaseFields &phase, const stepSettings &settings, const double &timestep);
    ~gpuStepSolverRK() noexcept override;
    gpuStepSolverRK(const gpuStepSolverRK &) = delete;
    gpuStepSolverRK &operator=(const gpuStepSolverRK &) = delete;

    void calculateStep() override;
    // Courant step computed on the device from the state after the last step
    // (Schemi.cpp:807-833).
    double nextTimestep() const noexcept { return dtNext; }
    // Marks the host fields as changed by the caller (uploaded before the next
    // step).
    void invalidate() noexcept { uploaded = false; }

private:
    phaseFields &gasPhase;
    const double &dt;
    schemi_gpu_ctx *ctx = nullptr;
    std::size_t n, ns;
    bool uploaded = false;
    double dtNext = 0;
    std::vector<double> cons, prim;

    void upload();
    void download();
};

// createStepSolver (Schemi/src/interfaces/abstractStepSolver.cpp:56-146) with
// the GPU keys "SecondOrderRK_GPU" and "SecondOrder_GPU" (they set
// settings.scheme.step_solver).  Other keys throw initialisationError.
std::unique_ptr<abstractStepSolver> createStepSolver(const std::string &key, phaseFields &phase,
                                                     const stepSettings &settings,
                                                     const double &timestep);

}  // namespace schemi
