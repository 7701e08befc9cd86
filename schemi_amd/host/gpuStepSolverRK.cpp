// gpuStepSolverRK.cpp -- see gpuStepSolverRK.hpp.
#include "gpuStepSolverRK.hpp"

#include <cstring>
#include <typeinfo>

#include "HLLCSolver.hpp"
#include "HLLSolver.hpp"
#include "exception.hpp"
#include "kEpsModels.hpp"
#include "linearLimiter.hpp"
#include "mesh.hpp"
#include "minmodLimiter.hpp"
#include "mixtureIdeal.hpp"
#include "vanAlbadaLimiter.hpp"
#include "vanLeerLimiter.hpp"
#include "zeroLimiter.hpp"

namespace schemi {

namespace {

[[noreturn]] void unsupported(const std::string &what)
{
    throw exception("GPU step solver: " + what + " is not supported on the GPU.",
                    errors::initialisationError);
}

// The iteration limit of a matrix solver (abstractMatrixSolver.hpp:27, a
// protected member), read through a pointer to member formed in a derived
// class.
struct solverIterations : abstractMatrixSolver {
    static std::size_t of(const abstractMatrixSolver &m)
    {
        return m.*(&solverIterations::maxIterationNumber);
    }
};

// The constant transport model's coefficients (abstractTransportModel.hpp:
// 24-30, protected accessors), read the same way.
struct transportConstants : abstractTransportModel {
    static scalar mu(const abstractTransportModel &t) { return (t.*(&transportConstants::muConst))(); }
    static scalar D(const abstractTransportModel &t) { return (t.*(&transportConstants::DConst))(); }
    static scalar kappa(const abstractTransportModel &t)
    {
        return (t.*(&transportConstants::kappaConst))();
    }
};

// A turbulence model's constants as read from set/turbulentParameters.txt
// (abstractTurbulenceModel.hpp:55, a protected accessor; shearModel.cpp:
// 56-58, kEpsAModel.cpp:73-80, BHRModel.cpp:73-84), read the same way.
struct turbulenceConstants : abstractTurbulenceModel {
    static const std::valarray<scalar> &of(const abstractTurbulenceModel &t)
    {
        return (t.*(&turbulenceConstants::modelParameters))();
    }
};

// First surface of each patch in the reference's surface numbering
// (mesh.cpp:1301-1790: tail, inner, point, bottom, right, left, top), in the
// C ABI's patch order 0 tail, 1 point, 2 bottom, 3 right, 4 left, 5 top, and
// the number of surfaces of each.
struct patchRanges {
    std::size_t begin[6], count[6];
    explicit patchRanges(const mesh &m)
    {
        std::size_t at = 0;
        begin[0] = at;
        count[0] = m.tailNumber();
        at += m.tailNumber() + m.innerNumber();
        const std::size_t rest[5] = {m.pointNumber(), m.bottomNumber(), m.rightNumber(),
                                     m.leftNumber(), m.topNumber()};
        for (int p = 1; p < 6; ++p) {
            begin[p] = at;
            count[p] = rest[p - 1];
            at += rest[p - 1];
        }
    }
};

// One boundary condition per patch: the reference allows sub-patches with
// their own conditions (field.hpp:159-300); the GPU takes one per patch.
template <class T>
std::pair<boundaryConditionType, T> patchCondition(
    const std::vector<std::pair<boundaryConditionType, T>> &info, const patchRanges &pr, int p,
    const char *name)
{
    const auto &first = info[pr.begin[p]];
    for (std::size_t j = pr.begin[p]; j < pr.begin[p] + pr.count[p]; ++j) {
        const auto &c = info[j];
        bool same = c.first == first.first;
        if constexpr (std::is_same_v<T, vector>)
            same = same && c.second() == first.second();
        else
            same = same && c.second == first.second;
        if (!same)
            unsupported(std::string("a patch with several boundary conditions (") + name + ")");
    }
    return first;
}

int bcCode(boundaryConditionType t, const char *name)
{
    switch (t) {
    case boundaryConditionType::freeBoundary:
        return SCHEMI_BC_FREE;
    case boundaryConditionType::slip:
        return SCHEMI_BC_SLIP;
    case boundaryConditionType::fixedValueCell:
        return SCHEMI_BC_FIXED_CELL;
    case boundaryConditionType::fixedValueSurface:
        return SCHEMI_BC_FIXED_SURFACE;
    default:
        unsupported(std::string("this boundary condition type (") + name + ")");
    }
}

struct descriptors {
    schemi_mesh_desc mesh{};
    schemi_mix_desc mix{};
    schemi_scheme_desc scheme{};
};

descriptors describe(int stepSolver, int precision, const homogeneousPhase<cubicCell> &gasPhase,
                     const abstractLimiter &limiter, const abstractFlowSolver &fsolver,
                     bool gravitation, const vector &g, bool diffusion,
                     const abstractMatrixSolver &msolver,
                     const std::pair<scalar, scalar> &timestepCoeffs, enthalpyFlow enthalpyFlag,
                     bool linear, timestep sourceTimeFlag, bool molMassDiffusion,
                     const chemicalKinetics::abstractChemicalKinetics &chemKin)
{
    descriptors d;
    const mesh &m = gasPhase.pressure.meshRef();
    // mesh (mesh.hpp:143-149)
    d.mesh.dim = m.taskDimension() == dimensions::task1D   ? 1
                 : m.taskDimension() == dimensions::task2D ? 2
                                                           : 3;
    for (int a = 0; a < 3; ++a) {
        d.mesh.n[a] = static_cast<int64_t>(m.nCells()[a]);
        d.mesh.extent[a] = m.delta()()[a];
        d.mesh.origin[a] = m.zero()()[a];
    }
    // mixture (abstractMixtureThermodynamics.hpp:40-46)
    const abstractMixtureThermodynamics &mix = *gasPhase.phaseThermodynamics;
    if (typeid(mix) != typeid(mixtureIdeal))
        unsupported("this mixture model (the drop-in maps the ideal mixture)");
    const std::size_t ns = mix.Mv().size();
    if (ns < 1 || ns > 3)
        unsupported("a mixture of more than 3 species");
    d.mix.ns = static_cast<int32_t>(ns);
    for (std::size_t k = 0; k < ns; ++k) {
        d.mix.M[k] = mix.Mv()[k];
        d.mix.Cv[k] = mix.Cvv()[k];
    }
    d.mix.R = mix.Rv();
    d.mix.gas_model = SCHEMI_GAS_IDEAL;
    // flux and limiter: their dynamic types (abstractFlowSolver.cpp,
    // abstractLimiter.cpp factories)
    if (typeid(fsolver) == typeid(HLLCSolver))
        d.scheme.flux = SCHEMI_FLUX_HLLC;
    else if (typeid(fsolver) == typeid(HLLSolver))
        d.scheme.flux = SCHEMI_FLUX_HLL;
    else
        unsupported("this flux scheme");
    if (typeid(limiter) == typeid(minmodLimiter))
        d.scheme.limiter = SCHEMI_LIM_MINMOD;
    else if (typeid(limiter) == typeid(vanLeerLimiter))
        d.scheme.limiter = SCHEMI_LIM_VANLEER;
    else if (typeid(limiter) == typeid(vanAlbadaLimiter))
        d.scheme.limiter = SCHEMI_LIM_VANALBADA;
    else if (typeid(limiter) == typeid(linearLimiter))
        d.scheme.limiter = SCHEMI_LIM_LINEAR;
    else if (typeid(limiter) == typeid(zeroLimiter))
        d.scheme.limiter = SCHEMI_LIM_ZERO;
    else
        unsupported("this limiter");
    d.scheme.precision = precision;
    d.scheme.step_solver = stepSolver;
    d.scheme.gravity = gravitation ? 1 : 0;
    for (int a = 0; a < 3; ++a)
        d.scheme.g[a] = g()[a];
    // the main loop forms the time step itself (Schemi.cpp:807-868); the
    // device's own Courant step is reported by nextTimestep()
    d.scheme.courant = 0.5;
    d.scheme.min_time = 0.0;
    d.scheme.const_dt = -1.0;
    // boundary conditions per field and patch
    const patchRanges pr(m);
    for (int p = 0; p < 6; ++p) {
        if (pr.count[p] == 0) {
            for (std::size_t k = 0; k < ns; ++k)
                d.scheme.bc_c[k][p].type = SCHEMI_BC_FREE;
            d.scheme.bc_v[p].type = SCHEMI_BC_FREE;
            d.scheme.bc_p[p].type = SCHEMI_BC_FREE;
            continue;
        }
        for (std::size_t k = 0; k < ns; ++k) {
            const auto c = patchCondition(gasPhase.concentration.v[k + 1].boundCond(), pr, p,
                                          "concentration");
            d.scheme.bc_c[k][p].type = bcCode(c.first, "concentration");
            d.scheme.bc_c[k][p].value[0] = c.second;
        }
        const auto v = patchCondition(gasPhase.velocity.boundCond(), pr, p, "velocity");
        d.scheme.bc_v[p].type = bcCode(v.first, "velocity");
        for (int a = 0; a < 3; ++a)
            d.scheme.bc_v[p].value[a] = v.second()[a];
        const auto pc = patchCondition(gasPhase.pressure.boundCond(), pr, p, "pressure");
        d.scheme.bc_p[p].type = bcCode(pc.first, "pressure");
        d.scheme.bc_p[p].value[0] = pc.second;
    }
    // k-epsilon turbulence (abstractTurbulenceModel.cpp:235-320): the model,
    // its constants and the boundary conditions of k, epsilon, a and b
    const abstractTurbulenceModel &tm = *gasPhase.turbulence;
    if (tm.turbulence()) {
        const auto *ke = dynamic_cast<const kEpsModels *>(&tm);
        int model = -1;
        switch (tm.model()) {
        case turbulenceModel::zeroModel:
            model = SCHEMI_TURB_ZERO;
            break;
        case turbulenceModel::shearModel:
            model = SCHEMI_TURB_SHEAR;
            break;
        case turbulenceModel::kEpsAModel:
            model = SCHEMI_TURB_KEPSA;
            break;
        case turbulenceModel::BHRModel:
            model = SCHEMI_TURB_BHR;
            break;
        default:
            break;
        }
        if (model < 0 || ke == nullptr)
            unsupported("this turbulence model");
        if (tm.isInitialisationModelUsed())
            unsupported("the Goncharov initialisation model");
        d.scheme.turbulence = 1;
        d.scheme.turbulence_model = model;
        d.scheme.turb_Cmu = tm.Cmu();
        d.scheme.turb_sigma_sc = tm.sigmaSc();
        d.scheme.turb_sigma_t = tm.sigmaT();
        d.scheme.turb_sigma_e = tm.sigmaE();
        d.scheme.turb_sigma_k = tm.sigmaK();
        d.scheme.turb_sigma_eps = tm.sigmaEps();
        d.scheme.turb_CMS_R = ke->CMS_R();
        d.scheme.turb_CMS_D = ke->CMS_D();
        d.scheme.turb_min_k = tm.mink();
        d.scheme.turb_min_eps = tm.minepsilon();
        const std::valarray<scalar> &C = turbulenceConstants::of(tm);
        if (model == SCHEMI_TURB_SHEAR)
            for (std::size_t j = 0; j < 3 && j < C.size(); ++j)
                d.scheme.turb_C[j] = C[j];
        if (model == SCHEMI_TURB_KEPSA) {
            d.scheme.turb_sigma_a = tm.sigmaa();
            for (std::size_t j = 0; j < 8 && j < C.size(); ++j)
                d.scheme.turb_kepsa_C[j] = C[j];
        }
        if (model == SCHEMI_TURB_BHR) {
            d.scheme.turb_sigma_a = tm.sigmaa();
            d.scheme.turb_sigma_b = tm.sigmab();
            d.scheme.turb_min_b = tm.minb();
            for (std::size_t j = 0; j < 12 && j < C.size(); ++j)
                d.scheme.turb_bhr_C[j] = C[j];
        }
        for (int p = 0; p < 6; ++p) {
            if (pr.count[p] == 0) {
                d.scheme.bc_k[p].type = d.scheme.bc_eps[p].type = SCHEMI_BC_FREE;
                d.scheme.bc_a[p].type = d.scheme.bc_b[p].type = SCHEMI_BC_FREE;
                continue;
            }
            const auto k = patchCondition(gasPhase.kTurb.boundCond(), pr, p, "k");
            d.scheme.bc_k[p].type = bcCode(k.first, "k");
            d.scheme.bc_k[p].value[0] = k.second;
            const auto e = patchCondition(gasPhase.epsTurb.boundCond(), pr, p, "epsilon");
            d.scheme.bc_eps[p].type = bcCode(e.first, "epsilon");
            d.scheme.bc_eps[p].value[0] = e.second;
            if (tm.aField()) {
                const auto a = patchCondition(gasPhase.aTurb.boundCond(), pr, p, "a");
                d.scheme.bc_a[p].type = bcCode(a.first, "a");
                for (int c = 0; c < 3; ++c)
                    d.scheme.bc_a[p].value[c] = a.second()[c];
            }
            if (tm.bField()) {
                const auto b = patchCondition(gasPhase.bTurb.boundCond(), pr, p, "b");
                d.scheme.bc_b[p].type = bcCode(b.first, "b");
                d.scheme.bc_b[p].value[0] = b.second;
            }
        }
    }
    if (chemKin.chemicalReaction)
        unsupported("chemical kinetics");
    // diffusion stage (Diffusion.cpp:20-1061)
    d.scheme.diffusion = diffusion ? 1 : 0;
    if (diffusion) {
        const abstractTransportModel &tm = *gasPhase.transportModel;
        if (typeid(tm) != typeid(abstractTransportModel))
            unsupported("this transport model (the constant one is)");
        d.scheme.transport_mu = transportConstants::mu(tm);
        d.scheme.transport_D = transportConstants::D(tm);
        d.scheme.transport_kappa = transportConstants::kappa(tm);
        // #Reconstruction (Schemi.cpp:152-171): linear or star
        d.scheme.reconstruction =
            linear ? SCHEMI_RECONSTRUCTION_LINEAR : SCHEMI_RECONSTRUCTION_STAR;
        switch (msolver.solverType) {
        case matrixSolver::explicitSolver:
            d.scheme.diffusion_solver = SCHEMI_SLE_EXPLICIT;
            break;
        case matrixSolver::Jacobi:
            d.scheme.diffusion_solver = SCHEMI_SLE_JACOBI;
            break;
        case matrixSolver::GaussSeidel:
            d.scheme.diffusion_solver = SCHEMI_SLE_GAUSS_SEIDEL;
            break;
        default:
            unsupported("this matrix solver");
        }
        d.scheme.diffusion_iterations = static_cast<int32_t>(solverIterations::of(msolver));
        d.scheme.diffusion_molar_mass = molMassDiffusion ? 1 : 0;
        if (enthalpyFlag == enthalpyFlow::explicitSolve)
            d.scheme.diffusion_enthalpy_flow = 1;
        else if (enthalpyFlag != enthalpyFlow::noSolve)
            unsupported("the implicit enthalpy flow");
        if (sourceTimeFlag == timestep::CourantAndSourceTimeStep)
            d.scheme.timestep_source = 1;
        else if (sourceTimeFlag != timestep::CourantTimeStep)
            unsupported("the Courant-Source-Diffusion time step");
        d.scheme.source_coeff = timestepCoeffs.first;
    }
    return d;
}

}  // namespace

gpuStepSolverRK::gpuStepSolverRK(
    int stepSolver, int precision, homogeneousPhase<cubicCell> &gasPhase_in,
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
    const bool &nonLinearityIteratonsFlag_in)
    : abstractStepSolver(gasPhase_in, limiter_in, fsolver_in, gravitationFlag_in, g_in,
                         boundaryConditionValueCalc_in, timeForTVD_in, timeForHancock_in,
                         timeForFlowCalculation_in, timeForTimeIntegration_in, parallelism_in,
                         diffusionFlag_in, msolver_in, msolverEnthFl_in, timestepCoeffs_in,
                         timeForDiffusion_in, commonConditions_in, enthalpyFlowFlag_in,
                         linearFlag_in, minimalLengthScale_in, sourceTimeFlag_in,
                         molMassDiffusionFlag_in, chemKin_in, nonLinearityIteratonsFlag_in)
{
    const descriptors d =
        describe(stepSolver, precision, gasPhase_in, limiter_in, fsolver_in, gravitationFlag_in,
                 g_in, diffusionFlag_in, msolver_in, timestepCoeffs_in, enthalpyFlowFlag_in,
                 linearFlag_in, sourceTimeFlag_in, molMassDiffusionFlag_in, chemKin_in);
    try {
        core = std::make_unique<schemi_gpu_host::stepCore>(d.mesh, d.mix, d.scheme);
    } catch (const schemi_gpu_host::gpu_error &e) {
        throw exception(e.what(), static_cast<errors>(e.code()));
    }
}

gpuStepSolverRK::~gpuStepSolverRK() noexcept
{
}

double gpuStepSolverRK::nextTimestep() const noexcept
{
    return core->nextTimestep();
}

void gpuStepSolverRK::calculateStep()
{
    const std::size_t n = core->cells(), ns = core->species();
    double *cons = core->conserved();
    // conserved vector rho_1..rho_Ns, rho, m, E (bunchOfFields.hpp:28-45)
    for (std::size_t k = 0; k < ns; ++k)
        std::memcpy(cons + k * n, &gasPhase.density[k + 1].cval()[0], n * sizeof(double));
    std::memcpy(cons + ns * n, &gasPhase.density[0].cval()[0], n * sizeof(double));
    for (std::size_t i = 0; i < n; ++i) {
        const auto &m = gasPhase.momentum.cval()[i]();
        for (std::size_t a = 0; a < 3; ++a)
            cons[(ns + 1 + a) * n + i] = m[a];
    }
    std::memcpy(cons + (ns + 4) * n, &gasPhase.totalEnergy.cval()[0], n * sizeof(double));
    // turbulent conserved fields after E (bunchOfFields.hpp:40-43)
    const abstractTurbulenceModel &tm = *gasPhase.turbulence;
    if (tm.turbulence()) {
        std::memcpy(cons + (ns + 5) * n, &gasPhase.rhokTurb.cval()[0], n * sizeof(double));
        std::memcpy(cons + (ns + 6) * n, &gasPhase.rhoepsTurb.cval()[0], n * sizeof(double));
        if (tm.aField())
            for (std::size_t i = 0; i < n; ++i)
                for (std::size_t a = 0; a < 3; ++a)
                    cons[(ns + 7 + a) * n + i] = gasPhase.rhoaTurb.cval()[i]()[a];
        if (tm.bField())
            std::memcpy(cons + (ns + 10) * n, &gasPhase.rhobTurb.cval()[0], n * sizeof(double));
    }
    try {
        core->upload();
        core->step(gasPhase.pressure.meshRef().timestep());
        core->download();
    } catch (const schemi_gpu_host::gpu_error &e) {
        throw exception(e.what(), static_cast<errors>(e.code()));
    }
    // the fields the reference's stage leaves consistent
    // (Advection2dOrder.cpp:578-608)
    const double *U = core->conserved();
    const double *W = core->primitives();
    for (std::size_t k = 0; k < ns; ++k) {
        std::memcpy(&gasPhase.density[k + 1].val()[0], U + k * n, n * sizeof(double));
        std::memcpy(&gasPhase.concentration.v[k + 1].val()[0], W + k * n, n * sizeof(double));
    }
    std::memcpy(&gasPhase.density[0].val()[0], U + ns * n, n * sizeof(double));
    std::memcpy(&gasPhase.concentration.v[0].val()[0], W + ns * n, n * sizeof(double));
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t a = 0; a < 3; ++a) {
            gasPhase.momentum.val()[i].wr()[a] = U[(ns + 1 + a) * n + i];
            gasPhase.velocity.val()[i].wr()[a] = W[(ns + 1 + a) * n + i];
        }
    std::memcpy(&gasPhase.totalEnergy.val()[0], U + (ns + 4) * n, n * sizeof(double));
    std::memcpy(&gasPhase.pressure.val()[0], W + (ns + 4) * n, n * sizeof(double));
    std::memcpy(&gasPhase.internalEnergy.val()[0], W + (ns + 5) * n, n * sizeof(double));
    std::memcpy(&gasPhase.temperature.val()[0], W + (ns + 6) * n, n * sizeof(double));
    // the turbulent fields and their densities (Advection2dOrder.cpp:610-627,
    // Diffusion.cpp:999-1020): k = rho k / rho and so on
    if (tm.turbulence()) {
        const double *rho = U + ns * n;
        auto scatter = [&](std::size_t row, auto &rf, auto &f) {
            std::memcpy(&rf.val()[0], U + row * n, n * sizeof(double));
            for (std::size_t i = 0; i < n; ++i)
                f.val()[i] = U[row * n + i] / rho[i];
        };
        scatter(ns + 5, gasPhase.rhokTurb, gasPhase.kTurb);
        scatter(ns + 6, gasPhase.rhoepsTurb, gasPhase.epsTurb);
        if (tm.aField())
            for (std::size_t i = 0; i < n; ++i)
                for (std::size_t a = 0; a < 3; ++a) {
                    gasPhase.rhoaTurb.val()[i].wr()[a] = U[(ns + 7 + a) * n + i];
                    gasPhase.aTurb.val()[i].wr()[a] = U[(ns + 7 + a) * n + i] / rho[i];
                }
        if (tm.bField())
            scatter(ns + 10, gasPhase.rhobTurb, gasPhase.bTurb);
    }
    // the diffusion stage's turbulent transport coefficients, formed from the
    // k and epsilon it started from (Diffusion.cpp:129-135), and the kEpsA
    // model's algebraic b of that stage (Diffusion.cpp:525-527, 1006)
    if (tm.turbulence() && diffusionFlag) {
        std::vector<double> st;
        try {
            st = core->turbulenceStage();
        } catch (const schemi_gpu_host::gpu_error &e) {
            throw exception(e.what(), static_cast<errors>(e.code()));
        }
        auto kStage = gasPhase.kTurb, epsStage = gasPhase.epsTurb;
        std::memcpy(&kStage.val()[0], st.data(), n * sizeof(double));
        std::memcpy(&epsStage.val()[0], st.data() + n, n * sizeof(double));
        gasPhase.calculateCoefficients(kStage, epsStage, tm);
        if (tm.aField() && !tm.bField())
            std::memcpy(&gasPhase.bTurb.val()[0], st.data() + 2 * n, n * sizeof(double));
    }
    gasPhase.HelmholtzEnergy.val() =
        gasPhase.phaseThermodynamics->Fv(gasPhase.concentration.p, gasPhase.temperature.cval());
    gasPhase.entropy.val() =
        gasPhase.phaseThermodynamics->Sv(gasPhase.concentration.p, gasPhase.temperature.cval());
    // the diffusion stage's source step for the main loop's Courant-Source
    // time step (Diffusion.cpp:696-700, Schemi.cpp:828-833)
    // (the fields' mesh is the reference's one mesh object, created by main)
    if (diffusionFlag)
        const_cast<mesh &>(gasPhase.pressure.meshRef()).timestepSourceRef() =
            core->sourceTimestep();
}

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
    const bool &nonLinearityIteratonsFlag_in)
{
    // the key selects the step solver as in abstractStepSolver.cpp:77-118
    int stepSolver = -1, precision = SCHEMI_PREC_FP64;
    std::string base = key;
    const std::string mixed = "_MIXED";
    if (base.size() > mixed.size() &&
        base.compare(base.size() - mixed.size(), mixed.size(), mixed) == 0) {
        precision = SCHEMI_PREC_MIXED;
        base.resize(base.size() - mixed.size());
    }
    if (base == "SecondOrderRK_GPU")
        stepSolver = SCHEMI_STEP_SECOND_ORDER_RK;
    else if (base == "SecondOrder_GPU")
        stepSolver = SCHEMI_STEP_SECOND_ORDER;
    else
        throw exception("Unknown or CPU-only step solver: " + key, errors::initialisationError);
    return std::make_unique<gpuStepSolverRK>(
        stepSolver, precision, gasPhase_in, limiter_in, fsolver_in, gravitationFlag_in, g_in,
        boundaryConditionValueCalc_in, timeForTVD_in, timeForHancock_in,
        timeForFlowCalculation_in, timeForTimeIntegration_in, parallelism_in, diffusionFlag_in,
        msolver_in, msolverEnthFl_in, timestepCoeffs_in, timeForDiffusion_in,
        commonConditions_in, enthalpyFlowFlag_in, linearFlag_in, minimalLengthScale_in,
        sourceTimeFlag_in, molMassDiffusionFlag_in, chemKin_in, nonLinearityIteratonsFlag_in);
}

}  // namespace schemi
