// gpuStepSolverRK.cpp -- see gpuStepSolverRK.hpp.
#include "gpuStepSolverRK.hpp"

namespace schemi {

namespace {

void check(int code, const char *where)
{
    if (code != SCHEMI_OK)
        throw exception(std::string(where) + ": " + schemi_gpu_strerror(code),
                        static_cast<errors>(code));
}

}  // namespace

phaseFields::phaseFields(std::size_t ncells, std::size_t ns)
    : concentration(ns + 1, std::valarray<double>(0.0, ncells)),
      density(ns + 1, std::valarray<double>(0.0, ncells)),
      pressure(0.0, ncells), internalEnergy(0.0, ncells), totalEnergy(0.0, ncells),
      temperature(0.0, ncells)
{
    for (auto &v : velocity)
        v.resize(ncells, 0.0);
    for (auto &m : momentum)
        m.resize(ncells, 0.0);
}

gpuStepSolverRK::gpuStepSolverRK(phaseFields &phase, const stepSettings &s, const double &timestep)
    : gasPhase(phase), dt(timestep)
{
    check(schemi_gpu_create(&s.mesh, &s.mixture, &s.scheme, &ctx), "schemi_gpu_create");
    ns = static_cast<std::size_t>(s.mixture.ns);
    n = gasPhase.pressure.size();
    cons.resize((ns + 5) * n);
    prim.resize((ns + 8) * n);
}

gpuStepSolverRK::~gpuStepSolverRK() noexcept
{
    schemi_gpu_destroy(ctx);
}

void gpuStepSolverRK::upload()
{
    // conserved vector rho_1..rho_Ns, rho, m, E
    for (std::size_t k = 0; k < ns; ++k)
        for (std::size_t i = 0; i < n; ++i)
            cons[k * n + i] = gasPhase.density[k + 1][i];
    for (std::size_t i = 0; i < n; ++i)
        cons[ns * n + i] = gasPhase.density[0][i];
    for (std::size_t a = 0; a < 3; ++a)
        for (std::size_t i = 0; i < n; ++i)
            cons[(ns + 1 + a) * n + i] = gasPhase.momentum[a][i];
    for (std::size_t i = 0; i < n; ++i)
        cons[(ns + 4) * n + i] = gasPhase.totalEnergy[i];
    check(schemi_gpu_set_state(ctx, cons.data(), static_cast<int64_t>(n),
                               SCHEMI_NUMBERING_REFERENCE),
          "schemi_gpu_set_state");
    uploaded = true;
}

void gpuStepSolverRK::download()
{
    check(schemi_gpu_get_state(ctx, cons.data(), prim.data(), SCHEMI_NUMBERING_REFERENCE),
          "schemi_gpu_get_state");
    for (std::size_t k = 0; k < ns; ++k)
        for (std::size_t i = 0; i < n; ++i) {
            gasPhase.density[k + 1][i] = cons[k * n + i];
            gasPhase.concentration[k + 1][i] = prim[k * n + i];
        }
    for (std::size_t i = 0; i < n; ++i) {
        gasPhase.density[0][i] = cons[ns * n + i];
        gasPhase.concentration[0][i] = prim[ns * n + i];
        gasPhase.totalEnergy[i] = cons[(ns + 4) * n + i];
        gasPhase.pressure[i] = prim[(ns + 4) * n + i];
        gasPhase.internalEnergy[i] = prim[(ns + 5) * n + i];
        gasPhase.temperature[i] = prim[(ns + 6) * n + i];
    }
    for (std::size_t a = 0; a < 3; ++a)
        for (std::size_t i = 0; i < n; ++i) {
            gasPhase.momentum[a][i] = cons[(ns + 1 + a) * n + i];
            gasPhase.velocity[a][i] = prim[(ns + 1 + a) * n + i];
        }
}

void gpuStepSolverRK::calculateStep()
{
    if (!uploaded)
        upload();
    check(schemi_gpu_step(ctx, dt, &dtNext), "calculateStep");
    download();
}

std::unique_ptr<abstractStepSolver> createStepSolver(const std::string &key, phaseFields &phase,
                                                     const stepSettings &settings,
                                                     const double &timestep)
{
    // the key selects the step solver (abstractStepSolver.cpp:83-84, 97-118)
    stepSettings s = settings;
    if (key == "SecondOrderRK_GPU")
        s.scheme.step_solver = SCHEMI_STEP_SECOND_ORDER_RK;
    else if (key == "SecondOrder_GPU")
        s.scheme.step_solver = SCHEMI_STEP_SECOND_ORDER;
    else
        throw exception("Unknown or CPU-only step solver: " + key, errors::initialisationError);
    return std::make_unique<gpuStepSolverRK>(phase, s, timestep);
}

}  // namespace schemi
