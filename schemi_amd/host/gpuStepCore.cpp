// gpuStepCore.cpp -- see gpuStepCore.hpp.
#include "gpuStepCore.hpp"

namespace schemi_gpu_host {

namespace {

void check(int code, const char *where)
{
    if (code != SCHEMI_OK)
        throw gpu_error(where, code);
}

}  // namespace

gpu_error::gpu_error(const std::string &where, int code)
    : std::runtime_error(where + ": " + schemi_gpu_strerror(code)), code_(code)
{
}

stepCore::stepCore(const schemi_mesh_desc &mesh, const schemi_mix_desc &mix,
                   const schemi_scheme_desc &scheme)
{
    check(schemi_gpu_create(&mesh, &mix, &scheme, &ctx), "schemi_gpu_create");
    ns = static_cast<std::size_t>(mix.ns);
    n = static_cast<std::size_t>(mesh.n[0] * mesh.n[1] * mesh.n[2]);
    nf = ns + 5;
    if (scheme.turbulence) {
        nf += 2;
        if (scheme.turbulence_model == SCHEMI_TURB_KEPSA)
            nf += 3;
        else if (scheme.turbulence_model == SCHEMI_TURB_BHR)
            nf += 4;
    }
    cons.assign(nf * n, 0.0);
    if (scheme.turbulence && scheme.diffusion)
        nstage = scheme.turbulence_model == SCHEMI_TURB_KEPSA ? 3 : 2;
    prim.assign((ns + 8) * n, 0.0);
}

stepCore::~stepCore() noexcept
{
    schemi_gpu_destroy(ctx);
}

void stepCore::upload()
{
    check(schemi_gpu_set_state(ctx, cons.data(), static_cast<int64_t>(n),
                               SCHEMI_NUMBERING_REFERENCE),
          "schemi_gpu_set_state");
}

void stepCore::step(double dt)
{
    check(schemi_gpu_step(ctx, dt, &dtNext), "calculateStep");
    check(schemi_gpu_source_timestep(ctx, &dtSource), "schemi_gpu_source_timestep");
}

void stepCore::download()
{
    check(schemi_gpu_get_state(ctx, cons.data(), prim.data(), SCHEMI_NUMBERING_REFERENCE),
          "schemi_gpu_get_state");
}

std::vector<double> stepCore::turbulenceStage() const
{
    std::vector<double> out(nstage * n);
    if (nstage)
        check(schemi_gpu_get_turbulence_stage(ctx, out.data(), SCHEMI_NUMBERING_REFERENCE),
              "schemi_gpu_get_turbulence_stage");
    return out;
}

}  // namespace schemi_gpu_host
