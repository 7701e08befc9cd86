// drive_shock_tube.cpp -- the reference's main loop (Schemi/main/Schemi.cpp:
// 767-940) around the GPU step core, without the reference's types: the 1-D
// two-component shock tube with the Courant time step, in drop-in mode (the
// state is uploaded once and downloaded after every step, as the reference's
// loop reads it for its time step and output).
//
// Usage: drive_shock_tube <nsteps> <key> <out>; key "SecondOrderRK_GPU" or
// "SecondOrder_GPU".  Prints the final time and the number of steps and
// writes the final conserved state (Ns + 5 rows of n doubles, reference cell
// order) to <out>, so the test harness can compare it with the oracle field
// by field.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpuStepCore.hpp"

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: drive_shock_tube <nsteps> <key> <out>\n");
        return 2;
    }
    const int nsteps = std::atoi(argv[1]);
    const std::string key = argv[2];
    const std::size_t N = 1000, ns = 2;
    schemi_mesh_desc mesh{};
    schemi_mix_desc mix{};
    schemi_scheme_desc scheme{};
    mesh.dim = 1;
    mesh.n[0] = N;
    mesh.n[1] = mesh.n[2] = 1;
    mesh.extent[0] = mesh.extent[1] = mesh.extent[2] = 1.0;
    mix.ns = 2;
    mix.M[0] = 29.0;
    mix.M[1] = 4.0;
    mix.Cv[0] = 1.0 / (1.4 - 1.0);
    mix.Cv[1] = 1.0 / (1.67 - 1.0);
    mix.R = 1.0;
    scheme.flux = SCHEMI_FLUX_HLLC;
    scheme.limiter = SCHEMI_LIM_MINMOD;
    scheme.precision = SCHEMI_PREC_FP64;
    scheme.courant = 0.4;
    scheme.min_time = 1e-15;
    scheme.const_dt = -1.0;
    for (int p = 0; p < 6; ++p) {
        for (std::size_t k = 0; k < ns; ++k)
            scheme.bc_c[k][p].type = SCHEMI_BC_FREE;
        scheme.bc_v[p].type = SCHEMI_BC_FREE;
        scheme.bc_p[p].type = SCHEMI_BC_FREE;
    }
    // the key selects the step solver (abstractStepSolver.cpp:77-118)
    if (key == "SecondOrderRK_GPU") {
        scheme.step_solver = SCHEMI_STEP_SECOND_ORDER_RK;
    } else if (key == "SecondOrder_GPU") {
        scheme.step_solver = SCHEMI_STEP_SECOND_ORDER;
    } else {
        std::fprintf(stderr, "Unknown or CPU-only step solver: %s\n", key.c_str());
        return static_cast<int>(SCHEMI_ERR_INIT);
    }
    try {
        schemi_gpu_host::stepCore core(mesh, mix, scheme);
        // initial fields as phaseInitialization (phaseInitialization.cpp:1827-1883):
        // c_k per zone, rho_k = c_k M_k, U = p / (R / Cv_mix), E = U + 0 + 0
        double *U = core.conserved();
        for (std::size_t i = 0; i < N; ++i) {
            const double x = (i + 0.5) / N;
            const bool left = x < 0.5;
            const double c[2] = {left ? 1.0 / 29.0 : 0.0, left ? 0.0 : 0.125 / 4.0};
            const double p = left ? 1.0 : 0.1;
            double c0 = 0, r0 = 0, cv = 0;
            for (std::size_t k = 0; k < ns; ++k) {
                c0 += c[k];
                U[k * N + i] = c[k] * mix.M[k];
            }
            for (std::size_t k = 0; k < ns; ++k)
                r0 += U[k * N + i];
            for (std::size_t k = 0; k < ns; ++k)
                cv += (c[k] / c0) * mix.Cv[k];
            U[ns * N + i] = r0;
            for (std::size_t a = 0; a < 3; ++a)
                U[(ns + 1 + a) * N + i] = 0.0;
            U[(ns + 4) * N + i] = (p / (mix.R / cv) + r0 * 0.0 * 0.5) + 0.0;
        }
        core.upload();
        double timestep = scheme.min_time, time = 0.0;
        for (int step = 0; step < nsteps; ++step) {
            core.step(timestep);
            core.download();
            time += timestep;
            timestep = core.nextTimestep();
        }
        std::printf("time %.17g\nsteps %d\n", time, nsteps);
        FILE *f = std::fopen(argv[3], "wb");
        if (!f)
            return 3;
        const std::size_t count = (ns + 5) * N;
        const bool ok = std::fwrite(core.conserved(), sizeof(double), count, f) == count;
        std::fclose(f);
        return ok ? 0 : 3;
    } catch (const schemi_gpu_host::gpu_error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return e.code();
    }
}
