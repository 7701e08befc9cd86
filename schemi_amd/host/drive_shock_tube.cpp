// drive_shock_tube.cpp -- the reference's main loop (Schemi/main/Schemi.cpp:
// 767-940) driving the GPU step through createStepSolver.
//
// Runs the 1-D two-component shock tube for a number of steps with the
// Courant time step and prints the final time, the number of steps and a
// checksum of every conserved field, one per line, so the test harness can
// compare them with the oracle.  Usage: drive_shock_tube <nsteps> [<key>],
// key "SecondOrderRK_GPU" (default) or "SecondOrder_GPU".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gpuStepSolverRK.hpp"

int main(int argc, char **argv)
{
    const int nsteps = argc > 1 ? std::atoi(argv[1]) : 10;
    const std::string key = argc > 2 ? argv[2] : "SecondOrderRK_GPU";
    const std::size_t N = 1000, ns = 2;
    schemi::stepSettings s;
    s.mesh.dim = 1;
    s.mesh.n[0] = N;
    s.mesh.n[1] = s.mesh.n[2] = 1;
    s.mesh.extent[0] = s.mesh.extent[1] = s.mesh.extent[2] = 1.0;
    s.mixture.ns = 2;
    s.mixture.M[0] = 29.0;
    s.mixture.M[1] = 4.0;
    s.mixture.Cv[0] = 1.0 / (1.4 - 1.0);
    s.mixture.Cv[1] = 1.0 / (1.67 - 1.0);
    s.mixture.R = 1.0;
    s.scheme.flux = SCHEMI_FLUX_HLLC;
    s.scheme.limiter = SCHEMI_LIM_MINMOD;
    s.scheme.precision = SCHEMI_PREC_FP64;
    s.scheme.courant = 0.4;
    s.scheme.min_time = 1e-15;
    s.scheme.const_dt = -1.0;
    for (int p = 0; p < 6; ++p) {
        for (std::size_t k = 0; k < ns; ++k)
            s.scheme.bc_c[k][p].type = SCHEMI_BC_FREE;
        s.scheme.bc_v[p].type = SCHEMI_BC_FREE;
        s.scheme.bc_p[p].type = SCHEMI_BC_FREE;
    }
    // initial fields as phaseInitialization (phaseInitialization.cpp:1827-1883)
    schemi::phaseFields ph(N, ns);
    for (std::size_t i = 0; i < N; ++i) {
        const double x = (i + 0.5) / N;
        const bool left = x < 0.5;
        ph.concentration[1][i] = left ? 1.0 / 29.0 : 0.0;
        ph.concentration[2][i] = left ? 0.0 : 0.125 / 4.0;
        ph.pressure[i] = left ? 1.0 : 0.1;
        double c0 = 0, r0 = 0;
        for (std::size_t k = 1; k <= ns; ++k) {
            c0 += ph.concentration[k][i];
            ph.density[k][i] = ph.concentration[k][i] * s.mixture.M[k - 1];
        }
        for (std::size_t k = 1; k <= ns; ++k)
            r0 += ph.density[k][i];
        ph.concentration[0][i] = c0;
        ph.density[0][i] = r0;
        double cv = 0;
        for (std::size_t k = 1; k <= ns; ++k)
            cv += (ph.concentration[k][i] / c0) * s.mixture.Cv[k - 1];
        ph.internalEnergy[i] = ph.pressure[i] / (s.mixture.R / cv);
        ph.totalEnergy[i] = (ph.internalEnergy[i] + r0 * 0.0 * 0.5) + 0.0;
    }
    double timestep = s.scheme.min_time, time = 0.0;
    try {
        auto solver = schemi::createStepSolver(key, ph, s, timestep);
        auto *gpu = static_cast<schemi::gpuStepSolverRK *>(solver.get());
        for (int step = 0; step < nsteps; ++step) {
            solver->calculateStep();
            time += timestep;
            timestep = gpu->nextTimestep();
        }
    } catch (const schemi::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return static_cast<int>(e.errCode());
    }
    std::printf("time %.17g\nsteps %d\n", time, nsteps);
    for (std::size_t k = 0; k <= ns; ++k) {
        double s_ = 0;
        for (std::size_t i = 0; i < N; ++i)
            s_ += ph.density[k][i];
        std::printf("density%zu %.17g\n", k, s_);
    }
    double se = 0;
    for (std::size_t i = 0; i < N; ++i)
        se += ph.totalEnergy[i];
    std::printf("totalEnergy %.17g\n", se);
    return 0;
}
