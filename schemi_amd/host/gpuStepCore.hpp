// gpuStepCore.hpp -- host-side core of the GPU step solver, independent of
// the reference's types: one C-ABI context (include/schemi_gpu.h), a host
// staging copy of the conserved state for the upload, and of the conserved
// state and the primitives for the download.
//
// gpuStepSolverRK (the reference's step-solver plugin, gpuStepSolverRK.hpp)
// maps homogeneousPhase<cubicCell> onto it; drive_shock_tube runs it under a
// copy of the reference's time loop.  Layouts (structure of arrays, cells in
// the reference's order i + Nx (j + Ny k)):
//   conserved:  rho_1 .. rho_Ns, rho, m_x, m_y, m_z, E          (Ns + 5 rows)
//               with turbulence then rho k, rho epsilon; with the kEpsA and
//               BHR models rho a_x, rho a_y, rho a_z; with BHR rho b
//   primitives: c_1 .. c_Ns, c, v_x, v_y, v_z, p, U, T, sound    (Ns + 8 rows)
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/schemi_gpu.h"

namespace schemi_gpu_host {

// A failed C-ABI call: its code is the reference's errors value.
class gpu_error : public std::runtime_error {
    int code_;

public:
    gpu_error(const std::string &where, int code);
    int code() const noexcept { return code_; }
};

class stepCore {
public:
    stepCore(const schemi_mesh_desc &mesh, const schemi_mix_desc &mix,
             const schemi_scheme_desc &scheme);
    ~stepCore() noexcept;
    stepCore(const stepCore &) = delete;
    stepCore &operator=(const stepCore &) = delete;

    std::size_t cells() const noexcept { return n; }
    std::size_t species() const noexcept { return ns; }
    std::size_t fields() const noexcept { return nf; }
    // host staging of the conserved state: fill it, then upload()
    double *conserved() noexcept { return cons.data(); }
    const double *conserved() const noexcept { return cons.data(); }
    const double *primitives() const noexcept { return prim.data(); }

    void upload();
    // one step of size dt on the device-resident state
    void step(double dt);
    // the device state into conserved() and primitives()
    void download();
    // Courant step of the state after the last step (Schemi.cpp:807-833) and
    // the source step of its diffusion stage (veryBig when none)
    // the k and epsilon the last diffusion stage started from and, with the
    // kEpsA model, its algebraic b (2 or 3 rows of cells(), turbulence only)
    std::vector<double> turbulenceStage() const;
    double nextTimestep() const noexcept { return dtNext; }
    double sourceTimestep() const noexcept { return dtSource; }

private:
    schemi_gpu_ctx *ctx = nullptr;
    std::size_t n = 0, ns = 0, nf = 0, nstage = 0;
    double dtNext = 0.0, dtSource = 0.0;
    std::vector<double> cons, prim;
};

}  // namespace schemi_gpu_host
