// fv_launch.hpp -- host-side launch interface of the stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_device.hpp"

namespace sg {

struct StageLaunch {
    const DevParams *P;
    const double *geo;
    const double *Uin;
    double *Uout;
    const double *Un;
    const double *dt;
    unsigned long long *sigmax;
    int *err;
    int dim, ns, lin, stage2;
    int nx, ny, nz;
    int kchunk;
    hipStream_t stream;
};

// Returns 0 on success, -1 when the species count has no instantiation,
// otherwise a hipError_t value.
int launch_stage_f64(const StageLaunch &L);
int launch_stage_mixed(const StageLaunch &L);

}  // namespace sg
