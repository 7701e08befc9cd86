// fv_star_launch.hpp -- host-side launch interface of star_kernel (fv_star.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "fv_launch.hpp"

namespace sg {

struct TurbParams;

struct StarLaunch {
    const DevParams *P;
    const TurbParams *TP;  // zero-filled for a laminar state
    const double *geo;
    const double *U;       // input state of the stage
    double *S;             // star buffer, (5 + na) x P.star_nft doubles (a, b)
    const double *sc;      // scalar block (SC_DONE)
    int avg;               // 0: first stage (store), 1: second (average)
    int dim, ns, gas_model;
    int slopes;            // 1-D mesh with a TVD limiter
    int nt, na;            // rho k, rho eps rows (0 or 2); rho a, rho b rows (0, 3, 4)
    long long n;
    hipStream_t stream;
};

// Returns 0 on success, -1 when the combination has no instantiation,
// otherwise a hipError_t value.
int launch_star(const StarLaunch &L);

}  // namespace sg
