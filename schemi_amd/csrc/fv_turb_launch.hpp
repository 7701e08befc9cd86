// fv_turb_launch.hpp -- host-side launch interface of the turbulent stage
// kernels (fv_turb.hpp), instantiated per flux precision by fv_turb_f64.hip
// and fv_turb_mixed.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_launch.hpp"

namespace sg {

struct TurbParams;

struct TurbStageLaunch {
    const DevParams *P;
    const TurbParams *TP;
    const double *geo;
    const double *Uin;
    double *Uout;
    const double *Un;
    const double *dt;
    unsigned long long *sigmax;
    int *err;
    int dim, ns, variant, gas_model;  // variant: SV_RK1 or SV_RK2
    int slopes;                       // 1-D mesh with a TVD limiter
    int na;                           // passive fields after rho eps (0, 3: a, 4: a and b)
    double *b2;                       // BHR: the second stage's b (n), else null
    long long n;
    hipStream_t stream;
};

// Returns 0 on success, -1 when the combination has no instantiation,
// otherwise a hipError_t value.
int launch_turb_stage_f64(const TurbStageLaunch &L);
int launch_turb_stage_mixed(const TurbStageLaunch &L);

// Explicit diffusion stage with k-epsilon turbulence (fv_diffusion.hip): W
// holds (2 ns + 9) n doubles, (2 ns + 13) n with the kEpsA model and
// (2 ns + 16) n with BHR; epsmax is the bit pattern of the largest epsilon of
// the state (one word, zero before each stage); b2 the BHR model's second
// stage b.
struct TurbDiffusion {
    const TurbParams *TP;
    unsigned long long *epsmax;
    int amodel;  // 0: no a field, 1: kEpsA, 2: BHR
    const double *b2;
    double *bk;  // kEpsA: the stage's algebraic b (n) is written here when not null
};
int launch_diffusion_turb(const DiffusionLaunch &L, const TurbDiffusion &T);

}  // namespace sg
