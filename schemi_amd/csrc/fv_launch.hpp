// fv_launch.hpp -- host-side launch interface of the stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_device.hpp"

namespace sg {

// Tile of columns per workgroup in 2-D/3-D (x fastest) and in 1-D, for the
// fused slope kernel (SG_TILE_*) and the first-order kernel (SG_TILE0_*).
#ifndef SG_TILE_X
#define SG_TILE_X 32
#endif
#ifndef SG_TILE_Y
#define SG_TILE_Y 8
#endif
#ifndef SG_TILE0_X
#define SG_TILE0_X 64
#endif
#ifndef SG_TILE0_Y
#define SG_TILE0_Y 8
#endif
// Rows (wavefronts) per workgroup of the lean mixed kernel (64 columns).
#ifndef SG_LEAN_TY
#define SG_LEAN_TY 8
#endif
template <int DIM, bool FIRST_ORDER = false> struct TileShape {
    static constexpr int TX = (DIM == 1) ? 256 : (FIRST_ORDER ? SG_TILE0_X : SG_TILE_X);
    static constexpr int TY = (DIM == 1) ? 1 : (FIRST_ORDER ? SG_TILE0_Y : SG_TILE_Y);
};

// Kernel family of a context.
enum : int {
    MODE_FIRST_ORDER = 0,  // slopes vanish identically: stage0_kernel
    MODE_SLOPES = 1,       // fused slopes (TVD in 1-D, linear on separable meshes)
    MODE_GENERAL = 2,      // linear limiter with transverse face offsets
    MODE_LEAN = 3,         // slopes vanish, mixed path, ideal mixture, free/slip/periodic
                           // walls, whole 64 x SG_LEAN_TY tiles: stage_lean_kernel
};

struct StageLaunch {
    const DevParams *P;
    const double *geo;
    const double *Uin;
    double *Uout;
    const double *Un;
    const double *dt;
    unsigned long long *sigmax;
    int *err;
    int dim, ns, lin, variant, mode, gas_model;  // variant: SV_*
    int nx, ny, nz;
    int kchunk;
    void *W, *Lg;  // MODE_GENERAL scratch: primitives and gradients
    void *Hs;      // MODE_GENERAL with MUSCL-Hancock: predicted face sides
    hipStream_t stream;
};

// Returns 0 on success, -1 when the combination has no instantiation,
// otherwise a hipError_t value.
int launch_stage_f64(const StageLaunch &L);
int launch_stage_mixed(const StageLaunch &L);
int launch_stage_lean(const StageLaunch &L);  // MODE_LEAN (fv_lean.hip)

// Adjoint of one first-order-exact stage linearised at U (FP64):
// out = a (dL/dU)^T lam + b add (add may be null).
struct AdjointLaunch {
    const DevParams *P;
    const double *geo;
    const double *U;
    const double *lam;
    double *out;
    const double *add;
    double dt, a, b;
    int dim, ns, gas_model;
    int slopes, lin;  // 1-D stage with limited slopes (TVD or linear limiter)
    long long n;
    hipStream_t stream;
};
int launch_adjoint_stage(const AdjointLaunch &L);

// Explicit molecular diffusion stage on the state after a step (FP64):
// primitives into W ((2 ns + 5) n doubles), then the update in place of the
// buffer holding the new state (MUSCL-Hancock: U or Ualt by the scalar
// block's parity), and the Courant maximum of the result.
struct DiffusionLaunch {
    const DevParams *P;
    const double *geo;
    double *U, *Ualt;
    int hk;
    double *W;
    double *sc;
    int *err;
    double mu, D, kappa;
    int dim, ns;
    long long n;
    hipStream_t stream;
    hipEvent_t mid;  // recorded between the two passes when not null
    // Jacobi matrix solver (null Mb: explicit): matrices and iterates
    // (diffusion_matrix_doubles(ns) x n), per-iteration maxima
    // ((ns + 4) x (maxit + 2), zeroed), finished iterations per field (device
    // and a host copy, ns + 4, zeroed)
    double *Mb;
    double *dmax;
    int *done, *done_host;
    int maxit;
    int mm, en;  // molar-mass correction, explicit enthalpy flow
    // #Reconstruction star: the step's star states (fv_star.hpp), or null
    // for linear reconstruction
    const double *star;
};
int launch_diffusion(const DiffusionLaunch &L);
size_t diffusion_matrix_doubles(int ns);

// Mixing width sum_i V_i Y (1 - Y), Y = rho_sp / rho, into *J (device), and
// dJ/dU into grad (device, nullable); partial: (n + 255) / 256 doubles.
int launch_mixing_width(const DevParams *P, const double *geo, const double *U, int ns, long long n,
                        int sp, double *partial, double *J, double *grad, hipStream_t stream);

}  // namespace sg
