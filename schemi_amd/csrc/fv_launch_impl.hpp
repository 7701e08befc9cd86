// fv_launch_impl.hpp -- instantiates and launches the stage kernels for one
// flux precision (included by fv_f64.hip and fv_mixed.hip).
#pragma once

#include "fv_general.hpp"
#include "fv_launch.hpp"
#include "fv_stage.hpp"

namespace sg {

template <int DIM, bool FO = false> static dim3 tile_grid(const StageLaunch &L)
{
    constexpr int TX = TileShape<DIM, FO>::TX, TY = TileShape<DIM, FO>::TY;
    return dim3((L.nx + TX - 1) / TX, (DIM >= 2) ? (L.ny + TY - 1) / TY : 1,
                (DIM == 3) ? (L.nz + L.kchunk - 1) / L.kchunk : 1);
}

template <class T, int NS, int DIM, int SV, bool MIXED>
static int launch_mode(const StageLaunch &L)
{
    constexpr int TX = TileShape<DIM>::TX, TY = TileShape<DIM>::TY;
    constexpr int TX0 = TileShape<DIM, true>::TX, TY0 = TileShape<DIM, true>::TY;
    switch (L.mode) {
    case MODE_FIRST_ORDER:
        if (L.gas_model == GAS_VDW)
            hipLaunchKernelGGL((stage0_kernel<T, NS, DIM, SV, MIXED, TX0, TY0, GAS_VDW>),
                               (tile_grid<DIM, true>(L)), dim3(TX0 * TY0), 0, L.stream, L.P,
                               L.geo, L.Uin, L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
        else if (L.gas_model == GAS_RK)
            hipLaunchKernelGGL((stage0_kernel<T, NS, DIM, SV, MIXED, TX0, TY0, GAS_RK>),
                               (tile_grid<DIM, true>(L)), dim3(TX0 * TY0), 0, L.stream, L.P,
                               L.geo, L.Uin, L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
        else if (L.gas_model == GAS_STIFF)
            hipLaunchKernelGGL((stage0_kernel<T, NS, DIM, SV, MIXED, TX0, TY0, GAS_STIFF>),
                               (tile_grid<DIM, true>(L)), dim3(TX0 * TY0), 0, L.stream, L.P,
                               L.geo, L.Uin, L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
        else if (L.gas_model == GAS_KATAOKA)
            hipLaunchKernelGGL((stage0_kernel<T, NS, DIM, SV, MIXED, TX0, TY0, GAS_KATAOKA>),
                               (tile_grid<DIM, true>(L)), dim3(TX0 * TY0), 0, L.stream, L.P,
                               L.geo, L.Uin, L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
        else
            hipLaunchKernelGGL((stage0_kernel<T, NS, DIM, SV, MIXED, TX0, TY0, GAS_IDEAL>),
                               (tile_grid<DIM, true>(L)), dim3(TX0 * TY0), 0, L.stream, L.P,
                               L.geo, L.Uin, L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
        break;
    case MODE_SLOPES:
        // the slope kernel runs MUSCL-Hancock on 1-D meshes only
        if constexpr (SV == SV_HANCOCK && DIM != 1)
            return -1;
        else if (L.lin)
            hipLaunchKernelGGL((stage_kernel<T, NS, DIM, true, SV, MIXED, TX, TY>),
                               tile_grid<DIM>(L), dim3(TX * TY), 0, L.stream, L.P, L.geo, L.Uin,
                               L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
        else
            hipLaunchKernelGGL((stage_kernel<T, NS, DIM, false, SV, MIXED, TX, TY>),
                               tile_grid<DIM>(L), dim3(TX * TY), 0, L.stream, L.P, L.geo, L.Uin,
                               L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
        break;
    case MODE_GENERAL: {
        if constexpr (DIM == 1) {
            return -1;
        } else if constexpr (SV == SV_HANCOCK) {
            const long long n = (long long)L.nx * L.ny * L.nz;
            const dim3 g((unsigned)((n + 255) / 256));
            T *W = static_cast<T *>(L.W), *Lg = static_cast<T *>(L.Lg);
            T *Hs = static_cast<T *>(L.Hs);
            hipLaunchKernelGGL((gen_grad_kernel<T, NS, DIM, MIXED, true>), g, dim3(256), 0,
                               L.stream, L.P, L.geo, L.Uin, W, Lg, L.dt, L.Uout);
            hipLaunchKernelGGL((gen_hancock_sides_kernel<T, NS, DIM, MIXED>), g, dim3(256), 0,
                               L.stream, L.P, L.geo, W, Lg, Hs, L.dt);
            hipLaunchKernelGGL((gen_hancock_update_kernel<T, NS, DIM, MIXED>), g, dim3(256), 0,
                               L.stream, L.P, L.geo, L.Uin, L.Uout, Hs, L.dt, L.sigmax, L.err);
        } else {
            const long long n = (long long)L.nx * L.ny * L.nz;
            const dim3 g((unsigned)((n + 255) / 256));
            T *W = static_cast<T *>(L.W), *Lg = static_cast<T *>(L.Lg);
            hipLaunchKernelGGL((gen_grad_kernel<T, NS, DIM, MIXED>), g, dim3(256), 0, L.stream,
                               L.P, L.geo, L.Uin, W, Lg, L.dt);
            hipLaunchKernelGGL((gen_update_kernel<T, NS, DIM, SV == SV_RK2, MIXED>), g, dim3(256), 0,
                               L.stream, L.P, L.geo, L.Uin, L.Uout, L.Un, W, Lg, L.dt, L.sigmax,
                               L.err);
        }
        break;
    }
    default:
        return -1;
    }
    return static_cast<int>(hipGetLastError());
}

template <class T, int NS, bool MIXED>
static int launch_ns(const StageLaunch &L)
{
#define SG_CASE(D)                                                               \
    if (L.dim == D) {                                                            \
        switch (L.variant) {                                                     \
        case SV_RK1:                                                             \
            return launch_mode<T, NS, D, SV_RK1, MIXED>(L);                      \
        case SV_RK2:                                                             \
            return launch_mode<T, NS, D, SV_RK2, MIXED>(L);                      \
        case SV_HANCOCK:                                                         \
            return launch_mode<T, NS, D, SV_HANCOCK, MIXED>(L);                  \
        default:                                                                 \
            return -1;                                                           \
        }                                                                        \
    }
    SG_CASE(1)
    SG_CASE(2)
    SG_CASE(3)
#undef SG_CASE
    return -1;
}

template <class T, bool MIXED>
static int launch_stage_impl(const StageLaunch &L)
{
    switch (L.ns) {
    case 1:
        return launch_ns<T, 1, MIXED>(L);
    case 2:
        return launch_ns<T, 2, MIXED>(L);
    case 3:
        return launch_ns<T, 3, MIXED>(L);
    default:
        return -1;
    }
}

}  // namespace sg
