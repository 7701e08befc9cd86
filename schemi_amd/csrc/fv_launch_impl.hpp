// fv_launch_impl.hpp -- instantiates and launches the stage kernels for one
// flux precision (included by fv_f64.hip and fv_mixed.hip).
#pragma once

#include "fv_launch.hpp"
#include "fv_stage.hpp"

namespace sg {

template <class T, int NS, int DIM, bool LIN, bool S2, bool MIXED>
static int launch_one(const StageLaunch &L)
{
    constexpr int TX = (DIM == 1) ? 256 : 32;
    constexpr int TY = (DIM == 1) ? 1 : 8;
    dim3 grid((L.nx + TX - 1) / TX, (DIM >= 2) ? (L.ny + TY - 1) / TY : 1,
              (DIM == 3) ? (L.nz + L.kchunk - 1) / L.kchunk : 1);
    hipLaunchKernelGGL((stage_kernel<T, NS, DIM, LIN, S2, MIXED, TX, TY>), grid, dim3(TX * TY), 0,
                       L.stream, L.P, L.geo, L.Uin, L.Uout, L.Un, L.dt, L.sigmax, L.err,
                       L.kchunk);
    return static_cast<int>(hipGetLastError());
}

template <class T, int NS, bool MIXED>
static int launch_ns(const StageLaunch &L)
{
#define SG_CASE(D)                                                               \
    if (L.dim == D) {                                                            \
        if (L.lin)                                                               \
            return L.stage2 ? launch_one<T, NS, D, true, true, MIXED>(L)         \
                            : launch_one<T, NS, D, true, false, MIXED>(L);       \
        return L.stage2 ? launch_one<T, NS, D, false, true, MIXED>(L)            \
                        : launch_one<T, NS, D, false, false, MIXED>(L);          \
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
