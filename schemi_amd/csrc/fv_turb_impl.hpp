// fv_turb_impl.hpp -- instantiates and launches the turbulent stage kernels
// for one flux precision (included by fv_turb_f64.hip and fv_turb_mixed.hip).
#pragma once

#include "fv_turb.hpp"
#include "fv_turb_launch.hpp"

namespace sg {

template <class T, int NS, int DIM, int SV, bool MIXED, bool SLOPES, int NA>
static int launch_turb_na(const TurbStageLaunch &L)
{
    const unsigned blocks = (unsigned)((L.n + 255) / 256);
    if (L.gas_model == GAS_VDW)
        hipLaunchKernelGGL((turb_stage_kernel<T, NS, DIM, SV, MIXED, SLOPES, GAS_VDW, NA>),
                           dim3(blocks), dim3(256), 0, L.stream, L.P, L.TP, L.geo, L.Uin, L.Uout,
                           L.Un, L.dt, L.sigmax, L.err, L.b2);
    else if (L.gas_model == GAS_IDEAL)
        hipLaunchKernelGGL((turb_stage_kernel<T, NS, DIM, SV, MIXED, SLOPES, GAS_IDEAL, NA>),
                           dim3(blocks), dim3(256), 0, L.stream, L.P, L.TP, L.geo, L.Uin, L.Uout,
                           L.Un, L.dt, L.sigmax, L.err, L.b2);
    else
        return -1;
    return static_cast<int>(hipGetLastError());
}

// na: passive fields after rho epsilon (0; 3 kEpsA: rho a; 4 BHR: rho a, rho b)
template <class T, int NS, int DIM, int SV, bool MIXED, bool SLOPES>
static int launch_turb_gas(const TurbStageLaunch &L)
{
    if (L.na == 4)
        return launch_turb_na<T, NS, DIM, SV, MIXED, SLOPES, 4>(L);
    if (L.na == 3)
        return launch_turb_na<T, NS, DIM, SV, MIXED, SLOPES, 3>(L);
    if (L.na == 0)
        return launch_turb_na<T, NS, DIM, SV, MIXED, SLOPES, 0>(L);
    return -1;
}

template <class T, int NS, int DIM, bool MIXED>
static int launch_turb_sv(const TurbStageLaunch &L)
{
    if constexpr (DIM == 1) {
        if (L.slopes)
            return L.variant == SV_RK2 ? launch_turb_gas<T, NS, 1, SV_RK2, MIXED, true>(L)
                                       : launch_turb_gas<T, NS, 1, SV_RK1, MIXED, true>(L);
    } else {
        if (L.slopes)
            return -1;
    }
    return L.variant == SV_RK2 ? launch_turb_gas<T, NS, DIM, SV_RK2, MIXED, false>(L)
                               : launch_turb_gas<T, NS, DIM, SV_RK1, MIXED, false>(L);
}

template <class T, bool MIXED> static int launch_turb_impl(const TurbStageLaunch &L)
{
    if (L.variant != SV_RK1 && L.variant != SV_RK2)
        return -1;
    auto by_dim = [&](auto ns) -> int {
        constexpr int NS = decltype(ns)::value;
        switch (L.dim) {
        case 1:
            return launch_turb_sv<T, NS, 1, MIXED>(L);
        case 2:
            return launch_turb_sv<T, NS, 2, MIXED>(L);
        case 3:
            return launch_turb_sv<T, NS, 3, MIXED>(L);
        }
        return -1;
    };
    switch (L.ns) {
    case 1:
        return by_dim(std::integral_constant<int, 1>{});
    case 2:
        return by_dim(std::integral_constant<int, 2>{});
    case 3:
        return by_dim(std::integral_constant<int, 3>{});
    }
    return -1;
}

}  // namespace sg
