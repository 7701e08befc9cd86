// fv_lean.hip -- the lean mixed-precision stage kernel (MODE_LEAN): ideal
// mixture, SSP-RK2, slopes that vanish, free/slip/periodic walls.
#include "fv_launch.hpp"
#include "fv_stage_lean.hpp"

namespace sg {

template <int NS, int DIM, int SV>
static int launch_lean(const StageLaunch &L)
{
    constexpr int TY = SG_LEAN_TY;
    const dim3 g((unsigned)(L.nx / 64), (unsigned)(L.ny / TY),
                 (DIM == 3) ? (unsigned)((L.nz + L.kchunk - 1) / L.kchunk) : 1u);
    hipLaunchKernelGGL((stage_lean_kernel<NS, DIM, SV, TY>), g, dim3(64 * TY), 0, L.stream, L.P,
                       L.geo, L.Uin, L.Uout, L.Un, L.dt, L.sigmax, L.err, L.kchunk);
    return static_cast<int>(hipGetLastError());
}

template <int NS>
static int launch_lean_ns(const StageLaunch &L)
{
    if (L.variant != SV_RK1 && L.variant != SV_RK2)
        return -1;
    const bool rk2 = L.variant == SV_RK2;
    if (L.dim == 2)
        return rk2 ? launch_lean<NS, 2, SV_RK2>(L) : launch_lean<NS, 2, SV_RK1>(L);
    if (L.dim == 3)
        return rk2 ? launch_lean<NS, 3, SV_RK2>(L) : launch_lean<NS, 3, SV_RK1>(L);
    return -1;
}

int launch_stage_lean(const StageLaunch &L)
{
    switch (L.ns) {
    case 1:
        return launch_lean_ns<1>(L);
    case 2:
        return launch_lean_ns<2>(L);
    case 3:
        return launch_lean_ns<3>(L);
    default:
        return -1;
    }
}

}  // namespace sg
