// fv_star.hpp -- #Reconstruction star: the Riemann star state of a face
// (starFields: rho, v, p, a; HLLCSolver.cpp:390-402, HLLSolver.cpp:302-314)
// and its place in the star buffer.  The diffusion stage takes the cell
// gradients of v, rho, p (and a) from these face values averaged over the two
// SSP-RK2 stages (Diffusion.cpp:176-193, 529-537,
// secondOrderStepSolverRK.cpp:73-85) instead of the interpolated cell values.
//
// Star buffer: value q (0 rho, 1..3 v, 4 p, 5..7 a) of face f at
// S[q * P.star_nft + f].  Faces normal to axis a are numbered
// P.star_off[a] + f_x + F_x (f_y + F_y f_z) on the grid of their positions:
// f_a = 0..N_a the position along a (F_a = N_a + 1), the cell indices along
// the other two axes (F = N), so that x is the fastest index for every axis
// and the cells of a wavefront (consecutive x) write consecutive faces; the
// wrapped face of a periodic axis (owner N_a - 1, neighbour 0) is f_a = N_a,
// and f_a = 0 is unused there.
#pragma once

#include "fv_turb.hpp"

namespace sg {

// Star-buffer index of the face of cell idx on axis a at position f along it
// (f = idx[a] for the low face, idx[a] + 1 for the high face).
__device__ __forceinline__ long long star_face(const DevParams &P, int a, int f, const int *idx)
{
    if (f == 0 && P.periodic[a] != 0)
        f = P.N[a];
    const long long c0 = a == 0 ? f : idx[0], c1 = a == 1 ? f : idx[1], c2 = a == 2 ? f : idx[2];
    const long long F0 = P.N[0] + (a == 0), F1 = P.N[1] + (a == 1);
    return P.star_off[a] + c0 + F0 * (c1 + F1 * c2);
}

// pressureStarClass::pressureStar (pressureStarClass.cpp:12-37): pFromUv of
// the star concentrations rho_k / M_k and internal energy
// (E - 0.5 (m & m) / rho) - rho k.
template <class T, int NS, int GAS>
__device__ __forceinline__ T pressure_star(const DevParams &P, T ds0, const T *ds, const T *ms,
                                           T Es, T rks)
{
    const T mm = ((T(0) + ms[0] * ms[0]) + ms[1] * ms[1]) + ms[2] * ms[2];
    const T kin = T(0.5) * mm / ds0;
    const T U = Es - kin - rks;
    T c[NS], c0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        c[k] = ds[k] / mixM<T>(P, k);
        c0 += c[k];
    }
    return eos_p<T, NS, GAS>(P, c, c0, gamma1<T, NS>(P, c, c0), U);
}

// Star state st[0..4+NA] (rho, v, p, a) of a face with owner side O and
// neighbour side B, normal along axis ax with sign sg, selected like the
// flux (HLLCSolver.cpp:52-388, HLLSolver.cpp:52-300): a side as it is on a
// supersonic face, the HLLC star state of the upwind side, or the HLL
// average state.  NaN where the face has no valid wave pattern.
template <class T, int NS, int NA, int GAS>
__device__ __forceinline__ void star_state(const DevParams &P, const TSide<T, NS, NA> &O,
                                           const TSide<T, NS, NA> &B, int ax, T sg, T *st)
{
    const T uO = sg * O.v[ax], uN = sg * B.v[ax];
    const T ptO = T(2. / 3.) * O.rk, ptN = T(2. / 3.) * B.rk;
    const T PO = O.p + ptO, PN = B.p + ptN;
    const T SO = tmin(uO - O.cs, uN - B.cs);
    const T SN = tmax(uO + O.cs, uN + B.cs);
    auto side = [&](const TSide<T, NS, NA> &X) {
        st[0] = X.r0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            st[1 + b] = X.v[b];
        st[4] = X.p;
#pragma unroll
        for (int c = 0; c < NA; ++c)
            st[5 + c] = X.ra[c] / X.r0;
    };
    auto fail = [&]() {
#pragma unroll
        for (int q = 0; q < 5 + NA; ++q)
            st[q] = SO - SO + (SN - SN) / T(0);
    };
    if (P.flux == FLUX_HLL) {
        if (SO >= T(0)) {
            side(O);
        } else if (SN > T(0)) {
            const T den = T(1) / (SN - SO);
            const T fN = SN - uN, fO = SO - uO;
            T ds[NS], ms[3];
            const T ds0 = (B.r0 * fN - O.r0 * fO) * den;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                ds[k] = (B.rho[k] * fN - O.rho[k] * fO) * den;
#pragma unroll
            for (int b = 0; b < 3; ++b)
                ms[b] = (b == ax) ? (B.m[b] * fN - O.m[b] * fO - sg * (PN - PO)) * den
                                  : (B.m[b] * fN - O.m[b] * fO) * den;
            const T Es = (B.E * fN - O.E * fO - (PN * uN - PO * uO)) * den;
            const T rks = (B.rk * fN - O.rk * fO) * den;
            st[0] = ds0;
#pragma unroll
            for (int b = 0; b < 3; ++b)
                st[1 + b] = ms[b] / ds0;
            st[4] = pressure_star<T, NS, GAS>(P, ds0, ds, ms, Es, rks);
#pragma unroll
            for (int c = 0; c < NA; ++c)
                st[5 + c] = ((B.ra[c] * fN - O.ra[c] * fO) * den) / ds0;
        } else if (SN <= T(0)) {
            side(B);
        } else {
            fail();
        }
        return;
    }
    const T rO = O.r0, rN = B.r0;
    const T Sc = (PN - PO + rO * uO * (SO - uO) - rN * uN * (SN - uN)) /
                 (rO * (SO - uO) - rN * (SN - uN));
    const T PLR = T(0.5) * (PO + PN + rO * (SO - uO) * (Sc - uO) + rN * (SN - uN) * (Sc - uN));
    // HLLC star state of side X (HLLCSolver.cpp:150-185, 245-280)
    auto star = [&](const TSide<T, NS, NA> &X, T u, T S, T PX) {
        const T factor = S - u;
        const T den = T(1) / (S - Sc);
        T ds[NS], ms[3];
        const T ds0 = X.r0 * factor * den;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            ds[k] = X.rho[k] * factor * den;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            ms[b] = (b == ax) ? (X.m[b] * factor - sg * (PX - PLR)) * den : (X.m[b] * factor) * den;
        const T Es = (X.E * factor - PX * u + PLR * Sc) * den;
        const T rks = X.rk * factor * den;
        st[0] = ds0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            st[1 + b] = ms[b] / ds0;
        st[4] = pressure_star<T, NS, GAS>(P, ds0, ds, ms, Es, rks);
#pragma unroll
        for (int c = 0; c < NA; ++c)
            st[5 + c] = (X.ra[c] * factor * den) / ds0;
    };
    if (SO >= T(0))
        side(O);
    else if (Sc >= T(0))
        star(O, uO, SO, PO);
    else if (SN > T(0))
        star(B, uN, SN, PN);
    else if (SN <= T(0))
        side(B);
    else
        fail();
}

}  // namespace sg
