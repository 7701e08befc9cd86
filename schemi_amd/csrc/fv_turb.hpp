// fv_turb.hpp -- k-epsilon turbulence on the step: the conserved state
// carries rho k and rho epsilon after E (Ns + 7 fields; the kEpsA model adds
// rho a, Ns + 10), the face closure adds rho k to the total energy and
// (2/3) k (1 + dpdUv) to the squared sound speed, the HLLC / HLL flux carries
// the turbulent pressure (2/3) rho k and the fluxes of rho k, rho epsilon and
// rho a (HLLCSolver.cpp:26-388, HLLSolver.cpp:26-300,
// Advection2dOrder.cpp:77-627), and the internal energy is
// E - rho v^2 / 2 - rho k.  The diffusion stage's turbulent terms live in
// fv_diffusion.hip and use TurbParams.
//
// NA is the number of passive fields after rho epsilon: 0, 3 (rho a of the
// kEpsA model) or 4 (rho a and rho b of the BHR model).  Each is fluxed
// with the arithmetic of rho k (HLLCSolver.cpp:145-147, HLLSolver.cpp:
// 117-119, 237-251), so the a and b transport adds passive fields; they
// differ only in their boundary conditions and slopes (a: vector, b: scalar).
//
// turb_stage_kernel: one thread per cell.  The cell evaluates its 2 DIM faces
// in the reference's order (x-, x+, y+, y-, z-, z+) with the face's
// owner/neighbour roles, so a face gives both of its cells the same bits, and
// writes only its own state (no atomics but the Courant maximum).  Face
// values are the cell values where the limited slopes vanish identically
// (TVD limiters in 2-D and 3-D, the zero limiter); on 1-D meshes with a TVD
// limiter every side is reconstructed from the limited slopes of c_k, v, p,
// k, epsilon and a of its cell (stencil i-2 .. i+2; of v and a only the
// component along the axis has a slope, TVDLimiter.cpp:127-273 gives the
// transverse components an exact zero limiter entry).
#pragma once

#include "fv_stage.hpp"

namespace sg {

// turbulenceModel (turbulenceModelEnum.hpp) values built here
enum : int { TURB_ZERO = 0, TURB_SHEAR = 2, TURB_KEPSA = 6, TURB_BHR = 7 };

// Parameters of set/turbulentParameters.txt in the k-epsilon models' meaning
// (abstractTurbulenceModel.hpp:26-57, kEpsModels.hpp, shearModel.cpp:35-59,
// kEpsAModel.cpp:29-121) and the boundary conditions of k (0) and epsilon (1)
// per patch, and of the vector a (kEpsA).
struct TurbParams {
    int model;
    double Cmu, sigmaSc, sigmaT, sigmaE, sigmaK, sigmaEps, CMS_R, CMS_D, mink, mineps;
    double C1, C2, C3;
    double ts_coeff;  // first coefficient of #Source_timestep_correction
    int bct[6][2];
    double bcv[6][2];
    // kEpsA: sigma_a, C0, Ca, CaMax, CMSA, chi (C1..C3 above)
    double sigmaa, C0, Ca, CaMax, CMSA, chi;
    int bcta[6];
    double bcva[6][3];
    // BHR (BHRModel.cpp:86-137; C0..C3, Ca, sigma_a above): C4, Cb1, Cb2,
    // alpha2, alpha3, alpha4, CMS_M, sigma_b, the minimal b, set/b.txt
    double C4, Cb1, Cb2, alpha2, alpha3, alpha4, CMSM, sigmab, minb;
    int bctb[6];
    double bcvb[6];
};

// kEpsModels::thetaS (kEpsModels.cpp:13-26): realisability factor of the
// turbulent viscosity for the velocity divergence divV.
__device__ __forceinline__ double theta_s(const TurbParams &TP, double divV, double k, double eps,
                                          double CMS)
{
    const double Cmu2 = TP.Cmu * TP.Cmu;
    const double k2 = k * k;
    const double eps2 = eps * eps;
    const double CMS2 = CMS * CMS;
    const double divV2 = divV * divV;
    return 1. / __builtin_sqrt(1 + 16. / 9. * Cmu2 * k2 / (CMS2 * eps2) * divV2);
}

// Primitives of a turbulent conserved state: c_k = rho_k / M_k, v = m / rho,
// U = (E - rho v^2 / 2) - rho k, p = pFromUv, k = rho k / rho,
// epsilon = rho epsilon / rho, a = rho a / rho (Advection2dOrder.cpp:578-627).
// u: [rho_1..rho_Ns, rho, m (3), E, rho k, rho eps, rho a (NA)];
// w: [c (Ns), v (3), p, k, eps, a (NA)]
template <class T, int NS, int GAS, int NA = 0>
__device__ __forceinline__ void prims_turb(const DevParams &P, const T *u, T *w)
{
    T c0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        w[k] = u[k] / mixM<T>(P, k);
        c0 += w[k];
    }
    const T rho = u[NS];
    const T vx = u[NS + 1] / rho, vy = u[NS + 2] / rho, vz = u[NS + 3] / rho;
    w[NS] = vx;
    w[NS + 1] = vy;
    w[NS + 2] = vz;
    const T v2 = ((T(0) + vx * vx) + vy * vy) + vz * vz;
    const T U = (u[NS + 4] - rho * v2 * T(0.5)) - u[NS + 5];
    w[NS + 3] = eos_p<T, NS, GAS>(P, w, c0, gamma1<T, NS>(P, w, c0), U);
    w[NS + 4] = u[NS + 5] / rho;
    w[NS + 5] = u[NS + 6] / rho;
#pragma unroll
    for (int c = 0; c < NA; ++c)
        w[NS + 6 + c] = u[NS + 7 + c] / rho;
}

// Face side with turbulence (Advection2dOrder.cpp:233-289, 301-372, 436-515):
// rho k = k rho, rho eps = eps rho, rho a = a rho, E = (U + rho v^2 / 2) +
// rho k, and the sound speed sqrt(c^2 + (2/3) k (1 + dpdUv)) of the flux
// (HLLCSolver.cpp:26-50).
template <class T, int NS, int NA = 0> struct TSide {
    T rho[NS];
    T r0;
    T v[3];
    T m[3];
    T p, U, E, cs, rk, reps;
    T ra[NA > 0 ? NA : 1];
};

template <class T, int NS, int GAS, int NA = 0>
__device__ __forceinline__ void close_side_turb(const DevParams &P, const T *c, const T *v, T p,
                                                T k, T eps, const T *a, TSide<T, NS, NA> &s)
{
    T c0 = 0, r0 = 0;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        c0 += c[q];
        s.rho[q] = c[q] * mixM<T>(P, q);
        r0 += s.rho[q];
    }
    s.r0 = r0;
    s.rk = k * r0;
    s.reps = eps * r0;
#pragma unroll
    for (int q = 0; q < NA; ++q)
        s.ra[q] = a[q] * r0;
    const T g1 = gamma1<T, NS>(P, c, c0);
    s.p = p;
    s.U = eos_U<T, NS, GAS>(P, c, c0, g1, p);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.v[a] = v[a];
        s.m[a] = v[a] * r0;
    }
    const T v2 = ((T(0) + v[0] * v[0]) + v[1] * v[1]) + v[2] * v[2];
    s.E = (s.U + r0 * v2 * T(0.5)) + s.rk;
    const T c2 = eos_c2<T, NS, GAS>(P, c, c0, g1, r0, s.U, p);
    // dpdUv: ideal g1, van der Waals g1 / (1 - c0 b)
    T dpdUv = g1;
    if (is_vdw<GAS>(P)) {
        T a, b;
        vdw_ab<T, NS>(P, c, c0, a, b);
        dpdUv = g1 / (T(1) - c0 * b);
    }
    s.cs = Real<T>::sqrt_(c2 + T(2. / 3.) * k * (T(1) + dpdUv));
}

// HLLC flux F & N of the turbulent state (HLLCSolver.cpp:52-388): the
// pressure P = p + (2/3) rho k in the wave speeds, the star pressure, the
// momentum and the energy ((E + p) + (2/3) rho k) u; rho k, rho epsilon and
// the components of rho a carried like the partial densities.  g: NS + 7 + NA
// fluxes.
template <class T, int NS, int NA = 0>
__device__ __forceinline__ bool hllc_flux_turb(const TSide<T, NS, NA> &O,
                                               const TSide<T, NS, NA> &B, int ax, T sg, T *g)
{
    constexpr int NF = NS + 7 + NA;
    const T uO = sg * O.v[ax], uN = sg * B.v[ax];
    const T ptO = T(2. / 3.) * O.rk, ptN = T(2. / 3.) * B.rk;
    const T PO = O.p + ptO, PN = B.p + ptN;
    const T SO = tmin(uO - O.cs, uN - B.cs);
    const T SN = tmax(uO + O.cs, uN + B.cs);
    const T rO = O.r0, rN = B.r0;
    const T Sc = (PN - PO + rO * uO * (SO - uO) - rN * uN * (SN - uN)) /
                 (rO * (SO - uO) - rN * (SN - uN));
    const T PLR = T(0.5) * (PO + PN + rO * (SO - uO) * (Sc - uO) + rN * (SN - uN) * (Sc - uN));
    const bool own = (SO >= T(0)) || (Sc >= T(0));
    const bool star = own ? !(SO >= T(0)) : (SN > T(0));
    if (!own && !(SN > T(0)) && !(SN <= T(0))) {
#pragma unroll
        for (int q = 0; q < NF; ++q)
            g[q] = SO;  // NaN
        return false;
    }
    const TSide<T, NS, NA> &X = own ? O : B;
    const T u = own ? uO : uN;
    const T S = own ? SO : SN;
    const T pt = own ? ptO : ptN;
    const T PX = own ? PO : PN;
#pragma unroll
    for (int k = 0; k < NS; ++k)
        g[k] = X.rho[k] * u;
    g[NS] = X.r0 * u;
#pragma unroll
    for (int b = 0; b < 3; ++b)
        g[NS + 1 + b] = (b == ax) ? X.m[b] * u + sg * PX : X.m[b] * u;
    g[NS + 4] = ((X.E + X.p) + pt) * u;
    g[NS + 5] = X.rk * u;
    g[NS + 6] = X.reps * u;
#pragma unroll
    for (int q = 0; q < NA; ++q)
        g[NS + 7 + q] = X.ra[q] * u;
    if (star) {
        const T factor = S - u;
        const T den = T(1) / (S - Sc);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const T rs = X.rho[k] * factor * den;
            g[k] = g[k] + S * (rs - X.rho[k]);
        }
        {
            const T rs = X.r0 * factor * den;
            g[NS] = g[NS] + S * (rs - X.r0);
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const T ms = (b == ax) ? (X.m[b] * factor - sg * (PX - PLR)) * den
                                   : (X.m[b] * factor) * den;
            g[NS + 1 + b] = g[NS + 1 + b] + (ms - X.m[b]) * S;
        }
        const T Es = (X.E * factor - PX * u + PLR * Sc) * den;
        g[NS + 4] = g[NS + 4] + S * (Es - X.E);
        const T ks = X.rk * factor * den, es = X.reps * factor * den;
        g[NS + 5] = g[NS + 5] + S * (ks - X.rk);
        g[NS + 6] = g[NS + 6] + S * (es - X.reps);
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const T as = X.ra[q] * factor * den;
            g[NS + 7 + q] = g[NS + 7 + q] + S * (as - X.ra[q]);
        }
    }
    return true;
}

// HLL flux of the turbulent state (HLLSolver.cpp:52-300).
template <class T, int NS, int NA = 0>
__device__ __forceinline__ bool hll_flux_turb(const TSide<T, NS, NA> &O,
                                              const TSide<T, NS, NA> &B, int ax, T sg, T *g)
{
    constexpr int NF = NS + 7 + NA;
    const T uO = sg * O.v[ax], uN = sg * B.v[ax];
    const T ptO = T(2. / 3.) * O.rk, ptN = T(2. / 3.) * B.rk;
    const T PO = O.p + ptO, PN = B.p + ptN;
    const T SO = tmin(uO - O.cs, uN - B.cs);
    const T SN = tmax(uO + O.cs, uN + B.cs);
    if (SO >= T(0) || !(SN > T(0))) {
        if (!(SO >= T(0)) && !(SN <= T(0))) {
#pragma unroll
            for (int q = 0; q < NF; ++q)
                g[q] = SO;
            return false;
        }
        const bool own = SO >= T(0);
        const TSide<T, NS, NA> &X = own ? O : B;
        const T u = own ? uO : uN;
        const T PX = own ? PO : PN;
        const T pt = own ? ptO : ptN;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            g[k] = u * X.rho[k];
        g[NS] = u * X.r0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            g[NS + 1 + b] = (b == ax) ? X.m[b] * u + sg * PX : X.m[b] * u;
        g[NS + 4] = u * ((X.E + X.p) + pt);
        g[NS + 5] = u * X.rk;
        g[NS + 6] = u * X.reps;
#pragma unroll
        for (int q = 0; q < NA; ++q)
            g[NS + 7 + q] = u * X.ra[q];
        return true;
    }
    const T den = T(1) / (SN - SO);
#pragma unroll
    for (int k = 0; k < NS; ++k)
        g[k] = (SN * uO * O.rho[k] - SO * uN * B.rho[k] + SN * SO * (B.rho[k] - O.rho[k])) * den;
    g[NS] = (SN * uO * O.r0 - SO * uN * B.r0 + SN * SO * (B.r0 - O.r0)) * den;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        if (b == ax)
            g[NS + 1 + b] = (((O.m[b] * uO + sg * PO) * SN - (B.m[b] * uN + sg * PN) * SO) +
                             (B.m[b] - O.m[b]) * (SN * SO)) * den;
        else
            g[NS + 1 + b] = (((O.m[b] * uO) * SN - (B.m[b] * uN) * SO) +
                             (B.m[b] - O.m[b]) * (SN * SO)) * den;
    }
    g[NS + 4] = (SN * (uO * ((O.E + O.p) + ptO)) - SO * (uN * ((B.E + B.p) + ptN)) +
                 SN * SO * (B.E - O.E)) * den;
    g[NS + 5] = (SN * uO * O.rk - SO * uN * B.rk + SN * SO * (B.rk - O.rk)) * den;
    g[NS + 6] = (SN * uO * O.reps - SO * uN * B.reps + SN * SO * (B.reps - O.reps)) * den;
#pragma unroll
    for (int q = 0; q < NA; ++q)
        g[NS + 7 + q] =
            (SN * uO * O.ra[q] - SO * uN * B.ra[q] + SN * SO * (B.ra[q] - O.ra[q])) * den;
    return true;
}

// Limited slope along axis a of cell i with neighbour values qm, qp
// (TVDLimiter.cpp:12-273 with the one-axis stencil of slopes_axis): boundary
// conditions tlo / thi with values vlo / vhi on the low / high patch of the
// axis, comp the vector component of the value (-1 for scalars).
template <class T>
__device__ __forceinline__ T slope_turb_bc(const DevParams &P, const AxisGeo &G, int a, int i,
                                           int tlo, T vlo, int thi, T vhi, int comp, T qm, T q0,
                                           T qp, T S, T V)
{
    const int Na = P.N[a];
    const bool per = P.periodic[a] != 0;
    const bool has_lo = (i > 0) || per;
    const bool has_hi = (i < Na - 1) || per;
    const bool flip = !has_hi && P.hi_flip[a] != 0;
    const T elo = has_lo ? -((qm - q0) / to_T<T>(G.dn[i]))
                         : -((bc_cell<T>(tlo, vlo, q0, comp, a) - q0) / to_T<T>(P.ghost_lo[a]));
    T ehi = has_hi ? (qp - q0) / to_T<T>(G.dn[i + 1])
                   : (bc_cell<T>(thi, vhi, q0, comp, a) - q0) / to_T<T>(P.ghost_hi[a]);
    if (flip)
        ehi = -ehi;
    T mn = tmin(elo, ehi), mx = tmax(elo, ehi);
    if (P.zero_entry[a] != 0) {
        mn = tmin(mn, T(0));
        mx = tmax(mx, T(0));
    }
    if (mn <= T(0) && mx + T(STAB) > T(0))
        return T(0);
    const T flo = has_lo ? qm * to_T<T>(G.wo[i]) + q0 * to_T<T>(G.wn[i])
                         : bc_surface<T>(tlo, vlo, q0, comp, a);
    const T fhi = has_hi ? q0 * to_T<T>(G.wo[i + 1]) + qp * to_T<T>(G.wn[i + 1])
                         : bc_surface<T>(thi, vhi, q0, comp, a);
    const T grad = ((flip ? -fhi : fhi) * S - flo * S) / V;
    const T r = mn / (mx + T(STAB));
    return phi<T>(P.limiter, r) * grad;
}

// Limited slope of a turbulent scalar (k: f = 0, epsilon: f = 1).
template <class T>
__device__ __forceinline__ T slope_turb(const DevParams &P, const TurbParams &TP,
                                        const AxisGeo &G, int a, int i, int f, T qm, T q0, T qp,
                                        T S, T V)
{
    const int plo = patch_of(a, 0), phi_ = patch_of(a, 1);
    return slope_turb_bc<T>(P, G, a, i, TP.bct[plo][f], T(TP.bcv[plo][f]), TP.bct[phi_][f],
                            T(TP.bcv[phi_][f]), -1, qm, q0, qp, S, V);
}

// Limited slope of the BHR model's b.
template <class T>
__device__ __forceinline__ T slope_turb_b(const DevParams &P, const TurbParams &TP,
                                          const AxisGeo &G, int a, int i, T qm, T q0, T qp, T S,
                                          T V)
{
    const int plo = patch_of(a, 0), phi_ = patch_of(a, 1);
    return slope_turb_bc<T>(P, G, a, i, TP.bctb[plo], T(TP.bcvb[plo]), TP.bctb[phi_],
                            T(TP.bcvb[phi_]), -1, qm, q0, qp, S, V);
}

// Limited slope along axis a of the a component along that axis (the
// transverse components have none).
template <class T>
__device__ __forceinline__ T slope_turb_a(const DevParams &P, const TurbParams &TP,
                                          const AxisGeo &G, int a, int i, T qm, T q0, T qp, T S,
                                          T V)
{
    const int plo = patch_of(a, 0), phi_ = patch_of(a, 1);
    return slope_turb_bc<T>(P, G, a, i, TP.bcta[plo], T(TP.bcva[plo][a]), TP.bcta[phi_],
                            T(TP.bcva[phi_][a]), a, qm, q0, qp, S, V);
}

// Ghost side of a boundary face from the owner side (Advection2dOrder.cpp:398-521);
// side values c (NS), v (3), p, k, eps, a (NA).
template <class T, int NS, int NA = 0>
__device__ __forceinline__ void ghost_turb(const DevParams &P, const TurbParams &TP, int a,
                                           int side, const T *o, T *g)
{
    ghost_side<T, NS>(P, a, side, o, o + NS, o[NS + 3], g, g + NS, g[NS + 3]);
    const int p = patch_of(a, side);
    g[NS + 4] = bc_cell<T>(TP.bct[p][0], T(TP.bcv[p][0]), o[NS + 4], -1, a);
    g[NS + 5] = bc_cell<T>(TP.bct[p][1], T(TP.bcv[p][1]), o[NS + 5], -1, a);
#pragma unroll
    for (int c = 0; c < NA && c < 3; ++c)
        g[NS + 6 + c] = bc_cell<T>(TP.bcta[p], T(TP.bcva[p][c]), o[NS + 6 + c], c, a);
    if constexpr (NA == 4)
        g[NS + 9] = bc_cell<T>(TP.bctb[p], T(TP.bcvb[p]), o[NS + 9], -1, a);
}

template <class T, int NS, int GAS, int NA = 0>
__device__ __forceinline__ bool flux_turb(const DevParams &P, int a, T sg, const T *o,
                                          const T *b, T *g)
{
    TSide<T, NS, NA> O, B;
    close_side_turb<T, NS, GAS, NA>(P, o, o + NS, o[NS + 3], o[NS + 4], o[NS + 5], o + NS + 6,
                                    O);
    close_side_turb<T, NS, GAS, NA>(P, b, b + NS, b[NS + 3], b[NS + 4], b[NS + 5], b + NS + 6,
                                    B);
    if (P.flux == FLUX_HLL)
        return hll_flux_turb<T, NS, NA>(O, B, a, sg, g);
    return hllc_flux_turb<T, NS, NA>(O, B, a, sg, g);
}

// Courant signal speed (|v| + c) / min edge with U = E - rho v^2 / 2 - rho k
// (Schemi.cpp:810-828: the squared sound speed without the turbulent term).
template <class T, int NS, int GAS>
__device__ __forceinline__ T signal_speed_turb(const DevParams &P, const T *u, T minlen)
{
    T w[NS + 6];
    prims_turb<T, NS, GAS>(P, u, w);
    T c0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
        c0 += w[k];
    const T vx = w[NS], vy = w[NS + 1], vz = w[NS + 2], p = w[NS + 3];
    const T v2 = ((T(0) + vx * vx) + vy * vy) + vz * vz;
    const T U = (u[NS + 4] - u[NS] * v2 * T(0.5)) - u[NS + 5];
    const T c2 = eos_c2<T, NS, GAS>(P, w, c0, gamma1<T, NS>(P, w, c0), u[NS], U, p);
    return (hypot3<T>(vx, vy, vz) + Real<T>::sqrt_(c2)) * minlen;
}

// One forward-Euler stage (SV_RK1) or the second SSP-RK2 stage with the
// average (SV_RK2) of a turbulent state.  SLOPES: 1-D mesh with a TVD limiter.

// XCD-aware block order (SG_TURB_XCD=1): the hardware deals consecutive
// workgroups round-robin to the 8 XCDs; renumbering them so that each XCD
// takes one contiguous eighth of the cells keeps the y and z neighbours a
// block reads in that XCD's L2.  A bijection when the grid is a multiple of 8.
// Measured on the 256x128x128 BHR workload: the mixed stage kernels within
// 2 % either way, the FP64 ones 6-7 % slower, the star kernel unchanged; off.
#ifndef SG_TURB_XCD
#define SG_TURB_XCD 0
#endif
__device__ __forceinline__ long long turb_block(unsigned b, unsigned nb)
{
#if SG_TURB_XCD
    if (nb % 8u == 0u)
        b = (b % 8u) * (nb / 8u) + b / 8u;
#endif
    return (long long)b;
}

template <class T, int NS, int DIM, int SV, bool MIXED, bool SLOPES, int GAS, int NA>
__global__ void __launch_bounds__(256)
turb_stage_kernel(const DevParams *__restrict__ Pp, const TurbParams *__restrict__ Tp,
                  const double *__restrict__ geo, const double *Uin, double *Uout,
                  const double *Un, const double *__restrict__ dt_ptr,
                  unsigned long long *sigmax, int *errflag, double *__restrict__ b2)
{
    constexpr int NF = NS + 7 + NA, NQ = NS + 6 + NA;
    __shared__ double sRed[4];
    const DevParams &P = *Pp;
    const TurbParams &TP = *Tp;
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = turb_block(blockIdx.x, gridDim.x) * 256LL + threadIdx.x;
    const bool act = id < n;
    const long long idc = act ? id : n - 1;
    const int idx[3] = {int(idc % P.N[0]), int((idc / P.N[0]) % P.N[1]),
                        int(idc / P.stride[2])};
    const double dt = *dt_ptr;
    double sig = 0.0;
    int err = 0;

    auto prims = [&](long long c, T *w) {
        T ut[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q)
            ut[q] = T(Uin[q * n + c]);
        prims_turb<T, NS, GAS, NA>(P, ut, w);
    };
    // neighbour of cell id along a at offset dlt (periodic wrap; the caller
    // checks existence on non-periodic axes)
    auto nbr = [&](int a, int dlt) -> long long {
        const int Na = P.N[a];
        int j = idx[a] + dlt;
        j = (j % Na + Na) % Na;
        return idc + (long long)(j - idx[a]) * P.stride[a];
    };

    T g[2 * DIM][NF];
    T w0[NQ];
    prims(idc, w0);
    if constexpr (SLOPES) {
        // 1-D: primitives of cells i-2..i+2, slopes of i-1, i, i+1
        const int Nx = P.N[0];
        const bool per = P.periodic[0] != 0;
        const int i = idx[0];
        T w[5][NQ];
#pragma unroll
        for (int d = -2; d <= 2; ++d) {
            const bool ok = per || (i + d >= 0 && i + d < Nx);
            if (d == 0) {
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    w[2][q] = w0[q];
            } else {
                prims(ok ? nbr(0, d) : idc, w[d + 2]);
            }
        }
        const AxisGeo G = axis_geo(P, geo, 0);
        using SL = SlopeLayout<NS, false>;
        T L[3][SL::NL], Lt[3][2], La[3], Lb[3];
#pragma unroll
        for (int d = -1; d <= 1; ++d) {
            const bool ok = per || (i + d >= 0 && i + d < Nx);
            const int j = ok ? ((i + d) % Nx + Nx) % Nx : i;
            const CellGeo cgj = cell_geo(P, geo, j, 0, 0);
            slopes_axis<T, NS, false>(P, G, 0, j, w[d + 1], w[d + 2], w[d + 3], T(cgj.S[0]),
                                      T(cgj.V), L[d + 1]);
#pragma unroll
            for (int f = 0; f < 2; ++f)
                Lt[d + 1][f] = slope_turb<T>(P, TP, G, 0, j, f, w[d + 1][NS + 4 + f],
                                             w[d + 2][NS + 4 + f], w[d + 3][NS + 4 + f],
                                             T(cgj.S[0]), T(cgj.V));
            if constexpr (NA > 0)
                La[d + 1] = slope_turb_a<T>(P, TP, G, 0, j, w[d + 1][NS + 6], w[d + 2][NS + 6],
                                            w[d + 3][NS + 6], T(cgj.S[0]), T(cgj.V));
            if constexpr (NA == 4)
                Lb[d + 1] = slope_turb_b<T>(P, TP, G, 0, j, w[d + 1][NS + 9], w[d + 2][NS + 9],
                                            w[d + 3][NS + 9], T(cgj.S[0]), T(cgj.V));
        }
        // face side of cell i + d (d = -1, 0, 1) at its low (hi = false) or
        // high face
        auto side = [&](int d, bool hi, T *s) {
            const int j = ((i + d) % Nx + Nx) % Nx;
            const T off = to_T<T>(hi ? G.dhi[j] : G.dlo[j]);
            T c[NS], v[3], p;
            reconstruct<T, NS, false>(w[d + 2], L[d + 1], 0, off, c, v, p);
#pragma unroll
            for (int k = 0; k < NS; ++k)
                s[k] = c[k];
#pragma unroll
            for (int b = 0; b < 3; ++b)
                s[NS + b] = v[b];
            s[NS + 3] = p;
            s[NS + 4] = Lt[d + 1][0] * off + w[d + 2][NS + 4];
            s[NS + 5] = Lt[d + 1][1] * off + w[d + 2][NS + 5];
            if constexpr (NA > 0) {
                s[NS + 6] = La[d + 1] * off + w[d + 2][NS + 6];
                s[NS + 7] = w[d + 2][NS + 7];
                s[NS + 8] = w[d + 2][NS + 8];
            }
            if constexpr (NA == 4)
                s[NS + 9] = Lb[d + 1] * off + w[d + 2][NS + 9];
        };
        T so[NQ], sb[NQ];
        // x- face
        if (!per && i == 0) {
            side(0, false, so);
            ghost_turb<T, NS, NA>(P, TP, 0, 0, so, sb);
            if (!flux_turb<T, NS, GAS, NA>(P, 0, T(-1), so, sb, g[0]))
                err |= ERRF_RIEMANN;
        } else {
            side(-1, true, so);
            side(0, false, sb);
            if (!flux_turb<T, NS, GAS, NA>(P, 0, T(1), so, sb, g[0]))
                err |= ERRF_RIEMANN;
        }
        // x+ face
        if (!per && i == Nx - 1) {
            side(0, true, so);
            ghost_turb<T, NS, NA>(P, TP, 0, 1, so, sb);
            if (!flux_turb<T, NS, GAS, NA>(P, 0, P.hi_flip[0] ? T(-1) : T(1), so, sb, g[1]))
                err |= ERRF_RIEMANN;
        } else {
            side(0, true, so);
            side(1, false, sb);
            if (!flux_turb<T, NS, GAS, NA>(P, 0, T(1), so, sb, g[1]))
                err |= ERRF_RIEMANN;
        }
    } else {
        // face values are the cell values
#pragma unroll
        for (int s = 0; s < 2 * DIM; ++s) {
            const int a = s / 2;
            const bool hi = (s == 1) || (s == 2) || (s == 5);
            const int Na = P.N[a];
            const bool per = P.periodic[a] != 0;
            const bool bnd = !per && (hi ? idx[a] == Na - 1 : idx[a] == 0);
            T wb[NQ];
            if (bnd) {
                ghost_turb<T, NS, NA>(P, TP, a, hi ? 1 : 0, w0, wb);
                const T sg = hi ? (P.hi_flip[a] ? T(-1) : T(1)) : T(-1);
                if (!flux_turb<T, NS, GAS, NA>(P, a, sg, w0, wb, g[s]))
                    err |= ERRF_RIEMANN;
            } else {
                prims(nbr(a, hi ? 1 : -1), wb);
                const bool ok = hi ? flux_turb<T, NS, GAS, NA>(P, a, T(1), w0, wb, g[s])
                                   : flux_turb<T, NS, GAS, NA>(P, a, T(1), wb, w0, g[s]);
                if (!ok)
                    err |= ERRF_RIEMANN;
            }
        }
    }

    if (act) {
        const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
        // the cell owns its low faces on a non-periodic boundary
        const bool lo[3] = {P.periodic[0] == 0 && idx[0] == 0, P.periodic[1] == 0 && idx[1] == 0,
                            P.periodic[2] == 0 && idx[2] == 0};
        double u[NF], un[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q)
            u[q] = Uin[q * n + id];
        const double dtV = MIXED ? dt / cg.V : 0.0;
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            double acc = 0;
            acc += (lo[0] ? double(g[0][q]) : -double(g[0][q])) * cg.S[0];
            acc += double(g[1][q]) * cg.S[0];
            if constexpr (DIM >= 2) {
                acc += double(g[2][q]) * cg.S[1];
                acc += (lo[1] ? double(g[3][q]) : -double(g[3][q])) * cg.S[1];
            }
            if constexpr (DIM == 3) {
                acc += (lo[2] ? double(g[4][q]) : -double(g[4][q])) * cg.S[2];
                acc += double(g[5][q]) * cg.S[2];
            }
            if constexpr (MIXED)
                un[q] = u[q] - acc * dtV;
            else
                un[q] = u[q] - (acc / cg.V) * dt;
        }
        if (P.gravity) {
            const double rho = u[NS];
            const double vx = u[NS + 1] / rho, vy = u[NS + 2] / rho, vz = u[NS + 3] / rho;
            const double vg = ((0. + vx * P.g[0]) + vy * P.g[1]) + vz * P.g[2];
            un[NS + 4] = un[NS + 4] + rho * vg * dt;
#pragma unroll
            for (int b = 0; b < 3; ++b)
                un[NS + 1 + b] = un[NS + 1 + b] + (P.g[b] * rho) * dt;
        }
        if constexpr (MIXED) {
            double sr = 0;
#pragma unroll
            for (int q = 0; q < NS; ++q)
                sr += un[q];
            const double f = un[NS] / sr;
#pragma unroll
            for (int q = 0; q < NS; ++q)
                un[q] *= f;
        }
        if constexpr (SV == SV_RK2) {
            // BHR: the RK average does not re-derive b (bunchOfFields.hpp:
            // 166-203); the diffusion stage reads the second stage's
            // rho b / rho (Advection2dOrder.cpp:623-625)
            if constexpr (NA == 4)
                b2[id] = un[NS + 10] / un[NS];
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const double o = Un[q * n + id];
                if (q >= NS + 1 && q <= NS + 3)
                    un[q] = un[q] * 0.5 + o * 0.5;
                else
                    un[q] = 0.5 * un[q] + 0.5 * o;
            }
            if constexpr (MIXED) {
                float uf[NF];
#pragma unroll
                for (int q = 0; q < NF; ++q)
                    uf[q] = float(un[q]);
                const float ml =
                    1.0f / tmin(tmin(float(cg.e[0]), float(cg.e[1])), float(cg.e[2]));
                sig = double(signal_speed_turb<float, NS, GAS>(P, uf, ml));
            } else {
                sig = signal_speed_turb<double, NS, GAS>(P, un, cg.minlen);
            }
        }
#pragma unroll
        for (int q = 0; q < NF; ++q)
            Uout[q * n + id] = un[q];
    }
    if (err && act)
        atomicOr(errflag, err);
    if constexpr (SV == SV_RK2)
        reduce_signal<256>(sig, sRed, sigmax);
}

}  // namespace sg
