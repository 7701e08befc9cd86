// fv_stage_lean.hpp -- the mixed-precision stage kernel of the ideal mixture
// for schemes whose face values are the cell values (zero limiter; the TVD
// limiters in 2-D/3-D, whose multi-dimensional ratio has an exact zero entry:
// Schemi/src/limiters/TVDLimiter.cpp:12-125), with free, slip or periodic
// boundaries.  One forward-Euler stage of Advection2dOrder
// (Schemi/src/Advections/Advection2dOrder.cpp:16-636): EOS closure of every
// cell, HLLC/HLL flux of every face, residual and update; stage 2 also forms
// the SSP-RK2 average (bunchOfFields.hpp:166-203) and the Courant signal
// speed (Schemi.cpp:807-833).
//
// Work layout.  A workgroup owns a 64 x TY tile of columns and marches along
// z.  Wavefront w is tile row w: lane l owns cell (x0 + l, y0 + w) of the
// current plane and evaluates the low x, low y and high z faces of its cell,
// so every face-flux evaluation of a wavefront is uniform (no task loop, no
// divergent boundary branch).  The tile's two extra face lines -- the high x
// seam (one face per row) and the high y seam (one face per column) -- go to
// wavefronts 1 and 0.  Closed states cross lanes and rows through LDS, the
// own column's z faces stay in registers; two workgroup barriers per plane.
//
// Precision: the closure and the Riemann solve are FP32; the residual sum
// (six faces in the reference's order x-, x+, y+, y-, z-, z+), the update,
// gravity, the species renormalisation and the RK average are FP64.
//
// Closed state (LeanZ): partial densities rho_k, rho, momentum m, velocity v,
// p, E and the sound speed, all straight from the conserved values: with the
// species renormalised (sum_k rho_k = rho) the reference's face quantities
// (rho = sum_k c_k M_k, U = p / g1, E = U + rho v^2 / 2) equal the cell's own
// to FP32 rounding.  Boundary ghosts (Advection2dOrder.cpp:398-521) of free
// and slip walls are the owner's closed state with the normal velocity and
// momentum reflected on slip walls (slipFunction.cpp:15-35): the closure of
// the ghost's primitives gives exactly these values.
#pragma once

#include "fv_stage.hpp"

namespace sg {

template <int NS> struct LeanZ {
    static constexpr int R0 = NS, M = NS + 1, V = NS + 4, PR = NS + 7, E = NS + 8,
                         CS = NS + 9, NV = NS + 10;
};

// Closed state of a cell from its FP64 conserved values (ideal mixture,
// mixtureIdeal.cpp:383-409, 495-501): c_0 = sum_k rho_k / M_k,
// g1 = R c_0 / sum_k rho_k Cv_k / M_k, U = E - rho v^2 / 2, p = g1 U,
// c^2 = g1 (U + p) / rho.
// The translation unit is compiled without contraction and the closure and
// the fluxes below place their multiply-adds explicitly: a cell closed by its
// own tile and as another tile's halo, and a seam face evaluated by both tiles
// that share it, then give the same bits, so every face flux leaves one cell
// exactly as it enters the other.
template <int NS>
__device__ __forceinline__ void lean_state(const DevParams &P, const double *u, float *z)
{
    using L = LeanZ<NS>;
    float c0 = 0.0f, ccv = 0.0f;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const float r = float(u[k]);
        z[k] = r;
        c0 = __builtin_fmaf(r, P.iMf[k], c0);
        ccv = __builtin_fmaf(r, P.CvMf[k], ccv);
    }
    const float r0 = float(u[NS]);
    const float ir = 1.0f / r0;
    const float mx = float(u[NS + 1]), my = float(u[NS + 2]), mz = float(u[NS + 3]);
    const float E = float(u[NS + 4]);
    const float vx = mx * ir, vy = my * ir, vz = mz * ir;
    const float v2 = __builtin_fmaf(vx, vx, __builtin_fmaf(vy, vy, vz * vz));
    const float U = __builtin_fmaf(-0.5f * r0, v2, E);
    const float g1 = P.Rf * c0 / ccv;
    const float p = g1 * U;
    z[L::R0] = r0;
    z[L::M] = mx;
    z[L::M + 1] = my;
    z[L::M + 2] = mz;
    z[L::V] = vx;
    z[L::V + 1] = vy;
    z[L::V + 2] = vz;
    z[L::PR] = p;
    z[L::E] = E;
    z[L::CS] = __builtin_sqrtf(g1 * (U + p) * ir);
}

// Slip-wall ghost: the normal velocity and momentum reflected where on.
template <int NS>
__device__ __forceinline__ void lean_reflect(float *z, int ax, bool on)
{
    using L = LeanZ<NS>;
    z[L::V + ax] = on ? -z[L::V + ax] : z[L::V + ax];
    z[L::M + ax] = on ? -z[L::M + ax] : z[L::M + ax];
}

// HLLC flux F & N of one face along axis ax with normal sign sg (+1 inner
// and high-side faces, -1 low-side boundary faces); O owner side, B
// neighbour side (HLLCSolver.cpp:52-388: Davis wave speeds, Toro's contact
// speed, averaged star pressure, four wave patterns).  The star flux of
// side X, F(X) + S (U*_X - U_X), is written as
//   rho_k w, m w + sg N (p + b (P_LR - p)), E w + p u + b (P_LR S_c - p u)
// with b = S / (S - S_c), w = u + b (S_c - u); b = 0 gives F(X).  Returns
// false when no wave pattern applies (NaN speeds).
template <int NS>
__device__ __forceinline__ bool lean_hllc(const float *O, const float *B, int ax, float sg,
                                          float *g)
{
    using L = LeanZ<NS>;
    const float uO = sg * O[L::V + ax], uB = sg * B[L::V + ax];
    const float SO = tmin(uO - O[L::CS], uB - B[L::CS]);
    const float SB = tmax(uO + O[L::CS], uB + B[L::CS]);
    const float dO = O[L::R0] * (SO - uO), dB = B[L::R0] * (SB - uB);
    const float num = __builtin_fmaf(-dB, uB, __builtin_fmaf(dO, uO, B[L::PR] - O[L::PR]));
    const float Sc = num / (dO - dB);
    const float PLR = 0.5f * __builtin_fmaf(dB, Sc - uB,
                                            __builtin_fmaf(dO, Sc - uO, O[L::PR] + B[L::PR]));
    const bool own = (SO >= 0.0f) || (Sc >= 0.0f);
    const bool star = own ? !(SO >= 0.0f) : (SB > 0.0f);
    const bool ok = own || (SB > 0.0f) || (SB <= 0.0f);
    const float u = own ? uO : uB;
    const float S = own ? SO : SB;
    const float b = star ? S / (S - Sc) : 0.0f;
    const float w = __builtin_fmaf(b, Sc - u, u);
    const float p = own ? O[L::PR] : B[L::PR];
    const float pn = __builtin_fmaf(b, PLR - p, p);
#pragma unroll
    for (int k = 0; k <= NS; ++k)
        g[k] = (own ? O[k] : B[k]) * w;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float m = own ? O[L::M + a] : B[L::M + a];
        g[NS + 1 + a] = (a == ax) ? __builtin_fmaf(m, w, sg * pn) : m * w;
    }
    const float E = own ? O[L::E] : B[L::E];
    const float pu = p * u;
    g[NS + 4] = __builtin_fmaf(E, w, __builtin_fmaf(b, __builtin_fmaf(PLR, Sc, -pu), pu));
    return ok;
}

// HLL flux (HLLSolver.cpp:65-296): F(O) for S_O >= 0, F(B) for S_B <= 0,
// otherwise (S_B F(O) - S_O F(B) + S_B S_O (U_B - U_O)) / (S_B - S_O),
// evaluated as one weighted sum.
template <int NS>
__device__ __forceinline__ bool lean_hll(const float *O, const float *B, int ax, float sg,
                                         float *g)
{
    using L = LeanZ<NS>;
    const float uO = sg * O[L::V + ax], uB = sg * B[L::V + ax];
    const float SO = tmin(uO - O[L::CS], uB - B[L::CS]);
    const float SB = tmax(uO + O[L::CS], uB + B[L::CS]);
    const bool lo = SO >= 0.0f, hi = SB <= 0.0f;
    const bool ok = lo || (SB > 0.0f) || hi;
    const float id = 1.0f / (SB - SO);
    const float wO = lo ? 1.0f : (hi ? 0.0f : SB * id);
    const float wB = lo ? 0.0f : (hi ? 1.0f : -SO * id);
    const float wd = (lo || hi) ? 0.0f : SB * SO * id;
    auto fl = [&](const float *X, float u, int q) -> float {
        if (q <= NS)
            return X[q] * u;
        if (q < NS + 4) {
            const int a = q - NS - 1;
            return (a == ax) ? __builtin_fmaf(X[L::M + a], u, sg * X[L::PR]) : X[L::M + a] * u;
        }
        return (X[L::E] + X[L::PR]) * u;
    };
    auto cons = [&](const float *X, int q) -> float {
        return (q <= NS) ? X[q] : (q < NS + 4 ? X[L::M + q - NS - 1] : X[L::E]);
    };
#pragma unroll
    for (int q = 0; q < NS + 5; ++q)
        g[q] = __builtin_fmaf(wd, cons(B, q) - cons(O, q),
                              __builtin_fmaf(wO, fl(O, uO, q), wB * fl(B, uB, q)));
    return ok;
}

template <int NS>
__device__ __forceinline__ bool lean_flux(int flux, const float *O, const float *B, int ax,
                                          float sg, float *g)
{
    if (flux == FLUX_HLL)
        return lean_hll<NS>(O, B, ax, sg, g);
    return lean_hllc<NS>(O, B, ax, sg, g);
}

// 1 / x in FP64 from the hardware reciprocal and two Newton steps (the mixed
// update's dt / V and the renormalisation factor; a few ulps).
__device__ __forceinline__ double rcp_f64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Streaming hints: the new state and U^n are touched once per stage, so
// their lines are marked non-temporal and leave L2 room for the rows that
// neighbouring tiles read again as their halo.
#ifndef SG_LEAN_NT_STORE
#define SG_LEAN_NT_STORE 1
#endif
#ifndef SG_LEAN_NT_UN
#define SG_LEAN_NT_UN 1
#endif

template <int NS, int DIM, int SV, int TY>
__global__ void __launch_bounds__(64 * TY, (TY == 8) ? 4 : 2)
stage_lean_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
                  const double *__restrict__ Uin, double *Uout, const double *Un,
                  const double *__restrict__ dt_ptr, unsigned long long *sigmax, int *errflag,
                  int kchunk)
{
    constexpr int TX = 64;
    constexpr int NT = TX * TY;
    using L = LeanZ<NS>;
    constexpr int NV = L::NV;
    constexpr int NF = NS + 5;
    // closed states of the plane: rows 0 and TY + 1 and columns 0 and TX + 1
    // are the halo
    __shared__ float sZ[NV][TY + 2][TX + 2];
    // face fluxes: x faces x0 + c (c = 0..TX) of each row, y faces y0 + r
    // (r = 0..TY) of each column
    __shared__ float sFx[NF][TY][TX + 1];
    __shared__ float sFy[NF][TY + 1][TX];
    __shared__ double sRed[NT / 64];
    // the own column's z+ flux waits here between its evaluation and the
    // update (fewer live registers)
    __shared__ float sFz[(DIM == 3) ? NF : 1][TY][TX];

    const DevParams &P = *Pp;
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    const int tid = threadIdx.x;
    const int tx = tid & (TX - 1), ty = tid / TX;
    const int Nx = P.N[0], Ny = P.N[1], Nz = P.N[2];
    const long long n = P.n;
    const long long sy = P.stride[1], sz = P.stride[2];
    int bx, by, bz;
    tile_coords(bx, by, bz);
    const int x0 = bx * TX, y0 = by * TY;
    const int kb = (DIM == 3) ? bz * kchunk : 0;
    const int ke = (DIM == 3) ? min(kb + kchunk, Nz) : 1;
    const int i = x0 + tx, j = y0 + ty;
    const double dt = *dt_ptr;
    const int flux = P.flux;
    const bool xper = P.periodic[0] != 0, yper = P.periodic[1] != 0;
    const bool zper = (DIM == 3) && P.periodic[2] != 0;
    // velocity slip on each patch (0 tail, 1 point, 2 bottom, 3 right, 4 left,
    // 5 top); the other fields are free or slip, both the identity
    const bool slip_x0 = P.bct[0][NS] == BC_SLIP, slip_x1 = P.bct[1][NS] == BC_SLIP;
    const bool slip_y0 = P.bct[3][NS] == BC_SLIP, slip_y1 = P.bct[4][NS] == BC_SLIP;
    const bool slip_z0 = P.bct[2][NS] == BC_SLIP, slip_z1 = P.bct[5][NS] == BC_SLIP;
    const bool xlo_tile = (x0 == 0) && !xper;
    const bool xhi_tile = (x0 + TX == Nx) && !xper;
    const bool ylo_tile = (y0 == 0) && !yper;
    const bool yhi_tile = (y0 + TY == Ny) && !yper;
    const float zsg_hi = P.hi_flip[2] ? -1.0f : 1.0f;
    int err = 0;
    double sig = 0.0;

    const double ex = geo[P.off_e[0] + i];
    const double ey = geo[P.off_e[1] + j];

    // halo of plane kk into LDS: rows y0 - 1 and y0 + TY (2 TX cells), then
    // columns x0 - 1 and x0 + TX (2 TY cells); wrapped on periodic axes,
    // absent at walls
    auto stage_halo = [&](int kk) {
        for (int h = tid; h < 2 * TX + 2 * TY; h += NT) {
            int row, col, gx, gy;
            bool ok = true;
            if (h < 2 * TX) {
                const int side = h / TX;
                col = h % TX + 1;
                row = side ? TY + 1 : 0;
                gx = x0 + col - 1;
                gy = side ? y0 + TY : y0 - 1;
                if (gy < 0 || gy >= Ny) {
                    if (yper)
                        gy = (gy < 0) ? Ny - 1 : 0;
                    else
                        ok = false;
                }
            } else {
                const int hh = h - 2 * TX;
                const int side = hh / TY;
                row = hh % TY + 1;
                col = side ? TX + 1 : 0;
                gx = side ? x0 + TX : x0 - 1;
                gy = y0 + row - 1;
                if (gx < 0 || gx >= Nx) {
                    if (xper)
                        gx = (gx < 0) ? Nx - 1 : 0;
                    else
                        ok = false;
                }
            }
            if (ok) {
                double u[NF];
                float z[NV];
                load_cons<NS>(Uin, n, gx + sy * gy + sz * (long long)kk, u);
                lean_state<NS>(P, u, z);
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    sZ[q][row][col] = z[q];
            }
        }
    };

    // prologue: the own cell of the first plane, its z- face, the first halo
    float Zc[NV];
    float fzm[NF];
    double uc[NF], unx[NF];
    load_cons<NS>(Uin, n, i + sy * j + sz * kb, uc);
    lean_state<NS>(P, uc, Zc);
    if constexpr (DIM == 3) {
        float zm[NV];
        if (kb > 0 || zper) {
            double u[NF];
            load_cons<NS>(Uin, n, i + sy * j + sz * ((kb > 0) ? kb - 1 : Nz - 1), u);
            lean_state<NS>(P, u, zm);
            if (!lean_flux<NS>(flux, zm, Zc, 2, 1.0f, fzm))
                err |= ERRF_RIEMANN;
        } else {
#pragma unroll
            for (int q = 0; q < NV; ++q)
                zm[q] = Zc[q];
            lean_reflect<NS>(zm, 2, slip_z0);
            if (!lean_flux<NS>(flux, Zc, zm, 2, -1.0f, fzm))
                err |= ERRF_RIEMANN;
        }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q)
        sZ[q][ty + 1][tx + 1] = Zc[q];
    stage_halo(kb);

    for (int k = kb; k < ke; ++k) {
        const bool znext = (DIM == 3) && (k + 1 < Nz || zper);
        if constexpr (DIM == 3) {
            if (k > kb) {
                // plane k was closed during the previous plane and parked in
                // the own slot of the LDS plane buffer
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    Zc[q] = sZ[q][ty + 1][tx + 1];
            }
            // conserved values of the next plane: in flight through the
            // face phase
            if (znext)
                load_cons<NS>(Uin, n, i + sy * j + sz * ((k + 1 < Nz) ? k + 1 : 0), unx);
        }
        __syncthreads();

        // ---- x- and y- faces of the own cell, the tile's seams -----------
        {
            float O[NV], B[NV], fx[NF];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                O[q] = sZ[q][ty + 1][tx];
                B[q] = Zc[q];
            }
            float sg = 1.0f;
            if (xlo_tile) {
                // lane 0: the low boundary face, owned by the cell, ghost
                // outside (normal -x)
                const bool b0 = tx == 0;
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    O[q] = b0 ? Zc[q] : O[q];
                lean_reflect<NS>(B, 0, b0 && slip_x0);
                sg = b0 ? -1.0f : 1.0f;
            }
            if (!lean_flux<NS>(flux, O, B, 0, sg, fx))
                err |= ERRF_RIEMANN;
#pragma unroll
            for (int q = 0; q < NF; ++q)
                sFx[q][ty][tx] = fx[q];
        }
        if constexpr (DIM >= 2) {
            float O[NV], B[NV], fy[NF];
            float sg = 1.0f;
            if (ylo_tile && ty == 0) {
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    O[q] = B[q] = Zc[q];
                lean_reflect<NS>(B, 1, slip_y0);
                sg = -1.0f;
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    O[q] = sZ[q][ty][tx + 1];
                    B[q] = Zc[q];
                }
            }
            if (!lean_flux<NS>(flux, O, B, 1, sg, fy))
                err |= ERRF_RIEMANN;
#pragma unroll
            for (int q = 0; q < NF; ++q)
                sFy[q][ty][tx] = fy[q];
            // high y seam (faces y0 + TY), wavefront 0
            if (ty == 0) {
                float g[NF];
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    O[q] = sZ[q][TY][tx + 1];
                    B[q] = yhi_tile ? O[q] : sZ[q][TY + 1][tx + 1];
                }
                lean_reflect<NS>(B, 1, yhi_tile && slip_y1);
                if (!lean_flux<NS>(flux, O, B, 1, 1.0f, g))
                    err |= ERRF_RIEMANN;
#pragma unroll
                for (int q = 0; q < NF; ++q)
                    sFy[q][TY][tx] = g[q];
            }
        }
        // high x seam (faces x0 + TX, one per row), lanes 0..TY-1 of
        // wavefront 1 (of wavefront 0 in 1-row tiles)
        if (ty == ((TY > 1) ? 1 : 0) && tx < TY) {
            const int r = tx;
            float O[NV], B[NV], g[NF];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                O[q] = sZ[q][r + 1][TX];
                B[q] = xhi_tile ? O[q] : sZ[q][r + 1][TX + 1];
            }
            lean_reflect<NS>(B, 0, xhi_tile && slip_x1);
            if (!lean_flux<NS>(flux, O, B, 0, 1.0f, g))
                err |= ERRF_RIEMANN;
#pragma unroll
            for (int q = 0; q < NF; ++q)
                sFx[q][r][TX] = g[q];
        }
        // ---- the next plane's closed state and the own z+ face -----------
        float Zn[NV];
        if constexpr (DIM == 3) {
            float fzp[NF];
            if (znext) {
                lean_state<NS>(P, unx, Zn);
                if (!lean_flux<NS>(flux, Zc, Zn, 2, 1.0f, fzp))
                    err |= ERRF_RIEMANN;
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    Zn[q] = Zc[q];
                lean_reflect<NS>(Zn, 2, slip_z1);
                if (!lean_flux<NS>(flux, Zc, Zn, 2, zsg_hi, fzp))
                    err |= ERRF_RIEMANN;
            }
#pragma unroll
            for (int q = 0; q < NF; ++q)
                sFz[q][ty][tx] = fzp[q];
        }
        __syncthreads();
        // no wavefront reads the plane buffer again before the next first
        // barrier: park the next plane's own cells and stage its halo
        if constexpr (DIM == 3) {
#pragma unroll
            for (int q = 0; q < NV; ++q)
                sZ[q][ty + 1][tx + 1] = Zn[q];
            if (k + 1 < ke)
                stage_halo(k + 1);
        }

        // ---- residual and update of the own cell (FP64) -------------------
        {
            const int kk = (DIM == 3) ? k : 0;
            const long long id = i + sy * j + sz * kk;
            // volume (ex ey) ez and face areas (mesh.cpp:1291-1293); in 2-D
            // ez is the extent of the single cell layer
            const double ez = geo[P.off_e[2] + kk];
            const double V = ex * ey * ez;
            const double S0 = ey * ez, S1 = ex * ez, S2 = ex * ey;
            const bool xlo = (i == 0) && !xper;
            const bool ylo = (DIM >= 2) && (j == 0) && !yper;
            const bool zlo = (DIM == 3) && (k == 0) && !zper;
            double un[NF];
            if constexpr (SV == SV_RK2) {
#pragma unroll
                for (int q = 0; q < NF; ++q)
                    un[q] = SG_LEAN_NT_UN ? __builtin_nontemporal_load(Un + q * n + id)
                                          : Un[q * n + id];
            }
            const double dtV = dt * rcp_f64(V);
            double unew[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const double f0 = double(sFx[q][ty][tx]);
                const double f1 = double(sFx[q][ty][tx + 1]);
                double acc = (f1 - (xlo ? -f0 : f0)) * S0;
                if constexpr (DIM >= 2) {
                    const double f2 = double(sFy[q][ty + 1][tx]);
                    const double f3 = double(sFy[q][ty][tx]);
                    acc = __builtin_fma(f2 - (ylo ? -f3 : f3), S1, acc);
                }
                if constexpr (DIM == 3) {
                    const double f4 = double(fzm[q]);
                    const double f5 = double(sFz[q][ty][tx]);
                    acc = __builtin_fma(f5 - (zlo ? -f4 : f4), S2, acc);
                }
                unew[q] = __builtin_fma(-acc, dtV, uc[q]);
            }
            if (P.gravity) {
                // pre-update rho and v (Advection2dOrder.cpp:547-554)
                const double rho = uc[NS];
                const double ir = rcp_f64(rho);
                const double vx = uc[NS + 1] * ir, vy = uc[NS + 2] * ir, vz = uc[NS + 3] * ir;
                const double vg = vx * P.g[0] + vy * P.g[1] + vz * P.g[2];
                unew[NS + 4] += rho * vg * dt;
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    unew[NS + 1 + b] += (P.g[b] * rho) * dt;
            }
            // species renormalisation: sum_k rho_k = rho
            double sr = 0.0;
#pragma unroll
            for (int q = 0; q < NS; ++q)
                sr += unew[q];
            const double f = unew[NS] * rcp_f64(sr);
#pragma unroll
            for (int q = 0; q < NS; ++q)
                unew[q] *= f;
            if constexpr (SV == SV_RK2) {
#pragma unroll
                for (int q = 0; q < NF; ++q)
                    unew[q] = __builtin_fma(0.5, unew[q], 0.5 * un[q]);
                // Courant signal speed (|v| + c) / min edge of the new state
                float z[NV];
                lean_state<NS>(P, unew, z);
                const float vm = __builtin_sqrtf(__builtin_fmaf(
                    z[L::V], z[L::V],
                    __builtin_fmaf(z[L::V + 1], z[L::V + 1], z[L::V + 2] * z[L::V + 2])));
                const float ml = 1.0f / tmin(tmin(float(ex), float(ey)), float(ez));
                sig = tmax(sig, double((vm + z[L::CS]) * ml));
            }
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                if (SG_LEAN_NT_STORE)
                    __builtin_nontemporal_store(unew[q], Uout + q * n + id);
                else
                    Uout[q * n + id] = unew[q];
            }
        }
        if constexpr (DIM == 3) {
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                uc[q] = unx[q];
                fzm[q] = sFz[q][ty][tx];
            }
        }
    }

    if (err)
        atomicOr(errflag, err);
    if constexpr (SV == SV_RK2)
        reduce_signal<NT>(sig, sRed, sigmax);
}

}  // namespace sg
