// fv_diffusion.hip -- molecular diffusion stage after each step
// (Diffusion.cpp:20-1061 with linear reconstruction, constant transport
// coefficients, no turbulence, ideal or van der Waals mixture): the explicit
// matrix solver or the implicit Laplacians with Jacobi sweeps, optionally
// with the molar-mass correction of the species fluxes and (explicit solver,
// ideal mixture) the enthalpy flow.
//
// Explicit solver: two passes over the cells, FP64 on both precision paths:
//  * diff_prims_kernel: mass fractions rho_k / rho renormalised by their sum
//    (diffusiveFields.cpp:30-47), concentrations, velocity, temperature and
//    pressure of every cell from the conserved state, into W;
//  * diff_update_kernel: per cell, the faces in the reference's order
//    (x-, x+, y+, y-, z-, z+): the face velocity (linearInterpolate) for
//    grad v and div v (gradient.hpp:19-58, divergence.hpp:19-58), the face
//    gradients surfGrad of Y_k, T and v and the face divergence of v
//    (gradient.hpp:96-165, divergence.hpp:84-150) with the boundary ghosts of
//    calculatedMassFraction and calculatedTemperature
//    (boundaryConditionValue.hpp:329-374), the face density sum_k c_k M_k
//    (Diffusion.cpp:156-166), then the explicit Laplacians
//    (SLEMatrix.cpp:2054-2142) of Y_k (rho D), v (mu, plus the divergence of
//    mu ((grad_s v)^T - 2/3 div_s v I)) and T (kappa, plus the viscous
//    heating (mu (grad v + grad v^T - 2/3 div v I)) : grad v), the clipping
//    and renormalisation of Y_k and the recalculation of the cell
//    (Diffusion.cpp:431-483, 795-828, 963-1029).  Each cell evaluates its
//    own faces with the face's owner/neighbour roles, so a face gives both
//    cells the same bits, and writes only itself: deterministic, no atomics
//    but the Courant maximum of the new state.
// Jacobi solver: the same primitives pass, diff_update_kernel<IMPL> assembling
// the implicit systems from the same face loop, batches of diff_jacobi_kernel
// sweeps, diff_finish_y_kernel (new concentrations, temperature system) and
// diff_finish_kernel (recalculated cells).
// Every face normal is +-1 along its axis; a face term of the reference
// (F & N) S then reduces to F_a N_a S plus signed zeros, which are dropped.
#include <cfloat>

#include "fv_launch.hpp"
#include "fv_stage.hpp"
#include "fv_general.hpp"
#include "fv_turb.hpp"
#include "fv_turb_launch.hpp"
#include "fv_star.hpp"

namespace sg {

namespace {

constexpr int NT = 256;

// Linear block id with each XCD (own L2) given a contiguous range of cells
// (SG_DIFF_XCD=1), so the y and z neighbours' primitives read by a block are
// mostly in its L2; measured 2 % slower on the 256x128x128 workload, off.
#ifndef SG_DIFF_XCD
#define SG_DIFF_XCD 0
#endif
// faces of a cell unrolled (all neighbour loads in flight at once) or one at
// a time (fewer registers, more waves; measured 45 % slower)
#ifndef SG_DIFF_UNROLL
#define SG_DIFF_UNROLL 1
#endif
__device__ __forceinline__ long long block_cell0()
{
    unsigned b = blockIdx.x;
#if SG_DIFF_XCD
    const unsigned nb = gridDim.x;
    if (nb % 8u == 0u)
        b = (b % 8u) * (nb / 8u) + b / 8u;
#endif
    return (long long)b * NT;
}

// W layout per cell: Y[NS], c[NS], v[3], T, p; with turbulence then k,
// epsilon, CvM = Cv / (rho / c0) and nonIdeality / rho (Diffusion.cpp:137-172);
// with the kEpsA and BHR models then a[3] and rho; with BHR then the second
// stage's b, the internal energy and the squared sound speed of the cell
template <int NS> struct DW {
    static constexpr int Y = 0, C = NS, V = 2 * NS, T = 2 * NS + 3, P = 2 * NS + 4,
                         NW = 2 * NS + 5, K = NW, EPS = NW + 1, CVM = NW + 2, NIR = NW + 3,
                         NWT = NW + 4, A = NWT, RHO = NWT + 3, NWA = NWT + 4, B = NWA,
                         UI = NWA + 1, CS2 = NWA + 2, NWB = NWA + 3;
};

// Cv_mix = sum_k (c_k / c0) Cv_k (mixtureIdeal.cpp:40-51)
template <int NS> __device__ __forceinline__ double cv_mix(const DevParams &P, const double *c,
                                                             double c0)
{
    double cv = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
        cv += (c[k] / c0) * P.Cv[k];
    return cv;
}

// Ideal or van der Waals closures of the stage: UvFromp, TFromUv
// (idealFluid.cpp:25-41; vanDerWaalsFluid.cpp:23-49), nonIdeality (0;
// -a c^2, vanDerWaalsFluid.cpp:78-82) and UvcFromT (Cv T; Cv T - a c,
// vanDerWaalsFluid.cpp:38-42) of a state with concentrations c, sum c0.
template <int NS>
__device__ __forceinline__ double u_from_p(const DevParams &P, const double *c, double c0,
                                           double p)
{
    const double cv = cv_mix<NS>(P, c, c0);
    if (P.gas_model == GAS_VDW) {
        double a, b;
        vdw_ab<double, NS>(P, c, c0, a, b);
        const double g1 = P.R / cv;
        const double acc = a * (c0 * c0);
        return (p + acc) * (1 - c0 * b) / g1 - acc;
    }
    return p / (P.R / cv);
}

template <int NS>
__device__ __forceinline__ double t_from_u(const DevParams &P, const double *c, double c0,
                                           double U)
{
    const double cv = cv_mix<NS>(P, c, c0);
    if (P.gas_model == GAS_VDW) {
        double a, b;
        vdw_ab<double, NS>(P, c, c0, a, b);
        return (U + a * (c0 * c0)) / ((c0 + STAB) * cv);
    }
    return U / ((c0 + STAB) * cv);
}

template <int NS>
__device__ __forceinline__ double non_ideality(const DevParams &P, const double *c, double c0)
{
    if (P.gas_model != GAS_VDW)
        return 0;
    double a, b;
    vdw_ab<double, NS>(P, c, c0, a, b);
    return -a * (c0 * c0);
}

template <int NS>
__device__ __forceinline__ double uvc_from_t(const DevParams &P, const double *c, double c0,
                                             double T)
{
    const double cv = cv_mix<NS>(P, c, c0);
    if (P.gas_model == GAS_VDW) {
        double a, b;
        vdw_ab<double, NS>(P, c, c0, a, b);
        return cv * T - a * c0;
    }
    return cv * T;
}

// Uses the buffer holding the new state: MUSCL-Hancock steps alternate
// between U and Ualt by the scalar block's parity.
__device__ __forceinline__ const double *state_of(const double *U, const double *Ualt,
                                                  const double *sc, int hk)
{
    return (hk && sc[SC_PARITY] == 0.0) ? Ualt : U;
}

// Positive doubles order like their bit patterns: atomicMax on the bits.
template <int NB>
__device__ __forceinline__ void reduce_max_bits(double r, unsigned long long *slot)
{
    __shared__ double sMax[NB / 64];
    const int tid = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1)
        r = tmax(r, __shfl_xor(r, off, 64));
    if ((tid & 63) == 0)
        sMax[tid >> 6] = r;
    __syncthreads();
    if (tid == 0) {
        double m = sMax[0];
        for (int w = 1; w < NB / 64; ++w)
            m = tmax(m, sMax[w]);
        atomicMax(slot, (unsigned long long)__double_as_longlong(m));
    }
}

// TURB: the state carries rho k and rho epsilon (U = E - rho v^2 / 2 - rho k);
// k, epsilon, CvM and nonIdeality / rho go to W, and the maximum of epsilon
// to *epsmax (shearModel.cpp:107-114; cleared after each step).  AM (a
// model: 1 kEpsA, 2 BHR): the state also carries rho a (and rho b); a and rho
// go to W, with BHR also the second stage's b (b2), U and c^2.
template <int NS, bool TURB = false, int AM = 0>
__global__ void __launch_bounds__(NT)
diff_prims_kernel(const DevParams *__restrict__ Pp, const double *U, const double *Ualt, int hk,
                  double *__restrict__ W, double *sc, unsigned long long *epsmax = nullptr,
                  const double *__restrict__ b2 = nullptr)
{
    const DevParams &P = *Pp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = blockIdx.x * (long long)NT + threadIdx.x;
    // the Courant maximum is taken again over the state after diffusion
    if (id == 0)
        reinterpret_cast<unsigned long long *>(sc)[SC_SIGMAX] = 0ull;
    if constexpr (TURB) {
        double e = 0.0;
        if (id < n)
            e = state_of(U, Ualt, sc, hk)[(NS + 6) * n + id] /
                state_of(U, Ualt, sc, hk)[NS * n + id];
        reduce_max_bits<NT>(e, epsmax);
    }
    if (id >= n)
        return;
    const double *Us = state_of(U, Ualt, sc, hk);
    constexpr int NA = AM == 2 ? 4 : AM == 1 ? 3 : 0;
    constexpr int NF = TURB ? NS + 7 + NA : NS + 5;
    double u[NF], w[NS + 6 + NA];
#pragma unroll
    for (int q = 0; q < NF; ++q)
        u[q] = Us[q * n + id];
    if constexpr (TURB)
        prims_turb<double, NS, -1, NA>(P, u, w);
    else
        prims_from_cons<double, NS>(P, u, w);
    using L = DW<NS>;
    double y[NS], sum = 0.0, c0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        y[k] = u[k] / u[NS];
        sum += y[k];
        c0 += w[k];
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        W[(L::Y + k) * n + id] = y[k] / sum;
        W[(L::C + k) * n + id] = w[k];
    }
#pragma unroll
    for (int b = 0; b < 3; ++b)
        W[(L::V + b) * n + id] = w[NS + b];
    const double vx = w[NS], vy = w[NS + 1], vz = w[NS + 2];
    const double v2 = ((0. + vx * vx) + vy * vy) + vz * vz;
    const double Uint = (u[NS + 4] - u[NS] * v2 * 0.5) - (TURB ? u[NS + 5] : 0.);
    W[L::T * n + id] = t_from_u<NS>(P, w, c0, Uint);
    W[L::P * n + id] = w[NS + 3];
    if constexpr (TURB) {
        W[L::K * n + id] = w[NS + 4];
        W[L::EPS * n + id] = w[NS + 5];
        const double avM = u[NS] / c0;
        W[L::CVM * n + id] = cv_mix<NS>(P, w, c0) / avM;
        W[L::NIR * n + id] = non_ideality<NS>(P, w, c0) / u[NS];
    }
    if constexpr (AM != 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            W[(L::A + c) * n + id] = w[NS + 6 + c];
        W[L::RHO * n + id] = u[NS];
    }
    if constexpr (AM == 2) {
        // sqSonicSpeed of the cell (BHRModel.cpp:182-184, 267-268)
        W[L::B * n + id] = b2[id];
        W[L::UI * n + id] = Uint;
        W[L::CS2 * n + id] =
            eos_c2<double, NS, -1>(P, w, c0, gamma1<double, NS>(P, w, c0), u[NS], Uint, w[NS + 3]);
    }
}

// Jacobi matrix solver (implicit Laplacians, SLEMatrix.cpp:410-547, 706-908):
// per cell, field-major (q n + id): the mass-fraction systems' diagonal
// (shared by the species: one rho D), right-hand sides and face coefficients
// (k_f / |d|) (N S / V & d / |d|) per face slot; the same for the velocity
// components (mu) and the temperature (kappa; its diagonal needs the new
// concentrations), the new concentrations, and two iterate buffers per field
// (Y_k, v_c, T).
template <int NS> struct MB {
    static constexpr int DY = 0, BY = NS, TY = 2 * NS, DV = 2 * NS + 6, BV = DV + 1, TV = DV + 4,
                         BT = TV + 6, TT = BT + 1, DT = TT + 6, CN = DT + 1, TB = CN + NS,
                         HT = TB + 6, X = HT + 1, NF = NS + 4, N = X + 2 * NF;
};

template <int NS, int NWX = DW<NS>::NW>
__device__ __forceinline__ void load_w(const double *W, long long n, long long id, double *w)
{
#pragma unroll
    for (int q = 0; q < NWX; ++q)
        w[q] = W[q * n + id];
}

// Molar-mass corrected species flux of one face component (caclulateDFluxes,
// transportCoefficients.hpp:285-395 with D_kl = D): the rows of all species
// but the least concentrated one balance the mole-fraction gradient, that
// row fixes sum_k c_k M_k flux_k = 0; GaussEliminationSolver
// (transportCoefficients.hpp:397-455: one row swap per column by the largest
// following entry, stabilised pivots, normalize); flux_k = phi_k M_k c_k.
template <int NS>
__device__ __forceinline__ void species_flux(const DevParams &P, const double *cs,
                                             const double *gXa, double Dc, double *DF)
{
    if constexpr (NS == 1) {
        DF[0] = 0;
        return;
    } else {
        double N0 = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            N0 += cs[k];
        double cmin = cs[0];
#pragma unroll
        for (int k = 1; k < NS; ++k)
            if (cs[k] < cmin)
                cmin = cs[k];
        int rep = 0;
#pragma unroll
        for (int k = NS - 1; k >= 0; --k)
            if (cmin == cs[k])
                rep = k + 1;
        double A[NS][NS], b[NS];
#pragma unroll
        for (int k1 = 0; k1 < NS; ++k1) {
            b[k1] = -gXa[k1];
#pragma unroll
            for (int k2 = 0; k2 < NS; ++k2)
                A[k1][k2] = 0;
        }
#pragma unroll
        for (int k1 = 1; k1 <= NS; ++k1) {
            if (k1 != rep) {
#pragma unroll
                for (int k2 = 1; k2 <= NS; ++k2) {
                    if (k1 == k2)
                        continue;
                    const double Aij = cs[k1 - 1] * cs[k2 - 1] / (N0 * N0 * Dc);
                    A[k1 - 1][k2 - 1] = -Aij;
                    A[k1 - 1][k1 - 1] += Aij;
                }
            } else {
                b[k1 - 1] = 0;
#pragma unroll
                for (int k2 = 1; k2 <= NS; ++k2)
                    A[k1 - 1][k2 - 1] = cs[k2 - 1] * P.M[k2 - 1];
            }
        }
#pragma unroll
        for (int k = 0; k < NS - 1; ++k) {
            double mv = A[k][k];
            int c = k;
#pragma unroll
            for (int r = k + 1; r < NS; ++r)
                if (A[r][k] > mv) {
                    mv = A[r][k];
                    c = r;
                }
            if (c != k) {
#pragma unroll
                for (int r = k + 1; r < NS; ++r)
                    if (r == c) {
#pragma unroll
                        for (int j = 0; j < NS; ++j) {
                            const double t = A[r][j];
                            A[r][j] = A[k][j];
                            A[k][j] = t;
                        }
                        const double t = b[r];
                        b[r] = b[k];
                        b[k] = t;
                    }
            }
        }
#pragma unroll
        for (int k = 0; k < NS - 1; ++k)
#pragma unroll
            for (int i = k + 1; i < NS; ++i) {
                const double ratio = A[i][k] / (A[k][k] + STAB);
#pragma unroll
                for (int j = k + 1; j < NS; ++j)
                    A[i][j] = A[i][j] - ratio * A[k][j];
                b[i] = b[i] - ratio * b[k];
            }
        double phi[NS];
        phi[NS - 1] = b[NS - 1] / (A[NS - 1][NS - 1] + STAB);
#pragma unroll
        for (int i = NS - 2; i >= 0; --i) {
            double term = 0.0;
#pragma unroll
            for (int j = i + 1; j < NS; ++j)
                term += A[i][j] * phi[j];
            phi[i] = (b[i] - term) / (A[i][i] + STAB);
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            DF[k] = (tabs(phi[k]) < DBL_EPSILON ? 0.0 : phi[k]) * P.M[k] * cs[k];
    }
}

// Block minimum of non-negative doubles (or +inf), merged into *slot as the
// bits of a double with atomicMin (integer order equals value order there).
template <int NB>
__device__ __forceinline__ void reduce_min_bits(double r, unsigned long long *slot)
{
    __shared__ double sMin[NB / 64];
    const int tid = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1)
        r = tmin(r, __shfl_xor(r, off, 64));
    if ((tid & 63) == 0)
        sMin[tid >> 6] = r;
    __syncthreads();
    if (tid == 0) {
        double m = sMin[0];
        for (int w = 1; w < NB / 64; ++w)
            m = tmin(m, sMin[w]);
        atomicMin(slot, (unsigned long long)__double_as_longlong(m));
    }
}

// SPEC: the molar-mass correction of the species fluxes (mm) and the
// explicit enthalpy flow (en) are evaluated.  TURB (explicit solver, no SPEC):
// k-epsilon turbulence (Diffusion.cpp:129-135, 242-286, 414-417, 580-598,
// 711-862, 905-917, 999-1028; transportCoefficients.hpp:65-83, 221-263): the
// face coefficients with the turbulent viscosity Cmu k^2 / eps of the face
// values (mu + rho nu_t, rho (D + nu_t / Sc_t), kappa + c Cv nu_t (1 / Pr_e +
// 1 / Pr_T)), the k and epsilon Laplacians (mu + rho nu_t / sigma, times
// thetaS_D of the face divergence), the non-ideality correction Laplacian,
// the dissipation rho eps and the filtered turbulent stress in the
// temperature equation, and the shear model's sources with its source step.
// AF (TURB, kEpsA; kEpsAModel.cpp:132-257, Diffusion.cpp:176-191, 288-305,
// 511-538, 574, 736-800, 869-887, 1005-1016): the algebraic b of the cell,
// grad a, grad p and grad rho (Green-Gauss; rho on a boundary face from the
// surface concentrations), the divergence of the molecular viscous stress
// mu (grad_s v + grad_s v^T - 2/3 div_s v I), the a Laplacian (mu +
// rho nu_t / sigma_a), the k, epsilon and a sources with the thetaA limiter,
// and the energy sink a . (grad p - div tau) distributed into the
// temperature equation.  AM: 0 no a field, 1 kEpsA, 2 BHR (BHRModel.cpp:
// 139-272, the Goncharov initialisation model off): also the b Laplacian
// (mu + rho nu_t / sigma_b) of the second stage's b, rho div(a a), the
// turbulent Mach number 2k/c^2 in the dissipation, Pi_K, the dissipation
// rho eps (1 + CMS_M M^2) in the temperature equation and the |U| / G source
// step; the explicit solver leaves b's own sources out.
template <int NS, int DIM, bool MIXED, bool IMPL, bool SPEC, bool TURB = false, int AM = 0>
__global__ void __launch_bounds__(NT)
diff_update_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo, double *U,
                   double *Ualt, int hk, const double *__restrict__ W, double *sc,
                   int *errflag, double mu, double Dc, double kappa, double *Mb, int mm, int en,
                   const TurbParams *__restrict__ Tp = nullptr,
                   const unsigned long long *__restrict__ epsmax = nullptr,
                   const double *__restrict__ star = nullptr, double *__restrict__ bk = nullptr)
{
    static_assert(!(TURB && (IMPL || SPEC)), "turbulence: explicit solver without species terms");
    constexpr bool AF = AM != 0;
    static_assert(TURB || !AF, "the a field needs turbulence");
    using M = MB<NS>;
    using L = DW<NS>;
    constexpr int NW = TURB ? (AM == 2 ? L::NWB : AF ? L::NWA : L::NWT) : L::NW;
    __shared__ double sRed[NT / 64];
    const DevParams &P = *Pp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = block_cell0() + threadIdx.x;
    double sig = 0.0;
    double rsrc = __builtin_inf();  // source-step ratio of the enthalpy flow
    double *Us = const_cast<double *>(state_of(U, Ualt, sc, hk));
    const double dt = sc[SC_DT];
    if (id < n) {
        const int idx[3] = {int(id % P.N[0]), int((id / P.N[0]) % P.N[1]),
                            int(id / P.stride[2])};
        double w0[NW];
        load_w<NS, NW>(W, n, id, w0);
        const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
        double gV[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dV = 0;
        double dY[NS], dT = 0, dFv[3] = {0, 0, 0}, dFs[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < NS; ++k)
            dY[k] = 0;
        int err = 0;
        double dMC[NS], dEF = 0, c0o = 0;
        double dK = 0, dE = 0, dNI = 0;  // turbulence: k, epsilon, non-ideality Laplacians
        // kEpsA: Green-Gauss sums of grad a, grad p, grad rho; the a Laplacian
        // and the divergence of the molecular viscous stress
        double gA[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gP[3] = {0, 0, 0}, gR[3] = {0, 0, 0};
        double dA[3] = {0, 0, 0}, dDP[3] = {0, 0, 0};
        double dB = 0, dAA[3] = {0, 0, 0};  // BHR: b Laplacian, div(a a)
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            dMC[k] = 0;
            c0o += w0[L::C + k];
        }
        // implicit Laplacians: diagonals and right-hand sides start from the
        // time derivative, rnew (1 / dt) and rold (1 / dt)
        double diagY = 0, diagV = 0, bY[NS], bV[3], bT = 0, tY[6], tV[6], tT[6], tB[6];
        double rdt = 0, rho_i = 0;
        if constexpr (IMPL) {
            rdt = 1 / dt;
            rho_i = Us[NS * n + id];
            diagY = rho_i * rdt;
            diagV = rho_i * rdt;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                bY[k] = (w0[L::Y + k] * rho_i) * rdt;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                bV[c] = (w0[L::V + c] * rho_i) * rdt;
            double c0o = 0;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                c0o += w0[L::C + k];
            // r_old = CCV_old T + nonIdeality_old - nonIdeality_new needs the new
            // concentrations: the boundary terms and the heating are kept
            // and the right-hand side is formed after the species solves
            bT = 0;
#pragma unroll
            for (int s = 0; s < 6; ++s)
                tY[s] = tV[s] = tT[s] = tB[s] = 0;
        }
#if SG_DIFF_UNROLL
#pragma unroll
#else
#pragma unroll 1
#endif
        for (int s = 0; s < 2 * DIM; ++s) {
            int a, f;
            bool hi;
            double d[3];
            slot_offset(P, geo, s, idx, d, a, hi, f);
            const int Na = P.N[a];
            const bool per = P.periodic[a] != 0;
            const bool inner = per || (hi ? (idx[a] < Na - 1) : (idx[a] > 0));
            // the wrapped face of a periodic axis (owner N-1, neighbour 0)
            const bool wrap = per && (hi ? (idx[a] == Na - 1) : (idx[a] == 0));
            // outward normal along a: the reference points a face normal away
            // from its owner's centre (mesh.cpp:34-53)
            const bool flip = hi && !inner && P.hi_flip[a] != 0;
            const double na = (hi && !flip) ? 1.0 : -1.0;
            const double S = cg.S[a];
            double vs[3], srho = 0;
            // owner / neighbour values (ghost for a boundary face)
            double qo[NW], qn[NW];
            double gY[NS][3], gT[3], gv[3][3], divs;
            double mf = 1, dotf = 0;  // |d| and (N S / V) & d / |d| of the Laplacian
            // species corrections: face concentrations, temperature, and the
            // normal component of surfGrad of the mole fractions
            double cs[NS], Ts = 0, gXa[NS];
            // turbulence: face k, epsilon, Cv and temperature, and the normal
            // components of surfGrad of k, epsilon, CvM and nonIdeality / rho
            double ks = 0, es = 0, cvs = 0, Tsf = 0, gk = 0, ge = 0, gcm = 0, gni = 0;
            // kEpsA: face a, p and rho, and the normal components of surfGrad a;
            // BHR: the interpolated face a of div(a a) and surfGrad b
            double as_[3] = {0, 0, 0}, ps_ = 0, rs_ = 0, ga_[3] = {0, 0, 0};
            double ai_[3] = {0, 0, 0}, gb_ = 0;
            if (inner) {
                const long long nb =
                    id + (wrap ? (hi ? -(long long)(Na - 1) : (long long)(Na - 1)) * P.stride[a]
                               : (hi ? P.stride[a] : -P.stride[a]));
                double wn[NW];
                load_w<NS, NW>(W, n, nb, wn);
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    qo[q] = hi ? w0[q] : wn[q];
                    qn[q] = hi ? wn[q] : w0[q];
                }
                const double wo = geo[P.off_wo[a] + f], wnn = geo[P.off_wn[a] + f];
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    vs[b] = qo[L::V + b] * wo + qn[L::V + b] * wnn;
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    cs[k] = qo[L::C + k] * wo + qn[L::C + k] * wnn;
                    srho += cs[k] * P.M[k];
                }
                // d = r_n - r_o along a; |d|^2 = pow<2>(hypot3(d)); a
                // periodic face spans its centre distance through the boundary
                const double da = wrap ? geo[P.off_dn[a] + f]
                                       : geo[P.off_xc[a] + f] - geo[P.off_xc[a] + f - 1];
                const double m = tabs(da);
                const double m2 = m * m;
                mf = m;
                dotf = cg.S[a] / cg.V;
                double dd[3] = {0.0, 0.0, 0.0};
                dd[a] = da;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        gY[k][b] = ((qn[L::Y + k] - qo[L::Y + k]) / m2) * dd[b];
                    gT[b] = ((qn[L::T] - qo[L::T]) / m2) * dd[b];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        gv[c][b] = ((qn[L::V + c] - qo[L::V + c]) / m2) * dd[b];
                }
                divs = ((qn[L::V + a] - qo[L::V + a]) / m2) * da;
                if constexpr (TURB) {
                    ks = qo[L::K] * wo + qn[L::K] * wnn;
                    es = qo[L::EPS] * wo + qn[L::EPS] * wnn;
                    double c0a = 0, c0b = 0;
#pragma unroll
                    for (int k = 0; k < NS; ++k) {
                        c0a += qo[L::C + k];
                        c0b += qn[L::C + k];
                    }
                    cvs = cv_mix<NS>(P, qo + L::C, c0a) * wo + cv_mix<NS>(P, qn + L::C, c0b) * wnn;
                    Tsf = qo[L::T] * wo + qn[L::T] * wnn;
                    gk = ((qn[L::K] - qo[L::K]) / m2) * da;
                    ge = ((qn[L::EPS] - qo[L::EPS]) / m2) * da;
                    gcm = ((qn[L::CVM] - qo[L::CVM]) / m2) * da;
                    gni = ((qn[L::NIR] - qo[L::NIR]) / m2) * da;
                }
                if constexpr (AF) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        as_[c] = qo[L::A + c] * wo + qn[L::A + c] * wnn;
                        ga_[c] = ((qn[L::A + c] - qo[L::A + c]) / m2) * da;
                    }
                    ps_ = qo[L::P] * wo + qn[L::P] * wnn;
                    rs_ = qo[L::RHO] * wo + qn[L::RHO] * wnn;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        ai_[c] = as_[c];
                }
                if constexpr (AM == 2)
                    gb_ = ((qn[L::B] - qo[L::B]) / m2) * da;
                if constexpr (SPEC) {
                    Ts = qo[L::T] * wo + qn[L::T] * wnn;
                    double c0a = 0, c0b = 0;
#pragma unroll
                    for (int k = 0; k < NS; ++k) {
                        c0a += qo[L::C + k];
                        c0b += qn[L::C + k];
                    }
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        gXa[k] = ((qn[L::C + k] / c0b - qo[L::C + k] / c0a) / m2) * da;
                }
            } else {
                const int p = patch_of(a, hi ? 1 : 0);
#pragma unroll
                for (int q = 0; q < NW; ++q)
                    qo[q] = w0[q];
                // face values: velocity and concentrations (boundary surface values)
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    vs[b] = bc_surface<double>(P.bct[p][NS], P.bcv[p][NS][b], w0[L::V + b], b, a);
                double cgh[NS], cg0 = 0, sumd = 0;
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    cs[k] = bc_surface<double>(P.bct[p][k], P.bcv[p][k][0], w0[L::C + k], -1, a);
                    srho += cs[k] * P.M[k];
                    cgh[k] = bc_cell<double>(P.bct[p][k], P.bcv[p][k][0], w0[L::C + k], -1, a);
                    cg0 += cgh[k];
                    sumd += cgh[k] * P.M[k];
                }
                // calculatedMassFraction: (c_k M_k) / sum_j c_j M_j of the ghost
#pragma unroll
                for (int k = 0; k < NS; ++k)
                    qn[L::Y + k] = (cgh[k] * P.M[k]) / sumd;
                // calculatedTemperature: TFromUv(c_g, UvFromp(c_g, p_g))
                const double pg =
                    bc_cell<double>(P.bct[p][NS + 1], P.bcv[p][NS + 1][0], w0[L::P], -1, a);
                qn[L::T] = t_from_u<NS>(P, cgh, cg0, u_from_p<NS>(P, cgh, cg0, pg));
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    qn[L::V + b] =
                        bc_cell<double>(P.bct[p][NS], P.bcv[p][NS][b], w0[L::V + b], b, a);
                if (P.bct[p][NS] == BC_FIXED_SURF || P.bct[p][NS + 1] == BC_FIXED_SURF)
                    err = 1;
#pragma unroll
                for (int k = 0; k < NS; ++k)
                    if (P.bct[p][k] == BC_FIXED_SURF)
                        err = 1;
                // D = 2 (r_f - r_o), |D| = hypot3(D), D / |D|
                double D[3];
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    D[b] = d[b] * 2.;
                const double m = hypot3<double>(D[0], D[1], D[2]);
                double nrm[3];
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    nrm[b] = D[b] / m;
                mf = m;
                dotf = ((na * cg.S[a]) / cg.V) * nrm[a];
                double dq[3];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    dq[c] = (qn[L::V + c] - qo[L::V + c]) / m;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        gY[k][b] = ((qn[L::Y + k] - qo[L::Y + k]) / m) * nrm[b];
                    gT[b] = ((qn[L::T] - qo[L::T]) / m) * nrm[b];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        gv[c][b] = dq[c] * nrm[b];
                }
                divs = ((0. + dq[0] * nrm[0]) + dq[1] * nrm[1]) + dq[2] * nrm[2];
                if constexpr (TURB) {
                    // surface values (linearInterpolate boundary values;
                    // calculatedCv and calculatedTemperature from the surface
                    // concentrations and pressure) and the ghosts (k, eps;
                    // calculatedCvM and calculatedNonidealityCorrectionPerDensity
                    // from the ghost concentrations, boundaryConditionValue.hpp:
                    // 415-484)
                    const TurbParams &TP = *Tp;
                    ks = bc_surface<double>(TP.bct[p][0], TP.bcv[p][0], w0[L::K], -1, a);
                    es = bc_surface<double>(TP.bct[p][1], TP.bcv[p][1], w0[L::EPS], -1, a);
                    double cs0 = 0;
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        cs0 += cs[k];
                    cvs = cv_mix<NS>(P, cs, cs0);
                    const double ps =
                        bc_surface<double>(P.bct[p][NS + 1], P.bcv[p][NS + 1][0], w0[L::P], -1, a);
                    Tsf = t_from_u<NS>(P, cs, cs0, u_from_p<NS>(P, cs, cs0, ps));
                    const double kg = bc_cell<double>(TP.bct[p][0], TP.bcv[p][0], w0[L::K], -1, a);
                    const double eg = bc_cell<double>(TP.bct[p][1], TP.bcv[p][1], w0[L::EPS], -1, a);
                    const double cmg = cv_mix<NS>(P, cgh, cg0) * cg0 / sumd;
                    const double nig = non_ideality<NS>(P, cgh, cg0) / sumd;
                    gk = ((kg - w0[L::K]) / m) * nrm[a];
                    ge = ((eg - w0[L::EPS]) / m) * nrm[a];
                    gcm = ((cmg - w0[L::CVM]) / m) * nrm[a];
                    gni = ((nig - w0[L::NIR]) / m) * nrm[a];
                    if (TP.bct[p][0] == BC_FIXED_SURF || TP.bct[p][1] == BC_FIXED_SURF)
                        err = 1;
                    if constexpr (AF) {
                        // a: surface values and ghosts of set/a.txt; p: its
                        // surface value; rho: calculatedDensity, sum_k c_k,f M_k
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            as_[c] = bc_surface<double>(TP.bcta[p], TP.bcva[p][c], w0[L::A + c], c, a);
                            const double ag =
                                bc_cell<double>(TP.bcta[p], TP.bcva[p][c], w0[L::A + c], c, a);
                            ga_[c] = ((ag - w0[L::A + c]) / m) * nrm[a];
                        }
                        ps_ = ps;
                        rs_ = srho;
                        if (TP.bcta[p] == BC_FIXED_SURF)
                            err = 1;
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            ai_[c] = as_[c];
                    }
                    if constexpr (AM == 2) {
                        const double bg = bc_cell<double>(TP.bctb[p], TP.bcvb[p], w0[L::B], -1, a);
                        gb_ = ((bg - w0[L::B]) / m) * nrm[a];
                        if (TP.bctb[p] == BC_FIXED_SURF)
                            err = 1;
                    }
                }
                if constexpr (SPEC) {
                    // calculatedTemperature on the surface; calculatedMolarFraction ghost
                    double cs0 = 0;
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        cs0 += cs[k];
                    const double ps =
                        bc_surface<double>(P.bct[p][NS + 1], P.bcv[p][NS + 1][0], w0[L::P], -1, a);
                    const double cvs = cv_mix<NS>(P, cs, cs0);
                    Ts = (ps / (P.R / cvs)) / ((cs0 + STAB) * cvs);
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        gXa[k] = ((cgh[k] / cg0 - w0[L::C + k] / c0o) / m) * nrm[a];
                }
            }
            if (star) {
                // #Reconstruction star: the step's star states of the face
                // for grad v, div v (and grad rho, grad p, grad a)
                const long long nft = P.star_nft;
                const long long fs = star_face(P, a, f, idx);
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    vs[c] = star[(1 + c) * nft + fs];
                if constexpr (AF) {
                    rs_ = star[fs];
                    ps_ = star[4 * nft + fs];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        as_[c] = star[(5 + c) * nft + fs];
                }
            }
            if constexpr (SPEC) {
                // molar-mass corrections (caclulateDFluxes with D_kl = D,
                // transportCoefficients.hpp:285-395) of the normal component,
                // and the enthalpy flow (Diffusion.cpp:379-421, 605-611)
                const double rhoD = srho * Dc;
                double Wa[NS], MCa[NS];
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    Wa[k] = (gY[k][a] * rhoD) * -1;
                    MCa[k] = 0;
                }
                if (mm) {
                    double DF[NS];
                    species_flux<NS>(P, cs, gXa, Dc, DF);
#pragma unroll
                    for (int k = 0; k < NS; ++k) {
                        MCa[k] = DF[k] - Wa[k];
                        dMC[k] += (MCa[k] * na) * S;
                    }
                }
                if (en) {
                    double e = 0;
#pragma unroll
                    for (int k = 0; k < NS; ++k) {
                        const double h = (P.Cv[k] * Ts + P.R * Ts) / Ts;
                        e += (Wa[k] + MCa[k]) * h / P.M[k];
                    }
                    dEF += ((e * Ts) * na) * S;
                }
            }
            // effective coefficients: with turbulence mu + rho nu_t,
            // rho (D + nu_t / Sc_t), kappa + c Cv nu_t / Pr_e + c Cv nu_t / Pr_T
            double mu_f = mu, kappa_f = kappa, tD = 0;
            if constexpr (TURB) {
                const TurbParams &TP = *Tp;
                double cs0 = 0;
#pragma unroll
                for (int k = 0; k < NS; ++k)
                    cs0 += cs[k];
                const double tnu = TP.Cmu * ks * ks / es;
                tD = tnu / TP.sigmaSc;
                const double tKap = tnu / TP.sigmaT;
                const double tLam = tnu / TP.sigmaE;
                mu_f = mu + srho * tnu;
                const double cC = cs0 * cvs;
                kappa_f = kappa + cC * tLam + cC * tKap;
                // k and epsilon Laplacians, thetaS_D of the face divergence
                const double th = (divs < 0.) ? theta_s(TP, divs, ks, es, TP.CMS_D) : 1.;
                const double kD = tnu / TP.sigmaK * th, eD = tnu / TP.sigmaEps * th;
                const double rhoDk = mu + srho * kD, rhoDe = mu + srho * eD;
                dK += (((gk * rhoDk) * -1) * na) * S;
                dE += (((ge * rhoDe) * -1) * na) * S;
                // -(surfGrad CvM T_f + surfGrad (nonIdeality / rho)) rho tLambda
                const double flow = -1 * (gcm * Tsf + gni);
                dNI += ((flow * srho * tLam) * na) * S;
                if constexpr (AF) {
                    // grad a (a_c N_b), grad p and grad rho (N_b q) sums; the a
                    // Laplacian; mu (grad_s v + grad_s v^T - 2/3 div_s v I) & N
                    const double rhoDa = mu + srho * (tnu / TP.sigmaa);
                    const double lam_s = (2. / 3.) * divs;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        gA[3 * c + a] += (as_[c] * na) * S;
                        dA[c] += (((ga_[c] * rhoDa) * -1) * na) * S;
                        double t = gv[c][a] + gv[a][c];
                        if (c == a)
                            t -= lam_s;
                        dDP[c] += ((t * mu) * na) * S;
                    }
                    gP[a] += (na * ps_) * S;
                    gR[a] += (na * rs_) * S;
                }
                if constexpr (AM == 2) {
                    // the b Laplacian with mu + rho nu_t / sigma_b; (a_f a_f) & N
                    const double rhoDb = mu + srho * (tnu / TP.sigmab);
                    dB += (((gb_ * rhoDb) * -1) * na) * S;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        dAA[c] += ((ai_[c] * ai_[a]) * na) * S;
                }
            }
            // Green-Gauss sums of grad v (v_c N_b) S and div v (v & N) S
#pragma unroll
            for (int c = 0; c < 3; ++c)
                gV[3 * c + a] += (vs[c] * na) * S;
            dV += (vs[a] * na) * S;
            // fluxes -coeff surfGrad, their normal component times S
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const double rhoD = srho * (Dc + tD);
                dY[k] += (((gY[k][a] * rhoD) * -1) * na) * S;
            }
            dT += (((gT[a] * kappa_f) * -1) * na) * S;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                dFv[c] += (((gv[c][a] * mu_f) * -1) * na) * S;
            // mu strainSurf1 & N: row c, column a of (grad_s v)^T - lam I
            const double lam = (2. / 3.) * divs;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double t = gv[a][c];
                if (c == a)
                    t -= lam;
                dFs[c] += ((t * mu_f) * na) * S;
            }
            if constexpr (IMPL) {
                // (k_f / |d|) ((N S / V) & (d / |d|)) into the diagonal, times
                // the ghost value into the right-hand side of a boundary face
                const double ty = (srho * Dc) / mf * dotf;
                const double tv = mu / mf * dotf;
                const double tt = kappa / mf * dotf;
                diagY += ty;
                diagV += tv;
                tY[s] = ty;
                tV[s] = tv;
                tT[s] = tt;
                tB[s] = 0;
                if (!inner) {
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        bY[k] += ty * qn[L::Y + k];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        bV[c] += tv * qn[L::V + c];
                    tB[s] = tt * qn[L::T];
                }
            }
        }
        if (err)
            atomicOr(errflag, ERRF_BOUNDARY);
        const double V = cg.V;
        double G[9];
#pragma unroll
        for (int q = 0; q < 9; ++q)
            G[q] = gV[q] / V;
        const double divV = dV / V;
        if constexpr (IMPL) {
            // explicit parts of the free terms: div(mu strainSurf1) for v,
            // the viscous heating for T (Diffusion.cpp:711-716, 737)
            double sv[9];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    sv[3 * a + b] = G[3 * b + a] + G[3 * a + b];
            const double lam = (2. / 3.) * divV;
            sv[0] -= lam;
            sv[4] -= lam;
            sv[8] -= lam;
            double heat = 0;
#pragma unroll
            for (int q = 0; q < 9; ++q)
                heat += (sv[q] * mu) * G[q];
            (void)bT;
            Mb[M::HT * n + id] = heat;
            Mb[M::DV * n + id] = diagV;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                double dk = diagY, bk = bY[k];
                if (SPEC && mm) {
                    const double src = (dMC[k] / V) * (-1.0);
                    if (src >= 0.)
                        bk += src;
                    else
                        dk -= src / (w0[L::Y + k] + STAB);
                }
                Mb[(M::DY + k) * n + id] = dk;
                Mb[(M::BY + k) * n + id] = bk;
                Mb[(M::X + 2 * k) * n + id] = w0[L::Y + k];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                Mb[(M::BV + c) * n + id] = bV[c] + dFs[c] / V;
                Mb[(M::X + 2 * (NS + c)) * n + id] = w0[L::V + c];
            }
            Mb[M::BT * n + id] = 0;
            Mb[(M::X + 2 * (NS + 3)) * n + id] = w0[L::T];
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                Mb[(M::TY + s) * n + id] = tY[s];
                Mb[(M::TV + s) * n + id] = tV[s];
                Mb[(M::TT + s) * n + id] = tT[s];
                Mb[(M::TB + s) * n + id] = tB[s];
            }
        } else {
        constexpr int NA = AM == 2 ? 4 : AM == 1 ? 3 : 0;
        constexpr int NF = TURB ? NS + 7 + NA : NS + 5;
        double u[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q)
            u[q] = Us[q * n + id];
        const double rho = u[NS];
        // mass fractions: (Y rho + (0 - div F) dt) / rho, clipped, renormalised
        double Yn[NS], sum = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            double ft = 0. - dY[k] / V;
            double diag = rho;
            if (SPEC && mm) {
                // distributeSourceTerm (SLEMatrix.cpp:2166-2175)
                const double src = (dMC[k] / V) * (-1.0);
                if (src >= 0.)
                    ft += src;
                else
                    diag -= src / (w0[L::Y + k] + STAB);
            }
            Yn[k] = (w0[L::Y + k] * rho + ft * dt) / diag;
            if (Yn[k] < 0.)
                Yn[k] = 0.0;
            sum += Yn[k];
        }
        double cn[NS], c0n = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            Yn[k] /= sum;
            cn[k] = Yn[k] * rho / P.M[k];
            c0n += cn[k];
        }
        double c0o = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            c0o += w0[L::C + k];
        const double CCVold = c0o * cv_mix<NS>(P, w0 + L::C, c0o);
        const double cvn = cv_mix<NS>(P, cn, c0n);
        const double CCVnew = cvn * c0n;
        // velocity
        double vn[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double ft = 0. - dFv[c] / V;
            ft += dFs[c] / V;
            vn[c] = (w0[L::V + c] * rho + ft * dt) / rho;
        }
        // temperature with the viscous heating
        double sv[9];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                sv[3 * a + b] = G[3 * b + a] + G[3 * a + b];
        const double lam = (2. / 3.) * divV;
        sv[0] -= lam;
        sv[4] -= lam;
        sv[8] -= lam;
        double heat = 0;
#pragma unroll
        for (int q = 0; q < 9; ++q)
            heat += (sv[q] * mu) * G[q];
        double ftT = 0. - dT / V;
        double diagT = CCVnew;
        if constexpr (TURB) {
            // non-ideality correction Laplacian, distributeSourceTerm
            // (Diffusion.cpp:604-611, SLEMatrix.cpp:2166-2175)
            const double src = (dNI / V) * (-1.0);
            if (src >= 0.)
                ftT += src;
            else
                diagT -= src / (w0[L::T] + STAB);
        }
        if (SPEC && en) {
            // source time step ratio C Cv_old T_old / |div + stab|
            // (Diffusion.cpp:696-700); the minimum over the cells times the
            // coefficient caps the next Courant step
            rsrc = (CCVold * w0[L::T]) / __builtin_fabs(dEF / V + STAB);
            const double src = (dEF / V) * -1;
            if (src >= 0.)
                ftT += src;
            else
                diagT -= src / (w0[L::T] + STAB);
        }
        ftT += heat;
        double kn = 0, en_ = 0, an[3] = {0, 0, 0}, bn = 0;
        if constexpr (TURB) {
            const TurbParams &TP = *Tp;
            const double k = w0[L::K], eps = w0[L::EPS];
            const double rk = u[NS + 5], reps = u[NS + 6];
            // dissipation rho eps (kEpsModels.cpp:76-82) and the filtered part
            // of the turbulent stress (Diffusion.cpp:586-598, 719-733)
            if constexpr (AM == 2)
                ftT += rho * eps * (1 + TP.CMSM * 2 * k / w0[L::CS2]);  // BHRModel.cpp:262-272
            else if (TP.model != TURB_ZERO)
                ftT += rho * eps;
            const double tnu = TP.Cmu * k * k / eps;
            const double thR = theta_s(TP, divV, k, eps, TP.CMS_R);
            const double tnu1 = (1. - thR) * tnu;
            double philt = 0, gsph = 0, gdev = 0;
#pragma unroll
            for (int q = 0; q < 9; ++q)
                philt += (sv[q] * rho * tnu1) * G[q];
            ftT += philt;
            // generation by the spherical (-(2/3) rho k I) and deviatoric
            // (rho nu_t thetaS_R strain) Reynolds stresses
            double sph[9], dev[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const double Iq = (q == 0 || q == 4 || q == 8) ? 1. : 0.;
                sph[q] = Iq * -(2. / 3.) * rk;
                dev[q] = sv[q] * rho * tnu * thR;
                gsph += sph[q] * G[q];
                gdev += dev[q] * G[q];
            }
            double sk = 0, se = 0;
            if (TP.model == TURB_SHEAR) {
                // shearModel.cpp:116-147
                const double ek = eps / k;
                const double dissip = -reps;
                const double Sk1 = gsph + gdev, Sk2 = dissip / k;
                const double Se1 = (TP.C1 * gdev + TP.C3 * gsph) * ek;
                const double Se2 = TP.C2 * dissip / k;
                sk = Sk1 + Sk2 * k;
                se = Se1 + Se2 * eps;
                // source step: eps below 1e-3 of its maximum is veryBig
                const double emax =
                    __longlong_as_double(static_cast<long long>(epsmax[0]));
                const double me = (eps < 1E-3 * emax) ? 1e-3 * __DBL_MAX__ : eps;
                rsrc = __builtin_fabs(TP.ts_coeff * me / ((Se1 + Se2 * eps) / rho + STAB));
            }
            if constexpr (AM == 2) {
                // BHRModel.cpp:188-252 with the generation flag 1
                const double av[3] = {w0[L::A], w0[L::A + 1], w0[L::A + 2]};
                const double b = w0[L::B];
                const double ek = eps / k;
                const double Mat2 = 2 * k / w0[L::CS2];
                const double Mat = __builtin_sqrt(Mat2);
                const double thS = theta_s(TP, G[0] + G[4] + G[8], k, eps, TP.CMS_R);
                double gdevA = 0, x[3], grav = 0;
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    gdevA += (dev[q] * thS) * G[q];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    x[c] = gP[c] / V - dDP[c] / V;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    grav += av[c] * x[c];
                grav = grav * 1.0;
                const double PiK =
                    (Mat * TP.alpha3 * reps - (TP.alpha2 * gdevA + 8 * Mat * TP.alpha4 * gsph)) * Mat;
                const double dissip = -reps;
                const double Sk1 = gsph + gdevA + grav + PiK;
                const double Sk2 = (1 + TP.CMSM * Mat2) * dissip / k;
                const double gpos = (grav < 0.0) ? 0.0 : grav;
                const double Se1 =
                    (TP.C1 * gdevA + TP.C3 * gsph + TP.C0 * gpos + TP.C4 * PiK) * ek;
                const double Se2 = TP.C2 * dissip / k;
                sk = Sk1 + Sk2 * k;
                se = Se1 + Se2 * eps;
                const double emax = __longlong_as_double(static_cast<long long>(epsmax[0]));
                const double me = (eps < 1E-3 * emax) ? 1e-3 * __DBL_MAX__ : eps;
                rsrc = __builtin_fabs(TP.ts_coeff * me / ((Se1 + Se2 * eps) / rho + STAB));
                const double rU =
                    __builtin_fabs(TP.ts_coeff * __builtin_fabs(w0[L::UI]) / (grav + STAB));
                rsrc = (rU < rsrc) ? rU : rsrc;
                double tau[9], D[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    tau[q] = 1.0 * (dev[q] * thS + sph[q]) / rho;
                    D[q] = gA[q] / V - G[q];
                }
                const double gr[3] = {gR[0] / V, gR[1] / V, gR[2] / V};
                const double Sa2 = -rho * ek * TP.Ca;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double bgp = x[c] * b;
                    const double tgr = tau[3 * c] * gr[0] + tau[3 * c + 1] * gr[1] + tau[3 * c + 2] * gr[2];
                    const double ragv =
                        D[c] * u[NS + 7] + D[3 + c] * u[NS + 8] + D[6 + c] * u[NS + 9];
                    const double redis = (dAA[c] / V) * rho;
                    const double sig = bgp + tgr + ragv + redis + Sa2 * av[c];
                    an[c] = (av[c] * rho + (0. - dA[c] / V) * dt) / rho + sig / rho * dt;
                }
                // b: the explicit Laplacian only (the explicit solver gives b
                // no sources, Diffusion.cpp:745-763), the minimal value
                bn = (b * rho + (0. - dB / V) * dt) / rho + 0. / rho * dt;
                if (bn < TP.minb)
                    bn = TP.minb;
                const double src = grav * (-1.0);
                if (src >= 0.)
                    ftT += src;
                else
                    diagT -= src / (w0[L::T] + STAB);
            }
            if constexpr (AM == 1) {
                // kEpsAModel.cpp:170-225.  b = rho sum x_k / (rho_k + chi rho) /
                // sum x_k rho_k / (rho_k + chi rho) - 1, clipped at 0
                // (kEpsAModel.cpp:232-257; the reference's field + scalar drops
                // the stabilizer there, fieldOperations.hpp:33-43)
                double num = 0, den = 0;
#pragma unroll
                for (int q = 0; q < NS; ++q) {
                    const double x = w0[L::C + q] / c0o;
                    const double r = u[q] + TP.chi * rho;
                    num += x / r;
                    den += x * u[q] / r;
                }
                double b = rho * num / den - 1.;
                if (b < 0.)
                    b = 0.0;
                // the phase keeps it as its b (Diffusion.cpp:525-527, 1006)
                if (bk)
                    bk[id] = b;
                const double av[3] = {w0[L::A], w0[L::A + 1], w0[L::A + 2]};
                const double ek = eps / k;
                // thetaS_R of trace(grad v); thetaA (kEpsAModel.cpp:16-27)
                const double thS = theta_s(TP, G[0] + G[4] + G[8], k, eps, TP.CMS_R);
                double thA = 1;
                if (TP.CMSA < 1E15) {
                    double aa = 0;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        aa += av[c] * av[c];
                    thA = (aa / (2 * k) < TP.CMSA && b < TP.CMSA && aa / (2 * k) < b)
                              ? 1
                              : TP.CaMax / TP.Ca;
                }
                double gdevA = 0, x[3], grav = 0;
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    gdevA += (dev[q] * thS) * G[q];
                // a . (grad p - div(mu strainSurf))
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    x[c] = gP[c] / V - dDP[c] / V;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    grav += av[c] * x[c];
                const double dissip = -reps;
                const double Sk1 = gsph + gdevA + grav, Sk2 = dissip / k;
                const double Se1 = (TP.C1 * gdevA + TP.C3 * gsph + TP.C0 * grav) * ek;
                const double Se2 = TP.C2 * dissip / k;
                sk = Sk1 + Sk2 * k;
                se = Se1 + Se2 * eps;
                const double emax = __longlong_as_double(static_cast<long long>(epsmax[0]));
                const double me = (eps < 1E-3 * emax) ? 1e-3 * __DBL_MAX__ : eps;
                rsrc = __builtin_fabs(TP.ts_coeff * me / ((Se1 + Se2 * eps) / rho + STAB));
                // a: explicit Laplacian plus b (grad p - div tau) + ((thetaS
                // devR + spherR) / rho) & grad rho + rho a & (grad a - grad v)
                // - rho (eps / k) Ca thetaA a (Diffusion.cpp:749-762, 869-887)
                double tau[9], D[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    tau[q] = (dev[q] * thS + sph[q]) / rho;
                    D[q] = gA[q] / V - G[q];
                }
                const double gr[3] = {gR[0] / V, gR[1] / V, gR[2] / V};
                const double Sa2 = -rho * ek * TP.Ca * thA;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double bgp = x[c] * b;
                    const double tgr = tau[3 * c] * gr[0] + tau[3 * c + 1] * gr[1] + tau[3 * c + 2] * gr[2];
                    const double ragv =
                        D[c] * u[NS + 7] + D[3 + c] * u[NS + 8] + D[6 + c] * u[NS + 9];
                    const double sig = bgp + tgr + ragv + Sa2 * av[c];
                    an[c] = (av[c] * rho + (0. - dA[c] / V) * dt) / rho + sig / rho * dt;
                }
                // the energy sink into the temperature, distributeSourceTerm
                // (Diffusion.cpp:780-799, SLEMatrix.cpp:2166-2175)
                const double src = grav * (-1.0);
                if (src >= 0.)
                    ftT += src;
                else
                    diagT -= src / (w0[L::T] + STAB);
            }
            // explicit Laplacians plus the sources (Diffusion.cpp:839-862),
            // minimal values (905-917)
            kn = (k * rho + (0. - dK / V) * dt) / rho + sk / rho * dt;
            en_ = (eps * rho + (0. - dE / V) * dt) / rho + se / rho * dt;
            if (kn < TP.mink)
                kn = TP.mink;
            if (en_ < TP.mineps)
                en_ = TP.mineps;
        }
        // oldEnergyField (Diffusion.cpp:484-485)
        const double oldE = CCVold * w0[L::T] + non_ideality<NS>(P, w0 + L::C, c0o) -
                            non_ideality<NS>(P, cn, c0n);
        const double Tn = (oldE + ftT * dt) / diagT;
        if (Tn < 0.)
            atomicOr(errflag, ERRF_NEG_T);
        // recalculated cell: rho_k = c_k M_k, m = v rho, U = Cv T c0,
        // E = (U + rho v.v / 2) + rho k
        double un[NF];
#pragma unroll
        for (int k = 0; k < NS; ++k)
            un[k] = cn[k] * P.M[k];
        un[NS] = rho;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            un[NS + 1 + c] = vn[c] * rho;
        const double Uin = uvc_from_t<NS>(P, cn, c0n, Tn) * c0n;
        const double v2 = ((0. + vn[0] * vn[0]) + vn[1] * vn[1]) + vn[2] * vn[2];
        double rkn = 0.0;
        if constexpr (TURB) {
            rkn = kn * rho;
            un[NS + 5] = rkn;
            un[NS + 6] = en_ * rho;
        }
#pragma unroll
        for (int c = 0; c < NA && c < 3; ++c)
            un[NS + 7 + c] = an[c] * rho;
        if constexpr (AM == 2)
            un[NS + 10] = bn * rho;
        un[NS + 4] = (Uin + rho * v2 * 0.5) + rkn;
#pragma unroll
        for (int q = 0; q < NF; ++q)
            Us[q * n + id] = un[q];
        (void)MIXED;
        if constexpr (TURB)
            sig = signal_speed_turb<double, NS, -1>(P, un, cg.minlen);
        else
            sig = signal_speed<double, NS>(P, un, cg.minlen);
        }
    }
    if constexpr (!IMPL) {
        reduce_signal<NT>(sig, sRed, reinterpret_cast<unsigned long long *>(sc) + SC_SIGMAX);
        if constexpr (SPEC)
            if (en)
                reduce_min_bits<NT>(rsrc, reinterpret_cast<unsigned long long *>(sc) + SC_SRCMIN);
        if constexpr (TURB)
            if (Tp->model == TURB_SHEAR || Tp->model == TURB_KEPSA || Tp->model == TURB_BHR)
                reduce_min_bits<NT>(rsrc, reinterpret_cast<unsigned long long *>(sc) + SC_SRCMIN);
    }
}

// Field f of the Jacobi solves: 0..NS-1 mass fractions, NS..NS+2 velocity
// components, NS+3 temperature.
template <int NS> struct JF {
    __device__ static int diag(int f) { return f < NS ? MB<NS>::DY + f : (f < NS + 3 ? MB<NS>::DV : MB<NS>::DT); }
    __device__ static int rhs(int f) { return f < NS ? MB<NS>::BY + f : (f < NS + 3 ? MB<NS>::BV + f - NS : MB<NS>::BT); }
    __device__ static int coef(int f) { return f < NS ? MB<NS>::TY : (f < NS + 3 ? MB<NS>::TV : MB<NS>::TT); }
};

// Field f converged with iteration it - 1 (relativeIterationDifference < 1e-12,
// abstractMatrixSolver.cpp:22-29) or reached the iteration limit.
__device__ __forceinline__ bool jacobi_stops(const double *dmax, int f, int it, int maxit)
{
    if (it <= 1)
        return false;
    if (it - 1 >= maxit)
        return true;
    return 100. * dmax[(long long)f * (maxit + 2) + it - 1] < 1E-12;
}

// Iteration it (from 1) of the Jacobi solves of fields f0 .. f1-1
// (JacobiSolver.cpp:12-84): x_i = (b_i + row_i) / a_ii with row_i summing
// -a_ij x_j over the lower triangle (neighbours of smaller index: z-, y-, x-
// from its last entry) and then the upper (z+, y+, x+); iterate it is
// written to buffer it & 1.  done[f] holds the last iteration of a finished
// field (0 while it runs).
template <int NS, int DIM>
__global__ void __launch_bounds__(NT)
diff_jacobi_kernel(const DevParams *__restrict__ Pp, double *Mb, double *dmax, int *done,
                   const double *sc, int f0, int f1, int it, int maxit)
{
    using M = MB<NS>;
    __shared__ double sRed[NT / 64];
    const DevParams &P = *Pp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = blockIdx.x * (long long)NT + threadIdx.x;
    const int Nx = P.N[0], Ny = P.N[1], Nz = P.N[2];
    const long long Nxy = (long long)Nx * Ny;
    for (int f = f0; f < f1; ++f) {
        if (done[f] != 0)
            continue;
        if (jacobi_stops(dmax, f, it, maxit)) {
            if (id == 0)
                done[f] = it - 1;
            continue;
        }
        double e = 0.0;
        if (id < n) {
            const int i = int(id % Nx), j = int((id / Nx) % Ny), k = int(id / Nxy);
            const double *xo = Mb + (M::X + 2 * f + ((it - 1) & 1)) * n;
            double *xn = Mb + (M::X + 2 * f + (it & 1)) * n;
            const double *t = Mb + JF<NS>::coef(f) * n;
            const double aii = 1. / Mb[JF<NS>::diag(f) * n + id];
            // neighbour of each face slot (x-, x+, y+, y-, z-, z+), -1 on a
            // boundary; periodic axes wrap
            long long nb[6];
            const int ii[3] = {i, j, k};
            const int NN[3] = {Nx, Ny, Nz};
            const long long st[3] = {1, Nx, Nxy};
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const int a = s >> 1;
                const bool hi = (s == 1 || s == 2 || s == 5);
                nb[s] = -1;
                if (a >= DIM)
                    continue;
                if (hi ? ii[a] < NN[a] - 1 : ii[a] > 0)
                    nb[s] = id + (hi ? st[a] : -st[a]);
                else if (P.periodic[a])
                    nb[s] = id + (hi ? -(long long)(NN[a] - 1) : (long long)(NN[a] - 1)) * st[a];
            }
            // lower triangle (smaller index), then upper, each from its last
            // entry (slot order)
            double row = 0;
#pragma unroll
            for (int s = 5; s >= 0; --s)
                if (nb[s] >= 0 && nb[s] < id)
                    row -= (-t[s * n + id]) * xo[nb[s]];
#pragma unroll
            for (int s = 5; s >= 0; --s)
                if (nb[s] > id)
                    row -= (-t[s * n + id]) * xo[nb[s]];
            const double x = (Mb[JF<NS>::rhs(f) * n + id] + row) * aii;
            xn[id] = x;
            e = tabs((x - xo[id]) / (tabs(x) + STAB));
        }
        reduce_signal<NT>(e, sRed,
                          reinterpret_cast<unsigned long long *>(dmax) +
                              (long long)f * (maxit + 2) + it);
        __syncthreads();
    }
}

// After a batch ending before iteration it: marks the fields that stopped
// with iteration it - 1 (every field, once the run has reached its end).
__global__ void diff_jacobi_check_kernel(const double *dmax, int *done, const double *sc, int f0,
                                         int f1, int it, int maxit)
{
    for (int f = f0; f < f1; ++f)
        if (done[f] == 0 && (sc[SC_DONE] != 0.0 || jacobi_stops(dmax, f, it, maxit)))
            done[f] = (sc[SC_DONE] != 0.0) ? 1 : it - 1;
}

// normalize (abstractMatrixSolver.cpp:31-38) of the solve's result
template <int NS>
__device__ __forceinline__ double jacobi_result(const double *Mb, const int *done, int f,
                                                long long n, long long id)
{
    const double x = Mb[(MB<NS>::X + 2 * f + (done[f] & 1)) * n + id];
    return tabs(x) < DBL_EPSILON ? 0 : x;
}

// Mass fractions from their solves: clipped at 0, renormalised, the new
// concentrations, and the temperature system's diagonal
// CCV_new / dt + sum of its face coefficients (Diffusion.cpp:431-479).
template <int NS, int DIM>
__global__ void __launch_bounds__(NT)
diff_finish_y_kernel(const DevParams *__restrict__ Pp, const double *U, const double *Ualt,
                     int hk, double *Mb, const int *done, const double *sc,
                     const double *__restrict__ W)
{
    using M = MB<NS>;
    const DevParams &P = *Pp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = blockIdx.x * (long long)NT + threadIdx.x;
    if (id >= n)
        return;
    const double *Us = state_of(U, Ualt, sc, hk);
    const double rho = Us[NS * n + id];
    double y[NS], sum = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        y[k] = jacobi_result<NS>(Mb, done, k, n, id);
        if (y[k] < 0.)
            y[k] = 0.0;
        sum += y[k];
    }
    double cn[NS], c0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        y[k] /= sum;
        cn[k] = y[k] * rho / P.M[k];
        c0 += cn[k];
        Mb[(M::CN + k) * n + id] = cn[k];
    }
    const double rdt = 1 / sc[SC_DT];
    double diag = (cv_mix<NS>(P, cn, c0) * c0) * rdt;
#pragma unroll
    for (int s = 0; s < 2 * DIM; ++s)
        diag += Mb[(M::TT + s) * n + id];
    Mb[M::DT * n + id] = diag;
    // temperature right-hand side: r_old / dt with r_old = CCV_old T +
    // nonIdeality_old - nonIdeality_new, the boundary terms, the heating
    const double *Wc = W;
    double co[NS], c0o = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        co[k] = Wc[(DW<NS>::C + k) * n + id];
        c0o += co[k];
    }
    const double T0 = Wc[DW<NS>::T * n + id];
    const double rold = (c0o * cv_mix<NS>(P, co, c0o)) * T0 + non_ideality<NS>(P, co, c0o) -
                        non_ideality<NS>(P, cn, c0);
    double b = rold * rdt;
#pragma unroll
    for (int s = 0; s < 2 * DIM; ++s)
        b += Mb[(M::TB + s) * n + id];
    b += Mb[M::HT * n + id];
    Mb[M::BT * n + id] = b;
}

// The recalculated cell from the solves (Diffusion.cpp:963-1029) and its
// Courant signal speed.
template <int NS>
__global__ void __launch_bounds__(NT)
diff_finish_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo, double *U,
                   double *Ualt, int hk, const double *Mb, const int *done, double *sc,
                   int *errflag)
{
    using M = MB<NS>;
    __shared__ double sRed[NT / 64];
    const DevParams &P = *Pp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = blockIdx.x * (long long)NT + threadIdx.x;
    double sig = 0.0;
    if (id < n) {
        double *Us = const_cast<double *>(state_of(U, Ualt, sc, hk));
        const int i = int(id % P.N[0]), j = int((id / P.N[0]) % P.N[1]),
                  k = int(id / P.stride[2]);
        const CellGeo cg = cell_geo(P, geo, i, j, k);
        const double rho = Us[NS * n + id];
        double cn[NS], c0 = 0;
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            cn[q] = Mb[(M::CN + q) * n + id];
            c0 += cn[q];
        }
        double vn[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            vn[c] = jacobi_result<NS>(Mb, done, NS + c, n, id);
        const double Tn = jacobi_result<NS>(Mb, done, NS + 3, n, id);
        if (Tn < 0.)
            atomicOr(errflag, ERRF_NEG_T);
        double un[NS + 5];
#pragma unroll
        for (int q = 0; q < NS; ++q)
            un[q] = cn[q] * P.M[q];
        un[NS] = rho;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            un[NS + 1 + c] = vn[c] * rho;
        const double Uin = uvc_from_t<NS>(P, cn, c0, Tn) * c0;
        const double v2 = ((0. + vn[0] * vn[0]) + vn[1] * vn[1]) + vn[2] * vn[2];
        un[NS + 4] = (Uin + rho * v2 * 0.5) + 0.0;
#pragma unroll
        for (int q = 0; q < NS + 5; ++q)
            Us[q * n + id] = un[q];
        sig = signal_speed<double, NS>(P, un, cg.minlen);
    }
    reduce_signal<NT>(sig, sRed, reinterpret_cast<unsigned long long *>(sc) + SC_SIGMAX);
}

// Turbulent explicit stage: the primitives pass with k, epsilon, CvM and
// nonIdeality / rho and the maximum of epsilon, then the update.
template <int NS, int DIM, int AM>
int launch_dim_turb(const DiffusionLaunch &L, const TurbDiffusion &T)
{
    const unsigned blocks = (unsigned)((L.n + NT - 1) / NT);
    hipLaunchKernelGGL((diff_prims_kernel<NS, true, AM>), dim3(blocks), dim3(NT), 0, L.stream,
                       L.P, L.U, L.Ualt, L.hk, L.W, L.sc, T.epsmax, T.b2);
    if (L.mid && hipEventRecord(L.mid, L.stream) != hipSuccess)
        return 1;
    hipLaunchKernelGGL((diff_update_kernel<NS, DIM, false, false, false, true, AM>),
                       dim3(blocks), dim3(NT), 0, L.stream, L.P, L.geo, L.U, L.Ualt, L.hk, L.W,
                       L.sc, L.err, L.mu, L.D, L.kappa, nullptr, 0, 0, T.TP, T.epsmax, L.star,
                       AM == 1 ? T.bk : nullptr);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <int NS, int DIM>
int launch_dim(const DiffusionLaunch &L)
{
    const unsigned blocks = (unsigned)((L.n + NT - 1) / NT);
    hipLaunchKernelGGL((diff_prims_kernel<NS>), dim3(blocks), dim3(NT), 0, L.stream, L.P, L.U,
                       L.Ualt, L.hk, L.W, L.sc);
    if (L.mid && hipEventRecord(L.mid, L.stream) != hipSuccess)
        return 1;
    const bool spec = L.mm || L.en;
    if (!L.Mb) {
        if (spec)
            hipLaunchKernelGGL((diff_update_kernel<NS, DIM, false, false, true>), dim3(blocks),
                               dim3(NT), 0, L.stream, L.P, L.geo, L.U, L.Ualt, L.hk, L.W, L.sc,
                               L.err, L.mu, L.D, L.kappa, nullptr, L.mm, L.en, nullptr, nullptr,
                               L.star);
        else
            hipLaunchKernelGGL((diff_update_kernel<NS, DIM, false, false, false>), dim3(blocks),
                               dim3(NT), 0, L.stream, L.P, L.geo, L.U, L.Ualt, L.hk, L.W, L.sc,
                               L.err, L.mu, L.D, L.kappa, nullptr, 0, 0, nullptr, nullptr,
                               L.star);
        return hipGetLastError() == hipSuccess ? 0 : 1;
    }
    // Jacobi: assemble, solve Y_k and v_c, finish Y, solve T, finish; the
    // host reads the solves' state after each batch of iterations
    if (spec)
        hipLaunchKernelGGL((diff_update_kernel<NS, DIM, false, true, true>), dim3(blocks),
                           dim3(NT), 0, L.stream, L.P, L.geo, L.U, L.Ualt, L.hk, L.W, L.sc, L.err,
                           L.mu, L.D, L.kappa, L.Mb, L.mm, 0, nullptr, nullptr, L.star);
    else
        hipLaunchKernelGGL((diff_update_kernel<NS, DIM, false, true, false>), dim3(blocks),
                           dim3(NT), 0, L.stream, L.P, L.geo, L.U, L.Ualt, L.hk, L.W, L.sc, L.err,
                           L.mu, L.D, L.kappa, L.Mb, 0, 0, nullptr, nullptr, L.star);
    auto solve = [&](int f0, int f1) -> int {
        int it = 1, batch = 4;
        for (;;) {
            for (int q = 0; q < batch; ++q, ++it)
                hipLaunchKernelGGL((diff_jacobi_kernel<NS, DIM>), dim3(blocks), dim3(NT), 0,
                                   L.stream, L.P, L.Mb, L.dmax, L.done, L.sc, f0, f1, it,
                                   L.maxit);
            hipLaunchKernelGGL(diff_jacobi_check_kernel, dim3(1), dim3(1), 0, L.stream, L.dmax,
                               L.done, L.sc, f0, f1, it, L.maxit);
            if (hipMemcpyAsync(L.done_host, L.done, sizeof(int) * (NS + 4),
                               hipMemcpyDeviceToHost, L.stream) != hipSuccess ||
                hipStreamSynchronize(L.stream) != hipSuccess)
                return 1;
            bool all = true;
            for (int f = f0; f < f1; ++f)
                all = all && L.done_host[f] != 0;
            if (all)
                return 0;
            batch = batch < 64 ? 2 * batch : 64;
        }
    };
    if (solve(0, NS + 3))
        return 1;
    hipLaunchKernelGGL((diff_finish_y_kernel<NS, DIM>), dim3(blocks), dim3(NT), 0, L.stream, L.P,
                       L.U, L.Ualt, L.hk, L.Mb, L.done, L.sc, L.W);
    if (solve(NS + 3, NS + 4))
        return 1;
    hipLaunchKernelGGL((diff_finish_kernel<NS>), dim3(blocks), dim3(NT), 0, L.stream, L.P, L.geo,
                       L.U, L.Ualt, L.hk, L.Mb, L.done, L.sc, L.err);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <int NS> int launch_ns(const DiffusionLaunch &L)
{
    switch (L.dim) {
    case 1:
        return launch_dim<NS, 1>(L);
    case 2:
        return launch_dim<NS, 2>(L);
    default:
        return launch_dim<NS, 3>(L);
    }
}

}  // namespace

size_t diffusion_matrix_doubles(int ns)
{
    switch (ns) {
    case 1:
        return MB<1>::N;
    case 2:
        return MB<2>::N;
    default:
        return MB<3>::N;
    }
}

int launch_diffusion_turb(const DiffusionLaunch &L, const TurbDiffusion &T)
{
    auto by_dim = [&](auto ns) -> int {
        constexpr int NS = decltype(ns)::value;
        auto by_model = [&](auto dim) -> int {
            constexpr int D = decltype(dim)::value;
            switch (T.amodel) {
            case 1:
                return launch_dim_turb<NS, D, 1>(L, T);
            case 2:
                return launch_dim_turb<NS, D, 2>(L, T);
            default:
                return launch_dim_turb<NS, D, 0>(L, T);
            }
        };
        switch (L.dim) {
        case 1:
            return by_model(std::integral_constant<int, 1>{});
        case 2:
            return by_model(std::integral_constant<int, 2>{});
        default:
            return by_model(std::integral_constant<int, 3>{});
        }
    };
    switch (L.ns) {
    case 1:
        return by_dim(std::integral_constant<int, 1>{});
    case 2:
        return by_dim(std::integral_constant<int, 2>{});
    default:
        return by_dim(std::integral_constant<int, 3>{});
    }
}

int launch_diffusion(const DiffusionLaunch &L)
{
    switch (L.ns) {
    case 1:
        return launch_ns<1>(L);
    case 2:
        return launch_ns<2>(L);
    default:
        return launch_ns<3>(L);
    }
}

}  // namespace sg
