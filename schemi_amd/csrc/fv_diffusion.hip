// fv_diffusion.hip -- explicit molecular diffusion stage after each step
// (Diffusion.cpp:20-1061 with the explicit matrix solver, linear
// reconstruction, constant transport coefficients, no turbulence, no
// enthalpy flow, no molar-mass correction; ideal mixture).
//
// Two passes over the cells, FP64 on both precision paths:
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
// Every face normal is +-1 along its axis; a face term of the reference
// (F & N) S then reduces to F_a N_a S plus signed zeros, which are dropped.
#include "fv_launch.hpp"
#include "fv_stage.hpp"
#include "fv_general.hpp"

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

// W layout per cell: Y[NS], c[NS], v[3], T, p
template <int NS> struct DW {
    static constexpr int Y = 0, C = NS, V = 2 * NS, T = 2 * NS + 3, P = 2 * NS + 4,
                         NW = 2 * NS + 5;
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

// Uses the buffer holding the new state: MUSCL-Hancock steps alternate
// between U and Ualt by the scalar block's parity.
__device__ __forceinline__ const double *state_of(const double *U, const double *Ualt,
                                                  const double *sc, int hk)
{
    return (hk && sc[SC_PARITY] == 0.0) ? Ualt : U;
}

template <int NS>
__global__ void __launch_bounds__(NT)
diff_prims_kernel(const DevParams *__restrict__ Pp, const double *U, const double *Ualt, int hk,
                  double *__restrict__ W, double *sc)
{
    const DevParams &P = *Pp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = blockIdx.x * (long long)NT + threadIdx.x;
    // the Courant maximum is taken again over the state after diffusion
    if (id == 0)
        reinterpret_cast<unsigned long long *>(sc)[SC_SIGMAX] = 0ull;
    if (id >= n)
        return;
    const double *Us = state_of(U, Ualt, sc, hk);
    double u[NS + 5], w[NS + 4];
#pragma unroll
    for (int q = 0; q < NS + 5; ++q)
        u[q] = Us[q * n + id];
    prims_from_cons<double, NS, GAS_IDEAL>(P, u, w);
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
    const double Uint = (u[NS + 4] - u[NS] * v2 * 0.5) - 0.;
    // TFromUv (idealFluid.cpp:37-41)
    W[L::T * n + id] = Uint / ((c0 + STAB) * cv_mix<NS>(P, w, c0));
    W[L::P * n + id] = w[NS + 3];
}

template <int NS>
__device__ __forceinline__ void load_w(const double *W, long long n, long long id, double *w)
{
#pragma unroll
    for (int q = 0; q < DW<NS>::NW; ++q)
        w[q] = W[q * n + id];
}

template <int NS, int DIM, bool MIXED>
__global__ void __launch_bounds__(NT)
diff_update_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo, double *U,
                   double *Ualt, int hk, const double *__restrict__ W, double *sc,
                   int *errflag, double mu, double Dc, double kappa)
{
    using L = DW<NS>;
    constexpr int NW = L::NW;
    __shared__ double sRed[NT / 64];
    const DevParams &P = *Pp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = block_cell0() + threadIdx.x;
    double sig = 0.0;
    double *Us = const_cast<double *>(state_of(U, Ualt, sc, hk));
    const double dt = sc[SC_DT];
    if (id < n) {
        const int idx[3] = {int(id % P.N[0]), int((id / P.N[0]) % P.N[1]),
                            int(id / P.stride[2])};
        double w0[NW];
        load_w<NS>(W, n, id, w0);
        const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
        double gV[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dV = 0;
        double dY[NS], dT = 0, dFv[3] = {0, 0, 0}, dFs[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < NS; ++k)
            dY[k] = 0;
        int err = 0;
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
            const bool inner = hi ? (idx[a] < Na - 1) : (idx[a] > 0);
            // outward normal along a: the reference points a face normal away
            // from its owner's centre (mesh.cpp:34-53)
            const bool flip = hi && !inner && P.hi_flip[a] != 0;
            const double na = (hi && !flip) ? 1.0 : -1.0;
            const double S = cg.S[a];
            double vs[3], srho = 0;
            // owner / neighbour values (ghost for a boundary face)
            double qo[NW], qn[NW];
            double gY[NS][3], gT[3], gv[3][3], divs;
            if (inner) {
                const long long nb = id + (hi ? P.stride[a] : -P.stride[a]);
                double wn[NW];
                load_w<NS>(W, n, nb, wn);
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
                for (int k = 0; k < NS; ++k)
                    srho += (qo[L::C + k] * wo + qn[L::C + k] * wnn) * P.M[k];
                // d = r_n - r_o along a; |d|^2 = pow<2>(hypot3(d))
                const double xc_lo = geo[P.off_xc[a] + f - 1], xc_hi = geo[P.off_xc[a] + f];
                const double da = xc_hi - xc_lo;
                const double m = tabs(da);
                const double m2 = m * m;
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
                    srho += bc_surface<double>(P.bct[p][k], P.bcv[p][k][0], w0[L::C + k], -1, a) *
                            P.M[k];
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
                const double cvg = cv_mix<NS>(P, cgh, cg0);
                const double Ug = pg / (P.R / cvg);
                qn[L::T] = Ug / ((cg0 + STAB) * cvg);
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
            }
            // Green-Gauss sums of grad v (v_c N_b) S and div v (v & N) S
#pragma unroll
            for (int c = 0; c < 3; ++c)
                gV[3 * c + a] += (vs[c] * na) * S;
            dV += (vs[a] * na) * S;
            // fluxes -coeff surfGrad, their normal component times S
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const double rhoD = srho * Dc;
                dY[k] += (((gY[k][a] * rhoD) * -1) * na) * S;
            }
            dT += (((gT[a] * kappa) * -1) * na) * S;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                dFv[c] += (((gv[c][a] * mu) * -1) * na) * S;
            // mu strainSurf1 & N: row c, column a of (grad_s v)^T - lam I
            const double lam = (2. / 3.) * divs;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double t = gv[a][c];
                if (c == a)
                    t -= lam;
                dFs[c] += ((t * mu) * na) * S;
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
        double u[NS + 5];
#pragma unroll
        for (int q = 0; q < NS + 5; ++q)
            u[q] = Us[q * n + id];
        const double rho = u[NS];
        // mass fractions: (Y rho + (0 - div F) dt) / rho, clipped, renormalised
        double Yn[NS], sum = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const double ft = 0. - dY[k] / V;
            Yn[k] = (w0[L::Y + k] * rho + ft * dt) / rho;
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
        ftT += heat;
        const double Tn = (CCVold * w0[L::T] + ftT * dt) / CCVnew;
        if (Tn < 0.)
            atomicOr(errflag, ERRF_NEG_T);
        // recalculated cell: rho_k = c_k M_k, m = v rho, U = Cv T c0,
        // E = (U + rho v.v / 2) + 0
        double un[NS + 5];
#pragma unroll
        for (int k = 0; k < NS; ++k)
            un[k] = cn[k] * P.M[k];
        un[NS] = rho;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            un[NS + 1 + c] = vn[c] * rho;
        const double Uin = (cvn * Tn) * c0n;
        const double v2 = ((0. + vn[0] * vn[0]) + vn[1] * vn[1]) + vn[2] * vn[2];
        un[NS + 4] = (Uin + rho * v2 * 0.5) + 0.0;
#pragma unroll
        for (int q = 0; q < NS + 5; ++q)
            Us[q * n + id] = un[q];
        (void)MIXED;
        sig = signal_speed<double, NS, GAS_IDEAL>(P, un, cg.minlen);
    }
    reduce_signal<NT>(sig, sRed, reinterpret_cast<unsigned long long *>(sc) + SC_SIGMAX);
}

template <int NS, int DIM>
int launch_dim(const DiffusionLaunch &L)
{
    const unsigned blocks = (unsigned)((L.n + NT - 1) / NT);
    hipLaunchKernelGGL((diff_prims_kernel<NS>), dim3(blocks), dim3(NT), 0, L.stream, L.P, L.U,
                       L.Ualt, L.hk, L.W, L.sc);
    if (L.mid && hipEventRecord(L.mid, L.stream) != hipSuccess)
        return 1;
    hipLaunchKernelGGL((diff_update_kernel<NS, DIM, false>), dim3(blocks), dim3(NT), 0, L.stream,
                       L.P, L.geo, L.U, L.Ualt, L.hk, L.W, L.sc, L.err, L.mu, L.D, L.kappa);
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
