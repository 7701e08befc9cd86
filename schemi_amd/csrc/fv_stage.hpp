// fv_stage.hpp -- the fused forward-Euler stage kernel (one Advection2dOrder
// call, Schemi/src/Advections/Advection2dOrder.cpp:16-636) and its helpers.
//
// One workgroup owns a TX x TY tile of columns and marches along z over a
// chunk of planes (2.5-D blocking).  For every plane it
//   1. stages the primitives of the plane (tile + 2-cell halo in x and y) in
//      LDS; the own column's primitives at k-2..k+2 stay in registers,
//   2. computes the limited slopes along x and y of the tile cells and their
//      face neighbours (Green-Gauss gradient, TVDLimiter ratio, phi) into LDS,
//      and the z slope of the own column in registers,
//   3. reconstructs both sides of every x, y and z face once, closes them with
//      the EOS and evaluates the HLLC/HLL flux (LDS for x/y faces, registers
//      for the z faces),
//   4. sums the six face contributions of every cell in the reference's order
//      [x-, x+, y+, y-, z-, z+] (divergence.hpp:315-350 over surfacesOfCells,
//      mesh.cpp:1809-1822) and updates the conserved state.  Each face is
//      written by one thread and read by its two cells: no atomics, the result
//      is deterministic.
// In stage 2 the epilogue also forms the SSP-RK2 average with U^n
// (bunchOfFields.hpp:166-203) and reduces the Courant signal speed of the new
// state (Schemi.cpp:807-833).
//
// Slopes along one axis only use the 1-D stencil along that axis: the
// reference's gradient components along an axis come only from that axis's
// faces (the normals are exactly axis-aligned), its limiter ratio along an
// axis only from the neighbours along it plus exact zero entries from the
// other neighbours (their centres share the coordinate), and the transverse
// velocity slopes have a limiter ratio with an exact zero entry or a zero
// gradient, so phi (r) G = 0 for them (TVD limiters).  This kernel serves the
// TVD limiters in 1-D and the linear limiter on meshes whose face-to-centre
// offsets across an axis are exactly zero; stage0_kernel (below) serves the
// schemes whose slopes vanish, fv_general.hpp the linear limiter elsewhere.
#pragma once

#include "fv_device.hpp"

namespace sg {

template <class T> __device__ __forceinline__ T to_T(double x) { return T(x); }

// Axis geometry read from the per-axis tables.
struct AxisGeo {
    const double *xc, *e, *xf, *wo, *wn, *dn, *dlo, *dhi;
};

__device__ __forceinline__ AxisGeo axis_geo(const DevParams &P, const double *geo, int a)
{
    AxisGeo g;
    g.xc = geo + P.off_xc[a];
    g.e = geo + P.off_e[a];
    g.xf = geo + P.off_xf[a];
    g.wo = geo + P.off_wo[a];
    g.wn = geo + P.off_wn[a];
    g.dn = geo + P.off_dn[a];
    g.dlo = geo + P.off_dlo[a];
    g.dhi = geo + P.off_dhi[a];
    return g;
}

// Number of slope components stored per cell and axis.  TVD limiters: the
// species, the pressure and the velocity component along the axis; the linear
// limiter also keeps the transverse velocity components.
template <int NS, bool LIN> struct SlopeLayout {
    static constexpr int NQ = NS + 4;
    static constexpr int NL = LIN ? NQ : NS + 2;
    // primitive index of slope entry s along axis a
    __device__ static int prim(int s, int a)
    {
        if constexpr (LIN) {
            (void)a;
            return s;
        } else {
            return s < NS ? s : (s == NS ? NS + 3 : NS + a);
        }
    }
    // velocity component of slope entry s (-1 for scalars)
    __device__ static int comp(int s, int a)
    {
        const int q = prim(s, a);
        return (q >= NS && q < NS + 3) ? q - NS : -1;
    }
};

// Boundary condition lookup for primitive q on patch p.
template <int NS>
__device__ __forceinline__ void bc_of(const DevParams &P, int p, int q, int &type, double &val)
{
    if (q < NS) {
        type = P.bct[p][q];
        val = P.bcv[p][q][0];
    } else if (q < NS + 3) {
        type = P.bct[p][NS];
        val = P.bcv[p][NS][q - NS];
    } else {
        type = P.bct[p][NS + 1];
        val = P.bcv[p][NS + 1][0];
    }
}

// Limited slopes along axis a of the cell with index i along that axis
// (Green-Gauss gradient, gradient.hpp:19-58 + linearInterpolate.hpp:214-251;
// TVDLimiter.cpp:12-125 and 127-273; limiter phi).  wm/w0/wp: primitives at
// i-1, i, i+1 (wm/wp unused at a physical boundary).  S: area of the cell's
// a-faces, V: volume.
template <class T, int NS, bool LIN>
__device__ __forceinline__ void slopes_axis(const DevParams &P, const AxisGeo &G, int a,
                                            int i, const T *wm, const T *w0,
                                            const T *wp, T S, T V, T *L)
{
    using SL = SlopeLayout<NS, LIN>;
    const int Na = P.N[a];
    const bool per = P.periodic[a] != 0;
    const bool has_lo = (i > 0) || per;
    const bool has_hi = (i < Na - 1) || per;
    const int plo = patch_of(a, 0), phi_ = patch_of(a, 1);
    const T wo_lo = to_T<T>(G.wo[i]), wn_lo = to_T<T>(G.wn[i]);
    const T wo_hi = to_T<T>(G.wo[i + 1]), wn_hi = to_T<T>(G.wn[i + 1]);
    const T dn_lo = to_T<T>(G.dn[i]), dn_hi = to_T<T>(G.dn[i + 1]);
    const T gh_lo = to_T<T>(P.ghost_lo[a]), gh_hi = to_T<T>(P.ghost_hi[a]);
    const bool zero = P.zero_entry[a] != 0;
    // normal of the high face: -1 on a flipped high boundary face
    const bool flip = !has_hi && P.hi_flip[a] != 0;
#pragma unroll
    for (int s = 0; s < SL::NL; ++s) {
        const int q = SL::prim(s, a);
        const int comp = SL::comp(s, a);
        const T q0 = w0[q];
        int tlo = BC_FREE, thi = BC_FREE;
        double vlo = 0, vhi = 0;
        if (!has_lo)
            bc_of<NS>(P, plo, q, tlo, vlo);
        if (!has_hi)
            bc_of<NS>(P, phi_, q, thi, vhi);
        // face values for the Green-Gauss gradient
        auto gradient = [&]() {
            const T flo = has_lo ? wm[q] * wo_lo + q0 * wn_lo
                                 : bc_surface<T>(tlo, T(vlo), q0, comp, a);
            const T fhi = has_hi ? q0 * wo_hi + wp[q] * wn_hi
                                 : bc_surface<T>(thi, T(vhi), q0, comp, a);
            return ((flip ? -fhi : fhi) * S - flo * S) / V;
        };
        if constexpr (LIN) {
            L[s] = gradient();
        } else {
            // limiter entries (d-hat = -1 below, +1 above the cell)
            const T elo = has_lo ? -((wm[q] - q0) / dn_lo)
                                 : -((bc_cell<T>(tlo, T(vlo), q0, comp, a) - q0) / gh_lo);
            T ehi = has_hi ? (wp[q] - q0) / dn_hi
                           : (bc_cell<T>(thi, T(vhi), q0, comp, a) - q0) / gh_hi;
            if (flip)
                ehi = -ehi;  // d-hat of the ghost is -1
            T mn = tmin(elo, ehi), mx = tmax(elo, ehi);
            if (zero) {
                mn = tmin(mn, T(0));
                mx = tmax(mx, T(0));
            }
            // r = mn / (mx + eps) <= 0 exactly when mn <= 0 and mx + eps > 0;
            // every limiter here has phi(r <= 0) = +-0, so phi (r) G is +-0 and
            // neither the ratio nor the gradient is needed (the reference
            // evaluates both and multiplies by the zero)
            if (mn <= T(0) && mx + T(STAB) > T(0)) {
                L[s] = T(0);
                continue;
            }
            const T r = mn / (mx + T(STAB));
            L[s] = phi<T>(P.limiter, r) * gradient();
        }
    }
}

// Reconstructed face value of primitive q from a cell with slopes L along
// axis a and face offset d (Advection2dOrder.cpp:110-137).
template <class T, int NS, bool LIN>
__device__ __forceinline__ void reconstruct(const T *w, const T *L, int a, T d, T *c, T *v, T &p)
{
    using SL = SlopeLayout<NS, LIN>;
    T f[NS + 4];
#pragma unroll
    for (int q = 0; q < NS + 4; ++q)
        f[q] = w[q];
#pragma unroll
    for (int s = 0; s < SL::NL; ++s) {
        const int q = SL::prim(s, a);
        f[q] = L[s] * d + w[q];
    }
#pragma unroll
    for (int k = 0; k < NS; ++k)
        c[k] = f[k];
    v[0] = f[NS];
    v[1] = f[NS + 1];
    v[2] = f[NS + 2];
    p = f[NS + 3];
}

// Ghost state of a boundary face of axis a (side 0 = low plane, 1 = high)
// from the owner's face value (Advection2dOrder.cpp:398-521).
template <class T, int NS>
__device__ __forceinline__ void ghost_side(const DevParams &P, int a, int side, const T *cO,
                                           const T *vO, T pO, T *cB, T *vB, T &pB)
{
    const int p = patch_of(a, side);
    int t;
    double val;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        bc_of<NS>(P, p, k, t, val);
        cB[k] = bc_cell<T>(t, T(val), cO[k], -1, a);
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        bc_of<NS>(P, p, NS + b, t, val);
        vB[b] = bc_cell<T>(t, T(val), vO[b], b, a);
    }
    bc_of<NS>(P, p, NS + 3, t, val);
    pB = bc_cell<T>(t, T(val), pO, -1, a);
}

// EOS closure of both face sides and the HLLC / HLL flux F & N of a face
// normal to axis a with sign sgn (Advection2dOrder.cpp:240-528); false when
// the face has no valid wave pattern.
template <class T, int NS>
__device__ __forceinline__ bool riemann(const DevParams &P, int a, T sgn, const T *cO,
                                        const T *vO, T pO, const T *cB, const T *vB, T pB, T *g)
{
    Side<T, NS> O, B;
    close_side<T, NS>(P, cO, vO, pO, O);
    close_side<T, NS>(P, cB, vB, pB, B);
    if (P.flux == FLUX_HLL)
        return hll_flux<T, NS>(O, B, a, sgn, g);
    return hllc_flux<T, NS>(O, B, a, sgn, g);
}

// Flux of face f (0..Na) along axis a.  Owner/neighbour data: primitives and
// slopes of cells f-1 (owner, lower) and f (neighbour); at the low boundary
// the owner is cell 0 (slopes Ln, primitives wn) and the normal points to -a;
// at the high boundary the owner is cell Na-1 (wo, Lo).
template <class T, int NS, bool LIN>
__device__ __forceinline__ bool face_flux(const DevParams &P, const AxisGeo &G, int a, int f,
                                          const T *wo, const T *Lo, const T *wn,
                                          const T *Ln, T *g)
{
    const int Na = P.N[a];
    const bool per = P.periodic[a] != 0;
    T cO[NS], vO[3], pO, cB[NS], vB[3], pB;
    T sgn = T(1);
    if (!per && f == 0) {
        // low boundary face: owner is cell 0, ghost from the owner face value
        reconstruct<T, NS, LIN>(wn, Ln, a, to_T<T>(G.dlo[0]), cO, vO, pO);
        ghost_side<T, NS>(P, a, 0, cO, vO, pO, cB, vB, pB);
        sgn = T(-1);
    } else if (!per && f == Na) {
        reconstruct<T, NS, LIN>(wo, Lo, a, to_T<T>(G.dhi[Na - 1]), cO, vO, pO);
        ghost_side<T, NS>(P, a, 1, cO, vO, pO, cB, vB, pB);
        if (P.hi_flip[a])
            sgn = T(-1);
    } else {
        const int io = (f == 0) ? Na - 1 : f - 1;  // periodic wrap
        const int in = (f == Na) ? 0 : f;
        reconstruct<T, NS, LIN>(wo, Lo, a, to_T<T>(G.dhi[io]), cO, vO, pO);
        reconstruct<T, NS, LIN>(wn, Ln, a, to_T<T>(G.dlo[in]), cB, vB, pB);
    }
    return riemann<T, NS>(P, a, sgn, cO, vO, pO, cB, vB, pB, g);
}

// Geometry of one cell: volume (ex ey) ez (mesh.cpp:1291-1293) and face areas.
struct CellGeo {
    double V, S[3], e[3], minlen;
};

__device__ __forceinline__ CellGeo cell_geo(const DevParams &P, const double *geo, int i, int j, int k)
{
    const double ex = geo[P.off_e[0] + i];
    const double ey = geo[P.off_e[1] + j];
    const double ez = geo[P.off_e[2] + k];
    CellGeo c;
    c.V = ex * ey * ez;
    c.S[0] = ey * ez;  // |rX0 - r00| |r0Y - r00| of an x-face
    c.S[1] = ex * ez;
    c.S[2] = ex * ey;
    c.e[0] = ex;
    c.e[1] = ey;
    c.e[2] = ez;
    c.minlen = 1. / tmin(tmin(ex, ey), ez);  // Schemi.cpp:649-662
    return c;
}

template <int NS> __device__ __forceinline__ void load_cons(const double *U, long long n, long long id, double *u)
{
#pragma unroll
    for (int q = 0; q < NS + 5; ++q)
        u[q] = U[q * n + id];
}

// Primitives of cell id in the flux precision T; the mixed path rounds the
// conserved values to FP32 first.
template <class T, int NS, bool MIXED>
__device__ __forceinline__ void cell_prims(const DevParams &P, const double *U, long long n,
                                           long long id, T *w)
{
    double u[NS + 5];
    load_cons<NS>(U, n, id, u);
    if constexpr (MIXED) {
        T ut[NS + 5];
#pragma unroll
        for (int q = 0; q < NS + 5; ++q)
            ut[q] = T(u[q]);
        prims_from_cons<T, NS>(P, ut, w);
    } else {
        prims_from_cons<T, NS>(P, u, w);
    }
}

// Courant signal speed (|v| + c) / min edge of one cell (Schemi.cpp:810-828),
// in precision T (FP64 on the parity path).
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ T signal_speed(const DevParams &P, const T *u, T minlen)
{
    T w[NS + 4];
    prims_from_cons<T, NS, GAS>(P, u, w);
    T c0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
        c0 += w[k];
    const T vx = w[NS], vy = w[NS + 1], vz = w[NS + 2], p = w[NS + 3];
    const T v2 = ((T(0) + vx * vx) + vy * vy) + vz * vz;
    const T U = (u[NS + 4] - u[NS] * v2 * T(0.5)) - T(0);
    const T c2 = eos_c2<T, NS, GAS>(P, w, c0, gamma1<T, NS>(P, w, c0), u[NS], U, p);
    const T cs = Real<T>::sqrt_(c2);
    return (hypot3<T>(vx, vy, vz) + cs) * minlen;
}

// Residual and update of one cell (Advection2dOrder.cpp:541-602 with the
// divergence of divergence.hpp:315-350).  F(s, q) is flux component q of the
// cell's face s in the order x-, x+, y+, y-, z-, z+ (F & N of the face's own
// normal); the cell owns its high faces and the low faces on a non-periodic
// boundary (xlo/ylo/zlo), and is the neighbour of its other low faces.  The
// contributions (s g) S are summed in that order in FP64.  RK stage 2 also
// forms the SSP-RK2 average with U^n (bunchOfFields.hpp:166-203); the last
// stage of a step (RK stage 2, the MUSCL-Hancock stage) reduces the Courant
// signal speed of the new state.  u: the cell's conserved input state, un:
// its U^n (RK stage 2), both supplied by the caller.
template <int NS, int DIM, int SV, bool MIXED, int GAS = -1, class FluxAt>
__device__ __forceinline__ void update_cell(const DevParams &P, const CellGeo &cg, long long n,
                                            long long id, const double *u, double *Uout,
                                            const double *un, double dt, bool xlo, bool ylo,
                                            bool zlo, const FluxAt &F, double &sig)
{
    constexpr int NF = NS + 5;
    double unew[NF];
    // mixed path: one division per cell instead of one per field
    const double dtV = MIXED ? dt / cg.V : 0.0;
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        double acc = 0;
        acc += (xlo ? F(0, q) : -F(0, q)) * cg.S[0];
        acc += F(1, q) * cg.S[0];
        if constexpr (DIM >= 2) {
            acc += F(2, q) * cg.S[1];
            acc += (ylo ? F(3, q) : -F(3, q)) * cg.S[1];
        }
        if constexpr (DIM == 3) {
            acc += (zlo ? F(4, q) : -F(4, q)) * cg.S[2];
            acc += F(5, q) * cg.S[2];
        }
        if constexpr (MIXED)
            unew[q] = u[q] - acc * dtV;
        else
            unew[q] = u[q] - (acc / cg.V) * dt;
    }
    if (P.gravity) {
        // pre-update rho and v (Advection2dOrder.cpp:547-554)
        const double rho = u[NS];
        const double vx = u[NS + 1] / rho, vy = u[NS + 2] / rho, vz = u[NS + 3] / rho;
        const double vg = ((0. + vx * P.g[0]) + vy * P.g[1]) + vz * P.g[2];
        unew[NS + 4] = unew[NS + 4] + rho * vg * dt;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            unew[NS + 1 + b] = unew[NS + 1 + b] + (P.g[b] * rho) * dt;
    }
    if constexpr (MIXED) {
        // species renormalisation: sum_k rho_k = rho
        double sr = 0;
#pragma unroll
        for (int q = 0; q < NS; ++q)
            sr += unew[q];
        const double f = unew[NS] / sr;
#pragma unroll
        for (int q = 0; q < NS; ++q)
            unew[q] *= f;
    }
    if constexpr (SV == SV_RK2) {
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            if (q >= NS + 1 && q <= NS + 3)
                unew[q] = unew[q] * 0.5 + un[q] * 0.5;
            else
                unew[q] = 0.5 * unew[q] + 0.5 * un[q];
        }
    }
    if constexpr (SV != SV_RK1) {
        if constexpr (MIXED) {
            // Courant speed of the FP64 state evaluated in FP32
            float uf[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q)
                uf[q] = float(unew[q]);
            const float ml = 1.0f / tmin(tmin(float(cg.e[0]), float(cg.e[1])), float(cg.e[2]));
            sig = tmax(sig, double(signal_speed<float, NS, GAS>(P, uf, ml)));
        } else {
            sig = tmax(sig, signal_speed<double, NS, GAS>(P, unew, cg.minlen));
        }
    }
#pragma unroll
    for (int q = 0; q < NF; ++q)
        Uout[q * n + id] = unew[q];
}

// Wavefront and workgroup maximum of the Courant signal speed, then one
// atomic per workgroup (positive doubles order like their bit patterns).
template <int NT>
__device__ __forceinline__ void reduce_signal(double sig, double *sRed,
                                              unsigned long long *sigmax)
{
    const int tid = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1)
        sig = tmax(sig, __shfl_xor(sig, off, 64));
    if ((tid & 63) == 0)
        sRed[tid >> 6] = sig;
    __syncthreads();
    if (tid == 0) {
        double m = sRed[0];
        for (int w = 1; w < NT / 64; ++w)
            m = tmax(m, sRed[w]);
        atomicMax(sigmax, (unsigned long long)__double_as_longlong(m));
    }
}

// Tile coordinates of this workgroup.  The hardware hands workgroups to the 8
// XCDs (each with its own L2) round-robin in linear order; with SG_XCD_REMAP
// the linear id is permuted so that each XCD receives a contiguous range of
// tiles (x fastest), whose shared halo rows then meet in one L2.
#ifndef SG_XCD_REMAP
#define SG_XCD_REMAP 1
#endif
__device__ __forceinline__ void tile_coords(int &bx, int &by, int &bz)
{
#if SG_XCD_REMAP
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (nb % 8u == 0u)
        lin = (lin % 8u) * (nb / 8u) + lin / 8u;
    bx = static_cast<int>(lin % gridDim.x);
    by = static_cast<int>((lin / gridDim.x) % gridDim.y);
    bz = static_cast<int>(lin / (gridDim.x * gridDim.y));
#else
    bx = blockIdx.x;
    by = blockIdx.y;
    bz = blockIdx.z;
#endif
}

// Cell state closed with the EOS once per cell, for schemes whose face
// values are the cell values.  Two layouts:
//   H = false (SSP-RK2 stages): c[NS], v[3], p, r0, U, E, cs -- primitives
//     plus the closure terms that need a division or a square root
//     (close_side); the partial densities c_k M_k and the momentum v r0 are
//     re-formed per face with the same products;
//   H = true (MUSCL-Hancock): the face side after the half-step predictor,
//     rho[NS], r0, m[3], E (conserved) and v[3], p, cs re-derived from them
//     (HancockStage.cpp:174-290); the ghost's concentrations are rho_k / M_k.
template <class T, int NS, bool H> struct Closed {
    static constexpr int NV = H ? NS + 10 : NS + 8;
};

template <class T, int NS, int GAS = -1>
__device__ __forceinline__ void close_cell(const DevParams &P, const T *w, T *z)
{
    Side<T, NS> s;
    close_side<T, NS, GAS>(P, w, w + NS, w[NS + 3], s);
#pragma unroll
    for (int q = 0; q < NS + 4; ++q)
        z[q] = w[q];
    z[NS + 4] = s.r0;
    z[NS + 5] = s.U;
    z[NS + 6] = s.E;
    z[NS + 7] = s.cs;
}

// Closed state of a cell straight from its conserved values, for the mixed
// path with the ideal mixture: c_k = rho_k / M_k, v = m / rho,
// U = E - rho v^2 / 2, g1 = R c0 / sum_k c_k Cv_k, p = g1 U, and the face
// quantities r0 = rho, U, E (the cell's own), cs^2 = g1 (U + p) / rho.  The
// reference recomputes U from p and r0 = sum_k rho_k, E from U; on the mixed
// path (species renormalised so that sum_k rho_k = rho) these agree with the
// cell values to FP32 rounding, so the round trip is skipped.
template <int NS>
__device__ __forceinline__ void close_cell_lean(const DevParams &P, const float *u, float *z)
{
    float c0 = 0.0f, ccv = 0.0f;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        z[k] = u[k] / P.Mf[k];
        c0 += z[k];
        ccv += z[k] * P.Cvf[k];
    }
    const float rho = u[NS];
    const float ir = 1.0f / rho;
    const float vx = u[NS + 1] * ir, vy = u[NS + 2] * ir, vz = u[NS + 3] * ir;
    const float E = u[NS + 4];
    const float U = E - 0.5f * rho * (vx * vx + vy * vy + vz * vz);
    const float g1 = P.Rf * c0 / ccv;
    const float p = g1 * U;
    z[NS] = vx;
    z[NS + 1] = vy;
    z[NS + 2] = vz;
    z[NS + 3] = p;
    z[NS + 4] = rho;
    z[NS + 5] = U;
    z[NS + 6] = E;
    z[NS + 7] = __builtin_sqrtf(g1 * (U + p) * ir);
}

// Predicted face side from its conserved values (rho_k, r0, m, E after the
// half step): c_k = rho_k / M_k, c_0 = sum c_k, v = m / r0,
// U = E - r0 v^2 / 2 - rho k, p = pFromUv, cs from sqSonicSpeed
// (HancockStage.cpp:174-290); stored in the H layout.
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ void hancock_rederive(const DevParams &P, const T *rho, T r0,
                                                 const T *m, T E, T *z)
{
    T c[NS], c0 = T(0);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        c[k] = rho[k] / mixM<T>(P, k);
        c0 += c[k];
    }
    T v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        v[a] = m[a] / r0;
    const T v2 = ((T(0) + v[0] * v[0]) + v[1] * v[1]) + v[2] * v[2];
    const T U = (E - r0 * v2 * T(0.5)) - T(0);
    const T g1 = gamma1<T, NS>(P, c, c0);
    const T p = eos_p<T, NS, GAS>(P, c, c0, g1, U);
    const T cs = Real<T>::sqrt_(eos_c2<T, NS, GAS>(P, c, c0, g1, r0, U, p));
#pragma unroll
    for (int k = 0; k < NS; ++k)
        z[k] = rho[k];
    z[NS] = r0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        z[NS + 1 + a] = m[a];
        z[NS + 5 + a] = v[a];
    }
    z[NS + 4] = E;
    z[NS + 8] = p;
    z[NS + 9] = cs;
}

// Cell closed for the MUSCL-Hancock stage when its face values are the cell
// values: the divergence of the physical fluxes of its own face sides
// cancels pairwise (opposite faces carry the same values, normals and areas,
// and are summed consecutively), so the predictor adds exactly zero and the
// face side is the closed cell state with its primitives re-derived.
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ void close_cell_hancock(const DevParams &P, const T *w, T *z)
{
    Side<T, NS> s;
    close_side<T, NS, GAS>(P, w, w + NS, w[NS + 3], s);
    hancock_rederive<T, NS, GAS>(P, s.rho, s.r0, s.m, s.E, z);
}

template <class T, int NS, bool H>
__device__ __forceinline__ void side_of(const DevParams &P, const T *z, Side<T, NS> &s)
{
    if constexpr (H) {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            s.rho[k] = z[k];
        s.r0 = z[NS];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s.m[a] = z[NS + 1 + a];
            s.v[a] = z[NS + 5 + a];
        }
        s.E = z[NS + 4];
        s.p = z[NS + 8];
        s.cs = z[NS + 9];
        s.U = T(0);  // not used by the flux
    } else {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            s.rho[k] = z[k] * mixM<T>(P, k);
        s.r0 = z[NS + 4];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s.v[a] = z[NS + a];
            s.m[a] = z[NS + a] * s.r0;
        }
        s.p = z[NS + 3];
        s.U = z[NS + 5];
        s.E = z[NS + 6];
        s.cs = z[NS + 7];
    }
}

// Primitives (c, v, p) of a closed face side, for its boundary ghost.
template <class T, int NS, bool H>
__device__ __forceinline__ void prims_of(const DevParams &P, const T *z, T *c, T *v, T &p)
{
    if constexpr (H) {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            c[k] = z[k] / mixM<T>(P, k);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            v[a] = z[NS + 5 + a];
        p = z[NS + 8];
    } else {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            c[k] = z[k];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            v[a] = z[NS + a];
        p = z[NS + 3];
    }
}

// Flux of face f along axis a from the closed states of cells f-1 (zo) and f
// (zn); on a non-periodic boundary only the owner's state is read and the
// ghost is formed from its primitives (Advection2dOrder.cpp:398-521,
// HancockStage.cpp:290-448).
template <class T, int NS, int GAS = -1, bool H = false>
__device__ __forceinline__ bool face_flux_closed(const DevParams &P, int a, int f, const T *zo,
                                                 const T *zn, T *g)
{
    constexpr int NV = Closed<T, NS, H>::NV;
    const int Na = P.N[a];
    const bool per = P.periodic[a] != 0;
    Side<T, NS> O, B;
    if (!per && (f == 0 || f == Na)) {
        const int side = (f == 0) ? 0 : 1;
        T z[NV];  // element-wise select keeps both in registers
#pragma unroll
        for (int q = 0; q < NV; ++q)
            z[q] = side ? zo[q] : zn[q];
        T c[NS], v[3], p, cB[NS], vB[3], pB;
        prims_of<T, NS, H>(P, z, c, v, p);
        ghost_side<T, NS>(P, a, side, c, v, p, cB, vB, pB);
        side_of<T, NS, H>(P, z, O);
        close_side<T, NS, GAS>(P, cB, vB, pB, B);
        const T sg = (side && !P.hi_flip[a]) ? T(1) : T(-1);
        if (P.flux == FLUX_HLL)
            return hll_flux<T, NS>(O, B, a, sg, g);
        return hllc_flux<T, NS>(O, B, a, sg, g);
    }
    side_of<T, NS, H>(P, zo, O);
    side_of<T, NS, H>(P, zn, B);
    if (P.flux == FLUX_HLL)
        return hll_flux<T, NS>(O, B, a, T(1), g);
    return hllc_flux<T, NS>(O, B, a, T(1), g);
}

// Physical flux of one face side through its face for the MUSCL-Hancock
// predictor (HancockStage.cpp:10-172): (F(s) . n) S added to the cell's sums,
// dr[0] for rho, dr[1 + k] for rho_k, dm for the momentum, dE for the energy,
// in the reference's operation order (dot products from 0, tensor rows
// (T_a0 n_0 + T_a1 n_1) + T_a2 n_2, the pressure plus 2/3 k with k = 0).
template <class T, int NS>
__device__ __forceinline__ void hancock_face_terms(const Side<T, NS> &s, const T *n, T S, T *dr,
                                                   T *dm, T &dE)
{
    auto dot = [&](const T *x) { return ((T(0) + x[0] * n[0]) + x[1] * n[1]) + x[2] * n[2]; };
#pragma unroll
    for (int k = 0; k <= NS; ++k) {
        const T rho = (k == 0) ? s.r0 : s.rho[(k > 0) ? k - 1 : 0];
        const T rv[3] = {s.v[0] * rho, s.v[1] * rho, s.v[2] * rho};
        dr[k] += dot(rv) * S;
    }
    const T pp = s.p + T(2.0 / 3.0) * T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T row = ((s.m[a] * s.v[0]) * n[0] + (s.m[a] * s.v[1]) * n[1]) +
                      (s.m[a] * s.v[2]) * n[2];
        dm[a] += (n[a] * pp + row) * S;
    }
    const T vE[3] = {s.v[0] * s.E, s.v[1] * s.E, s.v[2] * s.E};
    const T vp[3] = {s.v[0] * pp, s.v[1] * pp, s.v[2] * pp};
    dE += (dot(vE) + dot(vp)) * S;
}

// A face side advanced by half a step with the cell's flux sums (divided by
// V here) and re-derived into the H layout (HancockStage.cpp:174-290).
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ void hancock_advance(const DevParams &P, const Side<T, NS> &s,
                                                const T *dr, const T *dm, T dE, T half, T *z)
{
    T rho[NS], m[3];
#pragma unroll
    for (int k = 0; k < NS; ++k)
        rho[k] = s.rho[k] - dr[k + 1] * half;
    const T r0 = s.r0 - dr[0] * half;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        m[a] = s.m[a] - dm[a] * half;
    const T E = s.E - dE * half;
    hancock_rederive<T, NS, GAS>(P, rho, r0, m, E, z);
}

template <class T, int NS>
__device__ __forceinline__ void hancock_divide(T *dr, T *dm, T &dE, T V)
{
#pragma unroll
    for (int k = 0; k <= NS; ++k)
        dr[k] /= V;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        dm[a] /= V;
    dE /= V;
}

// MUSCL-Hancock face sides of a 1-D cell with slopes L (HancockStage.cpp:
// 10-448): both sides reconstructed (x-: d = x_f(i) - x_c(i), x+: d =
// x_f(i+1) - x_c(i)) and closed, the divergence of their physical fluxes
// (normals -e_x and +e_x, area S, the cell's faces in the order x-, x+)
// divided by V, each side advanced by -div dt / 2 and re-derived.  zl: the
// x- side, zr: the x+ side, both in the H layout.
template <class T, int NS, bool LIN, int GAS = -1>
__device__ __forceinline__ void hancock_sides_1d(const DevParams &P, const AxisGeo &G, int i,
                                                 const T *w, const T *L, T V, T S, T half,
                                                 T *zl, T *zr)
{
    T c[2][NS], v[2][3], p[2];
    reconstruct<T, NS, LIN>(w, L, 0, to_T<T>(G.dlo[i]), c[0], v[0], p[0]);
    reconstruct<T, NS, LIN>(w, L, 0, to_T<T>(G.dhi[i]), c[1], v[1], p[1]);
    Side<T, NS> sd[2];
    close_side<T, NS, GAS>(P, c[0], v[0], p[0], sd[0]);
    close_side<T, NS, GAS>(P, c[1], v[1], p[1], sd[1]);
    T dr[NS + 1], dm[3] = {T(0), T(0), T(0)}, dE = T(0);
#pragma unroll
    for (int k = 0; k <= NS; ++k)
        dr[k] = T(0);
    const T nl[3] = {T(-1), T(0), T(0)}, nr[3] = {T(1), T(0), T(0)};
    hancock_face_terms<T, NS>(sd[0], nl, S, dr, dm, dE);
    hancock_face_terms<T, NS>(sd[1], nr, S, dr, dm, dE);
    hancock_divide<T, NS>(dr, dm, dE, V);
    hancock_advance<T, NS, GAS>(P, sd[0], dr, dm, dE, half, zl);
    hancock_advance<T, NS, GAS>(P, sd[1], dr, dm, dE, half, zr);
}

// Top-layer cell of a 3-D mesh whose top face normal is reversed (the
// reference spaces z by Ly / Nz, so with Lz < Ly the top face plane lies
// below the last cell centres): with zero slopes every side is the cell
// state, the x and y pairs cancel as elsewhere, but the z- face (normal
// -e_z seen from this cell) and the top face (normal -e_z) add up, so the
// predictor is 2 (F(s) . -e_z) S_z / V; all six sides stay equal.
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ void close_cell_hancock_zflip(const DevParams &P, const T *w, T S,
                                                         T V, T half, T *z)
{
    Side<T, NS> s;
    close_side<T, NS, GAS>(P, w, w + NS, w[NS + 3], s);
    T dr[NS + 1], dm[3] = {T(0), T(0), T(0)}, dE = T(0);
#pragma unroll
    for (int k = 0; k <= NS; ++k)
        dr[k] = T(0);
    const T nz[3] = {T(0), T(0), T(-1)};
    hancock_face_terms<T, NS>(s, nz, S, dr, dm, dE);
    hancock_face_terms<T, NS>(s, nz, S, dr, dm, dE);
    hancock_divide<T, NS>(dr, dm, dE, V);
    hancock_advance<T, NS, GAS>(P, s, dr, dm, dE, half, z);
}

// Index helper for the x-y plane tile in LDS with a 2-cell halo.
template <int TX, int TY, int DIM> struct Tile {
    static constexpr int BX = TX + 4;
    static constexpr int BY = (DIM >= 2) ? TY + 4 : 1;
    static constexpr int OY = (DIM >= 2) ? 2 : 0;
};

template <class T, int NS, int DIM, bool LIN, int SV, bool MIXED, int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
stage_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
             const double *__restrict__ Uin, double *Uout, const double *Un,
             const double *__restrict__ dt_ptr, unsigned long long *sigmax, int *errflag,
             int kchunk)
{
    constexpr int NQ = NS + 4;
    constexpr int NF = NS + 5;
    using SL = SlopeLayout<NS, LIN>;
    constexpr int NL = SL::NL;
    constexpr int NT = TX * TY;
    using TL = Tile<TX, TY, DIM>;
    constexpr int BX = TL::BX, BY = TL::BY, OY = TL::OY;
    constexpr int SXW = TX + 2;
    constexpr int SYH = (DIM >= 2) ? TY + 2 : 1;
    constexpr int FYH = (DIM >= 2) ? TY + 1 : 1;

    __shared__ T sW[NQ][BY][BX];
    __shared__ T sSx[NL][TY][SXW];
    __shared__ T sSy[(DIM >= 2) ? NL : 1][SYH][TX];
    __shared__ T sFx[NF][TY][TX + 1];
    // MUSCL-Hancock (1-D): predicted x- and x+ face sides of every cell
    constexpr bool HK = (SV == SV_HANCOCK);
    constexpr int NVH = Closed<T, NS, true>::NV;
    static_assert(!HK || DIM == 1, "the slope kernel runs MUSCL-Hancock in 1-D only");
    __shared__ T sHl[HK ? NVH : 1][HK ? SXW : 1];
    __shared__ T sHr[HK ? NVH : 1][HK ? SXW : 1];
    __shared__ T sFy[(DIM >= 2) ? NF : 1][FYH][TX];
    __shared__ double sRed[NT / 64];

    const DevParams &P = *Pp;
    // dt_ptr points at the context's scalar block; once the run has reached
    // its end time, queued steps do nothing
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    // MUSCL-Hancock steps alternate between the two state buffers
    if constexpr (SV == SV_HANCOCK) {
        if (dt_ptr[SC_PARITY] != 0.0) {
            double *t = const_cast<double *>(Uin);
            Uin = Uout;
            Uout = t;
        }
    }
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int Nx = P.N[0], Ny = P.N[1], Nz = P.N[2];
    const long long n = P.n;
    const long long sy = P.stride[1], sz = P.stride[2];
    int bx, by, bz;
    tile_coords(bx, by, bz);
    const int x0 = bx * TX;
    const int y0 = (DIM >= 2) ? by * TY : 0;
    const int kb = (DIM == 3) ? bz * kchunk : 0;
    const int ke = (DIM == 3) ? min(kb + kchunk, Nz) : 1;
    const int i = x0 + tx, j = y0 + ty;
    const bool act = (i < Nx) && (j < Ny);
    const double dt = *dt_ptr;
    const AxisGeo GX = axis_geo(P, geo, 0);
    const AxisGeo GY = axis_geo(P, geo, 1);
    const AxisGeo GZ = axis_geo(P, geo, 2);
    int err = 0;
    double sig = 0.0;

    auto wrapx = [&](int x) { return (x < 0) ? x + Nx : (x >= Nx ? x - Nx : x); };
    auto wrapy = [&](int y) { return (y < 0) ? y + Ny : (y >= Ny ? y - Ny : y); };
    auto wrapz = [&](int z) { return (z < 0) ? z + Nz : (z >= Nz ? z - Nz : z); };
    auto prims_at = [&](long long id, T *w) { cell_prims<T, NS, MIXED>(P, Uin, n, id, w); };

    // ---- z-direction register window of the own column (DIM == 3) --------
    T wz[5][NQ];   // primitives at k-2 .. k+2
    T lz0[NL], lzp[NL];
    T fzm[NF], fzp[NF];
    const bool zper = (DIM == 3) && P.periodic[2] != 0;
    auto zvalid = [&](int z) { return (z >= 0 && z < Nz) || zper; };

    if constexpr (DIM == 3) {
        if (act) {
            for (int d = 0; d < 4; ++d) {
                const int z = kb - 2 + d;
                if (zvalid(z))
                    prims_at(i + sy * j + sz * wrapz(z), wz[d]);
            }
            const CellGeo cg0 = cell_geo(P, geo, i, j, kb);
            T lzm[NL];
            const int km = kb - 1;
            if (zvalid(km)) {
                const int kmw = wrapz(km);
                const CellGeo cgm = cell_geo(P, geo, i, j, kmw);
                slopes_axis<T, NS, LIN>(P, GZ, 2, kmw, wz[0], wz[1], wz[2], T(cgm.S[2]),
                                        T(cgm.V), lzm);
            }
            slopes_axis<T, NS, LIN>(P, GZ, 2, kb, wz[1], wz[2], wz[3], T(cg0.S[2]), T(cg0.V), lz0);
            // z- face of plane kb
            if (!face_flux<T, NS, LIN>(P, GZ, 2, kb, wz[1], lzm, wz[2], lz0, fzm))
                err |= ERRF_RIEMANN;
        }
    }

    for (int k = kb; k < ke; ++k) {
        // ---- 1. primitives of plane k (and k+2 for the own column) ---------
        if constexpr (DIM == 3) {
            if (act) {
                if (zvalid(k + 2))
                    prims_at(i + sy * j + sz * wrapz(k + 2), wz[4]);
                // z slope at k+1 and z+ face of plane k
                if (zvalid(k + 1)) {
                    const int kp = wrapz(k + 1);
                    const CellGeo cgp = cell_geo(P, geo, i, j, kp);
                    slopes_axis<T, NS, LIN>(P, GZ, 2, kp, wz[2], wz[3], wz[4], T(cgp.S[2]),
                                            T(cgp.V), lzp);
                }
                if (!face_flux<T, NS, LIN>(P, GZ, 2, k + 1, wz[2], lz0, wz[3], lzp, fzp))
                    err |= ERRF_RIEMANN;
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    sW[q][ty + OY][tx + 2] = wz[2][q];
            }
        } else {
            if (act) {
                T w[NQ];
                prims_at(i + sy * j, w);
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    sW[q][ty + OY][tx + 2] = w[q];
            }
        }
        // halo cells of the plane
        {
            constexpr int NHX = 4 * TY;
            constexpr int NH = NHX + ((DIM >= 2) ? 4 * TX : 0);
            for (int h = tid; h < NH; h += NT) {
                int bxc, byc;
                if (h < NHX) {
                    byc = h / 4 + OY;
                    const int c = h % 4;
                    bxc = (c < 2) ? c : TX + c;
                } else {
                    const int hh = h - NHX;
                    bxc = hh % TX + 2;
                    const int r = hh / TX;
                    byc = (r < 2) ? r : TY + r;
                }
                int gx = x0 - 2 + bxc;
                int gy = y0 - OY + byc;
                bool ok = true;
                if (gx < 0 || gx >= Nx) {
                    if (P.periodic[0])
                        gx = wrapx(gx);
                    else
                        ok = false;
                }
                if (DIM >= 2 && (gy < 0 || gy >= Ny)) {
                    if (P.periodic[1])
                        gy = wrapy(gy);
                    else
                        ok = false;
                }
                // halo columns of a partial tile that fall inside the domain
                // are regular cells; rows/columns past the tile edge inside
                // the domain are covered here too
                if (ok && (gx >= Nx || (DIM >= 2 && gy >= Ny)))
                    ok = false;
                if (ok) {
                    T w[NQ];
                    prims_at(gx + sy * gy + (DIM == 3 ? sz * (long long)k : 0), w);
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
                        sW[q][byc][bxc] = w[q];
                }
            }
        }
        __syncthreads();

        // ---- 2. slopes along x and y ------------------------------------
        {
            constexpr int NTX = SXW * TY;
            constexpr int NTY = (DIM >= 2) ? TX * SYH : 0;
            for (int t = tid; t < NTX + NTY; t += NT) {
                if (t < NTX) {
                    const int u = t % SXW, r = t / SXW;
                    int cx = x0 - 1 + u;
                    const int cy = y0 + r;
                    if (cy >= Ny)
                        continue;
                    if (cx < 0 || cx >= Nx) {
                        if (!P.periodic[0])
                            continue;
                        cx = wrapx(cx);
                    }
                    T wm[NQ], w0[NQ], wp[NQ], L[NL];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        wm[q] = sW[q][r + OY][u];
                        w0[q] = sW[q][r + OY][u + 1];
                        wp[q] = sW[q][r + OY][u + 2];
                    }
                    const CellGeo cg = cell_geo(P, geo, cx, (DIM >= 2) ? cy : 0, (DIM == 3) ? k : 0);
                    slopes_axis<T, NS, LIN>(P, GX, 0, cx, wm, w0, wp, T(cg.S[0]), T(cg.V), L);
                    if constexpr (HK) {
                        T zl[NVH], zr[NVH];
                        hancock_sides_1d<T, NS, LIN>(P, GX, cx, w0, L, T(cg.V), T(cg.S[0]),
                                                     T(0.5 * dt), zl, zr);
#pragma unroll
                        for (int q = 0; q < NVH; ++q) {
                            sHl[q][u] = zl[q];
                            sHr[q][u] = zr[q];
                        }
                    } else {
#pragma unroll
                        for (int s = 0; s < NL; ++s)
                            sSx[s][r][u] = L[s];
                    }
                } else if constexpr (DIM >= 2) {
                    const int tt = t - NTX;
                    const int c = tt % TX, v = tt / TX;
                    const int cx = x0 + c;
                    int cy = y0 - 1 + v;
                    if (cx >= Nx)
                        continue;
                    if (cy < 0 || cy >= Ny) {
                        if (!P.periodic[1])
                            continue;
                        cy = wrapy(cy);
                    }
                    T wm[NQ], w0[NQ], wp[NQ], L[NL];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        wm[q] = sW[q][v][c + 2];
                        w0[q] = sW[q][v + 1][c + 2];
                        wp[q] = sW[q][v + 2][c + 2];
                    }
                    const CellGeo cg = cell_geo(P, geo, cx, cy, (DIM == 3) ? k : 0);
                    slopes_axis<T, NS, LIN>(P, GY, 1, cy, wm, w0, wp, T(cg.S[1]), T(cg.V), L);
#pragma unroll
                    for (int s = 0; s < NL; ++s)
                        sSy[s][v][c] = L[s];
                }
            }
        }
        __syncthreads();

        // ---- 3. x and y face fluxes ------------------------------------
        {
            constexpr int NTX = (TX + 1) * TY;
            constexpr int NTY = (DIM >= 2) ? TX * FYH : 0;
            for (int t = tid; t < NTX + NTY; t += NT) {
                T gq[NF];
                if (t < NTX) {
                    const int u = t % (TX + 1), r = t / (TX + 1);
                    const int fx = x0 + u;
                    if (fx > Nx || y0 + r >= Ny)
                        continue;
                    if constexpr (HK) {
                        // owner: x+ side of cell fx-1, neighbour: x- side of cell fx
                        T zo[NVH], zn[NVH];
#pragma unroll
                        for (int q = 0; q < NVH; ++q) {
                            zo[q] = sHr[q][u];
                            zn[q] = sHl[q][u + 1];
                        }
                        if (!face_flux_closed<T, NS, -1, true>(P, 0, fx, zo, zn, gq))
                            err |= ERRF_RIEMANN;
                    } else {
                        T wo[NQ], wn[NQ], Lo[NL], Ln[NL];
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            wo[q] = sW[q][r + OY][u + 1];
                            wn[q] = sW[q][r + OY][u + 2];
                        }
#pragma unroll
                        for (int s = 0; s < NL; ++s) {
                            Lo[s] = sSx[s][r][u];
                            Ln[s] = sSx[s][r][u + 1];
                        }
                        if (!face_flux<T, NS, LIN>(P, GX, 0, fx, wo, Lo, wn, Ln, gq))
                            err |= ERRF_RIEMANN;
                    }
#pragma unroll
                    for (int q = 0; q < NF; ++q)
                        sFx[q][r][u] = gq[q];
                } else if constexpr (DIM >= 2) {
                    const int tt = t - NTX;
                    const int c = tt % TX, v = tt / TX;
                    const int fy = y0 + v;
                    if (fy > Ny || x0 + c >= Nx)
                        continue;
                    T wo[NQ], wn[NQ], Lo[NL], Ln[NL];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        wo[q] = sW[q][v + 1][c + 2];
                        wn[q] = sW[q][v + 2][c + 2];
                    }
#pragma unroll
                    for (int s = 0; s < NL; ++s) {
                        Lo[s] = sSy[s][v][c];
                        Ln[s] = sSy[s][v + 1][c];
                    }
                    if (!face_flux<T, NS, LIN>(P, GY, 1, fy, wo, Lo, wn, Ln, gq))
                        err |= ERRF_RIEMANN;
#pragma unroll
                    for (int q = 0; q < NF; ++q)
                        sFy[q][v][c] = gq[q];
                }
            }
        }
        __syncthreads();

        // ---- 4. residual and update of the own cell ---------------------
        if (act) {
            const int kk = (DIM == 3) ? k : 0;
            const long long id = i + sy * j + sz * kk;
            const CellGeo cg = cell_geo(P, geo, i, (DIM >= 2) ? j : 0, kk);
            const bool xlo = (i == 0) && !P.periodic[0];
            const bool ylo = (DIM >= 2) && (j == 0) && !P.periodic[1];
            const bool zlo = (DIM == 3) && (k == 0) && !zper;
            auto F = [&](int s, int q) -> double {
                switch (s) {
                case 0: return double(sFx[q][ty][tx]);
                case 1: return double(sFx[q][ty][tx + 1]);
                case 2: return double(sFy[q][ty + 1][tx]);
                case 3: return double(sFy[q][ty][tx]);
                case 4: return double(fzm[q]);
                default: return double(fzp[q]);
                }
            };
            double u[NF], un[NF];
            load_cons<NS>(Uin, n, id, u);
            if constexpr (SV == SV_RK2)
                load_cons<NS>(Un, n, id, un);
            update_cell<NS, DIM, SV, MIXED>(P, cg, n, id, u, Uout, un, dt, xlo, ylo, zlo, F,
                                            sig);
        }

        if constexpr (DIM == 3) {
#pragma unroll
            for (int d = 0; d < 4; ++d)
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    wz[d][q] = wz[d + 1][q];
#pragma unroll
            for (int s = 0; s < NL; ++s)
                lz0[s] = lzp[s];
#pragma unroll
            for (int q = 0; q < NF; ++q)
                fzm[q] = fzp[q];
        }
    }

    if (err)
        atomicOr(errflag, err);
    if constexpr (SV != SV_RK1)
        reduce_signal<NT>(sig, sRed, sigmax);
}

// Mixed path of the first-order kernel: keep the own cell's FP64 state in
// registers between its load and its update (a reload misses in L2: a plane
// of every resident workgroup has passed through it since), and park the
// own column's z+ flux in LDS to make room; measured 3 % faster than the
// reload on the 3-D headline case, with 12-20 bytes of spills per lane.
#ifndef SG_MIXED_HOLD_U
#define SG_MIXED_HOLD_U 1
#endif
#ifndef SG_MIXED_FZ_LDS
#define SG_MIXED_FZ_LDS 1
#endif
// Mixed path, ideal mixture: closure straight from the conserved values
// (close_cell_lean), 2 % faster than the reference-order closure.
#ifndef SG_MIXED_LEAN
#define SG_MIXED_LEAN 1
#endif


// Stage kernel for schemes whose limited slopes vanish identically (zero
// limiter; TVD limiters on 2-D and 3-D meshes: the reference's ratio along an
// axis takes its min/max over all neighbours of the cell, and the neighbours
// across the other axes contribute exact zero entries, so r <= 0 and
// phi (r) = 0 for minmod, van Leer and van Albada, TVDLimiter.cpp:12-125,
// minmodLimiter.cpp:14-57, vanLeerLimiter.cpp:14-53,
// vanAlbadaLimiter.cpp:15-55).  Every face value is then the cell value
// ((phi G) . d + q = q), so each cell is closed with the EOS once, its closed
// state staged in LDS with a one-cell halo, and no slopes are stored.  Same
// z-marching tile and update as stage_kernel.
// The mixed path asks for 4 waves per SIMD (two 512-thread workgroups per
// CU); its register budget then allows a few spills, which measured faster
// than one workgroup per CU.
template <class T, int NS, int DIM, int SV, bool MIXED, int TX, int TY, int GAS>
__global__ void __launch_bounds__(TX * TY, (MIXED && TX * TY == 512) ? 4 : 1)
stage0_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
              const double *__restrict__ Uin, double *Uout, const double *Un,
              const double *__restrict__ dt_ptr, unsigned long long *sigmax, int *errflag,
              int kchunk)
{
    constexpr int NQ = NS + 4;
    constexpr bool H = (SV == SV_HANCOCK);
    constexpr int NV = Closed<T, NS, H>::NV;
    constexpr int NF = NS + 5;
    constexpr int NT = TX * TY;
    constexpr int BX = TX + 2;
    constexpr int BY = (DIM >= 2) ? TY + 2 : 1;
    constexpr int OY = (DIM >= 2) ? 1 : 0;
    constexpr int FYH = (DIM >= 2) ? TY + 1 : 1;

    __shared__ T sZ[NV][BY][BX];
    __shared__ T sFx[NF][TY][TX + 1];
    __shared__ T sFy[(DIM >= 2) ? NF : 1][FYH][TX];
    __shared__ double sRed[NT / 64];
    // mixed path: the own cell's FP64 state stays in registers (HOLD_U) or is
    // reloaded for the update; the own column's z+ flux waits in LDS (FZ_LDS)
    constexpr bool HOLD_U = !MIXED || SG_MIXED_HOLD_U;
    constexpr bool FZ_LDS = (DIM == 3) && MIXED && SG_MIXED_FZ_LDS;
    // mixed path, ideal mixture, RK stages: closure straight from the
    // conserved values (close_cell_lean)
    constexpr bool LEAN = MIXED && SG_MIXED_LEAN && GAS == GAS_IDEAL && !H;
    // halo slots of a plane: two columns of TY cells, two rows of TX cells
    constexpr int NHX = 2 * TY;
    constexpr int NH = NHX + ((DIM >= 2) ? 2 * TX : 0);
    __shared__ T sFz[FZ_LDS ? NF : 1][FZ_LDS ? TY : 1][FZ_LDS ? TX : 1];

    const DevParams &P = *Pp;
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    // MUSCL-Hancock steps alternate between the two state buffers
    if constexpr (SV == SV_HANCOCK) {
        if (dt_ptr[SC_PARITY] != 0.0) {
            double *t = const_cast<double *>(Uin);
            Uin = Uout;
            Uout = t;
        }
    }
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int Nx = P.N[0], Ny = P.N[1], Nz = P.N[2];
    const long long n = P.n;
    const long long sy = P.stride[1], sz = P.stride[2];
    int bx, by, bz;
    tile_coords(bx, by, bz);
    const int x0 = bx * TX;
    const int y0 = (DIM >= 2) ? by * TY : 0;
    const int kb = (DIM == 3) ? bz * kchunk : 0;
    const int ke = (DIM == 3) ? min(kb + kchunk, Nz) : 1;
    const int i = x0 + tx, j = y0 + ty;
    const bool act = (i < Nx) && (j < Ny);
    const double dt = *dt_ptr;
    int err = 0;
    double sig = 0.0;
    const bool zper = (DIM == 3) && P.periodic[2] != 0;

    // closed state of cell (ci, cj, ck) from its conserved values u
    auto closed_from = [&](const double *u, T *z, int ci, int cj, int ck) {
        if constexpr (LEAN) {
            float ut[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q)
                ut[q] = float(u[q]);
            close_cell_lean<NS>(P, ut, z);
            return;
        }
        T w[NQ];
        if constexpr (MIXED) {
            T ut[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q)
                ut[q] = T(u[q]);
            prims_from_cons<T, NS, GAS>(P, ut, w);
        } else {
            prims_from_cons<T, NS, GAS>(P, u, w);
        }
        if constexpr (H) {
            if (DIM == 3 && P.hi_flip[2] && ck == Nz - 1) {
                const CellGeo g = cell_geo(P, geo, ci, cj, ck);
                close_cell_hancock_zflip<T, NS, GAS>(P, w, T(g.S[2]), T(g.V), T(0.5 * dt), z);
            } else {
                close_cell_hancock<T, NS, GAS>(P, w, z);
            }
        } else {
            close_cell<T, NS, GAS>(P, w, z);
        }
    };
    auto closed_at = [&](int ci, int cj, int ck, T *z) {
        double u[NF];
        load_cons<NS>(Uin, n, ci + sy * cj + sz * (long long)ck, u);
        closed_from(u, z, ci, cj, ck);
    };
    // halo slot h of the tile: its LDS position and global cell; false when
    // it lies outside a non-periodic mesh.  x: the halo columns wrap on a
    // periodic axis; y: the halo rows (h >= NHX) wrap, the tile rows of the x
    // halo never do.
    auto halo_cell = [&](int h, int &bxc, int &byc, int &gx, int &gy) -> bool {
        if (h < NHX) {
            byc = h / 2 + OY;
            bxc = (h % 2) ? TX + 1 : 0;
        } else {
            const int hh = h - NHX;
            bxc = hh % TX + 1;
            byc = (hh / TX) ? TY + 1 : 0;
        }
        gx = x0 - 1 + bxc;
        gy = y0 - OY + byc;
        bool okx = true, oky = true;
        if (gx < 0 || gx >= Nx) {
            if (P.periodic[0] && (gx == -1 || gx == Nx))
                gx = (gx < 0) ? Nx - 1 : 0;
            else
                okx = false;
        }
        if (gy < 0 || gy >= Ny) {
            if (DIM >= 2 && h >= NHX && P.periodic[1] && (gy == -1 || gy == Ny))
                gy = (gy < 0) ? Ny - 1 : 0;
            else
                oky = false;
        }
        return okx && oky;
    };

    // own column: closed states of planes k and k+1, fluxes of its z faces;
    // the conserved values of plane k stay in registers for the update (a
    // reload would miss in L2: the plane was read one iteration earlier)
    T zc[NV], znx[NV];
    T fzm[NF], fzp[NF];
    double ucur[NF], unx[NF];
    if constexpr (DIM == 3) {
        if (act) {
            load_cons<NS>(Uin, n, i + sy * j + sz * kb, ucur);
            closed_from(ucur, zc, i, j, kb);
            if (kb > 0 || zper) {
                T zm[NV];
                const int km = (kb > 0) ? kb - 1 : Nz - 1;
                closed_at(i, j, km, zm);
                if (!face_flux_closed<T, NS, GAS, H>(P, 2, kb, zm, zc, fzm))
                    err |= ERRF_RIEMANN;
            } else if (!face_flux_closed<T, NS, GAS, H>(P, 2, 0, zc, zc, fzm)) {
                err |= ERRF_RIEMANN;
            }
        }
    }

    for (int k = kb; k < ke; ++k) {
        // ---- 1. closed states of plane k (tile + one-cell halo) ----------
        if (act) {
            if constexpr (DIM == 3) {
                if ((k + 1 < Nz) || zper) {
                    const int kp = (k + 1 < Nz) ? k + 1 : 0;
                    load_cons<NS>(Uin, n, i + sy * j + sz * kp, unx);
                    closed_from(unx, znx, i, j, kp);
                } else {
#pragma unroll
                    for (int q = 0; q < NV; ++q)
                        znx[q] = zc[q];  // not read: the top face is a boundary face
                }
                if (!face_flux_closed<T, NS, GAS, H>(P, 2, k + 1, zc, znx, fzp))
                    err |= ERRF_RIEMANN;
                if constexpr (FZ_LDS) {
#pragma unroll
                    for (int q = 0; q < NF; ++q)
                        sFz[q][ty][tx] = fzp[q];
                }
            } else {
                load_cons<NS>(Uin, n, i + sy * j, ucur);
                closed_from(ucur, zc, i, j, 0);
            }
#pragma unroll
            for (int q = 0; q < NV; ++q)
                sZ[q][ty + OY][tx + 1] = zc[q];
        }
        for (int h = tid; h < NH; h += NT) {
            int bxc, byc, gx, gy;
            if (halo_cell(h, bxc, byc, gx, gy)) {
                T z[NV];
                closed_at(gx, gy, (DIM == 3) ? k : 0, z);
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    sZ[q][byc][bxc] = z[q];
            }
        }
        __syncthreads();

        // ---- 2. x and y face fluxes ------------------------------------
        {
            constexpr int NTX = (TX + 1) * TY;
            constexpr int NTY = (DIM >= 2) ? TX * FYH : 0;
            for (int t = tid; t < NTX + NTY; t += NT) {
                T g[NF];
                T zo[NV], zn[NV];
                if (t < NTX) {
                    const int u = t % (TX + 1), r = t / (TX + 1);
                    const int fx = x0 + u;
                    if (fx > Nx || y0 + r >= Ny)
                        continue;
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        zo[q] = sZ[q][r + OY][u];
                        zn[q] = sZ[q][r + OY][u + 1];
                    }
                    if (!face_flux_closed<T, NS, GAS, H>(P, 0, fx, zo, zn, g))
                        err |= ERRF_RIEMANN;
#pragma unroll
                    for (int q = 0; q < NF; ++q)
                        sFx[q][r][u] = g[q];
                } else if constexpr (DIM >= 2) {
                    const int tt = t - NTX;
                    const int c = tt % TX, v = tt / TX;
                    const int fy = y0 + v;
                    if (fy > Ny || x0 + c >= Nx)
                        continue;
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        zo[q] = sZ[q][v][c + 1];
                        zn[q] = sZ[q][v + 1][c + 1];
                    }
                    if (!face_flux_closed<T, NS, GAS, H>(P, 1, fy, zo, zn, g))
                        err |= ERRF_RIEMANN;
#pragma unroll
                    for (int q = 0; q < NF; ++q)
                        sFy[q][v][c] = g[q];
                }
            }
        }
        __syncthreads();

        // ---- 3. residual and update of the own cell ---------------------
        if (act) {
            const int kk = (DIM == 3) ? k : 0;
            const long long id = i + sy * j + sz * kk;
            const CellGeo cg = cell_geo(P, geo, i, (DIM >= 2) ? j : 0, kk);
            const bool xlo = (i == 0) && !P.periodic[0];
            const bool ylo = (DIM >= 2) && (j == 0) && !P.periodic[1];
            const bool zlo = (DIM == 3) && (k == 0) && !zper;
            auto F = [&](int s, int q) -> double {
                switch (s) {
                case 0: return double(sFx[q][ty][tx]);
                case 1: return double(sFx[q][ty][tx + 1]);
                case 2: return double(sFy[q][ty + 1][tx]);
                case 3: return double(sFy[q][ty][tx]);
                case 4: return double(fzm[q]);
                default: return double(FZ_LDS ? sFz[q][ty][tx] : fzp[q]);
                }
            };
            double un[NF];
            if constexpr (SV == SV_RK2)
                load_cons<NS>(Un, n, id, un);
            if constexpr (!HOLD_U)
                load_cons<NS>(Uin, n, id, ucur);

            update_cell<NS, DIM, SV, MIXED, GAS>(P, cg, n, id, ucur, Uout, un, dt, xlo, ylo,
                                                zlo, F, sig);
        }
        if constexpr (DIM == 3) {
#pragma unroll
            for (int q = 0; q < NF; ++q)
                ucur[q] = unx[q];
#pragma unroll
            for (int q = 0; q < NV; ++q)
                zc[q] = znx[q];
#pragma unroll
            for (int q = 0; q < NF; ++q)
                fzm[q] = FZ_LDS ? sFz[q][ty][tx] : fzp[q];
        }
    }

    if (err)
        atomicOr(errflag, err);
    if constexpr (SV != SV_RK1)
        reduce_signal<NT>(sig, sRed, sigmax);
}

}  // namespace sg
