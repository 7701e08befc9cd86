// fv_general.hpp -- stage kernels for the linear limiter on meshes whose face
// centres are offset from their cells' centres across the face normal.
//
// With a non-dyadic spacing the reference's face centres (4-vertex averages,
// mesh.cpp:1337-1338) and cell centres (8-vertex averages, mesh.cpp:1287-1289)
// round differently along the axes parallel to a face, so the reconstruction
// offset d = r_f - r_c has rounding-sized transverse components.  TVD limiters
// multiply them by zero slopes; the linear limiter keeps the full gradient
// (linearLimiter.cpp), so every face value is the full product
// ((0 + G_x d_x) + G_y d_y) + G_z d_z + q (Advection2dOrder.cpp:110-137).
// This path stores each cell's primitives and its gradient along every axis
// (one kernel), then updates every cell from its 2 D faces, each face
// evaluated by both of its cells with identical arithmetic (second kernel).
#pragma once

#include "fv_stage.hpp"

namespace sg {

// Primitives and Green-Gauss gradients (gradient.hpp:19-58,
// linearInterpolate.hpp:214-251) of every cell: W[q n + id],
// Lg[(a NQ + q) n + id] = d q / d x_a.
template <class T, int NS, int DIM, bool MIXED, bool HK = false>
__global__ void __launch_bounds__(256)
gen_grad_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
                const double *Uin, T *__restrict__ W, T *__restrict__ Lg,
                const double *__restrict__ dt_ptr, const double *Ualt = nullptr)
{
    constexpr int NQ = NS + 4;
    const DevParams &P = *Pp;
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    // MUSCL-Hancock steps alternate between the two state buffers
    if constexpr (HK) {
        if (dt_ptr[SC_PARITY] != 0.0)
            Uin = Ualt;
    }
    const long long n = P.n;
    const long long id = blockIdx.x * 256LL + threadIdx.x;
    if (id >= n)
        return;
    const int idx[3] = {int(id % P.N[0]), int((id / P.N[0]) % P.N[1]), int(id / P.stride[2])};
    T w0[NQ];
    cell_prims<T, NS, MIXED>(P, Uin, n, id, w0);
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        W[q * n + id] = w0[q];
    const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        const int Na = P.N[a];
        const bool per = P.periodic[a] != 0;
        T wm[NQ] = {}, wp[NQ] = {};
        if (idx[a] > 0 || per) {
            const long long nb = id + ((idx[a] > 0) ? -P.stride[a] : (Na - 1) * P.stride[a]);
            cell_prims<T, NS, MIXED>(P, Uin, n, nb, wm);
        }
        if (idx[a] < Na - 1 || per) {
            const long long nb = id + ((idx[a] < Na - 1) ? P.stride[a] : -(Na - 1) * P.stride[a]);
            cell_prims<T, NS, MIXED>(P, Uin, n, nb, wp);
        }
        T L[NQ];
        slopes_axis<T, NS, true>(P, axis_geo(P, geo, a), a, idx[a], wm, w0, wp, T(cg.S[a]),
                                 T(cg.V), L);
#pragma unroll
        for (int s = 0; s < NQ; ++s)
            Lg[(a * NQ + s) * n + id] = L[s];
    }
}

// Face value of a cell with primitives w and gradients L[b][q] at offset d:
// scalars dot3(L, d) + q with dot3 starting from 0, velocity components
// (T_a0 d0 + T_a1 d1 + T_a2 d2) + v_a (Advection2dOrder.cpp:110-137).
template <class T, int NS>
__device__ __forceinline__ void recon_full(const T *w, const T (*L)[NS + 4], const T *d, T *c,
                                           T *v, T &p)
{
    auto dot = [&](int q) { return ((T(0) + L[0][q] * d[0]) + L[1][q] * d[1]) + L[2][q] * d[2]; };
#pragma unroll
    for (int k = 0; k < NS; ++k)
        c[k] = dot(k) + w[k];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int q = NS + a;
        v[a] = ((L[0][q] * d[0] + L[1][q] * d[1]) + L[2][q] * d[2]) + w[q];
    }
    p = dot(NS + 3) + w[NS + 3];
}

template <class T, int NS, int DIM, bool STAGE2, bool MIXED>
__global__ void __launch_bounds__(256)
gen_update_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
                  const double *__restrict__ Uin, double *Uout, const double *Un,
                  const T *__restrict__ W, const T *__restrict__ Lg,
                  const double *__restrict__ dt_ptr, unsigned long long *sigmax, int *errflag)
{
    constexpr int NQ = NS + 4;
    constexpr int NF = NS + 5;
    __shared__ double sRed[4];
    const DevParams &P = *Pp;
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = blockIdx.x * 256LL + threadIdx.x;
    const bool act = id < n;
    const long long idc = act ? id : n - 1;
    const int idx[3] = {int(idc % P.N[0]), int((idc / P.N[0]) % P.N[1]),
                        int(idc / P.stride[2])};
    const double dt = *dt_ptr;
    int err = 0;
    double sig = 0.0;

    auto load = [&](long long c, T *w, T (*L)[NQ]) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            w[q] = W[q * n + c];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                L[a][q] = (a < DIM) ? Lg[(a * NQ + q) * n + c] : T(0);
    };
    T ws[NQ], Ls[3][NQ];
    load(idc, ws, Ls);

    // face slots x-, x+, y+, y-, z-, z+
    T g[2 * DIM][NF];
#pragma unroll
    for (int s = 0; s < 2 * DIM; ++s) {
        const int a = s >> 1;
        const int hi = (s == 1 || s == 2 || s == 5) ? 1 : 0;
        const int Na = P.N[a];
        const bool per = P.periodic[a] != 0;
        const int f = idx[a] + hi;
        const int cls = (f == 0 || f == Na) ? 1 : 0;
        T d[3];
#pragma unroll
        for (int b = 0; b < 3; ++b)
            d[b] = (b == a) ? T(0) : T(geo[P.off_td[a][b][cls] + idx[b]]);
        T cO[NS], vO[3], pO, cB[NS], vB[3], pB;
        T sgn = T(1);
        if (!per && (f == 0 || f == Na)) {
            // boundary face owned by this cell
            d[a] = T(f == 0 ? geo[P.off_dlo[a] + idx[a]] : geo[P.off_dhi[a] + idx[a]]);
            recon_full<T, NS>(ws, Ls, d, cO, vO, pO);
            ghost_side<T, NS>(P, a, f == 0 ? 0 : 1, cO, vO, pO, cB, vB, pB);
            sgn = (f == 0 || P.hi_flip[a]) ? T(-1) : T(1);
        } else {
            const int io = (f == 0) ? Na - 1 : f - 1;
            const int in = (f == Na) ? 0 : f;
            T wo[NQ], Lo[3][NQ], wn[NQ], Ln[3][NQ];
            load(idc + (long long)(io - idx[a]) * P.stride[a], wo, Lo);
            load(idc + (long long)(in - idx[a]) * P.stride[a], wn, Ln);
            d[a] = T(geo[P.off_dhi[a] + io]);
            recon_full<T, NS>(wo, Lo, d, cO, vO, pO);
            d[a] = T(geo[P.off_dlo[a] + in]);
            recon_full<T, NS>(wn, Ln, d, cB, vB, pB);
        }
        if (!riemann<T, NS>(P, a, sgn, cO, vO, pO, cB, vB, pB, g[s]))
            err |= ERRF_RIEMANN;
    }
    if (act) {
        const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
        const bool xlo = (idx[0] == 0) && !P.periodic[0];
        const bool ylo = (DIM >= 2) && (idx[1] == 0) && !P.periodic[1];
        const bool zlo = (DIM == 3) && (idx[2] == 0) && !P.periodic[2];
        auto F = [&](int s, int q) -> double { return double(g[s][q]); };
        double u[NF], un[NF];
        load_cons<NS>(Uin, n, id, u);
        if constexpr (STAGE2)
            load_cons<NS>(Un, n, id, un);
        update_cell<NS, DIM, STAGE2 ? SV_RK2 : SV_RK1, MIXED>(P, cg, n, id, u, Uout, un, dt,
                                                              xlo, ylo, zlo, F, sig);
    } else {
        err = 0;
    }
    if (err)
        atomicOr(errflag, err);
    if constexpr (STAGE2)
        reduce_signal<256>(sig, sRed, sigmax);
}

// ---- MUSCL-Hancock with full gradients (linear limiter, 2-D and 3-D) -----

// Offset of the centre of face slot s (x-, x+, y+, y-, z-, z+) of cell idx
// from the cell centre, its axis, whether the cell owns it (high side), and
// the face index along the axis.
__device__ __forceinline__ void slot_offset(const DevParams &P, const double *geo, int s,
                                            const int *idx, double *d, int &a, bool &hi,
                                            int &f)
{
    a = s >> 1;
    hi = (s == 1 || s == 2 || s == 5);
    const int Na = P.N[a];
    f = idx[a] + (hi ? 1 : 0);
    const int cls = (f == 0 || f == Na) ? 1 : 0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
        d[b] = (b == a) ? 0.0 : geo[P.off_td[a][b][cls] + idx[b]];
    d[a] = hi ? geo[P.off_dhi[a] + idx[a]] : geo[P.off_dlo[a] + idx[a]];
}

// Predicted face sides of every cell (HancockStage.cpp:10-448): the sides
// of all 2 DIM faces are reconstructed with the full gradient and closed,
// the divergence of their physical fluxes (normals out of the cell, the
// cell's faces in the order x-, x+, y+, y-, z-, z+) divided by V advances
// each side by half a step, and the sides are re-derived.  Hs holds slot s,
// value q of cell id at (s NVH + q) n + id.  Two passes over the sides (sum,
// then advance) keep one side live at a time.
template <class T, int NS, int DIM, bool MIXED>
__global__ void __launch_bounds__(256)
gen_hancock_sides_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
                         const T *__restrict__ W, const T *__restrict__ Lg, T *__restrict__ Hs,
                         const double *__restrict__ dt_ptr)
{
    constexpr int NQ = NS + 4;
    constexpr int NVH = Closed<T, NS, true>::NV;
    const DevParams &P = *Pp;
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = blockIdx.x * 256LL + threadIdx.x;
    if (id >= n)
        return;
    const int idx[3] = {int(id % P.N[0]), int((id / P.N[0]) % P.N[1]), int(id / P.stride[2])};
    T w[NQ], L[3][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        w[q] = W[q * n + id];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            L[a][q] = (a < DIM) ? Lg[(a * NQ + q) * n + id] : T(0);
    const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
    auto side = [&](int s, Side<T, NS> &sd, T *nrm, T &S) {
        double dd[3];
        int a, f;
        bool hi;
        slot_offset(P, geo, s, idx, dd, a, hi, f);
        const T d[3] = {T(dd[0]), T(dd[1]), T(dd[2])};
        T c[NS], v[3], p;
        recon_full<T, NS>(w, L, d, c, v, p);
        close_side<T, NS>(P, c, v, p, sd);
        // outward normal; the reversed top face of a 3-D mesh points inwards
        const bool flip = hi && f == P.N[a] && !P.periodic[a] && P.hi_flip[a];
#pragma unroll
        for (int b = 0; b < 3; ++b)
            nrm[b] = (b == a) ? ((hi && !flip) ? T(1) : T(-1)) : T(0);
        S = T(cg.S[a]);
    };
    T dr[NS + 1], dm[3] = {T(0), T(0), T(0)}, dE = T(0);
#pragma unroll
    for (int k = 0; k <= NS; ++k)
        dr[k] = T(0);
    for (int s = 0; s < 2 * DIM; ++s) {
        Side<T, NS> sd;
        T nrm[3], S;
        side(s, sd, nrm, S);
        hancock_face_terms<T, NS>(sd, nrm, S, dr, dm, dE);
    }
    hancock_divide<T, NS>(dr, dm, dE, T(cg.V));
    const T half = T(0.5 * *dt_ptr);
    for (int s = 0; s < 2 * DIM; ++s) {
        Side<T, NS> sd;
        T nrm[3], S, z[NVH];
        side(s, sd, nrm, S);
        hancock_advance<T, NS>(P, sd, dr, dm, dE, half, z);
#pragma unroll
        for (int q = 0; q < NVH; ++q)
            Hs[(s * NVH + q) * n + id] = z[q];
    }
}

// Flux of every face of a cell between the predicted sides (boundary faces:
// the owner's side and its ghost), evaluated by both cells of an inner face
// with the same arithmetic, then the update without RK average and the
// Courant speed of the new state.
template <class T, int NS, int DIM, bool MIXED>
__global__ void __launch_bounds__(256)
gen_hancock_update_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
                          const double *Uin, double *Uout, const T *__restrict__ Hs,
                          const double *__restrict__ dt_ptr, unsigned long long *sigmax,
                          int *errflag)
{
    constexpr int NF = NS + 5;
    constexpr int NVH = Closed<T, NS, true>::NV;
    __shared__ double sRed[4];
    const DevParams &P = *Pp;
    if (dt_ptr[SC_DONE] != 0.0)
        return;
    if (dt_ptr[SC_PARITY] != 0.0) {
        double *t = const_cast<double *>(Uin);
        Uin = Uout;
        Uout = t;
    }
    const long long n = P.n;
    const long long id = blockIdx.x * 256LL + threadIdx.x;
    const bool act = id < n;
    const long long idc = act ? id : n - 1;
    const int idx[3] = {int(idc % P.N[0]), int((idc / P.N[0]) % P.N[1]),
                        int(idc / P.stride[2])};
    const double dt = *dt_ptr;
    int err = 0;
    double sig = 0.0;
    T g[2 * DIM][NF];
    for (int s = 0; s < 2 * DIM; ++s) {
        const int a = s >> 1;
        const bool hi = (s == 1 || s == 2 || s == 5);
        const int Na = P.N[a];
        const bool per = P.periodic[a] != 0;
        const int f = idx[a] + (hi ? 1 : 0);
        T zo[NVH], zn[NVH];
#pragma unroll
        for (int q = 0; q < NVH; ++q)
            zo[q] = Hs[(s * NVH + q) * n + idc];
        bool ok;
        if (!per && (f == 0 || f == Na)) {
            ok = face_flux_closed<T, NS, -1, true>(P, a, f, zo, zo, g[s]);
        } else {
            int io = idx[a] + (hi ? 1 : -1);
            io = (io < 0) ? Na - 1 : ((io >= Na) ? 0 : io);
            const long long ido = idc + (long long)(io - idx[a]) * P.stride[a];
            // the other cell's side of this face is its opposite slot
#pragma unroll
            for (int q = 0; q < NVH; ++q)
                zn[q] = Hs[((s ^ 1) * NVH + q) * n + ido];
            ok = hi ? face_flux_closed<T, NS, -1, true>(P, a, f, zo, zn, g[s])
                    : face_flux_closed<T, NS, -1, true>(P, a, f, zn, zo, g[s]);
        }
        if (!ok)
            err |= ERRF_RIEMANN;
    }
    if (act) {
        const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
        const bool xlo = (idx[0] == 0) && !P.periodic[0];
        const bool ylo = (DIM >= 2) && (idx[1] == 0) && !P.periodic[1];
        const bool zlo = (DIM == 3) && (idx[2] == 0) && !P.periodic[2];
        auto F = [&](int s, int q) -> double { return double(g[s][q]); };
        double u[NF], un[NF];
        load_cons<NS>(Uin, n, id, u);
        update_cell<NS, DIM, SV_HANCOCK, MIXED>(P, cg, n, id, u, Uout, un, dt, xlo, ylo, zlo,
                                                F, sig);
    } else {
        err = 0;
    }
    if (err)
        atomicOr(errflag, err);
    reduce_signal<256>(sig, sRed, sigmax);
}

}  // namespace sg
