// fv_star.hip -- the star states of every face for #Reconstruction star
// (fv_star.hpp), FP64 on both precision paths; compiled with
// -ffp-contract=off so that every multiply and add is rounded separately.
//
// star_kernel: one thread per cell, run on the input state of each SSP-RK2
// stage.  The cell evaluates the faces it owns -- its high face along every
// axis, and its low face on a non-periodic low boundary -- with the face
// sides of the stage (face values = cell values where the limited slopes
// vanish; 1-D TVD: the limited slopes of c_k, v, p, k, epsilon and a, as in
// turb_stage_kernel) and writes the star state; in the second stage it
// writes the average (S + S') / 2 with the first stage's value.  NT: the
// state carries rho k and rho epsilon (2) or not (0; k = epsilon = 0, which
// leaves the closure, the wave speeds and the star state of a laminar face
// unchanged); NA: the passive fields after rho epsilon (rho a; rho a, rho b).
#include "fv_star.hpp"
#include "fv_star_launch.hpp"

namespace sg {

template <int NS, int DIM, bool SLOPES, int GAS, int NT, int NA, bool AVG>
__global__ void __launch_bounds__(256)
star_kernel(const DevParams *__restrict__ Pp, const TurbParams *__restrict__ Tp,
            const double *__restrict__ geo, const double *__restrict__ Uin, double *S,
            const double *__restrict__ sc)
{
    using T = double;
    constexpr int NFS = NS + 5 + NT + NA;  // rows of the state
    constexpr int NQ = NS + 6 + NA;        // c, v, p, k, eps, a
    constexpr int NV = 5 + NA;             // star values per face
    const DevParams &P = *Pp;
    const TurbParams &TP = *Tp;
    if (sc[SC_DONE] != 0.0)
        return;
    const long long n = P.n;
    const long long id = turb_block(blockIdx.x, gridDim.x) * 256LL + threadIdx.x;
    if (id >= n)
        return;
    const int idx[3] = {int(id % P.N[0]), int((id / P.N[0]) % P.N[1]), int(id / P.stride[2])};
    const long long nft = P.star_nft;

    auto prims = [&](long long c, T *w) {
        T u[NS + 7 + NA];
#pragma unroll
        for (int q = 0; q < NS + 5; ++q)
            u[q] = Uin[q * n + c];
        u[NS + 5] = NT ? Uin[(NS + 5) * n + c] : T(0);
        u[NS + 6] = NT ? Uin[(NS + 6) * n + c] : T(0);
#pragma unroll
        for (int q = 0; q < NA; ++q)
            u[NS + 7 + q] = Uin[(NFS - NA + q) * n + c];
        prims_turb<T, NS, GAS, NA>(P, u, w);
    };
    auto nbr = [&](int a, int dlt) -> long long {
        const int Na = P.N[a];
        int j = idx[a] + dlt;
        j = (j % Na + Na) % Na;
        return id + (long long)(j - idx[a]) * P.stride[a];
    };
    // the second stage averages with the first stage's values: the cell's
    // high faces (every cell owns its high face along each axis) are loaded
    // before any work so that their latency overlaps the closure and the
    // Riemann solves
    T pre[DIM][NV];
    if constexpr (AVG) {
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            const long long f = star_face(P, a, idx[a] + 1, idx);
#pragma unroll
            for (int q = 0; q < NV; ++q)
                pre[a][q] = S[q * nft + f];
        }
    }
    // star state of the face with owner-side values o and neighbour-side
    // values b, written at star index f (old: the face's first-stage values
    // when already loaded)
    auto emit = [&](int a, T sg, const T *o, const T *b, long long f, const T *old) {
        TSide<T, NS, NA> O, B;
        close_side_turb<T, NS, GAS, NA>(P, o, o + NS, o[NS + 3], o[NS + 4], o[NS + 5], o + NS + 6,
                                        O);
        close_side_turb<T, NS, GAS, NA>(P, b, b + NS, b[NS + 3], b[NS + 4], b[NS + 5], b + NS + 6,
                                        B);
        T st[NV];
        star_state<T, NS, NA, GAS>(P, O, B, a, sg, st);
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double *s = S + q * nft + f;
            if constexpr (AVG)
                *s = ((old ? old[q] : *s) + st[q]) / 2;
            else
                *s = st[q];
        }
    };

    T w0[NQ];
    prims(id, w0);
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
                prims(ok ? nbr(0, d) : id, w[d + 2]);
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
        if (!per && i == 0) {
            side(0, false, so);
            ghost_turb<T, NS, NA>(P, TP, 0, 0, so, sb);
            emit(0, T(-1), so, sb, star_face(P, 0, 0, idx), nullptr);
        }
        side(0, true, so);
        if (!per && i == Nx - 1) {
            ghost_turb<T, NS, NA>(P, TP, 0, 1, so, sb);
            emit(0, P.hi_flip[0] ? T(-1) : T(1), so, sb, star_face(P, 0, i + 1, idx),
                 AVG ? pre[0] : nullptr);
        } else {
            side(1, false, sb);
            emit(0, T(1), so, sb, star_face(P, 0, i + 1, idx), AVG ? pre[0] : nullptr);
        }
    } else {
        // face values are the cell values
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            const int Na = P.N[a];
            const bool per = P.periodic[a] != 0;
            T wb[NQ];
            if (!per && idx[a] == 0) {
                ghost_turb<T, NS, NA>(P, TP, a, 0, w0, wb);
                emit(a, T(-1), w0, wb, star_face(P, a, 0, idx), nullptr);
            }
            if (!per && idx[a] == Na - 1) {
                ghost_turb<T, NS, NA>(P, TP, a, 1, w0, wb);
                emit(a, P.hi_flip[a] ? T(-1) : T(1), w0, wb, star_face(P, a, Na, idx),
                     AVG ? pre[a] : nullptr);
            } else {
                prims(nbr(a, 1), wb);
                emit(a, T(1), w0, wb, star_face(P, a, idx[a] + 1, idx),
                     AVG ? pre[a] : nullptr);
            }
        }
    }
}

namespace {

template <int NS, int DIM, bool SLOPES, int NT, int NA>
int launch_star_gas(const StarLaunch &L)
{
    const unsigned blocks = (unsigned)((L.n + 255) / 256);
#define SG_STAR(GASV, AV)                                                                  \
    hipLaunchKernelGGL((star_kernel<NS, DIM, SLOPES, GASV, NT, NA, AV>), dim3(blocks), dim3(256), \
                       0, L.stream, L.P, L.TP, L.geo, L.U, L.S, L.sc)
    if (L.gas_model == GAS_VDW && L.avg)
        SG_STAR(GAS_VDW, true);
    else if (L.gas_model == GAS_VDW)
        SG_STAR(GAS_VDW, false);
    else if (L.gas_model == GAS_IDEAL && L.avg)
        SG_STAR(GAS_IDEAL, true);
    else if (L.gas_model == GAS_IDEAL)
        SG_STAR(GAS_IDEAL, false);
    else
        return -1;
#undef SG_STAR
    return static_cast<int>(hipGetLastError());
}

template <int NS, int DIM, bool SLOPES> int launch_star_fields(const StarLaunch &L)
{
    if (L.nt == 0 && L.na == 0)
        return launch_star_gas<NS, DIM, SLOPES, 0, 0>(L);
    if (L.nt == 2 && L.na == 0)
        return launch_star_gas<NS, DIM, SLOPES, 2, 0>(L);
    if (L.nt == 2 && L.na == 3)
        return launch_star_gas<NS, DIM, SLOPES, 2, 3>(L);
    if (L.nt == 2 && L.na == 4)
        return launch_star_gas<NS, DIM, SLOPES, 2, 4>(L);
    return -1;
}

template <int NS> int launch_star_dim(const StarLaunch &L)
{
    switch (L.dim) {
    case 1:
        return L.slopes ? launch_star_fields<NS, 1, true>(L) : launch_star_fields<NS, 1, false>(L);
    case 2:
        return L.slopes ? -1 : launch_star_fields<NS, 2, false>(L);
    case 3:
        return L.slopes ? -1 : launch_star_fields<NS, 3, false>(L);
    }
    return -1;
}

}  // namespace

int launch_star(const StarLaunch &L)
{
    switch (L.ns) {
    case 1:
        return launch_star_dim<1>(L);
    case 2:
        return launch_star_dim<2>(L);
    case 3:
        return launch_star_dim<3>(L);
    }
    return -1;
}

}  // namespace sg
