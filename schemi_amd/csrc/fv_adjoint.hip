// fv_adjoint.hip -- adjoint of the first-order-exact stage and the mixing
// width functional (no counterpart in the reference).
//
// A stage of the first-order-exact family (zero limiter; TVD limiters in 2-D
// and 3-D, whose slopes vanish identically) maps the FP64 state U to
//   L(U)_i = U_i - (dt / V_i) sum_f s_if S_i,f g_f(U_o(f), U_n(f)) + G_i(U_i),
// with g_f the HLLC/HLL flux between the closed states of the face's two
// cells (boundary faces: the owner and its ghost), s_if = +1 for the owner
// and -1 for the neighbour, G_i the gravity source.  Its adjoint is
//   (dL/dU)^T lam |_j = lam_j + (dG_j/dU_j)^T lam_j - sum_{f of j} (dg_f/dU_j)^T mu_f,
//   mu_f = sum_{i of f} s_if S_i,f (dt / V_i) lam_i.
// adj_stage_kernel evaluates it per cell: for each of its faces and each
// conserved component q of its own state it evaluates g_f with tangent
// numbers seeded along q (branches frozen at the forward state) and
// accumulates -mu_f . dg_f/dU_j[q]; the cell's result is written once, so the
// output is deterministic.  The RK2 step's adjoint composes two of them with
// the weights of U^{n+1} = (L(L(U^n)) + U^n) / 2 (capi.hip); out may then be
// the same array as add (an in-place step): each thread reads add and writes
// out for its own cell only.  On 1-D meshes
// with limited slopes adj_stage_1d_kernel does the same over the four faces
// whose fluxes a cell reaches through the slopes.
#include "fv_launch.hpp"
#include "fv_stage.hpp"
#include "fv_tangent.hpp"

namespace sg {

namespace {

// Closed state (close_cell layout) of a cell given as tangent numbers.
template <int NS, int GAS>
__device__ __forceinline__ void closed_tan(const DevParams &P, const Tan *u, Tan *z)
{
    Tan w[NS + 4];
    prims_from_cons<Tan, NS, GAS>(P, u, w);
    close_cell<Tan, NS, GAS>(P, w, z);
}

template <int NS, int DIM, int GAS>
__global__ void __launch_bounds__(256)
adj_stage_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
                 const double *__restrict__ U, const double *__restrict__ lam,
                 double *out, double dt, double a, const double *add,
                 double b)
{
    constexpr int NF = NS + 5;
    constexpr int NV = Closed<double, NS, false>::NV;
    const DevParams &P = *Pp;
    const long long n = P.n;
    const long long id = blockIdx.x * 256LL + threadIdx.x;
    if (id >= n)
        return;
    const int idx[3] = {int(id % P.N[0]), int((id / P.N[0]) % P.N[1]), int(id / P.stride[2])};
    const CellGeo cg = cell_geo(P, geo, idx[0], idx[1], idx[2]);
    double uj[NF], lj[NF], res[NF];
    load_cons<NS>(U, n, id, uj);
    load_cons<NS>(lam, n, id, lj);
#pragma unroll
    for (int q = 0; q < NF; ++q)
        res[q] = lj[q];

    if (P.gravity) {
        // E += rho (v . g) dt, m += (g rho) dt with pre-update rho, v
        // (Advection2dOrder.cpp:547-554)
        for (int q = 0; q < NF; ++q) {
            Tan u[NF];
#pragma unroll
            for (int r = 0; r < NF; ++r)
                u[r] = Tan(uj[r], r == q ? 1.0 : 0.0);
            const Tan rho = u[NS];
            const Tan vx = u[NS + 1] / rho, vy = u[NS + 2] / rho, vz = u[NS + 3] / rho;
            const Tan vg = ((Tan(0.0) + vx * P.g[0]) + vy * P.g[1]) + vz * P.g[2];
            double acc = lj[NS + 4] * (rho * vg * dt).d;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                acc += lj[NS + 1 + c] * ((P.g[c] * rho) * dt).d;
            res[q] += acc;
        }
    }

    // face slots x-, x+, y+, y-, z-, z+ (the residual's order)
    for (int s = 0; s < 2 * DIM; ++s) {
        const int ax = s >> 1;
        const bool hi = (s == 1 || s == 2 || s == 5);
        const int Na = P.N[ax];
        const bool per = P.periodic[ax] != 0;
        const int f = idx[ax] + (hi ? 1 : 0);
        const bool bnd = !per && (f == 0 || f == Na);
        const double Sj = cg.S[ax];
        double mu[NF];
        Tan zo[NV];  // the other cell's closed state (constant)
        if (bnd) {
#pragma unroll
            for (int q = 0; q < NF; ++q)
                mu[q] = Sj * (dt / cg.V) * lj[q];
        } else {
            int io = idx[ax] + (hi ? 1 : -1);
            if (io < 0)
                io = Na - 1;
            else if (io >= Na)
                io = 0;
            int oi[3] = {idx[0], idx[1], idx[2]};
            oi[ax] = io;
            const long long ido = id + (long long)(io - idx[ax]) * P.stride[ax];
            const CellGeo co = cell_geo(P, geo, oi[0], oi[1], oi[2]);
            double uo[NF], lo[NF];
            load_cons<NS>(U, n, ido, uo);
            load_cons<NS>(lam, n, ido, lo);
            // this cell is the owner of its high-side faces
            const double sj = hi ? 1.0 : -1.0;
#pragma unroll
            for (int q = 0; q < NF; ++q)
                mu[q] = sj * Sj * (dt / cg.V) * lj[q] - sj * co.S[ax] * (dt / co.V) * lo[q];
            Tan ut[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q)
                ut[q] = Tan(uo[q]);
            closed_tan<NS, GAS>(P, ut, zo);
        }
        for (int q = 0; q < NF; ++q) {
            Tan ut[NF], z[NV], g[NF];
#pragma unroll
            for (int r = 0; r < NF; ++r)
                ut[r] = Tan(uj[r], r == q ? 1.0 : 0.0);
            closed_tan<NS, GAS>(P, ut, z);
            if (bnd)
                face_flux_closed<Tan, NS, GAS, false>(P, ax, f, z, z, g);
            else if (hi)
                face_flux_closed<Tan, NS, GAS, false>(P, ax, f, z, zo, g);
            else
                face_flux_closed<Tan, NS, GAS, false>(P, ax, f, zo, z, g);
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < NF; ++r)
                acc += mu[r] * g[r].d;
            res[q] -= acc;
        }
    }
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        double o = a * res[q];
        if (add)
            o += b * add[q * n + id];
        out[q * n + id] = o;
    }
}

// Adjoint of a 1-D stage with limited slopes (TVD limiters, linear limiter):
// the flux of face f depends on the cells f-2 .. f+1 (the slopes of its two
// cells use their neighbours), so cell j seeds the fluxes of faces j-1 .. j+2.
// Each is evaluated with tangent numbers through the forward templates
// (primitives, slopes_axis with its limiter branches, reconstruction, EOS,
// boundary ghosts, HLLC/HLL) along every conserved component of cell j.
template <int NS, bool LIN>
__global__ void __launch_bounds__(256)
adj_stage_1d_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
                    const double *__restrict__ U, const double *__restrict__ lam,
                    double *out, double dt, double a, const double *add,
                    double b)
{
    constexpr int NF = NS + 5;
    constexpr int NQ = NS + 4;
    constexpr int NL = SlopeLayout<NS, LIN>::NL;
    const DevParams &P = *Pp;
    const long long n = P.n;
    const long long id = blockIdx.x * 256LL + threadIdx.x;
    if (id >= n)
        return;
    const int N = P.N[0];
    const bool per = P.periodic[0] != 0;
    const int j = int(id);
    const AxisGeo G = axis_geo(P, geo, 0);
    double lj[NF], res[NF];
    load_cons<NS>(lam, n, id, lj);
#pragma unroll
    for (int q = 0; q < NF; ++q)
        res[q] = lj[q];
    if (P.gravity) {
        double uj[NF];
        load_cons<NS>(U, n, id, uj);
        for (int q = 0; q < NF; ++q) {
            Tan u[NF];
#pragma unroll
            for (int r = 0; r < NF; ++r)
                u[r] = Tan(uj[r], r == q ? 1.0 : 0.0);
            const Tan rho = u[NS];
            const Tan vx = u[NS + 1] / rho, vy = u[NS + 2] / rho, vz = u[NS + 3] / rho;
            const Tan vg = ((Tan(0.0) + vx * P.g[0]) + vy * P.g[1]) + vz * P.g[2];
            double acc = lj[NS + 4] * (rho * vg * dt).d;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                acc += lj[NS + 1 + c] * ((P.g[c] * rho) * dt).d;
            res[q] += acc;
        }
    }
    auto wrap = [&](int i) { return per ? ((i % N) + N) % N : i; };
    auto exists = [&](int i) { return per || (i >= 0 && i < N); };
    for (int df = -1; df <= 2; ++df) {
        int f = j + df;  // face between cells f - 1 and f
        if (per)
            f = ((f % N) + N) % N;
        else if (f < 0 || f > N)
            continue;
        const bool lo_bnd = !per && f == 0, hi_bnd = !per && f == N;
        const int co = lo_bnd ? 0 : wrap(f - 1);  // owner
        const int cn = hi_bnd ? N - 1 : wrap(f);  // neighbour
        // mu_f: owner +S dt/V lam, neighbour -S dt/V lam (boundary: owner only)
        double mu[NF];
        {
            const CellGeo go = cell_geo(P, geo, co, 0, 0);
            double lo[NF];
            load_cons<NS>(lam, n, co, lo);
#pragma unroll
            for (int q = 0; q < NF; ++q)
                mu[q] = go.S[0] * (dt / go.V) * lo[q];
            if (!lo_bnd && !hi_bnd) {
                const CellGeo gn = cell_geo(P, geo, cn, 0, 0);
                double ln[NF];
                load_cons<NS>(lam, n, cn, ln);
#pragma unroll
                for (int q = 0; q < NF; ++q)
                    mu[q] -= gn.S[0] * (dt / gn.V) * ln[q];
            }
        }
        // cells f-2 .. f+1 around the face (owner co, neighbour cn)
        const int cells[4] = {co - 1, co, cn, cn + 1};
        for (int q = 0; q < NF; ++q) {
            Tan w[4][NQ];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int ci = wrap(cells[c]);
                if (!exists(cells[c])) {
#pragma unroll
                    for (int r = 0; r < NQ; ++r)
                        w[c][r] = Tan(0.0);
                    continue;
                }
                double uc[NF];
                load_cons<NS>(U, n, ci, uc);
                Tan u[NF];
#pragma unroll
                for (int r = 0; r < NF; ++r)
                    u[r] = Tan(uc[r], (ci == j && r == q) ? 1.0 : 0.0);
                prims_from_cons<Tan, NS>(P, u, w[c]);
            }
            Tan Lo[NL], Ln[NL], g[NF];
            {
                const CellGeo go = cell_geo(P, geo, co, 0, 0);
                slopes_axis<Tan, NS, LIN>(P, G, 0, co, w[0], w[1], w[2], Tan(go.S[0]),
                                          Tan(go.V), Lo);
                const CellGeo gn = cell_geo(P, geo, cn, 0, 0);
                slopes_axis<Tan, NS, LIN>(P, G, 0, cn, w[1], w[2], w[3], Tan(gn.S[0]),
                                          Tan(gn.V), Ln);
            }
            if (lo_bnd)  // owner is cell 0: its data sit in the neighbour slots
                face_flux<Tan, NS, LIN>(P, G, 0, f, w[2], Ln, w[2], Ln, g);
            else
                face_flux<Tan, NS, LIN>(P, G, 0, f, w[1], Lo, w[2], Ln, g);
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < NF; ++r)
                acc += mu[r] * g[r].d;
            res[q] -= acc;
        }
    }
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        double o = a * res[q];
        if (add)
            o += b * add[q * n + id];
        out[q * n + id] = o;
    }
}

// J = sum_i V_i Y_i (1 - Y_i), Y = rho_sp / rho: per-block partial sums in a
// fixed tree order, and dJ/dU when grad is given.
template <int NS>
__global__ void __launch_bounds__(256)
mixw_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ geo,
            const double *__restrict__ U, int sp, double *__restrict__ partial,
            double *__restrict__ grad)
{
    constexpr int NF = NS + 5;
    __shared__ double red[256];
    const DevParams &P = *Pp;
    const long long n = P.n;
    const long long id = blockIdx.x * 256LL + threadIdx.x;
    double v = 0.0;
    if (id < n) {
        const int i = int(id % P.N[0]), j = int((id / P.N[0]) % P.N[1]), k = int(id / P.stride[2]);
        const double V = cell_geo(P, geo, i, j, k).V;
        const double rs = U[sp * n + id], rho = U[NS * n + id];
        const double Y = rs / rho;
        v = V * (Y * (1.0 - Y));
        if (grad) {
            const double dY = V * (1.0 - 2.0 * Y);
            for (int q = 0; q < NF; ++q)
                grad[q * n + id] = 0.0;
            grad[sp * n + id] = dY / rho;
            grad[NS * n + id] = -dY * Y / rho;
        }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = red[0];
}

// Final sum of the partials, in index order, by one thread.
__global__ void mixw_final_kernel(const double *partial, long long nb, double *J)
{
    double s = 0.0;
    for (long long b = 0; b < nb; ++b)
        s += partial[b];
    *J = s;
}

template <int NS, int DIM>
int adj_dim(const AdjointLaunch &L)
{
    const dim3 g((unsigned)((L.n + 255) / 256));
    if constexpr (DIM == 1) {
        if (L.slopes) {
            if (L.lin)
                hipLaunchKernelGGL((adj_stage_1d_kernel<NS, true>), g, dim3(256), 0, L.stream,
                                   L.P, L.geo, L.U, L.lam, L.out, L.dt, L.a, L.add, L.b);
            else
                hipLaunchKernelGGL((adj_stage_1d_kernel<NS, false>), g, dim3(256), 0, L.stream,
                                   L.P, L.geo, L.U, L.lam, L.out, L.dt, L.a, L.add, L.b);
            return static_cast<int>(hipGetLastError());
        }
    }
    if (L.gas_model == GAS_STIFF)
        hipLaunchKernelGGL((adj_stage_kernel<NS, DIM, GAS_STIFF>), g, dim3(256), 0, L.stream,
                           L.P, L.geo, L.U, L.lam, L.out, L.dt, L.a, L.add, L.b);
    else if (L.gas_model == GAS_VDW)
        hipLaunchKernelGGL((adj_stage_kernel<NS, DIM, GAS_VDW>), g, dim3(256), 0, L.stream, L.P,
                           L.geo, L.U, L.lam, L.out, L.dt, L.a, L.add, L.b);
    else
        hipLaunchKernelGGL((adj_stage_kernel<NS, DIM, GAS_IDEAL>), g, dim3(256), 0, L.stream,
                           L.P, L.geo, L.U, L.lam, L.out, L.dt, L.a, L.add, L.b);
    return static_cast<int>(hipGetLastError());
}

template <int NS>
int adj_ns(const AdjointLaunch &L)
{
    switch (L.dim) {
    case 1:
        return adj_dim<NS, 1>(L);
    case 2:
        return adj_dim<NS, 2>(L);
    case 3:
        return adj_dim<NS, 3>(L);
    default:
        return -1;
    }
}

}  // namespace

int launch_adjoint_stage(const AdjointLaunch &L)
{
    switch (L.ns) {
    case 1:
        return adj_ns<1>(L);
    case 2:
        return adj_ns<2>(L);
    case 3:
        return adj_ns<3>(L);
    default:
        return -1;
    }
}

int launch_mixing_width(const DevParams *P, const double *geo, const double *U, int ns, long long n,
                        int sp, double *partial, double *J, double *grad, hipStream_t stream)
{
    const long long nb = (n + 255) / 256;
    const dim3 g((unsigned)nb);
    switch (ns) {
    case 1:
        hipLaunchKernelGGL(mixw_kernel<1>, g, dim3(256), 0, stream, P, geo, U, sp, partial, grad);
        break;
    case 2:
        hipLaunchKernelGGL(mixw_kernel<2>, g, dim3(256), 0, stream, P, geo, U, sp, partial, grad);
        break;
    case 3:
        hipLaunchKernelGGL(mixw_kernel<3>, g, dim3(256), 0, stream, P, geo, U, sp, partial, grad);
        break;
    default:
        return -1;
    }
    hipLaunchKernelGGL(mixw_final_kernel, dim3(1), dim3(1), 0, stream, partial, nb, J);
    return static_cast<int>(hipGetLastError());
}

}  // namespace sg
