// geometry.cpp -- per-axis geometry tables (see geometry.hpp).
#include "geometry.hpp"

#include <cmath>
#include <cstdio>

namespace sg {

// libstdc++ 11 std::hypot(x, y, z) (vector::mag, vector.cpp:37-43).
double hypot3_host(double x, double y, double z)
{
    x = std::fabs(x);
    y = std::fabs(y);
    z = std::fabs(z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a != 0.0)
        return a * std::sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
    return 0.0;
}

namespace {

struct AxisDef {
    int N;
    bool active;
    double o, d, L;  // origin, spacing, extent of the axis
    // vertex coordinate v (0..N) of the cells (mesh.cpp:311-341, 536-578,
    // 1229-1285): active axes o + d v, inactive axes 0 or the extent
    double vert(int v) const
    {
        if (!active)
            return v == 0 ? 0. : L;
        return o + d * static_cast<double>(v);
    }
    // coordinate of the face plane f (tail/right/bottom use o + 0., the inner
    // faces o + d f, point/left/top o + L; mesh.cpp:1313, 1378, 1526, ...)
    double face_plane(int f) const
    {
        if (f == 0)
            return o + 0.;
        if (f == N)
            return o + L;
        return o + d * static_cast<double>(f);
    }
};

// centre of a cell along an axis: the 8 vertices summed in the order
// r000, rX00, r0Y0, rXY0, r00Z, rX0Z, r0YZ, rXYZ, then / 8 (mesh.cpp:1287-1289)
double cell_centre(int axis, double lo, double hi)
{
    static const int pat[3][8] = {{0, 1, 0, 1, 0, 1, 0, 1},
                                  {0, 0, 1, 1, 0, 0, 1, 1},
                                  {0, 0, 0, 0, 1, 1, 1, 1}};
    double s = pat[axis][0] ? hi : lo;
    for (int q = 1; q < 8; ++q)
        s = s + (pat[axis][q] ? hi : lo);
    return s / 8.;
}

// centre of a face along a transverse axis from its vertex pattern over
// r00, rX0, r0Y, rXY, then / 4 (mesh.cpp:1337-1338)
double face_centre(const int pat[4], double lo, double hi)
{
    double s = pat[0] ? hi : lo;
    for (int q = 1; q < 4; ++q)
        s = s + (pat[q] ? hi : lo);
    return s / 4.;
}

}  // namespace

MeshTables build_tables(int dim, const int64_t n[3], const double extent[3],
                        const double origin[3], const bool periodic[3])
{
    MeshTables M;
    M.dim = dim;
    AxisDef A[3];
    for (int a = 0; a < 3; ++a) {
        A[a].active = a < dim;
        A[a].N = A[a].active ? static_cast<int>(n[a]) : 1;
        A[a].L = extent[a];
        A[a].o = A[a].active ? origin[a] : 0.;
        // spacing: dz uses the y extent in 3D (mesh.cpp:1219)
        const double Lspace = (dim == 3 && a == 2) ? extent[1] : extent[a];
        A[a].d = Lspace / static_cast<double>(A[a].N);
    }
    for (int a = 0; a < 3; ++a) {
        AxisTables &T = M.ax[a];
        const AxisDef &D = A[a];
        const int N = D.N;
        T.N = N;
        T.active = D.active;
        T.periodic = D.active && periodic[a];
        T.xc.resize(N);
        T.e.resize(N);
        T.xf.assign(N + 1, 0.);
        T.wo.assign(N + 1, 1.);
        T.wn.assign(N + 1, 0.);
        T.dn.assign(N + 1, 0.);
        T.dlo.resize(N);
        T.dhi.resize(N);
        for (int i = 0; i < N; ++i) {
            const double lo = D.vert(i), hi = D.vert(i + 1);
            T.xc[i] = cell_centre(a, lo, hi);
            T.e[i] = hypot3_host(hi - lo, 0., 0.);
        }
        if (!D.active)
            continue;
        for (int f = 0; f <= N; ++f) {
            const double X = D.face_plane(f);
            T.xf[f] = (((X + X) + X) + X) / 4.;
        }
        for (int i = 0; i < N; ++i) {
            T.dlo[i] = T.xf[i] - T.xc[i];
            T.dhi[i] = T.xf[i + 1] - T.xc[i];
        }
        // calculateWeights (mesh.cpp:55-86) for inner faces; limiter
        // neighbour distance |x_c(f) - x_c(f-1)| (TVDLimiter.cpp:35-38)
        for (int f = 1; f < N; ++f) {
            const double own = hypot3_host(T.xf[f] - T.xc[f - 1], 0., 0.);
            const double nei = hypot3_host(T.xf[f] - T.xc[f], 0., 0.);
            const double sum = own + nei;
            T.wo[f] = nei / sum;
            T.wn[f] = own / sum;
            T.dn[f] = hypot3_host(T.xc[f] - T.xc[f - 1], 0., 0.);
        }
        // boundary ghosts of the limiter: |2 (x_f - x_c)| (TVDLimiter.cpp:68-75)
        T.ghost_lo = hypot3_host((T.xf[0] - T.xc[0]) * 2, 0., 0.);
        T.ghost_hi = hypot3_host((T.xf[N] - T.xc[N - 1]) * 2, 0., 0.);
        if (T.periodic) {
            // periodic pairing (no reference counterpart): the face at the
            // low end is the face at the high end; owner N-1, neighbour 0
            const double own = hypot3_host(T.xf[N] - T.xc[N - 1], 0., 0.);
            const double nei = hypot3_host(T.xf[0] - T.xc[0], 0., 0.);
            const double sum = own + nei;
            T.wo[0] = T.wo[N] = nei / sum;
            T.wn[0] = T.wn[N] = own / sum;
            T.dn[0] = T.dn[N] = sum;
        }
    }
    // transverse face centres against cell centres
    auto check = [&](int fa, int b, const int pat[4], const char *blk) {
        const AxisTables &Tb = M.ax[b];
        for (int t = 0; t < Tb.N; ++t) {
            const double fc = face_centre(pat, A[b].vert(t), A[b].vert(t + 1));
            if (fc != Tb.xc[t]) {
                if (M.separable) {
                    char buf[160];
                    std::snprintf(buf, sizeof buf,
                                  "%s faces of axis %d: centre along axis %d at %d differs "
                                  "from the cell centre by %.3g",
                                  blk, fa, b, t, fc - Tb.xc[t]);
                    M.why = buf;
                }
                M.separable = false;
            }
        }
    };
    static const int LLHH[4] = {0, 0, 1, 1}, LHLH[4] = {0, 1, 0, 1};
    // transverse offsets per face class, from the vertex patterns below
    auto offsets = [&](int fa, int b, int cls, const int pat[4]) {
        const AxisTables &Tb = M.ax[b];
        std::vector<double> &v = M.td[fa][b][cls];
        v.resize(Tb.N);
        for (int t = 0; t < Tb.N; ++t)
            v[t] = face_centre(pat, A[b].vert(t), A[b].vert(t + 1)) - Tb.xc[t];
    };
    for (int cls = 0; cls < 2; ++cls) {
        offsets(0, 1, cls, LLHH);
        offsets(0, 2, cls, LHLH);
        offsets(1, 0, cls, cls ? LLHH : LHLH);
        offsets(1, 2, cls, cls ? LHLH : LLHH);
        offsets(2, 0, cls, LHLH);
        offsets(2, 1, cls, LLHH);
    }
    // x faces: y [lo,lo,hi,hi], z [lo,hi,lo,hi] in every dimension
    check(0, 1, LLHH, "x");
    check(0, 2, LHLH, "x");
    if (dim >= 2) {
        // y inner faces: x [lo,hi,lo,hi], z [lo,lo,hi,hi];
        // right/left faces: x [lo,lo,hi,hi], z [lo,hi,lo,hi]
        check(1, 0, LHLH, "y-inner");
        check(1, 2, LLHH, "y-inner");
        check(1, 0, LLHH, "y-boundary");
        check(1, 2, LHLH, "y-boundary");
    }
    if (dim == 3) {
        check(2, 0, LHLH, "z");
        check(2, 1, LLHH, "z");
    }
    return M;
}

std::vector<int64_t> face_permutation(int dim, const int64_t n[3])
{
    const int64_t Nx = n[0], Ny = dim >= 2 ? n[1] : 1, Nz = dim >= 3 ? n[2] : 1;
    const int64_t nfx = (Nx + 1) * Ny * Nz, nfy = dim >= 2 ? Nx * (Ny + 1) * Nz : 0;
    auto gx = [&](int64_t f, int64_t j, int64_t k) { return f + (Nx + 1) * (j + Ny * k); };
    auto gy = [&](int64_t i, int64_t f, int64_t k) { return nfx + i + Nx * (f + (Ny + 1) * k); };
    auto gz = [&](int64_t i, int64_t j, int64_t f) { return nfx + nfy + i + Nx * (j + Ny * f); };
    std::vector<int64_t> p;
    if (dim == 1) {
        for (int64_t f = 0; f <= Nx; ++f)
            p.push_back(f);
        return p;
    }
    if (dim == 2) {
        // tail, x-inner, y-inner, point, right, left (mesh.cpp:598-870)
        for (int64_t j = 0; j < Ny; ++j)
            p.push_back(gx(0, j, 0));
        for (int64_t j = 0; j < Ny; ++j)
            for (int64_t i = 0; i < Nx - 1; ++i)
                p.push_back(gx(i + 1, j, 0));
        for (int64_t j = 0; j < Ny - 1; ++j)
            for (int64_t i = 0; i < Nx; ++i)
                p.push_back(gy(i, j + 1, 0));
        for (int64_t j = 0; j < Ny; ++j)
            p.push_back(gx(Nx, j, 0));
        for (int64_t i = 0; i < Nx; ++i)
            p.push_back(gy(i, 0, 0));
        for (int64_t i = 0; i < Nx; ++i)
            p.push_back(gy(i, Ny, 0));
        return p;
    }
    // tail, x-inner, y-inner, z-inner, point, bottom, right, left, top
    // (mesh.cpp:1301-1790)
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t j = 0; j < Ny; ++j)
            p.push_back(gx(0, j, k));
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t j = 0; j < Ny; ++j)
            for (int64_t i = 0; i < Nx - 1; ++i)
                p.push_back(gx(i + 1, j, k));
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t j = 0; j < Ny - 1; ++j)
            for (int64_t i = 0; i < Nx; ++i)
                p.push_back(gy(i, j + 1, k));
    for (int64_t k = 0; k < Nz - 1; ++k)
        for (int64_t j = 0; j < Ny; ++j)
            for (int64_t i = 0; i < Nx; ++i)
                p.push_back(gz(i, j, k + 1));
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t j = 0; j < Ny; ++j)
            p.push_back(gx(Nx, j, k));
    for (int64_t j = 0; j < Ny; ++j)
        for (int64_t i = 0; i < Nx; ++i)
            p.push_back(gz(i, j, 0));
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t i = 0; i < Nx; ++i)
            p.push_back(gy(i, 0, k));
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t i = 0; i < Nx; ++i)
            p.push_back(gy(i, Ny, k));
    for (int64_t j = 0; j < Ny; ++j)
        for (int64_t i = 0; i < Nx; ++i)
            p.push_back(gz(i, j, Nz));
    return p;
}

}  // namespace sg
