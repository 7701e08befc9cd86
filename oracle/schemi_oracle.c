/*
 * schemi_oracle.c -- CPU oracle (test infrastructure only).
 *
 * Plain-C restatement of the reference's explicit SSP-RK2 finite-volume step.
 * Every function cites the reference file:line it restates.  Operation order
 * follows the reference exactly; build with -ffp-contract=off.  The oracle is
 * the checker for the GPU library; no product code path may use it.
 *
 * Reference types are restated as:
 *   vector  -> double[3]   (Schemi/headers/vector.hpp)
 *   tensor  -> double[9]   row-major, t[3a+b] (Schemi/headers/tensor.hpp)
 */
#include "schemi_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define STAB (5.0 * DBL_EPSILON) /* globalConstants.hpp:33 stabilizator */
#define VERYBIG (1E-3 * DBL_MAX)   /* globalConstants.hpp:35 veryBig */

enum { ERR_NONE = 0, ERR_RIEMANN = 1, ERR_NEG_T = 2, ERR_BC = 5, ERR_INIT = 11 };

struct or_solver {
    or_config cfg;
    long nc, nf;
    double *crc, *cV, *cminlen;      /* cells: centre (3), volume, 1/min edge */
    double *frc, *fN, *fS;           /* faces: centre (3), unit normal (3), area */
    long *own, *nei;
    int *fpatch;                     /* -1 inner, else patch 0..5 */
    double *wo, *wn;
    int *cnf;                        /* faces per cell */
    long *cf;                        /* 6 per cell, surfacesOfCells order */
    int *cnn;                        /* neighbours per cell */
    long *cn;                        /* 6 per cell, neighboursOfCells order */
    double *cfd;                     /* 6 x 3 per cell: face centre - cell centre */
    double *cnd;                     /* 6 x 3 per cell: neighbour centre - cell centre */
    double *fper;                    /* periodic faces: centre distance across the face */
    double dt_src;                   /* source time step (mesh::timestepSource) */
    or_state st;
    double *pool;
    /* star states (starFields) of the last stage and of the step, 8 per
     * face: rho, v (3), p, a (3) */
    double *star_stage, *star;
    int hancock_step; /* the diffusion stage follows a MUSCL-Hancock step */
    int keep_b;       /* recover_primitives leaves b as it is */
};

/* star values per face: rho, v (3), p, a (3), b */
#define NSTAR 9

/* the model transports a (kEpsA, BHR) / b (BHR) */
static int a_field(const or_config *c)
{
    return c->turb && (c->turb_model == OR_TURB_KEPSA || c->turb_model == OR_TURB_BHR);
}
static int b_field(const or_config *c) { return c->turb && c->turb_model == OR_TURB_BHR; }

/* ------------------------------------------------------------------------ */
/* primitives                                                                */

/* libstdc++ 11 std::hypot(x, y, z) (bits/... cmath:1854-1865), used by
 * vector::mag (Schemi/src/primitives/vector.cpp:37-43). */
double or_hypot3(double x, double y, double z)
{
    x = fabs(x);
    y = fabs(y);
    z = fabs(z);
    double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a != 0.0)
        return a * sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
    return 0.0;
}

/* vector & vector (vector.cpp:109-119): result starts at 0. */
static double dot3(const double *a, const double *b)
{
    double r = 0;
    for (int i = 0; i < 3; ++i)
        r += a[i] * b[i];
    return r;
}

/* std::min / std::max semantics */
static double smin(double a, double b) { return (b < a) ? b : a; }
static double smax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* mesh (Schemi/src/dataStructures/mesh.cpp)                                */

static void set_cell(or_solver *s, long id, double X0, double X1, double Y0,
                     double Y1, double Z0, double Z1)
{
    /* vertices r000, rX00, r0Y0, rXY0, r00Z, rX0Z, r0YZ, rXYZ
     * (mesh.cpp:1229-1285); centre = sum / 8 (mesh.cpp:1287-1289). */
    const double vx[8] = {X0, X1, X0, X1, X0, X1, X0, X1};
    const double vy[8] = {Y0, Y0, Y1, Y1, Y0, Y0, Y1, Y1};
    const double vz[8] = {Z0, Z0, Z0, Z0, Z1, Z1, Z1, Z1};
    double sx = vx[0], sy = vy[0], sz = vz[0];
    for (int q = 1; q < 8; ++q) {
        sx = sx + vx[q];
        sy = sy + vy[q];
        sz = sz + vz[q];
    }
    s->crc[3 * id + 0] = sx / 8.;
    s->crc[3 * id + 1] = sy / 8.;
    s->crc[3 * id + 2] = sz / 8.;
    /* V = |rX00-r000| |r0Y0-r000| |r00Z-r000| (mesh.cpp:1291-1293) */
    const double e1 = or_hypot3(X1 - X0, Y0 - Y0, Z0 - Z0);
    const double e2 = or_hypot3(X0 - X0, Y1 - Y0, Z0 - Z0);
    const double e3 = or_hypot3(X0 - X0, Y0 - Y0, Z1 - Z0);
    s->cV[id] = e1 * e2 * e3;
    /* minimalLengthScale = 1/min edge (Schemi.cpp:649-662) */
    const double me = smin(smin(e1, e2), e3);
    s->cminlen[id] = 1. / me;
}

/* Face from its four vertices r00, rX0, r0Y, rXY (mesh.cpp:1337-1340). */
static void set_face(or_solver *s, long id, const double *r00,
                     const double *rX0, const double *r0Y, const double *rXY,
                     long owner, long neigh, int patch)
{
    for (int a = 0; a < 3; ++a)
        s->frc[3 * id + a] = (((r00[a] + rX0[a]) + r0Y[a]) + rXY[a]) / 4.;
    double V1[3], V2[3];
    for (int a = 0; a < 3; ++a) {
        V1[a] = rX0[a] - r00[a];
        V2[a] = r0Y[a] - r00[a];
    }
    s->fS[id] = or_hypot3(V1[0], V1[1], V1[2]) * or_hypot3(V2[0], V2[1], V2[2]);
    /* normal = V1 ^ V2 normalised (mesh.cpp:34-53; vector.cpp:121-130) */
    double N[3];
    N[0] = V1[1] * V2[2] - V1[2] * V2[1];
    N[1] = -(V1[0] * V2[2] - V1[2] * V2[0]);
    N[2] = V1[0] * V2[1] - V1[1] * V2[0];
    const double mag = or_hypot3(N[0], N[1], N[2]);
    for (int a = 0; a < 3; ++a)
        N[a] = N[a] / mag;
    s->own[id] = owner;
    s->nei[id] = neigh;
    s->fpatch[id] = patch;
    memcpy(&s->fN[3 * id], N, sizeof N);
}

static void orient_and_weights(or_solver *s)
{
    for (long f = 0; f < s->nf; ++f) {
        double *N = &s->fN[3 * f];
        const long o = s->own[f];
        double d[3];
        for (int a = 0; a < 3; ++a)
            d[a] = s->frc[3 * f + a] - s->crc[3 * o + a];
        if (dot3(d, N) < 0.)
            for (int a = 0; a < 3; ++a)
                N[a] = N[a] * (-1.);
        /* calculateWeights (mesh.cpp:55-86) */
        if (s->fpatch[f] < 0) {
            const long n = s->nei[f];
            double dn[3];
            for (int a = 0; a < 3; ++a)
                dn[a] = s->frc[3 * f + a] - s->crc[3 * n + a];
            const double ro = or_hypot3(d[0], d[1], d[2]);
            const double rn = or_hypot3(dn[0], dn[1], dn[2]);
            const double sum = ro + rn;
            s->wo[f] = rn / sum;
            s->wn[f] = ro / sum;
        } else {
            s->wo[f] = 1.0;
            s->wn[f] = 0.0;
        }
    }
}

static void vtx(double *r, double x, double y, double z)
{
    r[0] = x;
    r[1] = y;
    r[2] = z;
}

/* Periodic boundaries (no counterpart in the reference): when both patches
 * of an axis are periodic for every field, the low-side boundary face of each
 * row is replaced in its cell's face list by the high-side face of the same
 * row, which becomes an inner face with owner = last cell and neighbour =
 * first cell.  Its weights use the two half-distances, the wrapped neighbours
 * are at distance own + nei along the axis, and the low-side cell keeps its own
 * face offset for the reconstruction.  The low-side faces are left unused. */
static int axis_periodic(const or_solver *s, int a)
{
    static const int lo[3] = {0, 3, 2}, hi[3] = {1, 4, 5};
    const or_config *c = &s->cfg;
    if (a >= c->dim)
        return 0;
    int n = 0, tot = 0;
    for (int side = 0; side < 2; ++side) {
        const int p = side ? hi[a] : lo[a];
        for (int k = 0; k < c->Ns; ++k) {
            n += c->bc_c[k][p].type == OR_BC_PERIODIC;
            ++tot;
        }
        n += c->bc_v[p].type == OR_BC_PERIODIC;
        n += c->bc_p[p].type == OR_BC_PERIODIC;
        tot += 2;
    }
    return n == tot ? 1 : (n == 0 ? 0 : -1);
}

static void pair_periodic(or_solver *s)
{
    const or_config *c = &s->cfg;
    const long Nx = c->N[0], Ny = c->dim >= 2 ? c->N[1] : 1, Nz = c->dim >= 3 ? c->N[2] : 1;
    const long NN[3] = {Nx, Ny, Nz};
    /* slot of the low/high face of axis a in surfacesOfCells: x-, x+, y+, y-, z-, z+ */
    static const int slot_lo[3] = {0, 3, 4}, slot_hi[3] = {1, 2, 5};
    int any = 0;
    for (int a = 0; a < c->dim; ++a) {
        if (axis_periodic(s, a) != 1)
            continue;
        any = 1;
        for (long id = 0; id < s->nc; ++id) {
            const long idx[3] = {id % Nx, (id / Nx) % Ny, id / (Nx * Ny)};
            if (idx[a] != 0)
                continue;
            long idh[3] = {idx[0], idx[1], idx[2]};
            idh[a] = NN[a] - 1;
            const long ch = idh[0] + Nx * (idh[1] + Ny * idh[2]);
            const long flo = s->cf[6 * id + slot_lo[a]];
            const long fhi = s->cf[6 * ch + slot_hi[a]];
            const double own = or_hypot3(s->frc[3 * fhi] - s->crc[3 * ch],
                                         s->frc[3 * fhi + 1] - s->crc[3 * ch + 1],
                                         s->frc[3 * fhi + 2] - s->crc[3 * ch + 2]);
            const double nei = or_hypot3(s->frc[3 * flo] - s->crc[3 * id],
                                         s->frc[3 * flo + 1] - s->crc[3 * id + 1],
                                         s->frc[3 * flo + 2] - s->crc[3 * id + 2]);
            const double sum = own + nei;
            s->wo[fhi] = nei / sum;
            s->wn[fhi] = own / sum;
            s->fpatch[fhi] = -1;
            s->nei[fhi] = id;
            s->fpatch[flo] = -3;
            s->cf[6 * id + slot_lo[a]] = fhi;
            s->fper[fhi] = sum;
        }
    }
    if (!any)
        return;
    /* neighbour lists in the order x-, x+, y+, y-, z-, z+ with wrapping */
    for (long id = 0; id < s->nc; ++id) {
        const long idx[3] = {id % Nx, (id / Nx) % Ny, id / (Nx * Ny)};
        static const int ax[6] = {0, 0, 1, 1, 2, 2}, dir[6] = {-1, 1, 1, -1, -1, 1};
        int n = 0;
        for (int q = 0; q < 2 * c->dim; ++q) {
            const int a = ax[q];
            long j[3] = {idx[0], idx[1], idx[2]};
            j[a] += dir[q];
            int wrapped = 0;
            if (j[a] < 0 || j[a] >= NN[a]) {
                if (axis_periodic(s, a) != 1)
                    continue;
                j[a] = (j[a] + NN[a]) % NN[a];
                wrapped = 1;
            }
            const long nb = j[0] + Nx * (j[1] + Ny * j[2]);
            s->cn[6 * id + n] = nb;
            for (int b = 0; b < 3; ++b)
                s->cnd[18 * id + 3 * n + b] = s->crc[3 * nb + b] - s->crc[3 * id + b];
            if (wrapped) {
                /* face slot of this direction holds the periodic face */
                const long f = s->cf[6 * id + (dir[q] < 0 ? (a == 0 ? 0 : a == 1 ? 3 : 4)
                                                         : (a == 0 ? 1 : a == 1 ? 2 : 5))];
                for (int b = 0; b < 3; ++b)
                    s->cnd[18 * id + 3 * n + b] = 0.0;
                s->cnd[18 * id + 3 * n + a] = dir[q] * s->fper[f];
            }
            ++n;
        }
        s->cnn[id] = n;
    }
}

static int build_mesh(or_solver *s)
{
    const or_config *c = &s->cfg;
    const long Nx = c->N[0], Ny = c->dim >= 2 ? c->N[1] : 1,
               Nz = c->dim >= 3 ? c->N[2] : 1;
    const double Lx = c->L[0], Ly = c->L[1], Lz = c->L[2];
    const double ox = c->origin[0], oy = c->origin[1], oz = c->origin[2];
    double r00[3], rX0[3], r0Y[3], rXY[3];
    const long NONE = -1;

    s->nc = Nx * Ny * Nz;
    if (c->dim == 1)
        s->nf = Nx + 1;
    else if (c->dim == 2)
        s->nf = Ny * (Nx + 1) + Nx * (Ny + 1);
    else
        s->nf = Ny * Nz * (Nx + 1) + Nx * Nz * (Ny + 1) + Nx * Ny * (Nz + 1);

    const long nc = s->nc, nf = s->nf;
    s->crc = calloc(3 * nc, sizeof(double));
    s->cV = calloc(nc, sizeof(double));
    s->cminlen = calloc(nc, sizeof(double));
    s->frc = calloc(3 * nf, sizeof(double));
    s->fN = calloc(3 * nf, sizeof(double));
    s->fS = calloc(nf, sizeof(double));
    s->own = calloc(nf, sizeof(long));
    s->nei = calloc(nf, sizeof(long));
    s->fpatch = calloc(nf, sizeof(int));
    s->wo = calloc(nf, sizeof(double));
    s->wn = calloc(nf, sizeof(double));
    s->cnf = calloc(nc, sizeof(int));
    s->cf = calloc(6 * nc, sizeof(long));
    s->cnn = calloc(nc, sizeof(int));
    s->cn = calloc(6 * nc, sizeof(long));
    s->fper = calloc(nf, sizeof(double));
    if (!s->crc || !s->cn)
        return ERR_INIT;

#define CID(i, j, k) ((i) + Nx * ((j) + Ny * (k)))
    if (c->dim == 1) {
        /* oneDParallelepiped (mesh.cpp:296-520) */
        const double dx = Lx / Nx;
        for (long i = 0; i < Nx; ++i)
            set_cell(s, i, ox + dx * i, ox + dx * (i + 1), 0., Ly, 0., Lz);
        vtx(r00, ox + 0., 0., 0.);
        vtx(rX0, ox + 0., 0., Lz);
        vtx(r0Y, ox + 0., Ly, 0.);
        vtx(rXY, ox + 0., Ly, Lz);
        set_face(s, 0, r00, rX0, r0Y, rXY, 0, NONE, 0);
        for (long i = 1; i < Nx; ++i) {
            const double x = ox + dx * i;
            vtx(r00, x, 0., 0.);
            vtx(rX0, x, 0., Lz);
            vtx(r0Y, x, Ly, 0.);
            vtx(rXY, x, Ly, Lz);
            set_face(s, i, r00, rX0, r0Y, rXY, i - 1, i, -1);
        }
        vtx(r00, ox + Lx, 0., 0.);
        vtx(rX0, ox + Lx, 0., Lz);
        vtx(r0Y, ox + Lx, Ly, 0.);
        vtx(rXY, ox + Lx, Ly, Lz);
        set_face(s, Nx, r00, rX0, r0Y, rXY, Nx - 1, NONE, 1);
        for (long i = 0; i < Nx; ++i) {
            s->cnf[i] = 2;
            s->cf[6 * i + 0] = i;
            s->cf[6 * i + 1] = i + 1;
            int n = 0;
            if (i > 0)
                s->cn[6 * i + n++] = i - 1;
            if (i < Nx - 1)
                s->cn[6 * i + n++] = i + 1;
            s->cnn[i] = n;
        }
    } else if (c->dim == 2) {
        /* twoDParallelepiped (mesh.cpp:522-1203) */
        const double dx = Lx / Nx, dy = Ly / Ny;
        for (long j = 0; j < Ny; ++j)
            for (long i = 0; i < Nx; ++i)
                set_cell(s, CID(i, j, 0), ox + dx * i, ox + dx * (i + 1),
                         oy + dy * j, oy + dy * (j + 1), 0., Lz);
        const long T0 = 0, X0 = Ny, Y0 = X0 + Ny * (Nx - 1), P0 = Y0 + Nx * (Ny - 1),
                   R0 = P0 + Ny, Le0 = R0 + Nx;
        for (long j = 0; j < Ny; ++j) {
            const double x = ox + 0.;
            vtx(r00, x, oy + dy * j, 0.);
            vtx(rX0, x, oy + dy * j, Lz);
            vtx(r0Y, x, oy + dy * (j + 1), 0.);
            vtx(rXY, x, oy + dy * (j + 1), Lz);
            set_face(s, T0 + j, r00, rX0, r0Y, rXY, CID(0, j, 0), NONE, 0);
        }
        for (long j = 0; j < Ny; ++j)
            for (long i = 0; i < Nx - 1; ++i) {
                const double x = ox + dx * (i + 1);
                vtx(r00, x, oy + dy * j, 0.);
                vtx(rX0, x, oy + dy * j, Lz);
                vtx(r0Y, x, oy + dy * (j + 1), 0.);
                vtx(rXY, x, oy + dy * (j + 1), Lz);
                set_face(s, X0 + i + (Nx - 1) * j, r00, rX0, r0Y, rXY,
                         CID(i, j, 0), CID(i + 1, j, 0), -1);
            }
        for (long j = 0; j < Ny - 1; ++j)
            for (long i = 0; i < Nx; ++i) {
                const double y = oy + dy * (j + 1);
                vtx(r00, ox + dx * i, y, 0.);
                vtx(rX0, ox + dx * (i + 1), y, 0.);
                vtx(r0Y, ox + dx * i, y, Lz);
                vtx(rXY, ox + dx * (i + 1), y, Lz);
                set_face(s, Y0 + i + Nx * j, r00, rX0, r0Y, rXY, CID(i, j, 0),
                         CID(i, j + 1, 0), -1);
            }
        for (long j = 0; j < Ny; ++j) {
            const double x = ox + Lx;
            vtx(r00, x, oy + dy * j, 0.);
            vtx(rX0, x, oy + dy * j, Lz);
            vtx(r0Y, x, oy + dy * (j + 1), 0.);
            vtx(rXY, x, oy + dy * (j + 1), Lz);
            set_face(s, P0 + j, r00, rX0, r0Y, rXY, CID(Nx - 1, j, 0), NONE, 1);
        }
        for (long i = 0; i < Nx; ++i) {
            const double y = oy + 0.;
            vtx(r00, ox + dx * i, y, 0.);
            vtx(rX0, ox + dx * i, y, Lz);
            vtx(r0Y, ox + dx * (i + 1), y, 0.);
            vtx(rXY, ox + dx * (i + 1), y, Lz);
            set_face(s, R0 + i, r00, rX0, r0Y, rXY, CID(i, 0, 0), NONE, 3);
        }
        for (long i = 0; i < Nx; ++i) {
            const double y = oy + Ly;
            vtx(r00, ox + dx * i, y, 0.);
            vtx(rX0, ox + dx * i, y, Lz);
            vtx(r0Y, ox + dx * (i + 1), y, 0.);
            vtx(rXY, ox + dx * (i + 1), y, Lz);
            set_face(s, Le0 + i, r00, rX0, r0Y, rXY, CID(i, Ny - 1, 0), NONE, 4);
        }
        for (long j = 0; j < Ny; ++j)
            for (long i = 0; i < Nx; ++i) {
                const long id = CID(i, j, 0);
                long *f = &s->cf[6 * id];
                f[0] = (i == 0) ? T0 + j : X0 + (i - 1) + (Nx - 1) * j;
                f[1] = (i == Nx - 1) ? P0 + j : X0 + i + (Nx - 1) * j;
                f[2] = (j == Ny - 1) ? Le0 + i : Y0 + i + Nx * j;
                f[3] = (j == 0) ? R0 + i : Y0 + i + Nx * (j - 1);
                s->cnf[id] = 4;
                int n = 0;
                long *nb = &s->cn[6 * id];
                if (i > 0)
                    nb[n++] = CID(i - 1, j, 0);
                if (i < Nx - 1)
                    nb[n++] = CID(i + 1, j, 0);
                if (j < Ny - 1)
                    nb[n++] = CID(i, j + 1, 0);
                if (j > 0)
                    nb[n++] = CID(i, j - 1, 0);
                s->cnn[id] = n;
            }
    } else {
        /* threeDParallelepiped (mesh.cpp:1205-3018); note dz uses the y
         * extent (mesh.cpp:1219). */
        const double dx = Lx / Nx, dy = Ly / Ny, dz = Ly / Nz;
        for (long k = 0; k < Nz; ++k)
            for (long j = 0; j < Ny; ++j)
                for (long i = 0; i < Nx; ++i)
                    set_cell(s, CID(i, j, k), ox + dx * i, ox + dx * (i + 1),
                             oy + dy * j, oy + dy * (j + 1), oz + dz * k,
                             oz + dz * (k + 1));
        const long T0 = 0, X0 = Ny * Nz, Y0 = X0 + Nz * Ny * (Nx - 1),
                   Z0 = Y0 + Nz * (Ny - 1) * Nx, P0 = Z0 + (Nz - 1) * Ny * Nx,
                   B0 = P0 + Ny * Nz, R0 = B0 + Nx * Ny, Le0 = R0 + Nx * Nz,
                   To0 = Le0 + Nx * Nz;
        for (long k = 0; k < Nz; ++k)
            for (long j = 0; j < Ny; ++j) {
                const double x = ox + 0.;
                vtx(r00, x, oy + dy * j, oz + dz * k);
                vtx(rX0, x, oy + dy * j, oz + dz * (k + 1));
                vtx(r0Y, x, oy + dy * (j + 1), oz + dz * k);
                vtx(rXY, x, oy + dy * (j + 1), oz + dz * (k + 1));
                set_face(s, T0 + j + Ny * k, r00, rX0, r0Y, rXY, CID(0, j, k),
                         NONE, 0);
            }
        for (long k = 0; k < Nz; ++k)
            for (long j = 0; j < Ny; ++j)
                for (long i = 0; i < Nx - 1; ++i) {
                    const double x = ox + dx * (i + 1);
                    vtx(r00, x, oy + dy * j, oz + dz * k);
                    vtx(rX0, x, oy + dy * j, oz + dz * (k + 1));
                    vtx(r0Y, x, oy + dy * (j + 1), oz + dz * k);
                    vtx(rXY, x, oy + dy * (j + 1), oz + dz * (k + 1));
                    set_face(s, X0 + i + (Nx - 1) * (j + Ny * k), r00, rX0, r0Y,
                             rXY, CID(i, j, k), CID(i + 1, j, k), -1);
                }
        for (long k = 0; k < Nz; ++k)
            for (long j = 0; j < Ny - 1; ++j)
                for (long i = 0; i < Nx; ++i) {
                    const double y = oy + dy * (j + 1);
                    vtx(r00, ox + dx * i, y, oz + dz * k);
                    vtx(rX0, ox + dx * (i + 1), y, oz + dz * k);
                    vtx(r0Y, ox + dx * i, y, oz + dz * (k + 1));
                    vtx(rXY, ox + dx * (i + 1), y, oz + dz * (k + 1));
                    set_face(s, Y0 + i + Nx * (j + (Ny - 1) * k), r00, rX0, r0Y,
                             rXY, CID(i, j, k), CID(i, j + 1, k), -1);
                }
        for (long k = 0; k < Nz - 1; ++k)
            for (long j = 0; j < Ny; ++j)
                for (long i = 0; i < Nx; ++i) {
                    const double z = oz + dz * (k + 1);
                    vtx(r00, ox + dx * i, oy + dy * j, z);
                    vtx(rX0, ox + dx * (i + 1), oy + dy * j, z);
                    vtx(r0Y, ox + dx * i, oy + dy * (j + 1), z);
                    vtx(rXY, ox + dx * (i + 1), oy + dy * (j + 1), z);
                    set_face(s, Z0 + i + Nx * (j + Ny * k), r00, rX0, r0Y, rXY,
                             CID(i, j, k), CID(i, j, k + 1), -1);
                }
        for (long k = 0; k < Nz; ++k)
            for (long j = 0; j < Ny; ++j) {
                const double x = ox + Lx;
                vtx(r00, x, oy + dy * j, oz + dz * k);
                vtx(rX0, x, oy + dy * j, oz + dz * (k + 1));
                vtx(r0Y, x, oy + dy * (j + 1), oz + dz * k);
                vtx(rXY, x, oy + dy * (j + 1), oz + dz * (k + 1));
                set_face(s, P0 + j + Ny * k, r00, rX0, r0Y, rXY, CID(Nx - 1, j, k),
                         NONE, 1);
            }
        for (long j = 0; j < Ny; ++j)
            for (long i = 0; i < Nx; ++i) {
                const double z = oz + 0.;
                vtx(r00, ox + dx * i, oy + dy * j, z);
                vtx(rX0, ox + dx * (i + 1), oy + dy * j, z);
                vtx(r0Y, ox + dx * i, oy + dy * (j + 1), z);
                vtx(rXY, ox + dx * (i + 1), oy + dy * (j + 1), z);
                set_face(s, B0 + i + Nx * j, r00, rX0, r0Y, rXY, CID(i, j, 0),
                         NONE, 2);
            }
        for (long k = 0; k < Nz; ++k)
            for (long i = 0; i < Nx; ++i) {
                const double y = oy + 0.;
                vtx(r00, ox + dx * i, y, oz + dz * k);
                vtx(rX0, ox + dx * i, y, oz + dz * (k + 1));
                vtx(r0Y, ox + dx * (i + 1), y, oz + dz * k);
                vtx(rXY, ox + dx * (i + 1), y, oz + dz * (k + 1));
                set_face(s, R0 + i + Nx * k, r00, rX0, r0Y, rXY, CID(i, 0, k),
                         NONE, 3);
            }
        for (long k = 0; k < Nz; ++k)
            for (long i = 0; i < Nx; ++i) {
                const double y = oy + Ly;
                vtx(r00, ox + dx * i, y, oz + dz * k);
                vtx(rX0, ox + dx * i, y, oz + dz * (k + 1));
                vtx(r0Y, ox + dx * (i + 1), y, oz + dz * k);
                vtx(rXY, ox + dx * (i + 1), y, oz + dz * (k + 1));
                set_face(s, Le0 + i + Nx * k, r00, rX0, r0Y, rXY,
                         CID(i, Ny - 1, k), NONE, 4);
            }
        for (long j = 0; j < Ny; ++j)
            for (long i = 0; i < Nx; ++i) {
                const double z = oz + Lz;
                vtx(r00, ox + dx * i, oy + dy * j, z);
                vtx(rX0, ox + dx * (i + 1), oy + dy * j, z);
                vtx(r0Y, ox + dx * i, oy + dy * (j + 1), z);
                vtx(rXY, ox + dx * (i + 1), oy + dy * (j + 1), z);
                set_face(s, To0 + i + Nx * j, r00, rX0, r0Y, rXY,
                         CID(i, j, Nz - 1), NONE, 5);
            }
        for (long k = 0; k < Nz; ++k)
            for (long j = 0; j < Ny; ++j)
                for (long i = 0; i < Nx; ++i) {
                    const long id = CID(i, j, k);
                    long *f = &s->cf[6 * id];
                    f[0] = (i == 0) ? T0 + j + Ny * k
                                    : X0 + (i - 1) + (Nx - 1) * (j + Ny * k);
                    f[1] = (i == Nx - 1) ? P0 + j + Ny * k
                                         : X0 + i + (Nx - 1) * (j + Ny * k);
                    f[2] = (j == Ny - 1) ? Le0 + i + Nx * k
                                         : Y0 + i + Nx * (j + (Ny - 1) * k);
                    f[3] = (j == 0) ? R0 + i + Nx * k
                                    : Y0 + i + Nx * ((j - 1) + (Ny - 1) * k);
                    f[4] = (k == 0) ? B0 + i + Nx * j
                                    : Z0 + i + Nx * (j + Ny * (k - 1));
                    f[5] = (k == Nz - 1) ? To0 + i + Nx * j
                                         : Z0 + i + Nx * (j + Ny * k);
                    s->cnf[id] = 6;
                    int n = 0;
                    long *nb = &s->cn[6 * id];
                    if (i > 0)
                        nb[n++] = CID(i - 1, j, k);
                    if (i < Nx - 1)
                        nb[n++] = CID(i + 1, j, k);
                    if (j < Ny - 1)
                        nb[n++] = CID(i, j + 1, k);
                    if (j > 0)
                        nb[n++] = CID(i, j - 1, k);
                    if (k > 0)
                        nb[n++] = CID(i, j, k - 1);
                    if (k < Nz - 1)
                        nb[n++] = CID(i, j, k + 1);
                    s->cnn[id] = n;
                }
    }
#undef CID
    orient_and_weights(s);
    s->cfd = calloc(18 * nc, sizeof(double));
    s->cnd = calloc(18 * nc, sizeof(double));
    for (long i = 0; i < nc; ++i) {
        for (int j = 0; j < s->cnf[i]; ++j)
            for (int a = 0; a < 3; ++a)
                s->cfd[18 * i + 3 * j + a] = s->frc[3 * s->cf[6 * i + j] + a] - s->crc[3 * i + a];
        for (int cx = 0; cx < s->cnn[i]; ++cx)
            for (int a = 0; a < 3; ++a)
                s->cnd[18 * i + 3 * cx + a] = s->crc[3 * s->cn[6 * i + cx] + a] - s->crc[3 * i + a];
    }
    pair_periodic(s);
    return ERR_NONE;
}

/* ------------------------------------------------------------------------ */
/* EOS: ideal mixture (mixtureIdeal.cpp, idealFluid.cpp,                     */
/* fractionCalculation.cpp)                                                  */

/* CvMixture = sum_k (c_k / c_0) Cv_k (mixtureIdeal.cpp:387-392) */
static double cv_mix(const or_config *cfg, const double *c /* c0..cNs */)
{
    double cv = 0;
    for (int k = 0; k < cfg->Ns; ++k)
        cv += (c[k + 1] / c[0]) * cfg->Cv[k];
    return cv;
}
/* van der Waals mixture (mixtureVanDerWaals.cpp:19-58): per species
 * b_kk = Tc R / (8 Pc), a_kk = Pc b_kk b_kk 27; across species
 * b_kl = (0.5 (cbrt b_kk + cbrt b_ll))^3, a_kl = sqrt(a_kk a_ll). */
static void vdw_matrices(or_config *cfg)
{
    const int n = cfg->Ns;
    for (int k = 0; k < n; ++k) {
        const double b = cfg->Tc[k] * cfg->R / (8 * cfg->Pc[k]);
        cfg->b_kl[k * OR_MAXS + k] = b;
        cfg->a_kl[k * OR_MAXS + k] = cfg->Pc[k] * b * b * 27;
    }
    for (int k = 0; k < n; ++k)
        for (int l = 0; l < n; ++l)
            if (k != l) {
                cfg->b_kl[k * OR_MAXS + l] = pow(
                    0.5 * (cbrt(cfg->b_kl[k * OR_MAXS + k]) + cbrt(cfg->b_kl[l * OR_MAXS + l])),
                    3);
                cfg->a_kl[k * OR_MAXS + l] =
                    sqrt(cfg->a_kl[k * OR_MAXS + k] * cfg->a_kl[l * OR_MAXS + l]);
            }
}

/* a_mix = sum_k sum_l X_k X_l a_kl, b_mix likewise, X_k = c_k / c_0
 * (mixtureVanDerWaals.cpp:104-109, fractionCalculation.cpp:35-44) */
static void vdw_mix(const or_config *cfg, const double *c, double *am, double *bm)
{
    double X[OR_MAXS];
    for (int k = 0; k < cfg->Ns; ++k)
        X[k] = c[k + 1] / c[0];
    double a = 0, b = 0;
    for (int k = 0; k < cfg->Ns; ++k)
        for (int l = 0; l < cfg->Ns; ++l) {
            a += X[k] * X[l] * cfg->a_kl[k * OR_MAXS + l];
            b += X[k] * X[l] * cfg->b_kl[k * OR_MAXS + l];
        }
    *am = a;
    *bm = b;
}

/* Redlich-Kwong mixture (mixtureRedlichKwong.cpp:19-67): chi = cbrt(2) - 1,
 * Sigma_a = 1 / (9 chi), Sigma_b = chi / 3; per species
 * b_kk = Sigma_b R Tc / Pc, a_kk = Sigma_a R R Tc Tc sqrt(Tc) / Pc; across
 * species b_kl = (0.5 (cbrt b_kk + cbrt b_ll))^3, a_kl = sqrt(a_kk a_ll).
 * The reference's chi is a constant expression under GCC, folded with the
 * correctly rounded cube root 0x1.428a2f98d728bp+0 (glibc's cbrt(2.0) is one
 * ulp above it). */
static void rk_matrices(or_config *cfg)
{
    const double chi = 0x1.428a2f98d728bp+0 - 1.;
    const double Sa = 1. / (9. * chi), Sb = chi / 3.;
    const int n = cfg->Ns;
    for (int k = 0; k < n; ++k) {
        const double R = cfg->R, Tc = cfg->Tc[k], Pc = cfg->Pc[k];
        cfg->b_kl[k * OR_MAXS + k] = Sb * R * Tc / Pc;
        cfg->a_kl[k * OR_MAXS + k] = Sa * R * R * Tc * Tc * sqrt(Tc) / Pc;
    }
    for (int k = 0; k < n; ++k)
        for (int l = 0; l < n; ++l)
            if (k != l) {
                cfg->b_kl[k * OR_MAXS + l] = pow(
                    0.5 * (cbrt(cfg->b_kl[k * OR_MAXS + k]) + cbrt(cfg->b_kl[l * OR_MAXS + l])),
                    3);
                cfg->a_kl[k * OR_MAXS + l] =
                    sqrt(cfg->a_kl[k * OR_MAXS + k] * cfg->a_kl[l * OR_MAXS + l]);
            }
}

/* stiffened mixture (mixtureStiffened.cpp:19-43): p0_kl = sqrt(p0_k p0_l),
 * gamma_kl = sqrt(gamma_k gamma_l) off the diagonal */
static void stiff_matrices(or_config *cfg)
{
    const int n = cfg->Ns;
    for (int k = 0; k < n; ++k) {
        cfg->a_kl[k * OR_MAXS + k] = cfg->p0[k];
        cfg->b_kl[k * OR_MAXS + k] = cfg->gam[k];
    }
    for (int k = 0; k < n; ++k)
        for (int l = 0; l < n; ++l)
            if (k != l) {
                cfg->a_kl[k * OR_MAXS + l] =
                    sqrt(cfg->a_kl[k * OR_MAXS + k] * cfg->a_kl[l * OR_MAXS + l]);
                cfg->b_kl[k * OR_MAXS + l] =
                    sqrt(cfg->b_kl[k * OR_MAXS + k] * cfg->b_kl[l * OR_MAXS + l]);
            }
}

/* Kataoka-van der Waals mixture (mixtureKataokaVanDerWaals.cpp:19-77):
 * b_kk = coeff R Tc / (8 Pc) ("critical") or coeff sigma^3 N_A ("Kataoka"),
 * V0_kk = sigma^3 / sqrt(2) N_A, eps_kk = eps; across species b_kl and V0_kl
 * = (0.5 (cbrt x_kk + cbrt x_ll))^3, eps_kl = sqrt(eps_kk eps_ll). */
#define NAVOGADRO 6.02214076E23
static void kataoka_matrices(or_config *cfg)
{
    const int n = cfg->Ns;
    for (int k = 0; k < n; ++k) {
        const double s3 = cfg->sigLJ[k] * cfg->sigLJ[k] * cfg->sigLJ[k];
        cfg->b_kl[k * OR_MAXS + k] =
            cfg->kb_critical ? cfg->kb_coeff * cfg->R * cfg->Tc[k] / (8 * cfg->Pc[k])
                             : cfg->kb_coeff * s3 * NAVOGADRO;
        cfg->a_kl[k * OR_MAXS + k] = s3 / sqrt(2.0) * NAVOGADRO;
        cfg->e_kl[k * OR_MAXS + k] = cfg->epsLJ[k];
    }
    for (int k = 0; k < n; ++k)
        for (int l = 0; l < n; ++l)
            if (k != l) {
                cfg->b_kl[k * OR_MAXS + l] = pow(
                    0.5 * (cbrt(cfg->b_kl[k * OR_MAXS + k]) + cbrt(cfg->b_kl[l * OR_MAXS + l])),
                    3);
                cfg->a_kl[k * OR_MAXS + l] = pow(
                    0.5 * (cbrt(cfg->a_kl[k * OR_MAXS + k]) + cbrt(cfg->a_kl[l * OR_MAXS + l])),
                    3);
                cfg->e_kl[k * OR_MAXS + l] =
                    sqrt(cfg->e_kl[k * OR_MAXS + k] * cfg->e_kl[l * OR_MAXS + l]);
            }
}

/* V0, eps, b, Cv and M of the mixture (mixtureKataokaVanDerWaals.cpp:
 * 641-663, 810-835): X_k X_l weighted, all three in one loop */
static void kat_mix(const or_config *cfg, const double *c, double *V0, double *eps,
                    double *b, double *cv, double *mm)
{
    double X[OR_MAXS];
    double cvs = 0, ms = 0;
    for (int k = 0; k < cfg->Ns; ++k) {
        X[k] = c[k + 1] / c[0];
        cvs += X[k] * cfg->Cv[k];
        ms += X[k] * cfg->M[k];
    }
    double v = 0, e = 0, bb = 0;
    for (int k = 0; k < cfg->Ns; ++k)
        for (int l = 0; l < cfg->Ns; ++l) {
            v += X[k] * X[l] * cfg->a_kl[k * OR_MAXS + l];
            e += X[k] * X[l] * cfg->e_kl[k * OR_MAXS + l];
            bb += X[k] * X[l] * cfg->b_kl[k * OR_MAXS + l];
        }
    *V0 = v;
    *eps = e;
    *b = bb;
    *cv = cvs;
    *mm = ms;
}

/* The association series of KataokaVanDerWaalsFluid.cpp:17-62 (M = 13):
 * qa = 1 + sum_{J=2..13} (V0 c)^((J-1)/3) exp((J-1) eps / (2 R T)),
 * UvA = 0.5 c eps (-sum (J-1) (...)) / qa, pA = c R T (-sum (J-1)/3 (...)) / qa */
static double kat_qa(double c, double T, double V0, double eps, double R)
{
    const double V0c = V0 * c, epsRT = 0.5 * eps / (R * T);
    double series = 1.0;
    for (int J = 2; J <= 13; ++J)
        series += pow(V0c, (J - 1.) / 3.) * exp((J - 1.) * epsRT);
    return series;
}
static double kat_UvA(double c, double T, double V0, double eps, double R)
{
    const double V0c = V0 * c, epsRT = 0.5 * eps / (R * T);
    double series = 0;
    for (int J = 2; J <= 13; ++J)
        series -= (J - 1.) * pow(V0c, (J - 1.) / 3.) * exp((J - 1.) * epsRT);
    return 0.5 * c * eps * series / kat_qa(c, T, V0, eps, R);
}
static double kat_pA(double c, double T, double V0, double eps, double R)
{
    const double V0c = V0 * c, epsRT = 0.5 * eps / (R * T);
    double series = 0;
    for (int J = 2; J <= 13; ++J)
        series -= (J - 1.) / 3. * pow(V0c, (J - 1.) / 3.) * exp((J - 1.) * epsRT);
    return c * R * T * series / kat_qa(c, T, V0, eps, R);
}

/* The derivative series of dPdC, dPdT, dUvdC and dUvdT run
 * "for (J = 2; M < J; ++J)" (KataokaVanDerWaalsFluid.cpp:64-219) and are
 * empty, so these reduce to their leading terms plus exact zeros:
 * dPdC = R T / (1 - c b) + c R T b / (1 - c b)^2, dPdT = c R / (1 - c b),
 * dUvdC = Cv T, dUvdT = c Cv (qa, positive and finite for these states,
 * only divides the zeros). */
static double kat_dPdC(double c, double T, double R, double b)
{
    return R * T / (1 - c * b) + c * R * T * b / ((1 - c * b) * (1 - c * b));
}
static double kat_dPdT(double c, double R, double b) { return c * R / (1 - c * b); }

/* NewtonMethod (NewtonMethod.cpp:13-40): x2 = x1 - f(x1) / f'(x1) until
 * |x2 - x1| <= 1e-12, at most 100 further iterations.  kind 0: T from U
 * (f = c Cv T + UvA - U, f' = c Cv); kind 1: T from p
 * (f = c R T / (1 - c b) + pA - p, f' = c R / (1 - c b)). */
static double kat_newton(int kind, double x0, double c, double cv, double target, double V0,
                         double eps, double R, double b)
{
    const double tol = 1E-12;
    int it = 1;
    double x1 = x0;
#define KAT_F(T)                                                                           \
    ((kind == 0) ? c * cv * (T) + kat_UvA(c, (T), V0, eps, R) - target                   \
                 : c * R * (T) / (1 - c * b) + kat_pA(c, (T), V0, eps, R) - target)
#define KAT_DF(T) ((kind == 0) ? c * cv : kat_dPdT(c, R, b))
    double x2 = x1 - KAT_F(x1) / KAT_DF(x1);
    if (fabs(x2 - x1) <= tol)
        return x2;
    while (fabs(x2 - x1) > tol && it <= 100) {
        x1 = x2;
        x2 = x1 - KAT_F(x1) / KAT_DF(x1);
        it++;
    }
#undef KAT_F
#undef KAT_DF
    return x2;
}

/* T from U (KataokaVanDerWaalsFluid.cpp:277-294): Newton from U / ((c + stab) Cv) */
static double kat_T(double c, double cv, double U, double V0, double eps, double R, double b)
{
    return kat_newton(0, U / ((c + STAB) * cv), c, cv, U, V0, eps, R, b);
}

/* signum with the reference's dead band (sign.cpp:12-24) */
static int rk_sign(double x)
{
    if (x > STAB)
        return 1;
    if (x < -STAB)
        return -1;
    return 0;
}

/* cubicEquationSolver (cubicEquationSolver.cpp:17-106): roots of
 * A x^3 + B x^2 + C x + D in long double (trigonometric / hyperbolic
 * Cardano), or of the quadratic B x^2 + C x + D when |A| <= 1e-9; missing
 * roots are NaN.  Constants onethirds, twothirds, Pi are doubles; the
 * coefficient quotients B / A etc. are double divisions. */
static void rk_cubic(double A, double B, double C, double D, double x[3])
{
    const double third = 1. / 3., two3pi = (2. / 3.) * 3.141592653589793;
    x[0] = x[1] = x[2] = NAN;
    if (A > 1e-9 || A < -1e-9) {
        const long double a = B / A, b = C / A, c = D / A;
        const long double Q = (a * a - 3 * b) / 9;
        const long double R = (a * (2 * (a * a) - 9 * b) + 27 * c) / 54;
        const long double S = Q * Q * Q - R * R;
        if (S >= STAB) {
            const long double phi = third * acosl(R / sqrtl(Q * Q * Q));
            x[0] = -2. * sqrtl(Q) * cosl(phi) - a * third;
            x[1] = -2. * sqrtl(Q) * cosl(phi - two3pi) - a * third;
            x[2] = -2. * sqrtl(Q) * cosl(phi + two3pi) - a * third;
        } else if (S < STAB && S >= -STAB) {
            x[0] = -2. * rk_sign((double)R) * sqrtl(Q) - a * third;
            x[1] = rk_sign((double)R) * sqrtl(Q) - a * third;
            x[2] = x[1];
        } else if (S < -STAB) {
            if (Q >= STAB) {
                const long double phi = third * acoshl(fabsl(R) / (sqrtl(Q * Q * Q) + STAB));
                x[0] = -2. * rk_sign((double)R) * sqrtl(Q) * coshl(phi) - a * third;
            } else if (Q < STAB && Q >= -STAB) {
                x[0] = -cbrtl(c - a * a * a / 27) - a * third;
            } else if (Q < -STAB) {
                const long double phi =
                    third * asinhl(fabsl(R) / (sqrtl(fabsl(Q * Q * Q)) + STAB));
                x[0] = -2. * rk_sign((double)R) * sqrtl(fabsl(Q)) * sinhl(phi) - a * third;
            }
        }
    } else {
        const long double a = B, b = C, c = D;
        const long double Discr = b * b - 4 * a * c;
        if (Discr >= 0) {
            const long double a2 = 0.5 / (a + STAB);
            x[0] = a2 * (-b + sqrtl(Discr));
            x[1] = a2 * (-b - sqrtl(Discr));
        }
    }
}

/* returnSinglePosValue (cubicEquationSolver.cpp:108-150): the one
 * non-negative root (NaN when the reference would throw) */
static double rk_single_pos(double x[3])
{
    int np = 0;
    for (int i = 0; i < 3; ++i) {
        if (x[i] >= 0)
            np++;
        else
            x[i] = -123.;
    }
    if (np == 1) {
        const double m = (x[0] < x[1]) ? x[1] : x[0];
        return (m < x[2]) ? x[2] : m;
    }
    if (np == 2) {
        if ((x[0] - x[1]) < STAB && x[2] < 0)
            return x[0];
        if ((x[1] - x[2]) < STAB && x[0] < 0)
            return x[1];
        if ((x[0] - x[2]) < STAB && x[1] < 0)
            return x[2];
        return NAN;
    }
    if (np == 3 && (x[0] - x[1]) < STAB && (x[1] - x[2]) < STAB)
        return x[0];
    return NAN;
}

/* sqrt(T) from U: the positive root of c Cv s^3 - U s - 1.5 a c / b log1p(c b)
 * (RedlichKwongFluid.cpp:46-56, 89-97) */
static double rk_sqrtT(double Cv, double U, double c, double a, double b)
{
    double x[3];
    rk_cubic(c * Cv, 0, -U, -1.5 * a * c / b * log1p(c * b), x);
    return rk_single_pos(x);
}

/* a_mix, b_mix, Cv_mix and M_mix of the Redlich-Kwong mixture
 * (mixtureRedlichKwong.cpp:604-625, 764-788) */
static void rk_mix(const or_config *cfg, const double *c, double *am, double *bm, double *cv,
                   double *mm)
{
    double X[OR_MAXS];
    double cvs = 0, ms = 0;
    for (int k = 0; k < cfg->Ns; ++k) {
        X[k] = c[k + 1] / c[0];
        cvs += X[k] * cfg->Cv[k];
        ms += X[k] * cfg->M[k];
    }
    double a = 0, b = 0;
    for (int k = 0; k < cfg->Ns; ++k)
        for (int l = 0; l < cfg->Ns; ++l) {
            a += X[k] * X[l] * cfg->a_kl[k * OR_MAXS + l];
            b += X[k] * X[l] * cfg->b_kl[k * OR_MAXS + l];
        }
    *am = a;
    *bm = b;
    *cv = cvs;
    *mm = ms;
}

/* RedlichKwongFluid.cpp:14-44: dP/dC, dP/dT, dUv/dC, dUv/dT at (c, T) */
static double rk_dPdC(double c, double T, double a, double b, double R)
{
    return R * T / (1 - c * b) + c * R * T * b / ((1 - c * b) * (1 - c * b)) -
           2 * a * c / (sqrt(T) * (1 + b * c)) +
           a * (c * c) * b / (sqrt(T) * ((1 + b * c) * (1 + b * c)));
}
static double rk_dPdT(double c, double T, double a, double b, double R)
{
    const double sT = sqrt(T);
    return c * R / (1 - c * b) + 0.5 * a * (c * c) / (sT * sT * sT * (1 + c * b));
}
static double rk_dUvdC(double c, double Cv, double T, double a, double b)
{
    return Cv * T - 1.5 * a * log1p(c * b) / (b * sqrt(T)) - 1.5 * a * c / (sqrt(T) * (1 + c * b));
}
static double rk_dUvdT(double c, double Cv, double T, double a, double b)
{
    const double sT = sqrt(T);
    return c * Cv + 0.75 * a * c * log1p(c * b) / (b * (sT * sT * sT));
}

/* pFromUv: ideal (R / Cv) U (mixtureIdeal.cpp:383-395, idealFluid.cpp:13-17);
 * vdW g1 (U + a c^2) / (1 - c b) - a c^2 (vanDerWaalsFluid.cpp:14-22) */
static double eos_p(const or_config *cfg, const double *c, double U)
{
    if (cfg->gas_model == OR_GAS_KATAOKA) {
        /* KataokaVanDerWaalsFluid.cpp:221-240 */
        double V0, eps, b, cv, mm;
        kat_mix(cfg, c, &V0, &eps, &b, &cv, &mm);
        const double T = kat_T(c[0], cv, U, V0, eps, cfg->R, b);
        return c[0] * cfg->R * T / (1 - c[0] * b) + kat_pA(c[0], T, V0, eps, cfg->R);
    }
    if (cfg->gas_model == OR_GAS_STIFFENED) {
        /* stiffenedFluid.cpp:13-17, mixtureStiffened.cpp:485-503 */
        double p0, g;
        vdw_mix(cfg, c, &p0, &g);
        return (g - 1.0) * U - p0 * g;
    }
    if (cfg->gas_model == OR_GAS_RK) {
        /* RedlichKwongFluid.cpp:46-57 */
        double a, b, cv, mm;
        rk_mix(cfg, c, &a, &b, &cv, &mm);
        const double s = rk_sqrtT(cv, U, c[0], a, b);
        return c[0] * cfg->R * (s * s) / (1 - c[0] * b) - a * (c[0] * c[0]) / (s * (1 + c[0] * b));
    }
    if (cfg->gas_model == OR_GAS_VDW) {
        double a, b;
        vdw_mix(cfg, c, &a, &b);
        const double g1 = cfg->R / cv_mix(cfg, c);
        const double acc = a * (c[0] * c[0]);
        return g1 * (U + acc) / (1 - c[0] * b) - acc;
    }
    return (cfg->R / cv_mix(cfg, c)) * U;
}
/* UvFromp: ideal p / (R / Cv) (mixtureIdeal.cpp:397-409, idealFluid.cpp:19-23);
 * vdW (p + a c^2) (1 - c b) / g1 - a c^2 (vanDerWaalsFluid.cpp:24-31) */
static double eos_U(const or_config *cfg, const double *c, double p)
{
    if (cfg->gas_model == OR_GAS_KATAOKA) {
        /* KataokaVanDerWaalsFluid.cpp:242-261: Newton from p / ((c + stab) R) */
        double V0, eps, b, cv, mm;
        kat_mix(cfg, c, &V0, &eps, &b, &cv, &mm);
        const double T =
            kat_newton(1, p / ((c[0] + STAB) * cfg->R), c[0], cv, p, V0, eps, cfg->R, b);
        return c[0] * cv * T + kat_UvA(c[0], T, V0, eps, cfg->R);
    }
    if (cfg->gas_model == OR_GAS_STIFFENED) {
        /* stiffenedFluid.cpp:19-23, mixtureStiffened.cpp:505-523 */
        double p0, g;
        vdw_mix(cfg, c, &p0, &g);
        return (p + p0 * g) / (g - 1.0);
    }
    if (cfg->gas_model == OR_GAS_RK) {
        /* RedlichKwongFluid.cpp:59-70 */
        double a, b, cv, mm, x[3];
        rk_mix(cfg, c, &a, &b, &cv, &mm);
        const double c0 = c[0];
        rk_cubic(c0 * cfg->R, 0, -p * (1 - c0 * b), -a * (c0 * c0) * (1 - c0 * b) / (1 + c0 * b), x);
        const double s = rk_single_pos(x);
        return c0 * cv * (s * s) - 1.5 * a * c0 / (b * s) * log1p(c0 * b);
    }
    if (cfg->gas_model == OR_GAS_VDW) {
        double a, b;
        vdw_mix(cfg, c, &a, &b);
        const double g1 = cfg->R / cv_mix(cfg, c);
        const double acc = a * (c[0] * c[0]);
        return (p + acc) * (1 - c[0] * b) / g1 - acc;
    }
    return p / (cfg->R / cv_mix(cfg, c));
}
/* TFromUv: ideal U / ((c0 + stab) Cv) (mixtureIdeal.cpp:443-455,
 * idealFluid.cpp:37-41); vdW (U + a c^2) / ((c0 + stab) Cv)
 * (vanDerWaalsFluid.cpp:45-49) */
static double eos_T(const or_config *cfg, const double *c, double U)
{
    if (cfg->gas_model == OR_GAS_KATAOKA) {
        double V0, eps, b, cv, mm;
        kat_mix(cfg, c, &V0, &eps, &b, &cv, &mm);
        return kat_T(c[0], cv, U, V0, eps, cfg->R, b);
    }
    if (cfg->gas_model == OR_GAS_STIFFENED) {
        /* stiffenedFluid.cpp:37-41, mixtureStiffened.cpp:578-596: Cv_mix
         * first, then p0_mix */
        const double cv = cv_mix(cfg, c);
        double p0, g;
        vdw_mix(cfg, c, &p0, &g);
        return (U - p0) / ((c[0] + STAB) * cv);
    }
    if (cfg->gas_model == OR_GAS_RK) {
        /* RedlichKwongFluid.cpp:89-97 */
        double a, b, cv, mm;
        rk_mix(cfg, c, &a, &b, &cv, &mm);
        const double s = rk_sqrtT(cv, U, c[0], a, b);
        return s * s;
    }
    if (cfg->gas_model == OR_GAS_VDW) {
        double a, b;
        vdw_mix(cfg, c, &a, &b);
        return (U + a * (c[0] * c[0])) / ((c[0] + STAB) * cv_mix(cfg, c));
    }
    return U / ((c[0] + STAB) * cv_mix(cfg, c));
}
/* dpdUv: ideal R / Cv; vdW g1 / (1 - c b) (vanDerWaalsFluid.cpp:79-83,
 * mixtureVanDerWaals.cpp:746-762) */
static double eos_dpdUv_rk(const or_config *cfg, const double *c, double U)
{
    /* RedlichKwongFluid.cpp:117-127 */
    double a, b, cv, mm;
    rk_mix(cfg, c, &a, &b, &cv, &mm);
    const double s = rk_sqrtT(cv, U, c[0], a, b);
    const double T = s * s;
    return rk_dPdT(c[0], T, a, b, cfg->R) / rk_dUvdT(c[0], cv, T, a, b);
}
static double eos_dpdUv(const or_config *cfg, const double *c)
{
    if (cfg->gas_model == OR_GAS_STIFFENED) {
        /* stiffenedFluid.cpp:55-58, mixtureStiffened.cpp:627-640 */
        double p0, g;
        vdw_mix(cfg, c, &p0, &g);
        return g - 1.0;
    }
    if (cfg->gas_model == OR_GAS_VDW) {
        double a, b;
        vdw_mix(cfg, c, &a, &b);
        return cfg->R / cv_mix(cfg, c) / (1 - c[0] * b);
    }
    return cfg->R / cv_mix(cfg, c);
}
/* sqSonicSpeed = dpdrho + dpdUv (U + p) / rho (mixtureIdeal.cpp:495-501,
 * mixtureVanDerWaals.cpp:779-785).  Ideal: dpdrho = 0.  vdW (vanDerWaalsFluid.
 * cpp:61-77, mixtureVanDerWaals.cpp:719-744): with m = 1 - c b, ac = a c,
 * dpdrho = (g1 (2 ac / m - b (U + ac c) / m^2) - 2 ac) / M_mix,
 * M_mix = sum_k X_k M_k. */
static double eos_c2(const or_config *cfg, const double *c, double rho, double U,
                     double p)
{
    if (cfg->gas_model == OR_GAS_KATAOKA) {
        /* KataokaVanDerWaalsFluid.cpp:311-353, mixtureKataokaVanDerWaals.cpp:
         * 810-886: dpdrho = (dPdC - dPdT dUvdC / dUvdT) / M_mix,
         * dpdUv = dPdT / dUvdT at T(U) */
        double V0, eps, b, cv, mm;
        kat_mix(cfg, c, &V0, &eps, &b, &cv, &mm);
        const double c0 = c[0], R = cfg->R;
        const double T = kat_T(c0, cv, U, V0, eps, R, b);
        const double dpdrho =
            (kat_dPdC(c0, T, R, b) - kat_dPdT(c0, R, b) * (cv * T) / (c0 * cv)) / mm;
        const double dpdUv = kat_dPdT(c0, R, b) / (c0 * cv);
        return dpdrho + dpdUv * (U + p) / rho;
    }
    if (cfg->gas_model == OR_GAS_STIFFENED) {
        /* mixtureStiffened.cpp:657-663: dpdrho = 0, dpdUv = gamma - 1 */
        const double dpdrho = 0;
        const double dpdUv = eos_dpdUv(cfg, c);
        return dpdrho + dpdUv * (U + p) / rho;
    }
    if (cfg->gas_model == OR_GAS_RK) {
        /* RedlichKwongFluid.cpp:99-115, mixtureRedlichKwong.cpp:764-838 */
        double a, b, cv, mm;
        rk_mix(cfg, c, &a, &b, &cv, &mm);
        const double c0 = c[0];
        const double s = rk_sqrtT(cv, U, c0, a, b);
        const double T = s * s;
        const double dpdrho = (rk_dPdC(c0, T, a, b, cfg->R) -
                               rk_dPdT(c0, T, a, b, cfg->R) * rk_dUvdC(c0, cv, T, a, b) /
                                   rk_dUvdT(c0, cv, T, a, b)) /
                              mm;
        const double dpdUv = eos_dpdUv_rk(cfg, c, U);
        return dpdrho + dpdUv * (U + p) / rho;
    }
    if (cfg->gas_model == OR_GAS_VDW) {
        double a, b;
        vdw_mix(cfg, c, &a, &b);
        double cv = 0, mm = 0;
        for (int k = 0; k < cfg->Ns; ++k) {
            const double X = c[k + 1] / c[0];
            cv += X * cfg->Cv[k];
            mm += X * cfg->M[k];
        }
        const double g1 = cfg->R / cv;
        const double m = 1 - c[0] * b;
        const double ac = a * c[0];
        const double dpdrho = (g1 * (2 * ac / m - b * (U + ac * c[0]) / (m * m)) - 2 * ac) / mm;
        const double dpdUv = eos_dpdUv(cfg, c);
        return dpdrho + dpdUv * (U + p) / rho;
    }
    const double dpdrho = 0;
    const double dpdUv = cfg->R / cv_mix(cfg, c);
    return dpdrho + dpdUv * (U + p) / rho;
}

/* ------------------------------------------------------------------------ */
/* state                                                                     */

static void copy_state(const or_state *src, or_state *dst, long n, int Ns)
{
    for (int k = 0; k <= Ns; ++k) {
        memcpy(dst->c[k], src->c[k], n * sizeof(double));
        memcpy(dst->rho[k], src->rho[k], n * sizeof(double));
    }
    for (int a = 0; a < 3; ++a) {
        memcpy(dst->v[a], src->v[a], n * sizeof(double));
        memcpy(dst->m[a], src->m[a], n * sizeof(double));
    }
    memcpy(dst->p, src->p, n * sizeof(double));
    memcpy(dst->U, src->U, n * sizeof(double));
    memcpy(dst->T, src->T, n * sizeof(double));
    memcpy(dst->E, src->E, n * sizeof(double));
    memcpy(dst->k, src->k, n * sizeof(double));
    memcpy(dst->eps, src->eps, n * sizeof(double));
    memcpy(dst->rk, src->rk, n * sizeof(double));
    memcpy(dst->reps, src->reps, n * sizeof(double));
    for (int a = 0; a < 3; ++a) {
        memcpy(dst->a[a], src->a[a], n * sizeof(double));
        memcpy(dst->ra[a], src->ra[a], n * sizeof(double));
    }
    memcpy(dst->b, src->b, n * sizeof(double));
    memcpy(dst->rb, src->rb, n * sizeof(double));
}

static double *state_pool_new(long n, int Ns, or_state *st)
{
    const int nfield = 2 * (Ns + 1) + 3 + 3 + 3 + 1 + 4 + 8;
    double *pool = calloc((size_t)nfield * n, sizeof(double));
    double *q = pool;
    st->n = n;
    st->Ns = Ns;
    for (int k = 0; k <= Ns; ++k) {
        st->c[k] = q;
        q += n;
    }
    for (int k = 0; k <= Ns; ++k) {
        st->rho[k] = q;
        q += n;
    }
    for (int a = 0; a < 3; ++a) {
        st->v[a] = q;
        q += n;
    }
    st->p = q;
    q += n;
    st->U = q;
    q += n;
    st->T = q;
    q += n;
    for (int a = 0; a < 3; ++a) {
        st->m[a] = q;
        q += n;
    }
    st->E = q;
    q += n;
    st->k = q;
    st->eps = q + n;
    st->rk = q + 2 * n;
    st->reps = q + 3 * n;
    q += 4 * n;
    for (int a = 0; a < 3; ++a) {
        st->a[a] = q;
        st->ra[a] = q + 3 * n;
        q += n;
    }
    q += 3 * n;
    st->b = q;
    st->rb = q + n;
    return pool;
}

or_solver *or_create(const or_config *cfg)
{
    if (cfg->dim < 1 || cfg->dim > 3 || cfg->Ns < 1 || cfg->Ns > OR_MAXS)
        return NULL;
    if (cfg->gas_model < OR_GAS_IDEAL || cfg->gas_model > OR_GAS_KATAOKA)
        return NULL;
    or_solver *s = calloc(1, sizeof(or_solver));
    s->cfg = *cfg;
    s->dt_src = VERYBIG; /* mesh.cpp:121 */
    if (cfg->gas_model == OR_GAS_VDW)
        vdw_matrices(&s->cfg);
    else if (cfg->gas_model == OR_GAS_RK)
        rk_matrices(&s->cfg);
    else if (cfg->gas_model == OR_GAS_STIFFENED)
        stiff_matrices(&s->cfg);
    else if (cfg->gas_model == OR_GAS_KATAOKA)
        kataoka_matrices(&s->cfg);
    for (int a = 0; a < cfg->dim; ++a)
        if (axis_periodic(s, a) < 0) {
            free(s);
            return NULL;
        }
    if (cfg->diffusion && ((cfg->gas_model != OR_GAS_IDEAL && cfg->gas_model != OR_GAS_VDW) ||
                           (cfg->enth_flow && cfg->gas_model != OR_GAS_IDEAL) ||
                           (cfg->enth_flow && cfg->sle_solver != OR_SLE_EXPLICIT) ||
                           cfg->sle_solver < OR_SLE_EXPLICIT ||
                           cfg->sle_solver > OR_SLE_GAUSS_SEIDEL || cfg->sle_iters < 1)) {
        free(s);
        return NULL;
    }
    /* turbulence: the shear, zero and kEpsA models, ideal or van der Waals mixture,
     * the diffusion stage (if on) with the explicit solver and neither the
     * molar-mass correction nor the enthalpy flow */
    if (cfg->turb && ((cfg->turb_model != OR_TURB_SHEAR && cfg->turb_model != OR_TURB_ZERO &&
                       cfg->turb_model != OR_TURB_KEPSA && cfg->turb_model != OR_TURB_BHR) ||
                      (cfg->gas_model != OR_GAS_IDEAL && cfg->gas_model != OR_GAS_VDW) ||
                      (cfg->diffusion && (cfg->sle_solver != OR_SLE_EXPLICIT || cfg->molmass ||
                                          cfg->enth_flow)))) {
        free(s);
        return NULL;
    }
    if (build_mesh(s) != ERR_NONE) {
        or_destroy(s);
        return NULL;
    }
    s->pool = state_pool_new(s->nc, cfg->Ns, &s->st);
    s->star_stage = calloc(NSTAR * (size_t)s->nf, sizeof(double));
    s->star = calloc(NSTAR * (size_t)s->nf, sizeof(double));
    return s;
}

void or_destroy(or_solver *s)
{
    if (!s)
        return;
    free(s->crc);
    free(s->cV);
    free(s->cminlen);
    free(s->frc);
    free(s->fN);
    free(s->fS);
    free(s->own);
    free(s->nei);
    free(s->fpatch);
    free(s->wo);
    free(s->wn);
    free(s->cnf);
    free(s->cf);
    free(s->cnn);
    free(s->cn);
    free(s->cfd);
    free(s->cnd);
    free(s->fper);
    free(s->pool);
    free(s->star_stage);
    free(s->star);
    free(s);
}

long or_ncells(const or_solver *s) { return s->nc; }
long or_nfaces(const or_solver *s) { return s->nf; }
or_state *or_get_state(or_solver *s) { return &s->st; }

/* primitive -> conserved initialisation (phaseInitialization.cpp:1827-1883) */
void or_init_primitive(or_solver *s, const double *c, const double *v,
                       const double *p)
{
    or_state *st = &s->st;
    const int Ns = s->cfg.Ns;
    const long n = s->nc;
    double cc[OR_MAXS + 1];
    for (long i = 0; i < n; ++i) {
        double c0 = 0;
        for (int k = 1; k <= Ns; ++k) {
            st->c[k][i] = c[(k - 1) * n + i];
            c0 += st->c[k][i];
        }
        st->c[0][i] = c0;
        double r0 = 0;
        for (int k = 1; k <= Ns; ++k) {
            st->rho[k][i] = st->c[k][i] * s->cfg.M[k - 1];
        }
        for (int k = 1; k <= Ns; ++k)
            r0 += st->rho[k][i];
        st->rho[0][i] = r0;
        for (int a = 0; a < 3; ++a) {
            st->v[a][i] = v[a * n + i];
            st->m[a][i] = st->v[a][i] * r0;
        }
        st->p[i] = p[i];
        for (int k = 0; k <= Ns; ++k)
            cc[k] = st->c[k][i];
        st->U[i] = eos_U(&s->cfg, cc, st->p[i]);
        st->T[i] = eos_T(&s->cfg, cc, st->U[i]);
        if (s->cfg.turb) {
            st->rk[i] = st->k[i] * r0;
            st->reps[i] = r0 * st->eps[i];
        }
        if (a_field(&s->cfg))
            for (int a = 0; a < 3; ++a) /* phaseInitialization.cpp:1850 */
                st->ra[a][i] = st->a[a][i] * r0;
        if (b_field(&s->cfg)) /* phaseInitialization.cpp:1852 */
            st->rb[i] = r0 * st->b[i];
        double vel[3] = {st->v[0][i], st->v[1][i], st->v[2][i]};
        const double v2 = dot3(vel, vel);
        st->E[i] = (st->U[i] + r0 * v2 * 0.5) + st->rk[i];
    }
}

/* minimal-value clip of the initial k and epsilon
 * (phaseInitialization.cpp:1814-1825); or_init_primitive then forms rho k and
 * rho epsilon (1846-1848) and adds rho k to the total energy (1873-1877) */
void or_init_turbulence(or_solver *s, const double *k, const double *eps)
{
    or_state *st = &s->st;
    for (long i = 0; i < s->nc; ++i) {
        if (!s->cfg.turb) {
            st->k[i] = st->eps[i] = st->rk[i] = st->reps[i] = 0.0;
            continue;
        }
        st->k[i] = (k[i] < s->cfg.mink) ? s->cfg.mink : k[i];
        st->eps[i] = (eps[i] < s->cfg.mineps) ? s->cfg.mineps : eps[i];
    }
}

void or_init_turbulence_a(or_solver *s, const double *a)
{
    or_state *st = &s->st;
    const long n = s->nc;
    const int on = a_field(&s->cfg);
    for (int c = 0; c < 3; ++c)
        for (long i = 0; i < n; ++i)
            st->a[c][i] = on ? a[c * n + i] : 0.0;
}

void or_init_turbulence_b(or_solver *s, const double *b)
{
    for (long i = 0; i < s->nc; ++i)
        s->st.b[i] = b_field(&s->cfg) ? b[i] : 0.0;
}

void or_set_turbulence_b_conserved(or_solver *s, const double *rb)
{
    memcpy(s->st.rb, rb, s->nc * sizeof(double));
}

void or_set_turbulence_a_conserved(or_solver *s, const double *ra)
{
    or_state *st = &s->st;
    const long n = s->nc;
    for (int c = 0; c < 3; ++c)
        memcpy(st->ra[c], ra + c * n, n * sizeof(double));
}

void or_set_turbulence_conserved(or_solver *s, const double *rk, const double *reps)
{
    or_state *st = &s->st;
    memcpy(st->rk, rk, s->nc * sizeof(double));
    memcpy(st->reps, reps, s->nc * sizeof(double));
}

void or_set_full(or_solver *s, const double *c0, const double *c,
                 const double *rho0, const double *rho, const double *v,
                 const double *p, const double *U, const double *T,
                 const double *m, const double *E)
{
    or_state *st = &s->st;
    const int Ns = s->cfg.Ns;
    const long n = s->nc;
    memcpy(st->c[0], c0, n * sizeof(double));
    memcpy(st->rho[0], rho0, n * sizeof(double));
    for (int k = 1; k <= Ns; ++k) {
        memcpy(st->c[k], c + (k - 1) * n, n * sizeof(double));
        memcpy(st->rho[k], rho + (k - 1) * n, n * sizeof(double));
    }
    for (int a = 0; a < 3; ++a) {
        memcpy(st->v[a], v + a * n, n * sizeof(double));
        memcpy(st->m[a], m + a * n, n * sizeof(double));
    }
    memcpy(st->p, p, n * sizeof(double));
    memcpy(st->U, U, n * sizeof(double));
    memcpy(st->T, T, n * sizeof(double));
    memcpy(st->E, E, n * sizeof(double));
}

/* Primitive recovery after the stage update (Advection2dOrder.cpp:578-602). */
static void recover_primitives(or_solver *s, or_state *st)
{
    const int Ns = s->cfg.Ns;
    double cc[OR_MAXS + 1];
    for (long i = 0; i < s->nc; ++i) {
        double c0 = 0;
        for (int k = 1; k <= Ns; ++k) {
            st->c[k][i] = st->rho[k][i] / s->cfg.M[k - 1];
            c0 += st->c[k][i];
        }
        st->c[0][i] = c0;
        for (int a = 0; a < 3; ++a)
            st->v[a][i] = st->m[a][i] / st->rho[0][i];
        double vel[3] = {st->v[0][i], st->v[1][i], st->v[2][i]};
        const double v2 = dot3(vel, vel);
        st->U[i] = (st->E[i] - st->rho[0][i] * v2 * 0.5) - st->rk[i];
        for (int k = 0; k <= Ns; ++k)
            cc[k] = st->c[k][i];
        st->p[i] = eos_p(&s->cfg, cc, st->U[i]);
        st->T[i] = eos_T(&s->cfg, cc, st->U[i]);
        if (s->cfg.turb) { /* Advection2dOrder.cpp:610-627 */
            st->k[i] = st->rk[i] / st->rho[0][i];
            st->eps[i] = st->reps[i] / st->rho[0][i];
        }
        if (a_field(&s->cfg))
            for (int a = 0; a < 3; ++a)
                st->a[a][i] = st->ra[a][i] / st->rho[0][i];
        if (b_field(&s->cfg) && !s->keep_b) /* Advection2dOrder.cpp:623-625 */
            st->b[i] = st->rb[i] / st->rho[0][i];
    }
}

void or_set_conserved(or_solver *s, const double *rho_k, const double *rho,
                      const double *m, const double *E)
{
    or_state *st = &s->st;
    const long n = s->nc;
    for (int k = 1; k <= s->cfg.Ns; ++k)
        memcpy(st->rho[k], rho_k + (k - 1) * n, n * sizeof(double));
    memcpy(st->rho[0], rho, n * sizeof(double));
    for (int a = 0; a < 3; ++a)
        memcpy(st->m[a], m + a * n, n * sizeof(double));
    memcpy(st->E, E, n * sizeof(double));
    recover_primitives(s, st);
}

/* ------------------------------------------------------------------------ */
/* boundary conditions (boundaryConditionValue.hpp:33-276, 278-511;         */
/* slipFunction.cpp:15-35)                                                   */

static int bc_surf_scalar(const or_bc *bc, double q, double *out)
{
    switch (bc->type) {
    case OR_BC_FREE:
        *out = q;
        return 0;
    case OR_BC_SLIP:
        *out = q; /* slipFunction(scalar) is the identity */
        return 0;
    case OR_BC_FIXED_CELL:
        *out = (q + bc->val[0]) * 0.5;
        return 0;
    case OR_BC_FIXED_SURF:
        *out = bc->val[0];
        return 0;
    default:
        return ERR_BC;
    }
}

static int bc_cell_scalar(const or_bc *bc, double q, double *out)
{
    switch (bc->type) {
    case OR_BC_FREE:
        *out = q;
        return 0;
    case OR_BC_SLIP:
        *out = q * 2 - q;
        return 0;
    case OR_BC_FIXED_CELL:
        *out = bc->val[0];
        return 0;
    default:
        return ERR_BC; /* fixedValueSurface throws for cell values */
    }
}

static void slip_vec(const double *v, const double *N, double *out)
{
    const double s = dot3(v, N);
    for (int a = 0; a < 3; ++a)
        out[a] = v[a] - N[a] * s;
}

static int bc_surf_vec(const or_bc *bc, const double *v, const double *N,
                       double *out)
{
    switch (bc->type) {
    case OR_BC_FREE:
        for (int a = 0; a < 3; ++a)
            out[a] = v[a];
        return 0;
    case OR_BC_SLIP:
        slip_vec(v, N, out);
        return 0;
    case OR_BC_FIXED_CELL:
        for (int a = 0; a < 3; ++a)
            out[a] = (v[a] + bc->val[a]) * 0.5;
        return 0;
    case OR_BC_FIXED_SURF:
        for (int a = 0; a < 3; ++a)
            out[a] = bc->val[a];
        return 0;
    default:
        return ERR_BC;
    }
}

static int bc_cell_vec(const or_bc *bc, const double *v, const double *N,
                       double *out)
{
    switch (bc->type) {
    case OR_BC_FREE:
        for (int a = 0; a < 3; ++a)
            out[a] = v[a];
        return 0;
    case OR_BC_SLIP: {
        double sl[3];
        slip_vec(v, N, sl);
        for (int a = 0; a < 3; ++a)
            out[a] = sl[a] * 2 - v[a];
        return 0;
    }
    case OR_BC_FIXED_CELL:
        for (int a = 0; a < 3; ++a)
            out[a] = bc->val[a];
        return 0;
    default:
        return ERR_BC;
    }
}

/* ------------------------------------------------------------------------ */
/* gradient (gradient.hpp:19-58) with linearInterpolate                      */
/* (linearInterpolate.hpp:214-251)                                           */

/* gradient of a scalar from its face values (gradient.hpp:19-58) */
static void grad_faces(const or_solver *s, const double *qf, int st, double *G /* 3 per cell */)
{
    for (long i = 0; i < s->nc; ++i) {
        double acc[3] = {0, 0, 0};
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            double n[3];
            if (s->own[f] == i)
                memcpy(n, &s->fN[3 * f], sizeof n);
            else
                for (int a = 0; a < 3; ++a)
                    n[a] = s->fN[3 * f + a] * -1;
            for (int a = 0; a < 3; ++a)
                acc[a] += (n[a] * qf[st * f]) * s->fS[f];
        }
        for (int a = 0; a < 3; ++a)
            G[3 * i + a] = acc[a] / s->cV[i];
    }
}

static int grad_scalar(or_solver *s, const double *q, const or_bc *bcs,
                       double *G /* 3 per cell */)
{
    double *qf = malloc(s->nf * sizeof(double));
    for (long f = 0; f < s->nf; ++f) {
        if (s->fpatch[f] == -3)
            continue;
        if (s->fpatch[f] < 0)
            qf[f] = q[s->own[f]] * s->wo[f] + q[s->nei[f]] * s->wn[f];
        else if (bc_surf_scalar(&bcs[s->fpatch[f]], q[s->own[f]], &qf[f])) {
            free(qf);
            return ERR_BC;
        }
    }
    grad_faces(s, qf, 1, G);
    free(qf);
    return 0;
}

/* gradient of a vector from its face values vf[st * f + a] (gradient.hpp:60-93) */
static void grad_vector_faces(const or_solver *s, const double *vf, int st, double *G)
{
    for (long i = 0; i < s->nc; ++i) {
        double acc[9] = {0};
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            double n[3];
            if (s->own[f] == i)
                memcpy(n, &s->fN[3 * f], sizeof n);
            else
                for (int a = 0; a < 3; ++a)
                    n[a] = s->fN[3 * f + a] * -1;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    acc[3 * a + b] += (vf[st * f + a] * n[b]) * s->fS[f];
        }
        for (int q = 0; q < 9; ++q)
            G[9 * i + q] = acc[q] / s->cV[i];
    }
}

static int grad_vector(or_solver *s, double *const *v, const or_bc *bcs,
                       double *G /* 9 per cell */)
{
    double *vf = malloc(3 * s->nf * sizeof(double));
    for (long f = 0; f < s->nf; ++f) {
        const long o = s->own[f];
        if (s->fpatch[f] == -3)
            continue;
        if (s->fpatch[f] < 0) {
            const long n = s->nei[f];
            for (int a = 0; a < 3; ++a)
                vf[3 * f + a] = v[a][o] * s->wo[f] + v[a][n] * s->wn[f];
        } else {
            double vo[3] = {v[0][o], v[1][o], v[2][o]};
            if (bc_surf_vec(&bcs[s->fpatch[f]], vo, &s->fN[3 * f], &vf[3 * f])) {
                free(vf);
                return ERR_BC;
            }
        }
    }
    grad_vector_faces(s, vf, 3, G);
    free(vf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* limiters (minmodLimiter.cpp:14-57, vanLeerLimiter.cpp:14-53,              */
/* vanAlbadaLimiter.cpp:15-55, linearLimiter.cpp, zeroLimiter.cpp)           */

static double phi_of(int lim, double r, double xi)
{
    switch (lim) {
    case OR_LIM_MINMOD:
        if (r <= 0.)
            return 0;
        else if (r <= 1.)
            return r;
        return smin(1.0, xi);
    case OR_LIM_VANLEER:
        if (r < 0.)
            return 0;
        return smin(2 * r / (1 + r), xi);
    case OR_LIM_VANALBADA:
        if (r <= 0.)
            return 0;
        return smin((r * r + r) / (1 + r * r), xi);
    default:
        return 0;
    }
}

double or_limiter_phi(int limiter, double r)
{
    return phi_of(limiter, r, 2 / (1 + r));
}

/* limited value = limiter.calculate(r, G) for n components */
static void limit_apply(int lim, const double *r, const double *G, double *L,
                        int n)
{
    if (lim == OR_LIM_LINEAR) {
        for (int q = 0; q < n; ++q)
            L[q] = G[q];
        return;
    }
    if (lim == OR_LIM_ZERO) {
        for (int q = 0; q < n; ++q)
            L[q] = 0;
        return;
    }
    double xi[9];
    for (int q = 0; q < n; ++q)
        xi[q] = 2 / (1 + r[q]);
    for (int q = 0; q < n; ++q)
        L[q] = phi_of(lim, r[q], xi[q]) * G[q];
}

/* TVDLimiter scalar overload (TVDLimiter.cpp:12-125) */
static int tvd_scalar(or_solver *s, const double *G, const double *q,
                      const or_bc *bcs, double *L)
{
    double ent[12][3];
    for (long i = 0; i < s->nc; ++i) {
        const double *ci = &s->crc[3 * i];
        int ne = 0;
        for (int cx = 0; cx < s->cnn[i]; ++cx) {
            const long n = s->cn[6 * i + cx];
            const double *d = &s->cnd[18 * i + 3 * cx];
            const double dm = or_hypot3(d[0], d[1], d[2]);
            double dn[3];
            for (int a = 0; a < 3; ++a)
                dn[a] = d[a] / dm;
            for (int a = 0; a < 3; ++a)
                ent[ne][a] = (q[n] - q[i]) / dm * dn[a];
            ++ne;
        }
        if (s->cnn[i] < s->cnf[i]) {
            for (int j = 0; j < s->cnf[i]; ++j) {
                const long f = s->cf[6 * i + j];
                if (s->fpatch[f] < 0)
                    continue;
                double d[3];
                for (int a = 0; a < 3; ++a)
                    d[a] = (s->frc[3 * f + a] - ci[a]) * 2;
                const double dm = or_hypot3(d[0], d[1], d[2]);
                double dn[3];
                for (int a = 0; a < 3; ++a)
                    dn[a] = d[a] / dm;
                double qb;
                if (bc_cell_scalar(&bcs[s->fpatch[f]], q[i], &qb))
                    return ERR_BC;
                for (int a = 0; a < 3; ++a)
                    ent[ne][a] = (qb - q[i]) / dm * dn[a];
                ++ne;
            }
        }
        double mn[3], mx[3], r[3];
        for (int a = 0; a < 3; ++a)
            mn[a] = mx[a] = ent[0][a];
        for (int e = 1; e < ne; ++e) {
            for (int a = 0; a < 3; ++a)
                mn[a] = smin(mn[a], ent[e][a]);
            for (int a = 0; a < 3; ++a)
                mx[a] = smax(mx[a], ent[e][a]);
        }
        for (int a = 0; a < 3; ++a)
            r[a] = mn[a] / (mx[a] + STAB);
        limit_apply(s->cfg.limiter, r, &G[3 * i], &L[3 * i], 3);
    }
    return 0;
}

/* TVDLimiter vector->tensor overload (TVDLimiter.cpp:127-273); entries are
 * stored at [3 j + c] (direction j, component c) while the gradient is
 * [3 a + b] (component a, direction b). */
static int tvd_vector(or_solver *s, const double *G, double *const *v,
                      const or_bc *bcs, double *L)
{
    double ent[12][9];
    for (long i = 0; i < s->nc; ++i) {
        const double *ci = &s->crc[3 * i];
        const double vi[3] = {v[0][i], v[1][i], v[2][i]};
        int ne = 0;
        for (int cx = 0; cx < s->cnn[i]; ++cx) {
            const long n = s->cn[6 * i + cx];
            const double *d = &s->cnd[18 * i + 3 * cx];
            const double dm = or_hypot3(d[0], d[1], d[2]);
            double dn[3];
            for (int a = 0; a < 3; ++a)
                dn[a] = d[a] / dm;
            double dv[3];
            for (int c = 0; c < 3; ++c)
                dv[c] = v[c][n] - vi[c];
            for (int j = 0; j < 3; ++j)
                for (int c = 0; c < 3; ++c)
                    ent[ne][3 * j + c] = dv[c] / dm * dn[j];
            ++ne;
        }
        if (s->cnn[i] < s->cnf[i]) {
            for (int jj = 0; jj < s->cnf[i]; ++jj) {
                const long f = s->cf[6 * i + jj];
                if (s->fpatch[f] < 0)
                    continue;
                double d[3];
                for (int a = 0; a < 3; ++a)
                    d[a] = (s->frc[3 * f + a] - ci[a]) * 2;
                const double dm = or_hypot3(d[0], d[1], d[2]);
                double dn[3];
                for (int a = 0; a < 3; ++a)
                    dn[a] = d[a] / dm;
                double vb[3];
                if (bc_cell_vec(&bcs[s->fpatch[f]], vi, &s->fN[3 * f], vb))
                    return ERR_BC;
                double dv[3];
                for (int c = 0; c < 3; ++c)
                    dv[c] = vb[c] - vi[c];
                for (int j = 0; j < 3; ++j)
                    for (int c = 0; c < 3; ++c)
                        ent[ne][3 * j + c] = dv[c] / dm * dn[j];
                ++ne;
            }
        }
        double mn[9], mx[9], r[9];
        for (int q = 0; q < 9; ++q)
            mn[q] = mx[q] = ent[0][q];
        for (int e = 1; e < ne; ++e) {
            for (int q = 0; q < 9; ++q)
                mn[q] = smin(mn[q], ent[e][q]);
            for (int q = 0; q < 9; ++q)
                mx[q] = smax(mx[q], ent[e][q]);
        }
        for (int q = 0; q < 9; ++q)
            r[q] = mn[q] / (mx[q] + STAB);
        limit_apply(s->cfg.limiter, r, &G[9 * i], &L[9 * i], 9);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* face states and numerical flux                                            */

typedef struct {
    double c[OR_MAXS + 1];
    double rho[OR_MAXS + 1];
    double v[3];
    double m[3];
    double p, U, T, E;
    double k, eps, rk, reps; /* turbulence (zero without it) */
    double a[3], ra[3];      /* kEpsA turbulent mass flux (zero otherwise) */
    double b, rb;            /* BHR b (zero otherwise) */
} face_state;

/* Face-side closure: Advection2dOrder.cpp:233-289 (owner side) and 301-372
 * (neighbour side); both give the same per-face arithmetic. */
static void close_face(const or_config *cfg, face_state *fs)
{
    double c0 = 0, r0 = 0;
    for (int k = 1; k <= cfg->Ns; ++k) {
        c0 += fs->c[k];
        fs->rho[k] = fs->c[k] * cfg->M[k - 1];
        r0 += fs->rho[k];
    }
    fs->c[0] = c0;
    fs->rho[0] = r0;
    if (cfg->turb) { /* Advection2dOrder.cpp:249-266, 326-347, 440-458 */
        fs->rk = fs->k * r0;
        fs->reps = fs->eps * r0;
        for (int a = 0; a < 3; ++a) /* Advection2dOrder.cpp:257-260, 336-340 */
            fs->ra[a] = fs->a[a] * r0;
        fs->rb = fs->b * r0; /* Advection2dOrder.cpp:262-264, 342-345, 480-482 */
    }
    fs->U = eos_U(cfg, fs->c, fs->p);
    fs->T = eos_T(cfg, fs->c, fs->U);
    for (int a = 0; a < 3; ++a)
        fs->m[a] = fs->v[a] * r0;
    const double v2 = dot3(fs->v, fs->v);
    fs->E = fs->U + r0 * v2 * 0.5 + fs->rk;
}

/* Flux vectors stored like conservativeFlows (conservativeFlows.hpp:22-34):
 * scalars as f N (3 comps), momentum as tensor (9 comps). */
typedef struct {
    double rho[OR_MAXS + 1][3];
    double m[9];
    double E[3];
    double rk[3], reps[3]; /* turbulence */
    double ra[9];          /* kEpsA: rho a as a tensor like the momentum */
    /* the face's star state (starFields): rho, v (3), p, a (3)
     * (HLLCSolver.cpp:390-402, HLLSolver.cpp:302-314) */
    double st[NSTAR];
    double rb[3]; /* BHR: rho b */
} face_flux;

/* pressureStarClass::pressureStar (pressureStarClass.cpp:12-37): pFromUv of
 * the star concentrations rho_k / M_k and the star internal energy
 * (E - 0.5 (m & m) / rho) - rho k.  ds: Ns + 1 densities, ds[0] the total. */
static double pressure_star(const or_config *cfg, const double *ds, const double *ms,
                            double Es, double rks)
{
    double mm = 0;
    for (int a = 0; a < 3; ++a)
        mm += ms[a] * ms[a];
    const double kin = 0.5 * mm / ds[0];
    const double U = Es - kin - rks;
    double cc[OR_MAXS + 1];
    cc[0] = 0;
    for (int k = 1; k <= cfg->Ns; ++k) {
        cc[k] = ds[k] / cfg->M[k - 1];
        cc[0] += cc[k];
    }
    return eos_p(cfg, cc, U);
}

/* star state of a face side taken as it is (supersonic branches) */
static void star_side(const face_state *X, double *st)
{
    st[0] = X->rho[0];
    for (int a = 0; a < 3; ++a)
        st[1 + a] = X->v[a];
    st[4] = X->p;
    for (int c = 0; c < 3; ++c)
        st[5 + c] = X->ra[c] / X->rho[0];
    st[8] = X->rb / X->rho[0];
}

/* HLLCSolver::calculateFlows per-face body (HLLCSolver.cpp:52-388) */
static int hllc_face(const or_config *cfg, const double *N, const face_state *O,
                     const face_state *Nb, face_flux *F)
{
    const int nd = cfg->Ns + 1;
    /* sqSonicSpeed + (2/3) k (1 + dpdUv) (HLLCSolver.cpp:26-50,
     * HLLSolver.cpp:26-50); turbulent pressure (2/3) rho k (:61-64) */
    const double cO = sqrt(eos_c2(cfg, O->c, O->rho[0], O->U, O->p) +
                           (2. / 3.) * O->k * (1. + eos_dpdUv(cfg, O->c)));
    const double cN = sqrt(eos_c2(cfg, Nb->c, Nb->rho[0], Nb->U, Nb->p) +
                           (2. / 3.) * Nb->k * (1. + eos_dpdUv(cfg, Nb->c)));
    const double uO = dot3(O->v, N);
    const double uN = dot3(Nb->v, N);
    const double ptO = (2. / 3.) * O->rk, ptN = (2. / 3.) * Nb->rk;
    const double SO = smin(uO - cO, uN - cN);
    const double SN = smax(uO + cO, uN + cN);
    const double rO = O->rho[0], rN = Nb->rho[0];
    const double Sc = ((Nb->p + ptN) - (O->p + ptO) + rO * uO * (SO - uO) -
                       rN * uN * (SN - uN)) /
                      (rO * (SO - uO) - rN * (SN - uN));
    const double PLR = 0.5 * ((O->p + ptO) + (Nb->p + ptN) +
                              rO * (SO - uO) * (Sc - uO) + rN * (SN - uN) * (Sc - uN));
    double IN[3]; /* tensor(1,0,0,0,1,0,0,0,1) & N (tensor.cpp:163-176) */
    IN[0] = 1 * N[0] + 0 * N[1] + 0 * N[2];
    IN[1] = 0 * N[0] + 1 * N[1] + 0 * N[2];
    IN[2] = 0 * N[0] + 0 * N[1] + 1 * N[2];

    if (SO >= 0) {
        for (int k = 0; k < nd; ++k)
            for (int a = 0; a < 3; ++a)
                F->rho[k][a] = (N[a] * uO) * O->rho[k];
        double w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = O->m[a] * uO + IN[a] * (O->p + ptO);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                F->m[3 * a + b] = w[a] * N[b];
        const double e = uO * (O->E + O->p + ptO);
        for (int a = 0; a < 3; ++a)
            F->E[a] = N[a] * e;
        for (int a = 0; a < 3; ++a) { /* HLLCSolver.cpp:126-133 */
            F->rk[a] = (uO * O->rk) * N[a];
            F->reps[a] = (uO * N[a]) * O->reps;
            F->rb[a] = (uO * N[a]) * O->rb; /* HLLCSolver.cpp:147-148 */
        }
        for (int c = 0; c < 3; ++c) /* HLLCSolver.cpp:145-146 */
            for (int b = 0; b < 3; ++b)
                F->ra[3 * c + b] = (O->ra[c] * uO) * N[b];
        star_side(O, F->st); /* HLLCSolver.cpp:107-121 */
    } else if (Sc >= 0) {
        const double factor = SO - uO;
        const double den = 1 / (SO - Sc);
        double rs[OR_MAXS + 1];
        for (int k = 0; k < nd; ++k)
            rs[k] = O->rho[k] * factor * den;
        double ms[3];
        for (int a = 0; a < 3; ++a)
            ms[a] = (O->m[a] * factor - N[a] * (O->p + ptO - PLR)) * den;
        const double Es = (O->E * factor - (O->p + ptO) * uO + PLR * Sc) * den;
        for (int k = 0; k < nd; ++k) {
            const double f = O->rho[k] * uO + SO * (rs[k] - O->rho[k]);
            for (int a = 0; a < 3; ++a)
                F->rho[k][a] = N[a] * f;
        }
        double w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = (O->m[a] * uO + IN[a] * (O->p + ptO)) + (ms[a] - O->m[a]) * SO;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                F->m[3 * a + b] = w[a] * N[b];
        const double e = (O->E + O->p + ptO) * uO + SO * (Es - O->E);
        for (int a = 0; a < 3; ++a)
            F->E[a] = N[a] * e;
        /* HLLCSolver.cpp:165-168, 215-228 */
        const double rks = O->rk * factor * den, reps_s = O->reps * factor * den;
        const double fk = O->rk * uO + SO * (rks - O->rk);
        const double fe = O->reps * uO + SO * (reps_s - O->reps);
        for (int a = 0; a < 3; ++a) {
            F->rk[a] = fk * N[a];
            F->reps[a] = fe * N[a];
        }
        for (int c = 0; c < 3; ++c) { /* HLLCSolver.cpp:175-176, 230-236 */
            const double ras = O->ra[c] * factor * den;
            const double fa = O->ra[c] * uO + SO * (ras - O->ra[c]);
            for (int b = 0; b < 3; ++b)
                F->ra[3 * c + b] = fa * N[b];
            F->st[5 + c] = ras / rs[0];
        }
        /* HLLCSolver.cpp:180-185 */
        F->st[0] = rs[0];
        for (int a = 0; a < 3; ++a)
            F->st[1 + a] = ms[a] / rs[0];
        F->st[4] = pressure_star(cfg, rs, ms, Es, rks);
        {
            /* HLLCSolver.cpp:177-178, 237-243 */
            const double rbs = O->rb * factor * den;
            const double fb = O->rb * uO + SO * (rbs - O->rb);
            for (int a = 0; a < 3; ++a)
                F->rb[a] = fb * N[a];
            F->st[8] = rbs / rs[0];
        }
    } else if (SN > 0) {
        const double factor = SN - uN;
        const double den = 1 / (SN - Sc);
        double rs[OR_MAXS + 1];
        for (int k = 0; k < nd; ++k)
            rs[k] = Nb->rho[k] * factor * den;
        double ms[3];
        for (int a = 0; a < 3; ++a)
            ms[a] = (Nb->m[a] * factor - N[a] * (Nb->p + ptN - PLR)) * den;
        const double Es = (Nb->E * factor - (Nb->p + ptN) * uN + PLR * Sc) * den;
        for (int k = 0; k < nd; ++k) {
            const double f = Nb->rho[k] * uN + SN * (rs[k] - Nb->rho[k]);
            for (int a = 0; a < 3; ++a)
                F->rho[k][a] = N[a] * f;
        }
        double w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = (Nb->m[a] * uN + IN[a] * (Nb->p + ptN)) + (ms[a] - Nb->m[a]) * SN;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                F->m[3 * a + b] = w[a] * N[b];
        const double e = (Nb->E + Nb->p + ptN) * uN + SN * (Es - Nb->E);
        for (int a = 0; a < 3; ++a)
            F->E[a] = N[a] * e;
        /* HLLCSolver.cpp:262-265, 312-325 */
        const double rks = Nb->rk * factor * den, reps_s = Nb->reps * factor * den;
        const double fk = Nb->rk * uN + SN * (rks - Nb->rk);
        const double fe = Nb->reps * uN + SN * (reps_s - Nb->reps);
        for (int a = 0; a < 3; ++a) {
            F->rk[a] = fk * N[a];
            F->reps[a] = fe * N[a];
        }
        for (int c = 0; c < 3; ++c) { /* HLLCSolver.cpp:270-271, 327-333 */
            const double ras = Nb->ra[c] * factor * den;
            const double fa = Nb->ra[c] * uN + SN * (ras - Nb->ra[c]);
            for (int b = 0; b < 3; ++b)
                F->ra[3 * c + b] = fa * N[b];
            F->st[5 + c] = ras / rs[0];
        }
        /* HLLCSolver.cpp:275-280 */
        F->st[0] = rs[0];
        for (int a = 0; a < 3; ++a)
            F->st[1 + a] = ms[a] / rs[0];
        F->st[4] = pressure_star(cfg, rs, ms, Es, rks);
        {
            /* HLLCSolver.cpp:272-273, 334-340 */
            const double rbs = Nb->rb * factor * den;
            const double fb = Nb->rb * uN + SN * (rbs - Nb->rb);
            for (int a = 0; a < 3; ++a)
                F->rb[a] = fb * N[a];
            F->st[8] = rbs / rs[0];
        }
    } else if (SN <= 0) {
        for (int k = 0; k < nd; ++k)
            for (int a = 0; a < 3; ++a)
                F->rho[k][a] = (N[a] * uN) * Nb->rho[k];
        double w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = Nb->m[a] * uN + IN[a] * (Nb->p + ptN);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                F->m[3 * a + b] = w[a] * N[b];
        const double e = uN * (Nb->E + Nb->p + ptN);
        for (int a = 0; a < 3; ++a)
            F->E[a] = N[a] * e;
        for (int a = 0; a < 3; ++a) { /* HLLCSolver.cpp:368-375 */
            F->rk[a] = (uN * Nb->rk) * N[a];
            F->reps[a] = (uN * N[a]) * Nb->reps;
            F->rb[a] = (uN * N[a]) * Nb->rb; /* HLLCSolver.cpp:382-383 */
        }
        for (int c = 0; c < 3; ++c) /* HLLCSolver.cpp:380-381 */
            for (int b = 0; b < 3; ++b)
                F->ra[3 * c + b] = (Nb->ra[c] * uN) * N[b];
        star_side(Nb, F->st); /* HLLCSolver.cpp:344-358 */
    } else {
        return ERR_RIEMANN;
    }
    return 0;
}

/* HLLSolver::calculateFlows per-face body (HLLSolver.cpp:52-300) */
static int hll_face(const or_config *cfg, const double *N, const face_state *O,
                    const face_state *Nb, face_flux *F)
{
    const int nd = cfg->Ns + 1;
    /* sqSonicSpeed + (2/3) k (1 + dpdUv) (HLLCSolver.cpp:26-50,
     * HLLSolver.cpp:26-50); turbulent pressure (2/3) rho k (:61-64) */
    const double cO = sqrt(eos_c2(cfg, O->c, O->rho[0], O->U, O->p) +
                           (2. / 3.) * O->k * (1. + eos_dpdUv(cfg, O->c)));
    const double cN = sqrt(eos_c2(cfg, Nb->c, Nb->rho[0], Nb->U, Nb->p) +
                           (2. / 3.) * Nb->k * (1. + eos_dpdUv(cfg, Nb->c)));
    const double uO = dot3(O->v, N);
    const double uN = dot3(Nb->v, N);
    const double ptO = (2. / 3.) * O->rk, ptN = (2. / 3.) * Nb->rk;
    const double SO = smin(uO - cO, uN - cN);
    const double SN = smax(uO + cO, uN + cN);
    double IN[3];
    IN[0] = 1 * N[0] + 0 * N[1] + 0 * N[2];
    IN[1] = 0 * N[0] + 1 * N[1] + 0 * N[2];
    IN[2] = 0 * N[0] + 0 * N[1] + 1 * N[2];

    if (SO >= 0) {
        for (int k = 0; k < nd; ++k)
            for (int a = 0; a < 3; ++a)
                F->rho[k][a] = N[a] * (uO * O->rho[k]);
        double w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = O->m[a] * uO + IN[a] * (O->p + ptO);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                F->m[3 * a + b] = w[a] * N[b];
        const double e = uO * (O->E + O->p + ptO);
        for (int a = 0; a < 3; ++a)
            F->E[a] = N[a] * e;
        for (int a = 0; a < 3; ++a) { /* HLLSolver.cpp:107-114 */
            F->rk[a] = (uO * O->rk) * N[a];
            F->reps[a] = (uO * O->reps) * N[a];
            F->rb[a] = (uO * O->rb) * N[a]; /* HLLSolver.cpp:119-120 */
        }
        for (int c = 0; c < 3; ++c) /* HLLSolver.cpp:117-118 */
            for (int b = 0; b < 3; ++b)
                F->ra[3 * c + b] = (O->ra[c] * uO) * N[b];
        star_side(O, F->st); /* HLLSolver.cpp:78-94 */
    } else if (SN > 0) {
        const double den = 1 / (SN - SO);
        for (int k = 0; k < nd; ++k) {
            const double f = (SN * uO * O->rho[k] - SO * uN * Nb->rho[k] +
                              SN * SO * (Nb->rho[k] - O->rho[k])) *
                             den;
            for (int a = 0; a < 3; ++a)
                F->rho[k][a] = N[a] * f;
        }
        double w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = (((O->m[a] * uO + N[a] * (O->p + ptO)) * SN -
                     (Nb->m[a] * uN + N[a] * (Nb->p + ptN)) * SO) +
                    (Nb->m[a] - O->m[a]) * (SN * SO)) *
                   den;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                F->m[3 * a + b] = w[a] * N[b];
        const double e = (SN * (uO * (O->E + O->p + ptO)) -
                          SO * (uN * (Nb->E + Nb->p + ptN)) +
                          SN * SO * (Nb->E - O->E)) *
                         den;
        for (int a = 0; a < 3; ++a)
            F->E[a] = N[a] * e;
        /* HLLSolver.cpp:209-224 */
        const double fk = (SN * uO * O->rk - SO * uN * Nb->rk + SN * SO * (Nb->rk - O->rk)) * den;
        const double fe =
            (SN * uO * O->reps - SO * uN * Nb->reps + SN * SO * (Nb->reps - O->reps)) * den;
        for (int a = 0; a < 3; ++a) {
            F->rk[a] = fk * N[a];
            F->reps[a] = fe * N[a];
        }
        for (int c = 0; c < 3; ++c) { /* HLLSolver.cpp:159-162, 237-243 */
            const double fa = (SN * uO * O->ra[c] - SO * uN * Nb->ra[c] +
                               SN * SO * (Nb->ra[c] - O->ra[c])) *
                              den;
            for (int b = 0; b < 3; ++b)
                F->ra[3 * c + b] = fa * N[b];
        }
        {
            /* the HLL average state (HLLSolver.cpp:125-173) */
            const double fN = SN - uN, fO = SO - uO;
            double ds[OR_MAXS + 1], ms[3];
            for (int k = 0; k < nd; ++k)
                ds[k] = (Nb->rho[k] * fN - O->rho[k] * fO) * den;
            for (int a = 0; a < 3; ++a)
                ms[a] = (Nb->m[a] * fN - O->m[a] * fO - N[a] * ((Nb->p + ptN) - (O->p + ptO))) * den;
            const double Es =
                (Nb->E * fN - O->E * fO - ((Nb->p + ptN) * uN - (O->p + ptO) * uO)) * den;
            const double rks = (Nb->rk * fN - O->rk * fO) * den;
            F->st[0] = ds[0];
            for (int a = 0; a < 3; ++a)
                F->st[1 + a] = ms[a] / ds[0];
            F->st[4] = pressure_star(cfg, ds, ms, Es, rks);
            for (int c = 0; c < 3; ++c)
                F->st[5 + c] = ((Nb->ra[c] * fN - O->ra[c] * fO) * den) / ds[0];
            F->st[8] = ((Nb->rb * fN - O->rb * fO) * den) / ds[0];
            /* HLLSolver.cpp:245-251 */
            const double fb = (SN * uO * O->rb - SO * uN * Nb->rb + SN * SO * (Nb->rb - O->rb)) * den;
            for (int a = 0; a < 3; ++a)
                F->rb[a] = fb * N[a];
        }
    } else if (SN <= 0) {
        for (int k = 0; k < nd; ++k)
            for (int a = 0; a < 3; ++a)
                F->rho[k][a] = N[a] * (uN * Nb->rho[k]);
        double w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = Nb->m[a] * uN + IN[a] * (Nb->p + ptN);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                F->m[3 * a + b] = w[a] * N[b];
        const double e = uN * (Nb->E + Nb->p + ptN);
        for (int a = 0; a < 3; ++a)
            F->E[a] = N[a] * e;
        for (int a = 0; a < 3; ++a) { /* HLLSolver.cpp:277-284 */
            F->rk[a] = (uN * Nb->rk) * N[a];
            F->reps[a] = (uN * Nb->reps) * N[a];
            F->rb[a] = (uN * Nb->rb) * N[a]; /* HLLSolver.cpp:289-290 */
        }
        for (int c = 0; c < 3; ++c) /* HLLSolver.cpp:286-287 */
            for (int b = 0; b < 3; ++b)
                F->ra[3 * c + b] = (Nb->ra[c] * uN) * N[b];
        star_side(Nb, F->st); /* HLLSolver.cpp:255-270 */
    } else {
        return ERR_RIEMANN;
    }
    return 0;
}

static int face_flux_dispatch(const or_config *cfg, const double *N,
                              const face_state *O, const face_state *Nb,
                              face_flux *F)
{
    if (cfg->flux == OR_FLUX_HLLC)
        return hllc_face(cfg, N, O, Nb, F);
    if (cfg->flux == OR_FLUX_HLL)
        return hll_face(cfg, N, O, Nb, F);
    return ERR_INIT;
}

int or_face_flux(const or_solver *s, const double *N, const double *cL,
                 const double *vL, double pL, const double *cR,
                 const double *vR, double pR, double *Frho, double *Fm,
                 double *FE)
{
    face_state O, Nb;
    memset(&O, 0, sizeof O);
    memset(&Nb, 0, sizeof Nb);
    for (int k = 1; k <= s->cfg.Ns; ++k) {
        O.c[k] = cL[k - 1];
        Nb.c[k] = cR[k - 1];
    }
    for (int a = 0; a < 3; ++a) {
        O.v[a] = vL[a];
        Nb.v[a] = vR[a];
    }
    O.p = pL;
    Nb.p = pR;
    close_face(&s->cfg, &O);
    close_face(&s->cfg, &Nb);
    face_flux F;
    const int err = face_flux_dispatch(&s->cfg, N, &O, &Nb, &F);
    if (err)
        return err;
    for (int k = 0; k <= s->cfg.Ns; ++k)
        Frho[k] = dot3(F.rho[k], N);
    for (int a = 0; a < 3; ++a)
        Fm[a] = (F.m[3 * a] * N[0] + F.m[3 * a + 1] * N[1]) + F.m[3 * a + 2] * N[2];
    *FE = dot3(F.E, N);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* MUSCL-Hancock predictor (HancockStage.cpp:10-448, HancockStage.hpp:23-89)  */

/* Recomputes a face side's primitives from its conserved values after the
 * predictor (HancockStage.cpp:174-290): c_k = rho_k / M_k, v = m / rho,
 * U = E - rho v^2 / 2 - rho k, p = pFromUv, T = TFromUv. */
static void rederive_face(const or_config *cfg, face_state *fs)
{
    double c0 = 0;
    for (int k = 1; k <= cfg->Ns; ++k) {
        fs->c[k] = fs->rho[k] / cfg->M[k - 1];
        c0 += fs->c[k];
    }
    fs->c[0] = c0;
    for (int a = 0; a < 3; ++a)
        fs->v[a] = fs->m[a] / fs->rho[0];
    const double v2 = dot3(fs->v, fs->v);
    fs->U = (fs->E - fs->rho[0] * v2 * 0.5) - 0.0;
    fs->p = eos_p(cfg, fs->c, fs->U);
    fs->T = eos_T(cfg, fs->c, fs->U);
}

/* Each cell's divergence of the physical fluxes of its own face sides
 * (HancockDivergence for rho_k and rho; the momentum and energy loops of
 * HancockStage.cpp:36-117), divided by V; every face side of the cell then
 * moves by -div dt / 2 (HancockTimeIntegration) and recovers its primitives.
 * The turbulent terms (2/3) rho k are zero. */
static void hancock_predict(or_solver *s, face_state *FO, face_state *FN, double dt)
{
    const or_config *cfg = &s->cfg;
    const int Ns = cfg->Ns;
    const double half = 0.5 * dt;
    for (long i = 0; i < s->nc; ++i) {
        double dr[OR_MAXS + 1] = {0}, dm[3] = {0, 0, 0}, dE = 0;
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            const face_state *fs = (s->own[f] == i) ? &FO[f] : &FN[f];
            double n[3];
            if (s->own[f] == i)
                memcpy(n, &s->fN[3 * f], sizeof n);
            else
                for (int a = 0; a < 3; ++a)
                    n[a] = s->fN[3 * f + a] * -1;
            const double S = s->fS[f];
            for (int k = 0; k <= Ns; ++k) {
                double rv[3];
                for (int a = 0; a < 3; ++a)
                    rv[a] = fs->v[a] * fs->rho[k];
                dr[k] += dot3(rv, n) * S;
            }
            const double pp = fs->p + (2. / 3.) * 0.0;
            for (int a = 0; a < 3; ++a) {
                const double T0 = fs->m[a] * fs->v[0], T1 = fs->m[a] * fs->v[1],
                             T2 = fs->m[a] * fs->v[2];
                const double row = (T0 * n[0] + T1 * n[1]) + T2 * n[2];
                dm[a] += (n[a] * pp + row) * S;
            }
            double vE[3], vp[3];
            for (int a = 0; a < 3; ++a) {
                vE[a] = fs->v[a] * fs->E;
                vp[a] = fs->v[a] * pp;
            }
            dE += (dot3(vE, n) + dot3(vp, n)) * S;
        }
        const double V = s->cV[i];
        for (int k = 0; k <= Ns; ++k)
            dr[k] /= V;
        for (int a = 0; a < 3; ++a)
            dm[a] /= V;
        dE /= V;
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            face_state *fs = (s->own[f] == i) ? &FO[f] : &FN[f];
            for (int k = 0; k <= Ns; ++k)
                fs->rho[k] -= dr[k] * half;
            for (int a = 0; a < 3; ++a)
                fs->m[a] -= dm[a] * half;
            fs->E -= dE * half;
        }
    }
    /* owner sides of every face, neighbour sides of inner faces */
    for (long f = 0; f < s->nf; ++f) {
        if (s->fpatch[f] == -3)
            continue;
        rederive_face(cfg, &FO[f]);
        if (s->fpatch[f] < 0)
            rederive_face(cfg, &FN[f]);
    }
}

/* ------------------------------------------------------------------------ */
/* one stage: Advection2dOrder (Advection2dOrder.cpp:16-636), or with the    */
/* Hancock predictor Advection (Advection.cpp:20-514)                         */

static int stage_impl(or_solver *s, double dt, int hancock);

int or_stage(or_solver *s, double dt)
{
    return stage_impl(s, dt, 0);
}

static int stage_impl(or_solver *s, double dt, int hancock)
{
    const or_config *cfg = &s->cfg;
    or_state *st = &s->st;
    const int Ns = cfg->Ns;
    const long nc = s->nc, nf = s->nf;
    int err = 0;

    double *Gc = malloc(3 * nc * Ns * sizeof(double));
    double *Lc = malloc(3 * nc * Ns * sizeof(double));
    double *Gv = malloc(9 * nc * sizeof(double));
    double *Lv = malloc(9 * nc * sizeof(double));
    double *Gp = malloc(3 * nc * sizeof(double));
    double *Lp = malloc(3 * nc * sizeof(double));
    /* turbulence: limited gradients of k and epsilon, and of a (kEpsA) */
    const int afield = a_field(cfg), bfield = b_field(cfg);
    double *Gt = calloc(3 * nc * 3, sizeof(double));
    double *Lt = calloc(3 * nc * 3, sizeof(double));
    double *Ga = calloc(afield ? 9 * nc : 1, sizeof(double));
    double *La = calloc(afield ? 9 * nc : 1, sizeof(double));
    face_state *FO = calloc(nf, sizeof(face_state));
    face_state *FN = calloc(nf, sizeof(face_state));
    face_flux *FF = calloc(nf, sizeof(face_flux));

    /* limited gradients (Advection2dOrder.cpp:54-75) */
    for (int k = 0; k < Ns && !err; ++k) {
        err = grad_scalar(s, st->c[k + 1], cfg->bc_c[k], &Gc[3 * nc * k]);
        if (!err)
            err = tvd_scalar(s, &Gc[3 * nc * k], st->c[k + 1], cfg->bc_c[k],
                             &Lc[3 * nc * k]);
    }
    if (!err)
        err = grad_vector(s, st->v, cfg->bc_v, Gv);
    if (!err)
        err = tvd_vector(s, Gv, st->v, cfg->bc_v, Lv);
    if (!err)
        err = grad_scalar(s, st->p, cfg->bc_p, Gp);
    if (!err)
        err = tvd_scalar(s, Gp, st->p, cfg->bc_p, Lp);
    if (!err && cfg->turb) { /* Advection2dOrder.cpp:77-90 */
        err = grad_scalar(s, st->k, cfg->bc_k, Gt);
        if (!err)
            err = tvd_scalar(s, Gt, st->k, cfg->bc_k, Lt);
        if (!err)
            err = grad_scalar(s, st->eps, cfg->bc_eps, &Gt[3 * nc]);
        if (!err)
            err = tvd_scalar(s, &Gt[3 * nc], st->eps, cfg->bc_eps, &Lt[3 * nc]);
    }
    if (!err && afield) { /* Advection2dOrder.cpp:91-97 */
        err = grad_vector(s, st->a, cfg->bc_a, Ga);
        if (!err)
            err = tvd_vector(s, Ga, st->a, cfg->bc_a, La);
    }
    if (!err && bfield) { /* Advection2dOrder.cpp:99-106 */
        err = grad_scalar(s, st->b, cfg->bc_b, &Gt[6 * nc]);
        if (!err)
            err = tvd_scalar(s, &Gt[6 * nc], st->b, cfg->bc_b, &Lt[6 * nc]);
    }
    if (err)
        goto done;

    /* reconstruction to faces (Advection2dOrder.cpp:110-231) */
    for (long i = 0; i < nc; ++i) {
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            const double *d = &s->cfd[18 * i + 3 * j];
            face_state *fs = (s->own[f] == i) ? &FO[f] : &FN[f];
            for (int k = 0; k < Ns; ++k)
                fs->c[k + 1] = dot3(&Lc[3 * nc * k + 3 * i], d) + st->c[k + 1][i];
            for (int a = 0; a < 3; ++a) {
                const double *T = &Lv[9 * i + 3 * a];
                fs->v[a] = (T[0] * d[0] + T[1] * d[1] + T[2] * d[2]) + st->v[a][i];
            }
            fs->p = dot3(&Lp[3 * i], d) + st->p[i];
            if (cfg->turb) { /* Advection2dOrder.cpp:143-150 */
                fs->k = dot3(&Lt[3 * i], d) + st->k[i];
                fs->eps = dot3(&Lt[3 * nc + 3 * i], d) + st->eps[i];
            }
            if (afield) /* Advection2dOrder.cpp:151-155 */
                for (int a = 0; a < 3; ++a) {
                    const double *T = &La[9 * i + 3 * a];
                    fs->a[a] = (T[0] * d[0] + T[1] * d[1] + T[2] * d[2]) + st->a[a][i];
                }
            if (bfield) /* Advection2dOrder.cpp:156-158 */
                fs->b = dot3(&Lt[6 * nc + 3 * i], d) + st->b[i];
        }
    }
    /* owner side everywhere, neighbour side on inner faces */
    for (long f = 0; f < nf; ++f) {
        if (s->fpatch[f] == -3)
            continue; /* low-side face of a periodic pair, unused */
        close_face(cfg, &FO[f]);
        if (s->fpatch[f] < 0)
            close_face(cfg, &FN[f]);
    }
    if (hancock)
        hancock_predict(s, FO, FN, dt);
    /* boundary ghosts from the owner face value (Advection2dOrder.cpp:398-521) */
    for (long f = 0; f < nf && !err; ++f) {
        if (s->fpatch[f] < 0)
            continue;
        const int pt = s->fpatch[f];
        face_state *g = &FN[f];
        for (int k = 1; k <= Ns && !err; ++k)
            err = bc_cell_scalar(&cfg->bc_c[k - 1][pt], FO[f].c[k], &g->c[k]);
        if (!err)
            err = bc_cell_vec(&cfg->bc_v[pt], FO[f].v, &s->fN[3 * f], g->v);
        if (!err)
            err = bc_cell_scalar(&cfg->bc_p[pt], FO[f].p, &g->p);
        if (!err && cfg->turb) { /* Advection2dOrder.cpp:440-458 */
            err = bc_cell_scalar(&cfg->bc_k[pt], FO[f].k, &g->k);
            if (!err)
                err = bc_cell_scalar(&cfg->bc_eps[pt], FO[f].eps, &g->eps);
            if (!err && afield) /* Advection2dOrder.cpp:460-470 */
                err = bc_cell_vec(&cfg->bc_a[pt], FO[f].a, &s->fN[3 * f], g->a);
            if (!err && bfield) /* Advection2dOrder.cpp:472-482 */
                err = bc_cell_scalar(&cfg->bc_b[pt], FO[f].b, &g->b);
        }
        if (!err)
            close_face(cfg, g);
    }
    if (err)
        goto done;

    /* numerical fluxes and the star states (Advection2dOrder.cpp:528) */
    for (long f = 0; f < nf && !err; ++f)
        if (s->fpatch[f] != -3) {
            err = face_flux_dispatch(cfg, &s->fN[3 * f], &FO[f], &FN[f], &FF[f]);
            memcpy(&s->star_stage[NSTAR * f], FF[f].st, sizeof FF[f].st);
        }
    if (err)
        goto done;

    /* time integration (Advection2dOrder.cpp:541-602) with divergence
     * (divergence.hpp:315-350) */
    for (long i = 0; i < nc; ++i) {
        double accE = 0, accm[3] = {0, 0, 0}, accr[OR_MAXS + 1] = {0}, acck = 0, acce = 0;
        double acca[3] = {0, 0, 0}, accb = 0;
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            double n[3];
            if (s->own[f] == i)
                memcpy(n, &s->fN[3 * f], sizeof n);
            else
                for (int a = 0; a < 3; ++a)
                    n[a] = s->fN[3 * f + a] * -1;
            const double S = s->fS[f];
            accE += dot3(FF[f].E, n) * S;
            for (int a = 0; a < 3; ++a) {
                const double *T = &FF[f].m[3 * a];
                accm[a] += (T[0] * n[0] + T[1] * n[1] + T[2] * n[2]) * S;
            }
            for (int k = 0; k <= Ns; ++k)
                accr[k] += dot3(FF[f].rho[k], n) * S;
            acck += dot3(FF[f].rk, n) * S;
            acce += dot3(FF[f].reps, n) * S;
            for (int a = 0; a < 3; ++a) {
                const double *T = &FF[f].ra[3 * a];
                acca[a] += (T[0] * n[0] + T[1] * n[1] + T[2] * n[2]) * S;
            }
            accb += dot3(FF[f].rb, n) * S;
        }
        const double V = s->cV[i];
        const double r0 = st->rho[0][i];
        const double vold[3] = {st->v[0][i], st->v[1][i], st->v[2][i]};
        st->E[i] = st->E[i] - (accE / V) * dt;
        for (int a = 0; a < 3; ++a)
            st->m[a][i] = st->m[a][i] - (accm[a] / V) * dt;
        if (cfg->gravity) {
            st->E[i] = st->E[i] + r0 * dot3(vold, cfg->g) * dt;
            for (int a = 0; a < 3; ++a)
                st->m[a][i] = st->m[a][i] + (cfg->g[a] * r0) * dt;
        }
        for (int k = 0; k <= Ns; ++k)
            st->rho[k][i] = st->rho[k][i] - (accr[k] / V) * dt;
        if (cfg->turb) { /* Advection2dOrder.cpp:560-564 */
            st->rk[i] = st->rk[i] - (acck / V) * dt;
            st->reps[i] = st->reps[i] - (acce / V) * dt;
        }
        if (afield) /* Advection2dOrder.cpp:567-570 */
            for (int a = 0; a < 3; ++a)
                st->ra[a][i] = st->ra[a][i] - (acca[a] / V) * dt;
        if (bfield) /* Advection2dOrder.cpp:572-574 */
            st->rb[i] = st->rb[i] - (accb / V) * dt;
    }
    recover_primitives(s, st);

done:
    free(Gc);
    free(Lc);
    free(Gv);
    free(Lv);
    free(Gp);
    free(Lp);
    free(Gt);
    free(Lt);
    free(Ga);
    free(La);
    free(FO);
    free(FN);
    free(FF);
    return err;
}

/* bunchOfFields::average + copyFrom (bunchOfFields.hpp:166-235) */
static void rk_average(or_solver *s, const or_state *Un)
{
    or_state *st = &s->st;
    const int Ns = s->cfg.Ns;
    const double w = 0.5, iw = 1 - w;
    double cc[OR_MAXS + 1];
    for (long i = 0; i < s->nc; ++i) {
        for (int k = 0; k <= Ns; ++k) {
            st->c[k][i] = w * st->c[k][i] + iw * Un->c[k][i];
            st->rho[k][i] = w * st->rho[k][i] + iw * Un->rho[k][i];
        }
        for (int a = 0; a < 3; ++a)
            st->m[a][i] = st->m[a][i] * w + Un->m[a][i] * iw;
        st->E[i] = w * st->E[i] + iw * Un->E[i];
        st->rk[i] = w * st->rk[i] + iw * Un->rk[i];
        st->reps[i] = w * st->reps[i] + iw * Un->reps[i];
        for (int a = 0; a < 3; ++a)
            st->ra[a][i] = st->ra[a][i] * w + Un->ra[a][i] * iw;
        /* rho b is averaged; b is not re-derived (bunchOfFields.hpp:186-192):
         * it keeps the second stage's rho b / rho */
        st->rb[i] = w * st->rb[i] + iw * Un->rb[i];
        for (int a = 0; a < 3; ++a)
            st->v[a][i] = st->m[a][i] / st->rho[0][i];
        if (s->cfg.turb) {
            st->k[i] = st->rk[i] / st->rho[0][i];
            st->eps[i] = st->reps[i] / st->rho[0][i];
        }
        if (a_field(&s->cfg))
            for (int a = 0; a < 3; ++a)
                st->a[a][i] = st->ra[a][i] / st->rho[0][i];
        double vel[3] = {st->v[0][i], st->v[1][i], st->v[2][i]};
        st->U[i] = st->E[i] - st->rk[i] - 0.5 * st->rho[0][i] * dot3(vel, vel);
        for (int k = 0; k <= Ns; ++k)
            cc[k] = st->c[k][i];
        st->p[i] = eos_p(&s->cfg, cc, st->U[i]);
        st->T[i] = eos_T(&s->cfg, cc, st->U[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* explicit molecular diffusion stage (Diffusion.cpp:20-1061) with constant  */
/* transport coefficients, the explicit matrix solver and linear             */
/* reconstruction; no turbulence, enthalpy flow or molar-mass correction     */

/* r_other - r_i across the inner face f seen from cell i; a periodic face
 * (no reference counterpart) spans its centre distance through the
 * boundary along its normal. */
static void face_delta(const or_solver *s, long f, long i, double *d)
{
    const long o = (s->own[f] == i) ? s->nei[f] : s->own[f];
    if (s->fper[f] > 0) {
        const double sgn = (s->own[f] == i) ? 1. : -1.;
        for (int b = 0; b < 3; ++b)
            d[b] = s->fN[3 * f + b] * s->fper[f] * sgn;
        return;
    }
    for (int b = 0; b < 3; ++b)
        d[b] = s->crc[3 * o + b] - s->crc[3 * i + b];
}

/* surfGrad (gradient.hpp:96-165) of cell values q[ncomp] of cell o across
 * face f: inner faces ((q_n - q_o) / |d|^2) d with d = r_n - r_o; boundary
 * faces ((q_g - q_o) / |D|) D / |D| with the ghost value q_g and
 * D = 2 (r_f - r_o).  out[3 c + b] = d(q_c) along b. */
static void surf_grad(const or_solver *s, long f, int ncomp, const double *qo,
                      const double *qn, double *out)
{
    const long o = s->own[f];
    if (s->fpatch[f] < 0) {
        double d[3];
        face_delta(s, f, o, d);
        const double m = or_hypot3(d[0], d[1], d[2]);
        const double m2 = m * m; /* pow<scalar, 2> (intExpPow.hpp) */
        for (int c = 0; c < ncomp; ++c)
            for (int b = 0; b < 3; ++b)
                out[3 * c + b] = ((qn[c] - qo[c]) / m2) * d[b];
    } else {
        double D[3];
        for (int b = 0; b < 3; ++b)
            D[b] = (s->frc[3 * f + b] - s->crc[3 * o + b]) * 2.;
        const double m = or_hypot3(D[0], D[1], D[2]);
        double nrm[3];
        for (int b = 0; b < 3; ++b)
            nrm[b] = D[b] / m;
        for (int c = 0; c < ncomp; ++c)
            for (int b = 0; b < 3; ++b)
                out[3 * c + b] = ((qn[c] - qo[c]) / m) * nrm[b];
    }
}

/* nonIdeality (mixtureIdeal.cpp:232-241: 0; mixtureVanDerWaals.cpp:369-387,
 * vanDerWaalsFluid.cpp:78-82: -a c^2) and UvcFromT (Cv T; vdW Cv T - a c,
 * vanDerWaalsFluid.cpp:38-42) of a cell with concentrations c (c0..cNs). */
static double non_ideality(const or_config *cfg, const double *c)
{
    if (cfg->gas_model != OR_GAS_VDW)
        return 0;
    double a, b;
    vdw_mix(cfg, c, &a, &b);
    return -a * (c[0] * c[0]);
}

static double uvc_from_T(const or_config *cfg, const double *c, double T)
{
    if (cfg->gas_model != OR_GAS_VDW)
        return cv_mix(cfg, c) * T;
    double a, b;
    vdw_mix(cfg, c, &a, &b);
    return cv_mix(cfg, c) * T - a * c[0];
}

/* outward normal of face f for cell i (owner: N, neighbour: N * -1) */
static void out_normal(const or_solver *s, long f, long i, double *n)
{
    for (int a = 0; a < 3; ++a)
        n[a] = (s->own[f] == i) ? s->fN[3 * f + a] : s->fN[3 * f + a] * -1;
}

/* ghost concentrations of a boundary face (boundaryConditionValueCell) */
static int ghost_conc(const or_solver *s, long f, long o, double *cg /* c0..cNs */)
{
    const int p = s->fpatch[f];
    cg[0] = 0;
    for (int k = 1; k <= s->cfg.Ns; ++k) {
        if (bc_cell_scalar(&s->cfg.bc_c[k - 1][p], s->st.c[k][o], &cg[k]))
            return ERR_BC;
        cg[0] += cg[k];
    }
    return 0;
}

/* generateDTimeLaplacian (SLEMatrix.cpp:410-547 scalar, 706-908 per vector
 * component): diag = rnew (1 / dt) + sum_f (k_f / |d|) (N S / V & d / |d|),
 * b = rold (1 / dt) + the same times the ghost value on boundary faces, and
 * off[6 i + j] = -(k_f / |d|) (...) for the inner face j of cell i
 * (surfacesOfCells order; 0 on boundary faces).  k: coefficient per face;
 * ghost: ghost value per face. */
static void laplacian_matrix(const or_solver *s, const double *k, const double *rnew,
                             const double *rold, const double *ghost, double dt, double *diag,
                             double *b, double *off)
{
    const double rdt = 1 / dt;
    for (long i = 0; i < s->nc; ++i) {
        diag[i] = rnew[i] * rdt;
        b[i] = rold[i] * rdt;
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            double on[3], dR[3];
            for (int a = 0; a < 3; ++a)
                on[a] = s->fN[3 * f + a] * s->fS[f] / s->cV[i];
            off[6 * i + j] = 0;
            if (s->fpatch[f] < 0) {
                const long o = (s->own[f] == i) ? s->nei[f] : s->own[f];
                if (s->own[f] != i)
                    for (int a = 0; a < 3; ++a)
                        on[a] = on[a] * -1;
                (void)o;
                face_delta(s, f, i, dR);
                const double m = or_hypot3(dR[0], dR[1], dR[2]);
                double dn[3];
                for (int a = 0; a < 3; ++a)
                    dn[a] = dR[a] / m;
                const double dt_ = dot3(on, dn);
                diag[i] += k[f] / m * dt_;
                off[6 * i + j] = -k[f] / m * dt_;
            } else {
                for (int a = 0; a < 3; ++a)
                    dR[a] = (s->frc[3 * f + a] - s->crc[3 * i + a]) * 2;
                const double m = or_hypot3(dR[0], dR[1], dR[2]);
                double dn[3];
                for (int a = 0; a < 3; ++a)
                    dn[a] = dR[a] / m;
                const double dt_ = dot3(on, dn);
                diag[i] += k[f] / m * dt_;
                b[i] += k[f] / m * dt_ * ghost[f];
            }
        }
    }
}

/* JacobiSolver::algorithm (JacobiSolver.cpp:12-84) with
 * relativeIterationDifference and normalize (abstractMatrixSolver.cpp:22-38):
 * the last sweep of each iteration runs the rows from the last to the first
 * and sums each row's lower, then upper, triangle from its last entry. */
static void jacobi(const or_solver *s, const double *diag, const double *b, const double *off,
                   const double *x0, long maxit, double *x)
{
    const long nc = s->nc;
    double *old = malloc(nc * sizeof(double));
    memcpy(old, x0, nc * sizeof(double));
    long it = 0;
    for (;;) {
        ++it;
        for (long i = 0; i < nc; ++i) {
            const double aii = 1. / diag[i];
            double row = 0;
            long lo[6], up[6];
            double alo[6], aup[6];
            int nl = 0, nu = 0;
            for (int j = 0; j < s->cnf[i]; ++j) {
                const long f = s->cf[6 * i + j];
                if (s->fpatch[f] >= 0)
                    continue;
                const long o = (s->own[f] == i) ? s->nei[f] : s->own[f];
                if (o < i) {
                    lo[nl] = o;
                    alo[nl++] = off[6 * i + j];
                } else {
                    up[nu] = o;
                    aup[nu++] = off[6 * i + j];
                }
            }
            for (int j = nl - 1; j >= 0; --j)
                row -= alo[j] * old[lo[j]];
            for (int j = nu - 1; j >= 0; --j)
                row -= aup[j] * old[up[j]];
            x[i] = (b[i] + row) * aii;
        }
        double mx = 0;
        for (long i = 0; i < nc; ++i) {
            const double e = fabs((x[i] - old[i]) / (fabs(x[i]) + STAB));
            if (i == 0 || mx < e)
                mx = e;
        }
        const double diff = 100. * mx;
        if (diff < 1E-12 || it >= maxit)
            break;
        memcpy(old, x, nc * sizeof(double));
    }
    for (long i = 0; i < nc; ++i)
        if (fabs(x[i]) < DBL_EPSILON)
            x[i] = 0;
    free(old);
}

/* GaussSeidelSolver::algorithm (GaussSeidelSolver.cpp:12-91): per iteration
 * a forward sweep (rows in order, lower triangle from the new iterate, upper
 * from the old) and a backward sweep (rows in reverse, lower from the old
 * iterate, upper from the new, each triangle from its last entry), every
 * term divided as a_ij x_j / a_ii; same stop test as Jacobi. */
static void gauss_seidel(const or_solver *s, const double *diag, const double *b,
                         const double *off, const double *x0, long maxit, double *x)
{
    const long nc = s->nc;
    double *old = malloc(nc * sizeof(double));
    memcpy(old, x0, nc * sizeof(double));
    memcpy(x, x0, nc * sizeof(double));
    long it = 0;
    for (;;) {
        ++it;
        for (int pass = 0; pass < 2; ++pass)
            for (long r = 0; r < nc; ++r) {
                const long i = pass == 0 ? r : nc - 1 - r;
                const double aii = 1. / diag[i];
                long lo[6], up[6];
                double alo[6], aup[6];
                int nl = 0, nu = 0;
                for (int j = 0; j < s->cnf[i]; ++j) {
                    const long f = s->cf[6 * i + j];
                    if (s->fpatch[f] >= 0)
                        continue;
                    const long o = (s->own[f] == i) ? s->nei[f] : s->own[f];
                    if (o < i) {
                        lo[nl] = o;
                        alo[nl++] = off[6 * i + j];
                    } else {
                        up[nu] = o;
                        aup[nu++] = off[6 * i + j];
                    }
                }
                double xi = b[i] * aii;
                if (pass == 0) {
                    for (int j = 0; j < nl; ++j)
                        xi -= alo[j] * x[lo[j]] * aii;
                    for (int j = 0; j < nu; ++j)
                        xi -= aup[j] * old[up[j]] * aii;
                } else {
                    for (int j = nl - 1; j >= 0; --j)
                        xi -= alo[j] * old[lo[j]] * aii;
                    for (int j = nu - 1; j >= 0; --j)
                        xi -= aup[j] * x[up[j]] * aii;
                }
                x[i] = xi;
            }
        double mx = 0;
        for (long i = 0; i < nc; ++i) {
            const double e = fabs((x[i] - old[i]) / (fabs(x[i]) + STAB));
            if (i == 0 || mx < e)
                mx = e;
        }
        if (100. * mx < 1E-12 || it >= maxit)
            break;
        memcpy(old, x, nc * sizeof(double));
    }
    for (long i = 0; i < nc; ++i)
        if (fabs(x[i]) < DBL_EPSILON)
            x[i] = 0;
    free(old);
}

static void sle_solve(const or_solver *s, const double *diag, const double *b, const double *off,
                      const double *x0, double *x)
{
    if (s->cfg.sle_solver == OR_SLE_GAUSS_SEIDEL)
        gauss_seidel(s, diag, b, off, x0, s->cfg.sle_iters, x);
    else
        jacobi(s, diag, b, off, x0, s->cfg.sle_iters, x);
}

/* The stage with the Jacobi matrix solver (Diffusion.cpp:318-330, 431-458,
 * 487-499, 711-760, 795-806, 818-847 with msolver.solve), from the explicit
 * face terms of diffusion_explicit: mass fractions first, then the new
 * concentrations, velocity (plus the explicit div(mu strainSurf1)) and
 * temperature (plus the viscous heating). */
static int diffusion_implicit(or_solver *s, double dt, const double *Y, const double *CCVold,
                              const double *gradV, const double *divV, const double *srho,
                              const double *FS, const double *GH, const double *MC)
{
    const or_config *cfg = &s->cfg;
    or_state *st = &s->st;
    const int Ns = cfg->Ns;
    const long nc = s->nc, nf = s->nf;
    double *kf = malloc(nf * sizeof(double));
    double *rn = malloc(nc * sizeof(double)), *ro = malloc(nc * sizeof(double));
    double *x0 = malloc(nc * sizeof(double));
    double *diag = malloc(nc * sizeof(double)), *b = malloc(nc * sizeof(double));
    double *off = malloc(6 * nc * sizeof(double));
    double *Yn = malloc(Ns * nc * sizeof(double));
    double *cn = malloc((Ns + 1) * nc * sizeof(double));
    double *CCVnew = malloc(nc * sizeof(double));
    double *vn = malloc(3 * nc * sizeof(double)), *Tn = malloc(nc * sizeof(double));
    int err = 0;
    /* mass fractions: rho Y_k^new = rho Y_k^old + dt div(rho D grad Y_k) */
    for (int k = 0; k < Ns; ++k) {
        for (long f = 0; f < nf; ++f)
            kf[f] = srho[f] * cfg->D;
        for (long i = 0; i < nc; ++i) {
            rn[i] = st->rho[0][i];
            ro[i] = Y[i * Ns + k] * st->rho[0][i];
            x0[i] = Y[i * Ns + k];
        }
        laplacian_matrix(s, kf, rn, ro, GH + k * nf, dt, diag, b, off);
        if (cfg->molmass)
            for (long i = 0; i < nc; ++i) {
                /* distributeSourceTerm of -div(corr_k) (Diffusion.cpp:356-364) */
                double dm = 0;
                for (int j = 0; j < s->cnf[i]; ++j) {
                    const long f = s->cf[6 * i + j];
                    double n[3];
                    out_normal(s, f, i, n);
                    dm += dot3(&MC[(k * nf + f) * 3], n) * s->fS[f];
                }
                const double src = (dm / s->cV[i]) * (-1.0);
                if (src >= 0.)
                    b[i] += src;
                else
                    diag[i] -= src / (x0[i] + STAB);
            }
        sle_solve(s, diag, b, off, x0, Yn + k * nc);
    }
    for (long i = 0; i < nc; ++i) {
        double sum = 0.0;
        for (int k = 0; k < Ns; ++k) {
            if (Yn[k * nc + i] < 0.)
                Yn[k * nc + i] = 0.0;
            sum += Yn[k * nc + i];
        }
        double c[OR_MAXS + 1];
        c[0] = 0;
        for (int k = 0; k < Ns; ++k) {
            Yn[k * nc + i] /= sum;
            c[k + 1] = Yn[k * nc + i] * st->rho[0][i] / cfg->M[k];
            c[0] += c[k + 1];
        }
        for (int k = 0; k <= Ns; ++k)
            cn[k * nc + i] = c[k];
        CCVnew[i] = cv_mix(cfg, c) * c[0];
    }
    /* velocity components */
    for (long f = 0; f < nf; ++f)
        kf[f] = cfg->mu;
    for (int a = 0; a < 3; ++a) {
        for (long i = 0; i < nc; ++i) {
            rn[i] = st->rho[0][i];
            ro[i] = st->v[a][i] * st->rho[0][i];
            x0[i] = st->v[a][i];
        }
        laplacian_matrix(s, kf, rn, ro, GH + (Ns + 1 + a) * nf, dt, diag, b, off);
        for (long i = 0; i < nc; ++i) {
            double dS = 0;
            for (int j = 0; j < s->cnf[i]; ++j) {
                const long f = s->cf[6 * i + j];
                double n[3];
                out_normal(s, f, i, n);
                const double *Ts = &FS[9 * f + 3 * a];
                dS += (Ts[0] * n[0] + Ts[1] * n[1] + Ts[2] * n[2]) * s->fS[f];
            }
            b[i] += dS / s->cV[i];
        }
        sle_solve(s, diag, b, off, x0, vn + a * nc);
    }
    /* temperature */
    for (long f = 0; f < nf; ++f)
        kf[f] = cfg->kappa;
    for (long i = 0; i < nc; ++i) {
        double co[OR_MAXS + 1], cw[OR_MAXS + 1];
        for (int k = 0; k <= Ns; ++k) {
            co[k] = st->c[k][i];
            cw[k] = cn[k * nc + i];
        }
        rn[i] = CCVnew[i];
        ro[i] = CCVold[i] * st->T[i] + non_ideality(cfg, co) - non_ideality(cfg, cw);
        x0[i] = st->T[i];
    }
    laplacian_matrix(s, kf, rn, ro, GH + Ns * nf, dt, diag, b, off);
    for (long i = 0; i < nc; ++i) {
        const double *G = &gradV[9 * i];
        double sv[9];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c)
                sv[3 * a + c] = G[3 * c + a] + G[3 * a + c];
        const double lam = (2. / 3.) * divV[i];
        sv[0] -= lam;
        sv[4] -= lam;
        sv[8] -= lam;
        double heat = 0;
        for (int q = 0; q < 9; ++q)
            heat += (sv[q] * cfg->mu) * G[q];
        b[i] += heat;
    }
    sle_solve(s, diag, b, off, x0, Tn);
    /* recalculation of the cell values (Diffusion.cpp:963-1029) */
    for (long i = 0; i < nc; ++i) {
        double c[OR_MAXS + 1];
        for (int k = 0; k <= Ns; ++k) {
            c[k] = cn[k * nc + i];
            st->c[k][i] = c[k];
        }
        const double rho = st->rho[0][i];
        for (int k = 1; k <= Ns; ++k)
            st->rho[k][i] = c[k] * cfg->M[k - 1];
        double v[3];
        for (int a = 0; a < 3; ++a) {
            v[a] = vn[a * nc + i];
            st->v[a][i] = v[a];
            st->m[a][i] = v[a] * rho;
        }
        st->T[i] = Tn[i];
        st->U[i] = uvc_from_T(cfg, c, Tn[i]) * c[0];
        st->p[i] = eos_p(cfg, c, st->U[i]);
        st->E[i] = (st->U[i] + rho * dot3(v, v) * 0.5) + 0.0;
        if (Tn[i] < 0.)
            err = ERR_NEG_T;
    }
    free(kf);
    free(rn);
    free(ro);
    free(x0);
    free(diag);
    free(b);
    free(off);
    free(Yn);
    free(cn);
    free(CCVnew);
    free(vn);
    free(Tn);
    return err;
}

/* GaussEliminationSolver of effectiveTransportCoefficients
 * (transportCoefficients.hpp:397-455): rows swapped once per column by the
 * largest following entry, elimination and back substitution with
 * stabilised pivots, normalize. */
static void gauss_elimination(int N, double A[OR_MAXS][OR_MAXS], double *b, double *phi)
{
    for (int k = 0; k < N - 1; ++k) {
        double mv = A[k][k];
        int c = k;
        for (int r = k + 1; r < N; ++r)
            if (A[r][k] > mv) {
                mv = A[r][k];
                c = r;
            }
        if (c != k) {
            for (int j = 0; j < N; ++j) {
                const double t = A[c][j];
                A[c][j] = A[k][j];
                A[k][j] = t;
            }
            const double t = b[c];
            b[c] = b[k];
            b[k] = t;
        }
    }
    for (int k = 0; k < N - 1; ++k)
        for (int i = k + 1; i < N; ++i) {
            const double ratio = A[i][k] / (A[k][k] + STAB);
            for (int j = k + 1; j < N; ++j)
                A[i][j] = A[i][j] - ratio * A[k][j];
            b[i] = b[i] - ratio * b[k];
        }
    phi[N - 1] = b[N - 1] / (A[N - 1][N - 1] + STAB);
    for (int i = N - 2; i >= 0; --i) {
        double term = 0.0;
        for (int j = i + 1; j < N; ++j)
            term += A[i][j] * phi[j];
        phi[i] = (b[i] - term) / (A[i][i] + STAB);
    }
    for (int k = 0; k < N; ++k)
        if (fabs(phi[k]) < DBL_EPSILON)
            phi[k] = 0;
}

/* Molar-mass corrections of the species fluxes (caclulateDFluxes,
 * transportCoefficients.hpp:285-395, constant D_kl = D; Diffusion.cpp:
 * 326-377) and the explicit enthalpy flow sum_k (corr_k + W_k) h_k / M_k
 * times the surface temperature (Diffusion.cpp:379-421, 605-611), with
 * W_k = -rho D_k surfGrad(Y_k) the face fluxes FY and h_k = (Cv_k T + R T) / T
 * (mixtureIdeal.cpp:291-298). */
static int species_corrections(or_solver *s, const double *FY, const double *CSf,
                               const double *TSf, double *MC, double *EF)
{
    const or_config *cfg = &s->cfg;
    const or_state *st = &s->st;
    const int Ns = cfg->Ns;
    const long nf = s->nf;
    for (long f = 0; f < nf; ++f) {
        const long o = s->own[f];
        double DF[OR_MAXS][3];
        for (int k = 0; k < Ns; ++k)
            for (int b = 0; b < 3; ++b)
                DF[k][b] = 0;
        if (cfg->molmass && Ns > 1) {
            /* mole fractions c_k / c_0, ghosts from the ghost concentrations
             * (calculatedMolarFraction) */
            double qo[OR_MAXS], qn[OR_MAXS], g[3 * OR_MAXS];
            for (int k = 0; k < Ns; ++k)
                qo[k] = st->c[k + 1][o] / st->c[0][o];
            if (s->fpatch[f] < 0) {
                const long n = s->nei[f];
                for (int k = 0; k < Ns; ++k)
                    qn[k] = st->c[k + 1][n] / st->c[0][n];
            } else {
                double cg[OR_MAXS + 1];
                if (ghost_conc(s, f, o, cg))
                    return ERR_BC;
                for (int k = 0; k < Ns; ++k)
                    qn[k] = cg[k + 1] / cg[0];
            }
            surf_grad(s, f, Ns, qo, qn, g);
            const double *N = &CSf[0];
            double c[OR_MAXS];
            for (int k = 0; k < Ns; ++k)
                c[k] = N[(k + 1) * nf + f];
            double cmin = c[0];
            for (int k = 1; k < Ns; ++k)
                if (c[k] < cmin)
                    cmin = c[k];
            int rep = 0;
            for (int k = 0; k < Ns; ++k)
                if (cmin == c[k]) {
                    rep = k + 1;
                    break;
                }
            const double N0 = N[f];
            for (int b = 0; b < 3; ++b) {
                double A[OR_MAXS][OR_MAXS] = {{0}}, rhs[OR_MAXS], phi[OR_MAXS];
                for (int k = 0; k < Ns; ++k)
                    rhs[k] = -g[3 * k + b];
                for (int k1 = 1; k1 <= Ns; ++k1)
                    if (k1 != rep) {
                        for (int k2 = 1; k2 <= Ns; ++k2) {
                            if (k1 == k2)
                                continue;
                            const double N1 = N[k1 * nf + f], N2 = N[k2 * nf + f];
                            const double Aij = N1 * N2 / (N0 * N0 * cfg->D);
                            A[k1 - 1][k2 - 1] = -Aij;
                            A[k1 - 1][k1 - 1] += Aij;
                        }
                    } else {
                        rhs[k1 - 1] = 0;
                        for (int k2 = 1; k2 <= Ns; ++k2)
                            A[k1 - 1][k2 - 1] = N[k2 * nf + f] * cfg->M[k2 - 1];
                    }
                gauss_elimination(Ns, A, rhs, phi);
                for (int k = 0; k < Ns; ++k)
                    DF[k][b] = phi[k] * cfg->M[k] * c[k];
            }
        }
        if (cfg->molmass)
            for (int k = 0; k < Ns; ++k)
                for (int b = 0; b < 3; ++b)
                    MC[(k * nf + f) * 3 + b] = DF[k][b] - FY[(k * nf + f) * 3 + b];
        if (cfg->enth_flow) {
            const double T = TSf[f];
            double e[3] = {0, 0, 0};
            for (int k = 0; k < Ns; ++k) {
                const double h = (cfg->Cv[k] * T + cfg->R * T) / T;
                for (int b = 0; b < 3; ++b)
                    e[b] += (FY[(k * nf + f) * 3 + b] + MC[(k * nf + f) * 3 + b]) * h /
                            cfg->M[k];
            }
            for (int b = 0; b < 3; ++b)
                EF[3 * f + b] = e[b] * T;
        }
    }
    return 0;
}

/* kEpsModels::thetaS (kEpsModels.cpp:13-26): the realisability factor of
 * the turbulent viscosity for the velocity divergence divV */
static double theta_s(const or_config *cfg, double divV, double k, double eps, double CMS)
{
    const double Cmu2 = cfg->Cmu * cfg->Cmu;
    const double k2 = k * k;
    const double eps2 = eps * eps;
    const double CMS2 = CMS * CMS;
    const double divV2 = divV * divV;
    return 1. / sqrt(1 + 16. / 9. * Cmu2 * k2 / (CMS2 * eps2) * divV2);
}

static int diffusion_explicit(or_solver *s, double dt)
{
    const or_config *cfg = &s->cfg;
    or_state *st = &s->st;
    const int Ns = cfg->Ns;
    const long nc = s->nc, nf = s->nf;
    int err = 0;
    /* cell work arrays: Y[Ns], CCVold, gradV[9], divV */
    double *Y = malloc(nc * Ns * sizeof(double));
    double *CCVold = malloc(nc * sizeof(double));
    double *gradV = malloc(9 * nc * sizeof(double));
    double *divV = malloc(nc * sizeof(double));
    double *vf = malloc(3 * nf * sizeof(double));
    double *srho = malloc(nf * sizeof(double));
    double *FY = malloc(3 * Ns * nf * sizeof(double)); /* -rhoD_k surfGrad(Y_k) */
    double *FT = malloc(3 * nf * sizeof(double));      /* -kappa surfGrad(T) */
    double *FV = malloc(9 * nf * sizeof(double));      /* -mu surfGrad(v) */
    double *FS = malloc(9 * nf * sizeof(double));      /* mu strainSurf1 */
    /* boundary ghost values per face: Y_k (Ns), T, v (3) */
    double *GH = malloc((Ns + 4) * nf * sizeof(double));
    /* surface concentrations (c_0 = sum, c_1..c_Ns) and temperature, the
     * molar-mass corrections of the species fluxes and the enthalpy flow */
    double *CSf = malloc((Ns + 1) * nf * sizeof(double));
    double *TSf = malloc(nf * sizeof(double));
    double *MC = calloc(3 * Ns * nf, sizeof(double));
    double *EF = calloc(3 * nf, sizeof(double));
    double cc[OR_MAXS + 1];
    /* turbulence (Diffusion.cpp:129-135, 137-183, 242-260, 414-417, 580-598,
     * 711-933, 999-1028): face k and epsilon, the cells' Cv, CvM = Cv / (rho
     * / c0) and nonIdeality / rho, the surface Cv, and the face fluxes of k,
     * epsilon and of the non-ideality correction Laplacian */
    const int turb = cfg->turb;
    const long tn = turb ? nc : 1, tf = turb ? nf : 1;
    double *Cvc = malloc(tn * sizeof(double));
    double *CvM = malloc(tn * sizeof(double));
    double *NIR = malloc(tn * sizeof(double));
    double *ksf = malloc(tf * sizeof(double));
    double *esf = malloc(tf * sizeof(double));
    double *Cvs = malloc(tf * sizeof(double));
    double *FKt = calloc(3 * tf, sizeof(double));
    double *FEt = calloc(3 * tf, sizeof(double));
    double *FNI = calloc(3 * tf, sizeof(double));
    /* kEpsA (Diffusion.cpp:176-191, 288-305, 511-538, 574-575, 736-800,
     * 869-887; kEpsAModel.cpp:132-257): the cells' b, grad a, grad p and
     * grad rho, the divergence of the molecular viscous stress at the faces,
     * and the face fluxes of the a Laplacian */
    const int afield = a_field(cfg), bfield = b_field(cfg);
    const long an = afield ? nc : 1, af = afield ? nf : 1;
    /* BHR: the face fluxes of the b Laplacian and a_f a_f of div(a a)
     * (BHRModel.cpp:185-186) */
    double *FB = calloc(3 * af, sizeof(double));
    double *AAf = calloc(9 * af, sizeof(double));
    double *bC = malloc(an * sizeof(double));
    double *Ga = malloc(9 * an * sizeof(double));
    double *GP = malloc(3 * an * sizeof(double));
    double *GR = malloc(3 * an * sizeof(double));
    double *rhof = malloc(af * sizeof(double));
    double *FA = calloc(9 * af, sizeof(double));
    double *FDP = calloc(9 * af, sizeof(double));

    /* mass fractions rho_k / rho, normalised (diffusiveFields.cpp:30-47) */
    for (long i = 0; i < nc; ++i) {
        double sum = 0.0;
        for (int k = 0; k < Ns; ++k) {
            Y[i * Ns + k] = st->rho[k + 1][i] / st->rho[0][i];
            sum += Y[i * Ns + k];
        }
        for (int k = 0; k < Ns; ++k)
            Y[i * Ns + k] /= sum;
        for (int k = 0; k <= Ns; ++k)
            cc[k] = st->c[k][i];
        CCVold[i] = st->c[0][i] * cv_mix(cfg, cc); /* Diffusion.cpp:148 */
        if (turb) { /* Diffusion.cpp:137-148, 168 */
            Cvc[i] = cv_mix(cfg, cc);
            const double avM = st->rho[0][i] / st->c[0][i];
            CvM[i] = Cvc[i] / avM;
            NIR[i] = non_ideality(cfg, cc) / st->rho[0][i];
        }
    }
    /* face velocity (linearInterpolate) for grad and divergence of v, and
     * the face density sum_k c_k,f M_k (Diffusion.cpp:156-166) */
    for (long f = 0; f < nf; ++f) {
        const long o = s->own[f];
        double cs[OR_MAXS + 1];
        if (s->fpatch[f] < 0) {
            const long n = s->nei[f];
            for (int a = 0; a < 3; ++a)
                vf[3 * f + a] = st->v[a][o] * s->wo[f] + st->v[a][n] * s->wn[f];
            for (int k = 1; k <= Ns; ++k)
                cs[k] = st->c[k][o] * s->wo[f] + st->c[k][n] * s->wn[f];
        } else {
            const int p = s->fpatch[f];
            double vo[3] = {st->v[0][o], st->v[1][o], st->v[2][o]};
            if (bc_surf_vec(&cfg->bc_v[p], vo, &s->fN[3 * f], &vf[3 * f]))
                err = ERR_BC;
            for (int k = 1; k <= Ns; ++k)
                if (bc_surf_scalar(&cfg->bc_c[k - 1][p], st->c[k][o], &cs[k]))
                    err = ERR_BC;
        }
        srho[f] = 0;
        for (int k = 1; k <= Ns; ++k)
            srho[f] += cs[k] * cfg->M[k - 1];
        cs[0] = 0;
        for (int k = 1; k <= Ns; ++k)
            cs[0] += cs[k];
        for (int k = 0; k <= Ns; ++k)
            CSf[k * nf + f] = cs[k];
        /* surface temperature (linearInterpolate; calculatedTemperature on
         * boundary faces from the surface concentrations and pressure) */
        if (turb) {
            /* linearInterpolate of k, epsilon and Cv (boundary: the surface
             * values; calculatedCv from the surface concentrations,
             * boundaryConditionValue.hpp:210-227) */
            if (s->fpatch[f] < 0) {
                const long n = s->nei[f];
                ksf[f] = st->k[o] * s->wo[f] + st->k[n] * s->wn[f];
                esf[f] = st->eps[o] * s->wo[f] + st->eps[n] * s->wn[f];
                Cvs[f] = Cvc[o] * s->wo[f] + Cvc[n] * s->wn[f];
            } else {
                const int p = s->fpatch[f];
                if (bc_surf_scalar(&cfg->bc_k[p], st->k[o], &ksf[f]) ||
                    bc_surf_scalar(&cfg->bc_eps[p], st->eps[o], &esf[f]))
                    err = ERR_BC;
                Cvs[f] = cv_mix(cfg, cs);
            }
        }
        if (s->fpatch[f] < 0) {
            const long n = s->nei[f];
            TSf[f] = st->T[o] * s->wo[f] + st->T[n] * s->wn[f];
        } else {
            double ps = 0;
            if (bc_surf_scalar(&cfg->bc_p[s->fpatch[f]], st->p[o], &ps))
                err = ERR_BC;
            TSf[f] = eos_T(cfg, cs, eos_U(cfg, cs, ps));
        }
    }
    if (err)
        goto done;
    /* #Reconstruction star: grad v and div v from the step's star velocity
     * (Diffusion.cpp:187-193); after a MUSCL-Hancock step with gravity the
     * reference replaces it by the interpolated velocity
     * (secondOrderStepSolver.cpp:74-76), which vf already holds */
    const int star_v = cfg->star_rec && !(s->hancock_step && cfg->gravity);
    if (star_v)
        for (long f = 0; f < nf; ++f)
            for (int a = 0; a < 3; ++a)
                vf[3 * f + a] = s->star[NSTAR * f + 1 + a];
    if (afield) {
        /* kEpsA: b = rho sum(x_k / (rho_k + chi rho)) / sum(x_k rho_k /
         * (rho_k + chi rho)) - 1, clipped at 0 (kEpsAModel.cpp:232-257).  The
         * source adds the stabilizer to rho_k as field + scalar, which
         * returns the field unchanged (fieldOperations.hpp:33-43), so it does
         * not enter.  BHR: the transported b. */
        const double chi = cfg->tC[7];
        for (long i = 0; i < nc && bfield; ++i)
            bC[i] = st->b[i];
        for (long i = 0; i < nc && !bfield; ++i) {
            double num = 0, den = 0;
            for (int k = 1; k <= Ns; ++k) {
                const double x = st->c[k][i] / st->c[0][i];
                const double q = st->rho[k][i] + chi * st->rho[0][i];
                num += x / q;
                den += x * st->rho[k][i] / q;
            }
            const double b = st->rho[0][i] * num / den - 1.;
            bC[i] = (b < 0.) ? 0.0 : b;
        }
        /* grad a, grad p (Diffusion.cpp:183-184, 531) and grad rho with the
         * face density sum_k c_k,f M_k on boundary faces (calculatedDensity);
         * with star reconstruction the star states' (Diffusion.cpp:189-191,
         * 536) */
        if (cfg->star_rec) {
            grad_vector_faces(s, &s->star[5], NSTAR, Ga);
            grad_faces(s, &s->star[4], NSTAR, GP);
            grad_faces(s, &s->star[0], NSTAR, GR);
        } else {
            err = grad_vector(s, st->a, cfg->bc_a, Ga);
            if (!err)
                err = grad_scalar(s, st->p, cfg->bc_p, GP);
            if (err)
                goto done;
            for (long f = 0; f < nf; ++f) {
                if (s->fpatch[f] == -3)
                    continue;
                rhof[f] = (s->fpatch[f] < 0)
                              ? st->rho[0][s->own[f]] * s->wo[f] + st->rho[0][s->nei[f]] * s->wn[f]
                              : srho[f];
            }
            grad_faces(s, rhof, 1, GR);
        }
    }
    /* grad v and div v of the cells (gradient.hpp:19-58,
     * divergence.hpp:19-58) */
    for (long i = 0; i < nc; ++i) {
        double acc[9] = {0}, dacc = 0;
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            double n[3];
            out_normal(s, f, i, n);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    acc[3 * a + b] += (vf[3 * f + a] * n[b]) * s->fS[f];
            dacc += dot3(&vf[3 * f], n) * s->fS[f];
        }
        for (int q = 0; q < 9; ++q)
            gradV[9 * i + q] = acc[q] / s->cV[i];
        divV[i] = dacc / s->cV[i];
    }
    /* face fluxes */
    for (long f = 0; f < nf; ++f) {
        const long o = s->own[f];
        const int bnd = s->fpatch[f] >= 0;
        const long n = bnd ? o : s->nei[f];
        double qo[OR_MAXS + 3], qn[OR_MAXS + 3], g[3 * (OR_MAXS + 3)];
        /* mass fractions: ghost rho_k / rho from the ghost concentrations
         * (calculatedMassFraction, boundaryConditionValue.hpp:356-374) */
        for (int k = 0; k < Ns; ++k) {
            qo[k] = Y[o * Ns + k];
            qn[k] = Y[n * Ns + k];
        }
        double cg[OR_MAXS + 1];
        if (bnd) {
            if (ghost_conc(s, f, o, cg)) {
                err = ERR_BC;
                goto done;
            }
            double sumd = 0;
            for (int k = 1; k <= Ns; ++k)
                sumd += cg[k] * cfg->M[k - 1];
            for (int k = 0; k < Ns; ++k)
                qn[k] = (cg[k + 1] * cfg->M[k]) / sumd;
            for (int k = 0; k < Ns; ++k)
                GH[k * nf + f] = qn[k];
        }
        /* effective coefficients (transportCoefficients.hpp:65-83, 221-263):
         * turbulent nu = Cmu k^2 / eps at the face (kEpsModels.cpp:64-74) */
        double mu_f = cfg->mu, kappa_f = cfg->kappa, tD = 0, tnu = 0, tLam = 0;
        if (turb) {
            tnu = cfg->Cmu * ksf[f] * ksf[f] / esf[f];
            tD = tnu / cfg->sigmaSc;
            const double tKap = tnu / cfg->sigmaT;
            tLam = tnu / cfg->sigmaE;
            mu_f = cfg->mu + srho[f] * tnu;
            const double cC = CSf[f] * Cvs[f];
            kappa_f = cfg->kappa + cC * tLam + cC * tKap;
        }
        surf_grad(s, f, Ns, qo, qn, g);
        for (int k = 0; k < Ns; ++k) {
            const double rhoD = srho[f] * (cfg->D + tD); /* transportCoefficients.hpp:246-261 */
            for (int b = 0; b < 3; ++b)
                FY[(k * nf + f) * 3 + b] = (g[3 * k + b] * rhoD) * -1;
        }
        /* temperature: ghost T from the ghost concentrations and pressure
         * (calculatedTemperature, boundaryConditionValue.hpp:329-354) */
        qo[0] = st->T[o];
        qn[0] = st->T[n];
        if (bnd) {
            double pg;
            if (bc_cell_scalar(&cfg->bc_p[s->fpatch[f]], st->p[o], &pg)) {
                err = ERR_BC;
                goto done;
            }
            qn[0] = eos_T(cfg, cg, eos_U(cfg, cg, pg));
            GH[Ns * nf + f] = qn[0];
        }
        surf_grad(s, f, 1, qo, qn, g);
        for (int b = 0; b < 3; ++b)
            FT[3 * f + b] = (g[b] * kappa_f) * -1;
        /* velocity: surfGrad (tensor v_a d_b) and surfDivergence
         * (divergence.hpp:84-150) */
        for (int a = 0; a < 3; ++a) {
            qo[a] = st->v[a][o];
            qn[a] = st->v[a][n];
        }
        if (bnd) {
            double vo[3] = {qo[0], qo[1], qo[2]};
            if (bc_cell_vec(&cfg->bc_v[s->fpatch[f]], vo, &s->fN[3 * f], qn)) {
                err = ERR_BC;
                goto done;
            }
            for (int a = 0; a < 3; ++a)
                GH[(Ns + 1 + a) * nf + f] = qn[a];
        }
        surf_grad(s, f, 3, qo, qn, g);
        double divs;
        {
            double dq[3], dd[3];
            if (!bnd) {
                double d[3];
                face_delta(s, f, o, d);
                const double m = or_hypot3(d[0], d[1], d[2]);
                for (int a = 0; a < 3; ++a)
                    dq[a] = (qn[a] - qo[a]) / (m * m);
                memcpy(dd, d, sizeof dd);
            } else {
                double D[3];
                for (int b = 0; b < 3; ++b)
                    D[b] = (s->frc[3 * f + b] - s->crc[3 * o + b]) * 2.;
                const double m = or_hypot3(D[0], D[1], D[2]);
                for (int a = 0; a < 3; ++a) {
                    dq[a] = (qn[a] - qo[a]) / m;
                    dd[a] = D[a] / m;
                }
            }
            divs = dot3(dq, dd);
        }
        for (int q = 0; q < 9; ++q)
            FV[9 * f + q] = (g[q] * mu_f) * -1;
        /* strainSurf1 = (surfGrad v)^T - 2/3 div_s v I, times mu
         * (Diffusion.cpp:193-230, 482-483) */
        const double lam = (2. / 3.) * divs;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double t = g[3 * b + a];
                if (a == b)
                    t -= lam;
                FS[9 * f + 3 * a + b] = t * mu_f;
            }
        if (turb) {
            /* k and epsilon Laplacians with thetaS_D of the face divergence
             * (Diffusion.cpp:247-255, 270-286; SLEMatrix.cpp:2054-2078) */
            const double th = (divs < 0.) ? theta_s(cfg, divs, ksf[f], esf[f], cfg->CMS_D) : 1.;
            const double kD = tnu / cfg->sigmaK * th, eD = tnu / cfg->sigmaEps * th;
            const double rhoDk = cfg->mu + srho[f] * kD, rhoDe = cfg->mu + srho[f] * eD;
            double to[2] = {st->k[o], st->eps[o]}, tg[2] = {st->k[n], st->eps[n]}, gt[6];
            if (bnd) {
                const int p = s->fpatch[f];
                if (bc_cell_scalar(&cfg->bc_k[p], to[0], &tg[0]) ||
                    bc_cell_scalar(&cfg->bc_eps[p], to[1], &tg[1])) {
                    err = ERR_BC;
                    goto done;
                }
            }
            surf_grad(s, f, 2, to, tg, gt);
            for (int b = 0; b < 3; ++b) {
                FKt[3 * f + b] = (gt[b] * rhoDk) * -1;
                FEt[3 * f + b] = (gt[3 + b] * rhoDe) * -1;
            }
            /* non-ideality correction Laplacian -(surfGrad(CvM) T_f +
             * surfGrad(nonIdeality / rho)) rho_f tLambda (Diffusion.cpp:
             * 170-172, 414-417); boundary ghosts calculatedCvM and
             * calculatedNonidealityCorrectionPerDensity from the ghost
             * concentrations (boundaryConditionValue.hpp:415-484) */
            double mo[2] = {CvM[o], NIR[o]}, mg[2] = {CvM[n], NIR[n]}, gm[6];
            if (bnd) {
                double dg = 0;
                for (int k = 1; k <= Ns; ++k)
                    dg += cg[k] * cfg->M[k - 1];
                mg[0] = cv_mix(cfg, cg) * cg[0] / dg;
                mg[1] = non_ideality(cfg, cg) / dg;
            }
            surf_grad(s, f, 2, mo, mg, gm);
            for (int b = 0; b < 3; ++b) {
                const double flow = -1 * (gm[b] * TSf[f] + gm[3 + b]);
                FNI[3 * f + b] = flow * srho[f] * tLam;
            }
        }
        if (afield) {
            /* molecular viscous stress (surfGrad v + (surfGrad v)^T - 2/3
             * div_s v I) mu (Diffusion.cpp:210-222, 574) */
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double t = g[3 * a + b] + g[3 * b + a];
                    if (a == b)
                        t -= lam;
                    FDP[9 * f + 3 * a + b] = t * cfg->mu;
                }
            /* a Laplacian with mu + rho_f nu_t / sigma_a
             * (transportCoefficients.hpp:81, 242; SLEMatrix.cpp:2080-2110) */
            const double rhoDa = cfg->mu + srho[f] * (tnu / cfg->sigmaa);
            double ao[3] = {st->a[0][o], st->a[1][o], st->a[2][o]};
            double ag[3] = {st->a[0][n], st->a[1][n], st->a[2][n]}, ga[9];
            if (bnd && bc_cell_vec(&cfg->bc_a[s->fpatch[f]], ao, &s->fN[3 * f], ag)) {
                err = ERR_BC;
                goto done;
            }
            surf_grad(s, f, 3, ao, ag, ga);
            for (int q = 0; q < 9; ++q)
                FA[9 * f + q] = (ga[q] * rhoDa) * -1;
            if (bfield) {
                /* b Laplacian with mu + rho_f nu_t / sigma_b
                 * (transportCoefficients.hpp:82, 244) */
                const double rhoDb = cfg->mu + srho[f] * (tnu / cfg->sigmab);
                double bo = st->b[o], bg = st->b[n], gb[3];
                if (bnd && bc_cell_scalar(&cfg->bc_b[s->fpatch[f]], bo, &bg)) {
                    err = ERR_BC;
                    goto done;
                }
                surf_grad(s, f, 1, &bo, &bg, gb);
                for (int q = 0; q < 3; ++q)
                    FB[3 * f + q] = (gb[q] * rhoDb) * -1;
                /* a_f a_f of the interpolated a (BHRModel.cpp:185-186) */
                double afv[3];
                if (!bnd) {
                    for (int c = 0; c < 3; ++c)
                        afv[c] = st->a[c][o] * s->wo[f] + st->a[c][n] * s->wn[f];
                } else if (bc_surf_vec(&cfg->bc_a[s->fpatch[f]], ao, &s->fN[3 * f], afv)) {
                    err = ERR_BC;
                    goto done;
                }
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        AAf[9 * f + 3 * r + c] = afv[r] * afv[c];
            }
        }
    }
    if (cfg->molmass || cfg->enth_flow) {
        err = species_corrections(s, FY, CSf, TSf, MC, EF);
        if (err)
            goto done;
    }
    if (cfg->sle_solver != OR_SLE_EXPLICIT) {
        err = diffusion_implicit(s, dt, Y, CCVold, gradV, divV, srho, FS, GH, MC);
        goto done;
    }
    /* cell updates (SLEMatrix.cpp:2054-2142 explicit Laplacians, then
     * Diffusion.cpp:431-458, 462-483, 599-615, 711-760, 795-806, 818-828,
     * 963-1059) */
    double dt_src = 0;
    /* the turbulence model's source step: epsilon below 1e-3 of its maximum
     * is replaced by veryBig (shearModel.cpp:107-114) */
    double epsmax = 0, tsrc = VERYBIG;
    if (turb)
        for (long i = 0; i < nc; ++i)
            epsmax = (i == 0 || epsmax < st->eps[i]) ? st->eps[i] : epsmax;
    for (long i = 0; i < nc; ++i) {
        double dY[OR_MAXS] = {0}, dT = 0, dV[3] = {0}, dS[3] = {0}, dA[3] = {0}, dP[3] = {0};
        double dB = 0, dAA[3] = {0, 0, 0};
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            double n[3];
            out_normal(s, f, i, n);
            const double S = s->fS[f];
            for (int k = 0; k < Ns; ++k)
                dY[k] += dot3(&FY[(k * nf + f) * 3], n) * S;
            dT += dot3(&FT[3 * f], n) * S;
            for (int a = 0; a < 3; ++a) {
                /* tensor & vector (tensor.cpp:163-176) */
                const double *Tv = &FV[9 * f + 3 * a], *Ts = &FS[9 * f + 3 * a];
                dV[a] += (Tv[0] * n[0] + Tv[1] * n[1] + Tv[2] * n[2]) * S;
                dS[a] += (Ts[0] * n[0] + Ts[1] * n[1] + Ts[2] * n[2]) * S;
            }
            if (afield)
                for (int a = 0; a < 3; ++a) {
                    const double *Ta = &FA[9 * f + 3 * a], *Tp = &FDP[9 * f + 3 * a];
                    dA[a] += (Ta[0] * n[0] + Ta[1] * n[1] + Ta[2] * n[2]) * S;
                    dP[a] += (Tp[0] * n[0] + Tp[1] * n[1] + Tp[2] * n[2]) * S;
                }
            if (bfield) {
                dB += dot3(&FB[3 * f], n) * S;
                for (int a = 0; a < 3; ++a) {
                    const double *T = &AAf[9 * f + 3 * a];
                    dAA[a] += (T[0] * n[0] + T[1] * n[1] + T[2] * n[2]) * S;
                }
            }
        }
        const double V = s->cV[i], rho = st->rho[0][i];
        /* divergences of the molar-mass corrections and of the enthalpy
         * flow, distributed into the free term (>= 0) or the diagonal
         * (SLEMatrix.cpp:2166-2175) */
        double dMC[OR_MAXS] = {0}, dEF = 0;
        for (int j = 0; j < s->cnf[i]; ++j) {
            const long f = s->cf[6 * i + j];
            double n[3];
            out_normal(s, f, i, n);
            for (int k = 0; k < Ns; ++k)
                dMC[k] += dot3(&MC[(k * nf + f) * 3], n) * s->fS[f];
            dEF += dot3(&EF[3 * f], n) * s->fS[f];
        }
        /* mass fractions: (Y rho + (0 - div F) dt) / rho, clipped at 0 and
         * renormalised */
        double Yn[OR_MAXS], sum = 0.0;
        for (int k = 0; k < Ns; ++k) {
            double free_term = 0. - dY[k] / V;
            double diag = rho;
            if (cfg->molmass) {
                const double src = (dMC[k] / V) * (-1.0);
                if (src >= 0.)
                    free_term += src;
                else
                    diag -= src / (Y[i * Ns + k] + STAB);
            }
            Yn[k] = (Y[i * Ns + k] * rho + free_term * dt) / diag;
            if (Yn[k] < 0.)
                Yn[k] = 0.0;
            sum += Yn[k];
        }
        double cn[OR_MAXS + 1];
        cn[0] = 0;
        for (int k = 0; k < Ns; ++k) {
            Yn[k] /= sum;
            cn[k + 1] = Yn[k] * rho / cfg->M[k];
            cn[0] += cn[k + 1];
        }
        const double CCVnew = cv_mix(cfg, cn) * cn[0];
        /* velocity: (rho v + ((0 - div F_v) + div(mu strainSurf1)) dt) / rho */
        double vn[3];
        for (int a = 0; a < 3; ++a) {
            double free_term = 0. - dV[a] / V;
            free_term += dS[a] / V;
            vn[a] = (st->v[a][i] * rho + free_term * dt) / rho;
        }
        /* temperature: (CCV_old T + ((0 - div F_T) + (strainVol mu) && gradV)
         * dt) / CCV_new */
        const double *G = &gradV[9 * i];
        double sv[9];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                sv[3 * a + b] = G[3 * b + a] + G[3 * a + b];
        const double lam = (2. / 3.) * divV[i];
        sv[0] -= lam;
        sv[4] -= lam;
        sv[8] -= lam;
        double heat = 0;
        for (int q = 0; q < 9; ++q)
            heat += (sv[q] * cfg->mu) * G[q];
        double free_t = 0. - dT / V;
        double diag_t = CCVnew;
        double dK = 0, dE = 0;
        if (turb) {
            double dNI = 0;
            for (int j = 0; j < s->cnf[i]; ++j) {
                const long f = s->cf[6 * i + j];
                double n[3];
                out_normal(s, f, i, n);
                dNI += dot3(&FNI[3 * f], n) * s->fS[f];
                dK += dot3(&FKt[3 * f], n) * s->fS[f];
                dE += dot3(&FEt[3 * f], n) * s->fS[f];
            }
            /* Diffusion.cpp:604-611 (distributeSourceTerm) */
            const double src = (dNI / V) * (-1.0);
            if (src >= 0.)
                free_t += src;
            else
                diag_t -= src / (st->T[i] + STAB);
        }
        if (cfg->enth_flow) {
            /* source time step coeff * min(CCV_old T_old / |div(flow) +
             * stab|) (Diffusion.cpp:696-700), reduced over the cells here */
            const double r = (CCVold[i] * st->T[i]) / fabs(dEF / V + STAB);
            dt_src = (i == 0 || r < dt_src) ? r : dt_src;
            const double src = (dEF / V) * -1;
            if (src >= 0.)
                free_t += src;
            else
                diag_t -= src / (st->T[i] + STAB);
        }
        free_t += heat;
        double kn = 0, en = 0;
        if (turb) {
            const double k = st->k[i], eps = st->eps[i];
            /* rho epsilon (kEpsModels.cpp:76-82) and the filtered part of the
             * turbulent stress (Diffusion.cpp:586-598, 719-733) */
            /* BHR: rho eps (1 + CMS_M 2 k / c^2) (BHRModel.cpp:262-272) */
            double cc_i[OR_MAXS + 1], a_s2 = 0;
            if (bfield) {
                for (int q = 0; q <= Ns; ++q)
                    cc_i[q] = st->c[q][i];
                a_s2 = eos_c2(cfg, cc_i, rho, st->U[i], st->p[i]);
            }
            if (bfield)
                free_t += rho * eps * (1 + cfg->tC[11] * 2 * k / a_s2);
            else if (cfg->turb_model != OR_TURB_ZERO)
                free_t += rho * eps;
            const double tnu = cfg->Cmu * k * k / eps;
            const double thR = theta_s(cfg, divV[i], k, eps, cfg->CMS_R);
            const double tnu1 = (1. - thR) * tnu;
            double philt = 0, gsph = 0, gdev = 0, sph[9], dev[9];
            for (int q = 0; q < 9; ++q)
                philt += (sv[q] * rho * tnu1) * G[q];
            free_t += philt;
            /* spherical and deviatoric Reynolds stresses and their
             * generation terms (Diffusion.cpp:592-597; shearModel.cpp:116-147) */
            for (int q = 0; q < 9; ++q) {
                const double Iq = (q == 0 || q == 4 || q == 8) ? 1. : 0.;
                sph[q] = Iq * -(2. / 3.) * st->rk[i];
                dev[q] = sv[q] * rho * tnu * thR;
                gsph += sph[q] * G[q];
                gdev += dev[q] * G[q];
            }
            double sk = 0, se = 0;
            if (cfg->turb_model == OR_TURB_SHEAR) {
                const double ek = eps / k;
                const double dissip = -st->reps[i];
                const double Sk1 = gsph + gdev, Sk2 = dissip / st->k[i];
                const double Se1 = (cfg->tC[0] * gdev + cfg->tC[2] * gsph) * ek;
                const double Se2 = cfg->tC[1] * dissip / st->k[i];
                sk = Sk1 + Sk2 * k;
                se = Se1 + Se2 * eps;
                const double me = (eps < 1E-3 * epsmax) ? VERYBIG : eps;
                const double r = fabs(cfg->ts_coeff * me / ((Se1 + Se2 * st->eps[i]) / rho + STAB));
                tsrc = (r < tsrc) ? r : tsrc;
            } else if (bfield) {
                /* BHRModel.cpp:188-252 (generation flag 1: the Goncharov
                 * initialisation model off); tC: C0, C1, C2, C3, C4, Ca, Cb1,
                 * Cb2, alpha2, alpha3, alpha4, CMS_M */
                const double *C = cfg->tC;
                const double *ga = &Ga[9 * i], *gp = &GP[3 * i], *gr = &GR[3 * i];
                const double ai[3] = {st->a[0][i], st->a[1][i], st->a[2][i]};
                const double ek = eps / k;
                const double Mat2 = 2 * k / a_s2;
                const double Mat = sqrt(Mat2);
                const double thS = theta_s(cfg, G[0] + G[4] + G[8], k, eps, cfg->CMS_R);
                double gdevA = 0, x[3], grav = 0;
                for (int q = 0; q < 9; ++q)
                    gdevA += (dev[q] * thS) * G[q];
                for (int c = 0; c < 3; ++c)
                    x[c] = gp[c] - dP[c] / V;
                for (int c = 0; c < 3; ++c)
                    grav += ai[c] * x[c];
                grav = grav * 1.0;
                const double PiK =
                    (Mat * C[9] * st->reps[i] - (C[8] * gdevA + 8 * Mat * C[10] * gsph)) * Mat;
                const double dissip = -st->reps[i];
                const double Sk1 = gsph + gdevA + grav + PiK;
                const double Sk2 = (1 + C[11] * Mat2) * dissip / st->k[i];
                const double gpos = (grav < 0.0) ? 0.0 : grav;
                const double Se1 = (C[1] * gdevA + C[3] * gsph + C[0] * gpos + C[4] * PiK) * ek;
                const double Se2 = C[2] * dissip / st->k[i];
                sk = Sk1 + Sk2 * k;
                se = Se1 + Se2 * eps;
                const double me = (eps < 1E-3 * epsmax) ? VERYBIG : eps;
                const double r = fabs(cfg->ts_coeff * me / ((Se1 + Se2 * st->eps[i]) / rho + STAB));
                const double rU = fabs(cfg->ts_coeff * fabs(st->U[i]) / (grav + STAB));
                tsrc = (r < tsrc) ? r : tsrc;
                tsrc = (rU < tsrc) ? rU : tsrc;
                /* a: (grad p - div tau) b + ((thetaS devR + spherR) / rho) &
                 * grad rho + rho a & (grad a - grad v) + rho div(a a) -
                 * rho (eps / k) Ca a */
                double tau[9], D[9];
                for (int q = 0; q < 9; ++q) {
                    tau[q] = 1.0 * (dev[q] * thS + sph[q]) / rho;
                    D[q] = ga[q] - G[q];
                }
                const double Sa2 = -rho * ek * C[5];
                double an[3];
                for (int c = 0; c < 3; ++c) {
                    const double bgp = x[c] * bC[i];
                    const double tgr = tau[3 * c] * gr[0] + tau[3 * c + 1] * gr[1] + tau[3 * c + 2] * gr[2];
                    const double ragv = D[c] * st->ra[0][i] + D[3 + c] * st->ra[1][i] +
                                        D[6 + c] * st->ra[2][i];
                    const double redis = (dAA[c] / V) * rho;
                    const double sig = bgp + tgr + ragv + redis + Sa2 * ai[c];
                    an[c] = (ai[c] * rho + (0. - dA[c] / V) * dt) / rho + sig / rho * dt;
                }
                /* b: the explicit Laplacian; the explicit solver leaves b's
                 * sources out (sigmaSourceb stays zero, Diffusion.cpp:745-763,
                 * 889-899), minimal value (917-923) */
                double bn = (bC[i] * rho + (0. - dB / V) * dt) / rho + 0. / rho * dt;
                if (bn < cfg->minb)
                    bn = cfg->minb;
                const double src = grav * (-1.0);
                if (src >= 0.)
                    free_t += src;
                else
                    diag_t -= src / (st->T[i] + STAB);
                for (int c = 0; c < 3; ++c) {
                    st->a[c][i] = an[c];
                    st->ra[c][i] = an[c] * rho;
                }
                st->b[i] = bn;
                st->rb[i] = bn * rho;
            } else if (afield) {
                /* kEpsAModel.cpp:170-225 with C0..C3, Ca, CaMax, CMSA in tC */
                const double *ga = &Ga[9 * i], *gp = &GP[3 * i], *gr = &GR[3 * i];
                const double ai[3] = {st->a[0][i], st->a[1][i], st->a[2][i]};
                const double ek = eps / k;
                /* thetaS_R of trace(grad v) (tensor.cpp:41-44) */
                const double thS = theta_s(cfg, G[0] + G[4] + G[8], k, eps, cfg->CMS_R);
                /* thetaA (kEpsAModel.cpp:16-27) */
                double thA = 1;
                if (cfg->tC[6] < 1E15) {
                    double aa = 0;
                    for (int c = 0; c < 3; ++c)
                        aa += ai[c] * ai[c];
                    thA = (aa / (2 * k) < cfg->tC[6] && bC[i] < cfg->tC[6] && aa / (2 * k) < bC[i])
                              ? 1
                              : cfg->tC[5] / cfg->tC[4];
                }
                double gdevA = 0, x[3], grav = 0;
                for (int q = 0; q < 9; ++q)
                    gdevA += (dev[q] * thS) * G[q];
                /* a & (grad p - div(mu strainSurf)) */
                for (int c = 0; c < 3; ++c)
                    x[c] = gp[c] - dP[c] / V;
                for (int c = 0; c < 3; ++c)
                    grav += ai[c] * x[c];
                const double dissip = -st->reps[i];
                const double Sk1 = gsph + gdevA + grav, Sk2 = dissip / st->k[i];
                const double Se1 =
                    (cfg->tC[1] * gdevA + cfg->tC[3] * gsph + cfg->tC[0] * grav) * ek;
                const double Se2 = cfg->tC[2] * dissip / st->k[i];
                sk = Sk1 + Sk2 * k;
                se = Se1 + Se2 * eps;
                const double me = (eps < 1E-3 * epsmax) ? VERYBIG : eps;
                const double r = fabs(cfg->ts_coeff * me / ((Se1 + Se2 * st->eps[i]) / rho + STAB));
                tsrc = (r < tsrc) ? r : tsrc;
                /* Sa1 = b (grad p - div(...)) + ((thetaS devR + spherR) / rho) &
                 * grad rho + rho a & (grad a - grad v); Sa2 = -rho (eps/k) Ca
                 * thetaA on every component (kEpsAModel.cpp:208-222), summed
                 * as Sa1 + Sa2 a (Diffusion.cpp:749-762) */
                double tau[9], D[9];
                for (int q = 0; q < 9; ++q) {
                    tau[q] = (dev[q] * thS + sph[q]) / rho;
                    D[q] = ga[q] - G[q];
                }
                const double Sa2 = -rho * ek * cfg->tC[4] * thA;
                double an[3];
                for (int c = 0; c < 3; ++c) {
                    const double bgp = x[c] * bC[i];
                    const double tgr = tau[3 * c] * gr[0] + tau[3 * c + 1] * gr[1] + tau[3 * c + 2] * gr[2];
                    const double ragv = D[c] * st->ra[0][i] + D[3 + c] * st->ra[1][i] +
                                        D[6 + c] * st->ra[2][i];
                    const double sig = bgp + tgr + ragv + Sa2 * ai[c];
                    /* explicit Laplacian plus the source (Diffusion.cpp:869-887) */
                    an[c] = (ai[c] * rho + (0. - dA[c] / V) * dt) / rho + sig / rho * dt;
                }
                /* gravEnSink into the temperature (Diffusion.cpp:789-799) */
                const double src = grav * (-1.0);
                if (src >= 0.)
                    free_t += src;
                else
                    diag_t -= src / (st->T[i] + STAB);
                for (int c = 0; c < 3; ++c) {
                    st->a[c][i] = an[c];
                    st->ra[c][i] = an[c] * rho;
                }
                st->b[i] = bC[i];
            }
            /* explicit Laplacians plus the sources (Diffusion.cpp:839-862),
             * minimal values (905-917) */
            kn = (k * rho + (0. - dK / V) * dt) / rho + sk / rho * dt;
            en = (eps * rho + (0. - dE / V) * dt) / rho + se / rho * dt;
            if (kn < cfg->mink)
                kn = cfg->mink;
            if (en < cfg->mineps)
                en = cfg->mineps;
        }
        double co[OR_MAXS + 1];
        for (int k = 0; k <= Ns; ++k)
            co[k] = st->c[k][i];
        /* oldEnergyField (Diffusion.cpp:484-485) */
        const double oldE = CCVold[i] * st->T[i] + non_ideality(cfg, co) - non_ideality(cfg, cn);
        const double Tn = (oldE + free_t * dt) / diag_t;
        /* recalculation of the cell values (Diffusion.cpp:963-1029) */
        for (int k = 0; k <= Ns; ++k)
            st->c[k][i] = cn[k];
        for (int k = 1; k <= Ns; ++k)
            st->rho[k][i] = cn[k] * cfg->M[k - 1];
        for (int a = 0; a < 3; ++a) {
            st->v[a][i] = vn[a];
            st->m[a][i] = vn[a] * rho;
        }
        st->T[i] = Tn;
        st->U[i] = uvc_from_T(cfg, cn, Tn) * cn[0];
        st->p[i] = eos_p(cfg, cn, st->U[i]);
        if (turb) { /* Diffusion.cpp:999-1011 */
            st->k[i] = kn;
            st->eps[i] = en;
            st->rk[i] = kn * rho;
            st->reps[i] = en * rho;
        }
        const double v2 = dot3(vn, vn);
        st->E[i] = (st->U[i] + rho * v2 * 0.5) + st->rk[i];
        if (Tn < 0.)
            err = ERR_NEG_T;
    }
    if (cfg->enth_flow)
        s->dt_src = cfg->ts_coeff * dt_src;
    if (turb) /* shearModel.cpp:150, zeroModel.cpp:90 */
        s->dt_src = (tsrc < s->dt_src) ? tsrc : s->dt_src;
done:
    free(Y);
    free(CCVold);
    free(gradV);
    free(divV);
    free(vf);
    free(srho);
    free(FY);
    free(FT);
    free(FV);
    free(FS);
    free(GH);
    free(CSf);
    free(TSf);
    free(MC);
    free(EF);
    free(Cvc);
    free(CvM);
    free(NIR);
    free(ksf);
    free(esf);
    free(Cvs);
    free(FKt);
    free(FEt);
    free(FNI);
    free(bC);
    free(Ga);
    free(GP);
    free(GR);
    free(rhof);
    free(FA);
    free(FDP);
    free(FB);
    free(AAf);
    return err;
}

/* secondOrderStepSolverRK::calculateStep (secondOrderStepSolverRK.cpp:48-71,
 * diffusion 102-109) */
int or_step(or_solver *s, double dt, int mode)
{
    if (mode == 1)
        recover_primitives(s, &s->st);
    s->hancock_step = 0;
    or_state Un;
    double *pool = state_pool_new(s->nc, s->cfg.Ns, &Un);
    copy_state(&s->st, &Un, s->nc, s->cfg.Ns);
    int err = or_stage(s, dt);
    if (!err) {
        memcpy(s->star, s->star_stage, NSTAR * (size_t)s->nf * sizeof(double));
        err = or_stage(s, dt);
    }
    if (!err) {
        /* star = (star + star1) / 2 (secondOrderStepSolverRK.cpp:73-85) */
        for (long q = 0; q < NSTAR * s->nf; ++q)
            s->star[q] = (s->star[q] + s->star_stage[q]) / 2;
        rk_average(s, &Un);
    }
    if (!err && mode == 1) {
        s->keep_b = 1; /* b stays the second stage's, as in the reference */
        recover_primitives(s, &s->st);
        s->keep_b = 0;
    }
    if (!err && s->cfg.diffusion) {
        err = diffusion_explicit(s, dt);
        if (!err && mode == 1)
            recover_primitives(s, &s->st);
    }
    free(pool);
    return err;
}

/* secondOrderStepSolver::calculateStep (secondOrderStepSolver.cpp:48-93):
 * one Advection with the Hancock predictor, then the diffusion stage when
 * enabled; chemistry off. */
int or_step_hancock(or_solver *s, double dt, int mode)
{
    if (s->cfg.turb)
        return ERR_INIT; /* turbulence is built for the SSP-RK2 step only */
    if (mode == 1)
        recover_primitives(s, &s->st);
    int err = stage_impl(s, dt, 1);
    if (!err)
        memcpy(s->star, s->star_stage, NSTAR * (size_t)s->nf * sizeof(double));
    s->hancock_step = 1;
    if (!err && mode == 1)
        recover_primitives(s, &s->st);
    if (!err && s->cfg.diffusion) {
        /* secondOrderStepSolver.cpp:72-82 */
        err = diffusion_explicit(s, dt);
        if (!err && mode == 1)
            recover_primitives(s, &s->st);
    }
    return err;
}

/* Courant time step (Schemi.cpp:807-833). */
double or_courant_dt(or_solver *s, double courant, double min_time)
{
    or_state *st = &s->st;
    double cc[OR_MAXS + 1];
    double mx = 0;
    for (long i = 0; i < s->nc; ++i) {
        for (int k = 0; k <= s->cfg.Ns; ++k)
            cc[k] = st->c[k][i];
        const double c = sqrt(eos_c2(&s->cfg, cc, st->rho[0][i], st->U[i], st->p[i]));
        const double sig =
            (or_hypot3(st->v[0][i], st->v[1][i], st->v[2][i]) + c) * s->cminlen[i];
        if (i == 0 || mx < sig)
            mx = sig;
    }
    double dt = courant / mx;
    if (s->cfg.ts_source)
        dt = (s->dt_src < dt) ? s->dt_src : dt; /* std::min(dt, dt_source) */
    dt = smax(dt, min_time);
    s->dt_src = VERYBIG;
    return dt;
}

void or_cell_geometry(const or_solver *s, double *rc, double *V)
{
    memcpy(rc, s->crc, 3 * s->nc * sizeof(double));
    memcpy(V, s->cV, s->nc * sizeof(double));
}

void or_face_geometry(const or_solver *s, double *rc, double *N, double *S,
                      long *owner, long *neigh, double *wo, double *wn)
{
    memcpy(rc, s->frc, 3 * s->nf * sizeof(double));
    memcpy(N, s->fN, 3 * s->nf * sizeof(double));
    memcpy(S, s->fS, s->nf * sizeof(double));
    memcpy(owner, s->own, s->nf * sizeof(long));
    memcpy(neigh, s->nei, s->nf * sizeof(long));
    memcpy(wo, s->wo, s->nf * sizeof(double));
    memcpy(wn, s->wn, s->nf * sizeof(double));
}

void or_cell_faces(const or_solver *s, long *faces)
{
    for (long i = 0; i < s->nc; ++i)
        for (int j = 0; j < 6; ++j)
            faces[6 * i + j] = (j < s->cnf[i]) ? s->cf[6 * i + j] : -1;
}
