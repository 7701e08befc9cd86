// geometry.hpp -- per-axis geometry tables of the structured mesh.
//
// The reference builds every cell and face of its parallelepiped from vertex
// coordinates (mesh.cpp:296-3018) and derives centres, volumes, areas,
// normals and interpolation weights from them (mesh.cpp:34-115).  On a
// Cartesian product mesh all of these factor into one-dimensional tables per
// axis; this builder evaluates the reference's formulas along each axis with
// the same floating-point operations, so the device reproduces the reference's
// geometry bit for bit without storing per-face data.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sg {

struct AxisTables {
    int N = 1;          // cells along the axis (1 for an inactive axis)
    bool active = false;
    bool periodic = false;
    std::vector<double> xc;   // cell centre coordinate            [N]
    std::vector<double> e;    // cell edge length                  [N]
    std::vector<double> xf;   // face coordinate along the axis    [N + 1]
    std::vector<double> wo;   // owner interpolation weight        [N + 1]
    std::vector<double> wn;   // neighbour interpolation weight    [N + 1]
    std::vector<double> dn;   // centre distance across the face   [N + 1]
    std::vector<double> dlo;  // x_f(i) - x_c(i)                   [N]
    std::vector<double> dhi;  // x_f(i + 1) - x_c(i)               [N]
    double ghost_lo = 0, ghost_hi = 0;
};

struct MeshTables {
    int dim = 1;
    AxisTables ax[3];
    bool separable = true;  // face centres share the cells' transverse coordinates
    std::string why;        // first transverse mismatch found
    // Offset of a face centre from its cell's centre across the face normal:
    // td[a][b][cls][t] for faces normal to axis a, along transverse axis b
    // (b != a), cell index t along b; cls 0 = inner faces, 1 = faces on the
    // two boundary planes of axis a.  Exactly zero on separable meshes.
    std::vector<double> td[3][3][2];
};

double hypot3_host(double x, double y, double z);

// Builds the tables of the reference's oneD/twoD/threeDParallelepiped mesh.
// periodic[a] marks axes whose two boundary patches are periodic.
MeshTables build_tables(int dim, const int64_t n[3], const double extent[3],
                        const double origin[3], const bool periodic[3]);

// Reference face id -> structured face id (axis-major: all x faces
// (i, j, k), i fastest, then y faces, then z faces).
std::vector<int64_t> face_permutation(int dim, const int64_t n[3]);

}  // namespace sg
