// fv_tangent.hpp -- tangent numbers for the adjoint kernels.
//
// Tan carries a value and one directional derivative in FP64.  The forward
// device templates (prims_from_cons, close_cell, face_flux_closed, the EOS)
// evaluated with T = Tan return the derivative of their outputs along the
// seeded input direction.  Every branch (comparisons, tmin/tmax, the HLLC
// wave pattern, the |v| scaling of hypot3) is decided on the values, so the
// derivative is that of the forward map with its switches frozen at the
// forward state; the values themselves are computed with the forward's FP64
// operations in the same order.
#pragma once

#include "fv_device.hpp"

namespace sg {

struct Tan {
    double v, d;
    Tan() = default;
    __device__ constexpr Tan(double x) : v(x), d(0.0) {}
    __device__ constexpr Tan(double x, double dx) : v(x), d(dx) {}
};

__device__ __forceinline__ Tan operator+(Tan a, Tan b) { return Tan(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Tan operator-(Tan a, Tan b) { return Tan(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Tan operator-(Tan a) { return Tan(-a.v, -a.d); }
__device__ __forceinline__ Tan operator*(Tan a, Tan b)
{
    return Tan(a.v * b.v, a.d * b.v + a.v * b.d);
}
__device__ __forceinline__ Tan operator/(Tan a, Tan b)
{
    const double q = a.v / b.v;
    return Tan(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Tan operator+(double a, Tan b) { return Tan(a) + b; }
__device__ __forceinline__ Tan operator+(Tan a, double b) { return a + Tan(b); }
__device__ __forceinline__ Tan operator-(double a, Tan b) { return Tan(a) - b; }
__device__ __forceinline__ Tan operator-(Tan a, double b) { return a - Tan(b); }
__device__ __forceinline__ Tan operator*(double a, Tan b) { return Tan(a * b.v, a * b.d); }
__device__ __forceinline__ Tan operator*(Tan a, double b) { return Tan(a.v * b, a.d * b); }
__device__ __forceinline__ Tan operator/(double a, Tan b) { return Tan(a) / b; }
__device__ __forceinline__ Tan operator/(Tan a, double b) { return Tan(a.v / b, a.d / b); }
__device__ __forceinline__ Tan &operator+=(Tan &a, Tan b) { return a = a + b; }
__device__ __forceinline__ Tan &operator-=(Tan &a, Tan b) { return a = a - b; }
__device__ __forceinline__ Tan &operator*=(Tan &a, Tan b) { return a = a * b; }
__device__ __forceinline__ Tan &operator/=(Tan &a, Tan b) { return a = a / b; }

__device__ __forceinline__ bool operator<(Tan a, Tan b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(Tan a, Tan b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(Tan a, Tan b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(Tan a, Tan b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(Tan a, Tan b) { return a.v == b.v; }
__device__ __forceinline__ bool operator!=(Tan a, Tan b) { return a.v != b.v; }

// |x| with the derivative of the branch taken (x >= 0: +1)
__device__ __forceinline__ Tan tabs(Tan x) { return (x.v < 0.0) ? -x : x; }

template <> struct Real<Tan> {
    __device__ static Tan sqrt_(Tan x)
    {
        const double s = __builtin_sqrt(x.v);
        return Tan(s, x.d / (2.0 * s));
    }
};

template <> __device__ __forceinline__ Tan mixM<Tan>(const DevParams &P, int k) { return Tan(P.M[k]); }
template <> __device__ __forceinline__ Tan mixCv<Tan>(const DevParams &P, int k) { return Tan(P.Cv[k]); }
template <> __device__ __forceinline__ Tan mixR<Tan>(const DevParams &P) { return Tan(P.R); }
template <> __device__ __forceinline__ Tan mixA<Tan>(const DevParams &P, int k, int l) { return Tan(P.akl[k][l]); }
template <> __device__ __forceinline__ Tan mixB<Tan>(const DevParams &P, int k, int l) { return Tan(P.bkl[k][l]); }

}  // namespace sg
