// fv_device.hpp -- device-side building blocks of the explicit FV step.
//
// Everything here is templated on the flux precision T (double for the FP64
// path, float for the mixed path).  In the FP64 path every expression keeps
// the reference's operation order, and the translation unit is compiled with
// -ffp-contract=off, so results are bit-identical to the reference's CPU step.
//
// Faces of the structured mesh are axis-aligned: the reference's unit normals
// are exactly +-1 along the axis and +-0 across it (mesh.cpp:34-53), so a dot
// product v & N reduces to +-v_axis, and a flux vector f N contracted with
// the cell-side normal s N reduces to s f.  The terms dropped by these
// reductions are exact signed zeros.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sg {

constexpr int MAXS = 9;
constexpr double STAB = 5.0 * 2.220446049250313080847e-16;  // globalConstants.hpp:33

enum : int { BC_FREE = 1, BC_SLIP = 2, BC_FIXED_CELL = 3, BC_FIXED_SURF = 4,
             BC_PERIODIC = 32 };
enum : int { LIM_ZERO = 0, LIM_MINMOD = 1, LIM_VANLEER = 2, LIM_LINEAR = 3,
             LIM_VANALBADA = 4 };
enum : int { FLUX_HLL = 0, FLUX_HLLC = 2 };
// gasModel (gasModelEnum.hpp:15-18): ideal, vanDerWaals, RedlichKwong,
// stiffened, KataokaVanDerWaals
enum : int { GAS_IDEAL = 0, GAS_VDW = 1, GAS_RK = 2, GAS_STIFF = 3, GAS_KATAOKA = 4 };
// stage variants: the two SSP-RK2 stages and the single MUSCL-Hancock stage
enum : int { SV_RK1 = 0, SV_RK2 = 1, SV_HANCOCK = 2 };
// Scalar block of a context (device memory, one double each): time step,
// time, end time, the signal-speed maximum (bits of a non-negative double,
// reduced with atomicMax), next time step, step count, end-of-run flag, and
// the MUSCL-Hancock buffer parity (1: the current state is in the second of
// the two state buffers), and the minimum over the cells of the source-step
// ratio of the last diffusion stage (bits of a non-negative double, reduced
// with atomicMin; +inf when no stage set it), and the source step of the
// last finished step (the reference's mesh::timestepSource(), veryBig when
// no stage set one).
enum : int { SC_DT = 0, SC_TIME, SC_TEND, SC_SIGMAX, SC_DTNEXT, SC_STEPS, SC_DONE, SC_PARITY,
             SC_SRCMIN, SC_SRCDT, SC_N };
constexpr unsigned long long SRCMIN_NONE = 0x7FF0000000000000ull;  // +inf
enum : int { ERRF_RIEMANN = 1, ERRF_BOUNDARY = 2, ERRF_NEG_T = 4 };

// Parameters shared by all kernels of a context, kept in device memory and
// read with uniform (scalar) loads.
struct DevParams {
    int Ns, dim;
    int N[3];
    long long n;
    long long stride[3];  // 1, Nx, Nx*Ny
    int periodic[3];
    int zero_entry[3];    // the limiter ratio along the axis has an exact 0 entry
    // the high boundary face of the axis lies below the last cell centre, so
    // the reference orients its normal to -axis (mesh.cpp:34-53); happens
    // along z in 3-D when Lz < Ly (the z spacing is Ly / Nz, mesh.cpp:1219)
    int hi_flip[3];
    int limiter, flux;
    double M[MAXS], Cv[MAXS], R;
    float Mf[MAXS], Cvf[MAXS], Rf;
    // mixed lean kernel: 1 / M_k and Cv_k / M_k (molar Cv per unit mass)
    float iMf[MAXS], CvMf[MAXS];
    // equation of state; van der Waals mixing matrices a_kl, b_kl
    int gas_model;
    double akl[MAXS][MAXS], bkl[MAXS][MAXS];
    double ekl[MAXS][MAXS];  // Kataoka: eps_kl (V0_kl in akl, b_kl in bkl)
    float aklf[MAXS][MAXS], bklf[MAXS][MAXS];
    int gravity;
    double g[3];
    // boundary conditions per patch (0 tail .. 5 top) and field
    // (0..Ns-1: c_k, Ns: v, Ns+1: p)
    int bct[6][MAXS + 2];
    double bcv[6][MAXS + 2][3];
    double ghost_lo[3], ghost_hi[3];  // |2 (x_f - x_c)| of the boundary ghosts
    // offsets (in doubles) of the per-axis tables inside geo[]
    int off_xc[3], off_e[3], off_xf[3], off_wo[3], off_wn[3], off_dn[3],
        off_dlo[3], off_dhi[3];
    // transverse face-centre offsets: faces normal to a, along b, class
    // (0 inner, 1 boundary plane), indexed by the cell index along b
    int off_td[3][3][2];
    double courant, min_time;
    int ts_source;     // 1: the Courant step is capped by the source step
    double ts_coeff;   // source step coefficient
    // #Reconstruction star: first face of each axis in the star buffer and
    // the number of faces (fv_star.hpp)
    long long star_off[3], star_nft;
};

// patch of axis a, side (0 = low, 1 = high): tail/point, right/left, bottom/top
__host__ __device__ inline int patch_of(int a, int side)
{
    return a == 0 ? side : (a == 1 ? (side ? 4 : 3) : (side ? 5 : 2));
}

template <class T> __device__ __forceinline__ T tmin(T a, T b) { return (b < a) ? b : a; }
template <class T> __device__ __forceinline__ T tmax(T a, T b) { return (a < b) ? b : a; }

template <class T> struct Real;
template <> struct Real<double> {
    __device__ static double sqrt_(double x) { return __builtin_sqrt(x); }
};
template <> struct Real<float> {
    __device__ static float sqrt_(float x) { return __builtin_sqrtf(x); }
};

__device__ __forceinline__ double tabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }

// libstdc++ 11 std::hypot(x, y, z), used by vector::mag (vector.cpp:37-43).
template <class T> __device__ inline T hypot3(T x, T y, T z)
{
    x = tabs(x);
    y = tabs(y);
    z = tabs(z);
    const T a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a != T(0))
        return a * Real<T>::sqrt_((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
    return T(0);
}

template <class T> __device__ __forceinline__ T mixM(const DevParams &P, int k);
template <> __device__ __forceinline__ double mixM<double>(const DevParams &P, int k) { return P.M[k]; }
template <> __device__ __forceinline__ float mixM<float>(const DevParams &P, int k) { return P.Mf[k]; }
template <class T> __device__ __forceinline__ T mixCv(const DevParams &P, int k);
template <> __device__ __forceinline__ double mixCv<double>(const DevParams &P, int k) { return P.Cv[k]; }
template <> __device__ __forceinline__ float mixCv<float>(const DevParams &P, int k) { return P.Cvf[k]; }
template <class T> __device__ __forceinline__ T mixR(const DevParams &P);
template <> __device__ __forceinline__ double mixR<double>(const DevParams &P) { return P.R; }
template <> __device__ __forceinline__ float mixR<float>(const DevParams &P) { return P.Rf; }

template <class T> __device__ __forceinline__ T mixA(const DevParams &P, int k, int l);
template <> __device__ __forceinline__ double mixA<double>(const DevParams &P, int k, int l) { return P.akl[k][l]; }
template <> __device__ __forceinline__ float mixA<float>(const DevParams &P, int k, int l) { return P.aklf[k][l]; }
template <class T> __device__ __forceinline__ T mixB(const DevParams &P, int k, int l);
template <> __device__ __forceinline__ double mixB<double>(const DevParams &P, int k, int l) { return P.bkl[k][l]; }
template <> __device__ __forceinline__ float mixB<float>(const DevParams &P, int k, int l) { return P.bklf[k][l]; }

// R / Cv_mix with Cv_mix = sum_k (c_k / c_0) Cv_k (mixtureIdeal.cpp:383-395,
// mixtureVanDerWaals.cpp:570-591, fractionCalculation.cpp:35-44).
template <class T, int NS>
__device__ __forceinline__ T gamma1(const DevParams &P, const T *c, T c0)
{
    T cv = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
        cv += (c[k] / c0) * mixCv<T>(P, k);
    return mixR<T>(P) / cv;
}

// van der Waals mixture parameters a = sum_k sum_l X_k X_l a_kl and b
// likewise, X_k = c_k / c_0 (mixtureVanDerWaals.cpp:580-587).
template <class T, int NS>
__device__ __forceinline__ void vdw_ab(const DevParams &P, const T *c, T c0, T &a, T &b)
{
    T X[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k)
        X[k] = c[k] / c0;
    a = T(0);
    b = T(0);
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int l = 0; l < NS; ++l) {
            a += X[k] * X[l] * mixA<T>(P, k, l);
            b += X[k] * X[l] * mixB<T>(P, k, l);
        }
}

// Gas model of a kernel: GAS >= 0 fixes it at compile time, -1 reads it from
// the parameter block.
template <int GAS> __device__ __forceinline__ bool is_vdw(const DevParams &P)
{
    if constexpr (GAS >= 0)
        return GAS == GAS_VDW;
    else
        return P.gas_model == GAS_VDW;
}
template <int GAS> __device__ __forceinline__ bool is_stiff(const DevParams &P)
{
    if constexpr (GAS >= 0)
        return GAS == GAS_STIFF;
    else
        return P.gas_model == GAS_STIFF;
}
// Redlich-Kwong code is compiled only where the model can occur (and only
// for plain floating-point T; the adjoint's tangent numbers never meet it).
template <int GAS> constexpr bool may_be_rk() { return GAS < 0 || GAS == GAS_RK; }
template <int GAS> constexpr bool may_be_kataoka() { return GAS < 0 || GAS == GAS_KATAOKA; }
template <int GAS> __device__ __forceinline__ bool is_kataoka(const DevParams &P)
{
    if constexpr (GAS >= 0)
        return GAS == GAS_KATAOKA;
    else
        return P.gas_model == GAS_KATAOKA;
}
template <int GAS> __device__ __forceinline__ bool is_rk(const DevParams &P)
{
    if constexpr (GAS >= 0)
        return GAS == GAS_RK;
    else
        return P.gas_model == GAS_RK;
}

// ---- Redlich-Kwong mixture, evaluated in FP64 on both paths --------------
// The reference solves its cubics with a long-double trigonometric Cardano
// formula (cubicEquationSolver.cpp:17-106) and keeps the single non-negative
// root.  Every cubic here is A s^3 + C s + D = 0 with A > 0 and D < 0, whose
// positive root is unique; it is found by Newton's method in FP64, starting
// from the root without D (C < 0) or without C (C >= 0), after which the
// iterates decrease monotonically (f is convex for s > 0).  The result is
// within a few ulps of the reference's root.
__device__ inline double rk_root(double A, double C, double D)
{
    double s = (C < 0.0) ? __builtin_sqrt(-C / A) : ::cbrt(-D / A);
    for (int it = 0; it < 60; ++it) {
        const double s2 = s * s;
        const double f = (A * s2 + C) * s + D;
        const double ds = f / (3.0 * A * s2 + C);
        s -= ds;
        if (!(__builtin_fabs(ds) > 1e-13 * s))
            break;
    }
    return s;
}

// a_mix, b_mix (the van der Waals mixing rule, mixtureRedlichKwong.cpp:
// 604-625) and Cv_mix = sum_k X_k Cv_k, M_mix = sum_k X_k M_k in FP64.
template <int NS>
__device__ __forceinline__ void rk_mix(const DevParams &P, const double *c, double c0, double &a,
                                       double &b, double &cv, double &mm)
{
    double X[NS];
    cv = 0.0;
    mm = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        X[k] = c[k] / c0;
        cv += X[k] * P.Cv[k];
        mm += X[k] * P.M[k];
    }
    a = 0.0;
    b = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int l = 0; l < NS; ++l) {
            a += X[k] * X[l] * P.akl[k][l];
            b += X[k] * X[l] * P.bkl[k][l];
        }
}

// sqrt(T) from U: root of c Cv s^3 - U s - 1.5 a c / b log1p(c b)
// (RedlichKwongFluid.cpp:46-56, 89-97)
__device__ __forceinline__ double rk_sqrtT(double cv, double U, double c0, double a, double b)
{
    return rk_root(c0 * cv, -U, -1.5 * a * c0 / b * ::log1p(c0 * b));
}

// pFromUv (RedlichKwongFluid.cpp:46-57)
template <int NS>
__device__ inline double rk_p(const DevParams &P, const double *c, double c0, double U)
{
    double a, b, cv, mm;
    rk_mix<NS>(P, c, c0, a, b, cv, mm);
    const double s = rk_sqrtT(cv, U, c0, a, b);
    return c0 * P.R * (s * s) / (1 - c0 * b) - a * (c0 * c0) / (s * (1 + c0 * b));
}

// UvFromp (RedlichKwongFluid.cpp:59-70)
template <int NS>
__device__ inline double rk_U(const DevParams &P, const double *c, double c0, double p)
{
    double a, b, cv, mm;
    rk_mix<NS>(P, c, c0, a, b, cv, mm);
    const double s = rk_root(c0 * P.R, -p * (1 - c0 * b),
                             -a * (c0 * c0) * (1 - c0 * b) / (1 + c0 * b));
    return c0 * cv * (s * s) - 1.5 * a * c0 / (b * s) * ::log1p(c0 * b);
}

// TFromUv (RedlichKwongFluid.cpp:89-97)
template <int NS>
__device__ inline double rk_T(const DevParams &P, const double *c, double c0, double U)
{
    double a, b, cv, mm;
    rk_mix<NS>(P, c, c0, a, b, cv, mm);
    const double s = rk_sqrtT(cv, U, c0, a, b);
    return s * s;
}

// sqSonicSpeed = dpdrho + dpdUv (U + p) / rho with dpdrho = (dP/dC - dP/dT
// dUv/dC / dUv/dT) / M_mix, dpdUv = dP/dT / dUv/dT at T(U)
// (RedlichKwongFluid.cpp:14-44, 99-127; mixtureRedlichKwong.cpp:764-838)
template <int NS>
__device__ inline double rk_c2(const DevParams &P, const double *c, double c0, double rho,
                               double U, double p)
{
    double a, b, cv, mm;
    rk_mix<NS>(P, c, c0, a, b, cv, mm);
    const double R = P.R;
    const double s = rk_sqrtT(cv, U, c0, a, b);
    const double T = s * s;
    const double sT = __builtin_sqrt(T);
    const double l1 = ::log1p(c0 * b);
    const double dPdC = R * T / (1 - c0 * b) + c0 * R * T * b / ((1 - c0 * b) * (1 - c0 * b)) -
                        2 * a * c0 / (sT * (1 + b * c0)) +
                        a * (c0 * c0) * b / (sT * ((1 + b * c0) * (1 + b * c0)));
    const double dPdT = c0 * R / (1 - c0 * b) + 0.5 * a * (c0 * c0) / (sT * sT * sT * (1 + c0 * b));
    const double dUdC = cv * T - 1.5 * a * l1 / (b * sT) - 1.5 * a * c0 / (sT * (1 + c0 * b));
    const double dUdT = c0 * cv + 0.75 * a * c0 * l1 / (b * (sT * sT * sT));
    const double dpdrho = (dPdC - dPdT * dUdC / dUdT) / mm;
    return dpdrho + dPdT / dUdT * (U + p) / rho;
}

// ---- Kataoka-van der Waals mixture, FP64 on both paths -------------------
// Association series (KataokaVanDerWaalsFluid.cpp:17-62, M = 13 terms) and
// the reference's Newton iteration for T (NewtonMethod.cpp:13-40: stop when
// a step is <= 1e-12, at most 101 steps); the derivative series of
// dPdC / dPdT / dUvdC / dUvdT never run in the reference ("J = 2; M < J"),
// leaving dPdC = R T / (1 - c b) + c R T b / (1 - c b)^2, dPdT = c R / (1 - c b),
// dUvdC = Cv T, dUvdT = c Cv.  pow and exp are the device's, a few ulps from
// the host library's.
template <int NS>
__device__ __forceinline__ void kat_mix(const DevParams &P, const double *c, double c0, double &V0,
                                        double &eps, double &b, double &cv, double &mm)
{
    double X[NS];
    cv = 0.0;
    mm = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        X[k] = c[k] / c0;
        cv += X[k] * P.Cv[k];
        mm += X[k] * P.M[k];
    }
    V0 = 0.0;
    eps = 0.0;
    b = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int l = 0; l < NS; ++l) {
            V0 += X[k] * X[l] * P.akl[k][l];
            eps += X[k] * X[l] * P.ekl[k][l];
            b += X[k] * X[l] * P.bkl[k][l];
        }
}

// qa, -sum (J-1) x_J and -sum (J-1)/3 x_J with x_J = (V0 c)^((J-1)/3) e^((J-1) eps/(2RT))
__device__ inline void kat_series(double c, double T, double V0, double eps, double R, double &qa,
                                  double &su, double &sp)
{
    const double V0c = V0 * c, epsRT = 0.5 * eps / (R * T);
    qa = 1.0;
    su = 0.0;
    sp = 0.0;
    for (int J = 2; J <= 13; ++J) {
        const double x = ::pow(V0c, (J - 1.) / 3.) * ::exp((J - 1.) * epsRT);
        qa += x;
        su -= (J - 1.) * x;
        sp -= (J - 1.) / 3. * x;
    }
}
__device__ __forceinline__ double kat_UvA(double c, double T, double V0, double eps, double R)
{
    double qa, su, sp;
    kat_series(c, T, V0, eps, R, qa, su, sp);
    return 0.5 * c * eps * su / qa;
}
__device__ __forceinline__ double kat_pA(double c, double T, double V0, double eps, double R)
{
    double qa, su, sp;
    kat_series(c, T, V0, eps, R, qa, su, sp);
    return c * R * T * sp / qa;
}
// kind 0: T with c Cv T + UvA(T) = target; kind 1: c R T / (1 - c b) + pA(T) = target
__device__ inline double kat_newton(int kind, double x0, double c, double cv, double target,
                                    double V0, double eps, double R, double b)
{
    auto f = [&](double T) {
        return (kind == 0) ? c * cv * T + kat_UvA(c, T, V0, eps, R) - target
                           : c * R * T / (1 - c * b) + kat_pA(c, T, V0, eps, R) - target;
    };
    const double df = (kind == 0) ? c * cv : c * R / (1 - c * b);
    double x1 = x0;
    double x2 = x1 - f(x1) / df;
    for (int it = 1; __builtin_fabs(x2 - x1) > 1e-12 && it <= 100; ++it) {
        x1 = x2;
        x2 = x1 - f(x1) / df;
    }
    return x2;
}
template <int NS>
__device__ inline double kat_T(const DevParams &P, const double *c, double c0, double U)
{
    double V0, eps, b, cv, mm;
    kat_mix<NS>(P, c, c0, V0, eps, b, cv, mm);
    return kat_newton(0, U / ((c0 + STAB) * cv), c0, cv, U, V0, eps, P.R, b);
}
template <int NS>
__device__ inline double kat_p(const DevParams &P, const double *c, double c0, double U)
{
    double V0, eps, b, cv, mm;
    kat_mix<NS>(P, c, c0, V0, eps, b, cv, mm);
    const double T = kat_newton(0, U / ((c0 + STAB) * cv), c0, cv, U, V0, eps, P.R, b);
    return c0 * P.R * T / (1 - c0 * b) + kat_pA(c0, T, V0, eps, P.R);
}
template <int NS>
__device__ inline double kat_U(const DevParams &P, const double *c, double c0, double p)
{
    double V0, eps, b, cv, mm;
    kat_mix<NS>(P, c, c0, V0, eps, b, cv, mm);
    const double T = kat_newton(1, p / ((c0 + STAB) * P.R), c0, cv, p, V0, eps, P.R, b);
    return c0 * cv * T + kat_UvA(c0, T, V0, eps, P.R);
}
template <int NS>
__device__ inline double kat_c2(const DevParams &P, const double *c, double c0, double rho,
                                double U, double p)
{
    double V0, eps, b, cv, mm;
    kat_mix<NS>(P, c, c0, V0, eps, b, cv, mm);
    const double R = P.R;
    const double T = kat_newton(0, U / ((c0 + STAB) * cv), c0, cv, U, V0, eps, R, b);
    const double dPdC = R * T / (1 - c0 * b) + c0 * R * T * b / ((1 - c0 * b) * (1 - c0 * b));
    const double dPdT = c0 * R / (1 - c0 * b);
    const double dpdrho = (dPdC - dPdT * (cv * T) / (c0 * cv)) / mm;
    return dpdrho + dPdT / (c0 * cv) * (U + p) / rho;
}

template <class T, int NS>
__device__ __forceinline__ void to_f64(const T *c, double *cd)
{
#pragma unroll
    for (int k = 0; k < NS; ++k)
        cd[k] = static_cast<double>(c[k]);
}

// pFromUv: ideal g1 U (idealFluid.cpp:13-17); van der Waals
// g1 (U + a c0^2) / (1 - c0 b) - a c0^2 (vanDerWaalsFluid.cpp:14-22);
// Redlich-Kwong rk_p.
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ T eos_p(const DevParams &P, const T *c, T c0, T g1, T U)
{
    if (is_stiff<GAS>(P)) {
        // stiffened mixture (stiffenedFluid.cpp:13-17, mixtureStiffened.cpp:
        // 485-503): p0 and gamma mixed like a and b, held in the same matrices
        T p0, g;
        vdw_ab<T, NS>(P, c, c0, p0, g);
        return (g - T(1)) * U - p0 * g;
    }
    if constexpr (may_be_rk<GAS>() && std::is_floating_point_v<T>) {
        if (is_rk<GAS>(P)) {
            double cd[NS];
            to_f64<T, NS>(c, cd);
            return T(rk_p<NS>(P, cd, double(c0), double(U)));
        }
    }
    if constexpr (may_be_kataoka<GAS>() && std::is_floating_point_v<T>) {
        if (is_kataoka<GAS>(P)) {
            double cd[NS];
            to_f64<T, NS>(c, cd);
            return T(kat_p<NS>(P, cd, double(c0), double(U)));
        }
    }
    if (is_vdw<GAS>(P)) {
        T a, b;
        vdw_ab<T, NS>(P, c, c0, a, b);
        const T acc = a * (c0 * c0);
        return g1 * (U + acc) / (T(1) - c0 * b) - acc;
    }
    return g1 * U;
}

// UvFromp: ideal p / g1 (idealFluid.cpp:19-23); van der Waals
// (p + a c0^2) (1 - c0 b) / g1 - a c0^2 (vanDerWaalsFluid.cpp:24-31);
// Redlich-Kwong rk_U.
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ T eos_U(const DevParams &P, const T *c, T c0, T g1, T p)
{
    if (is_stiff<GAS>(P)) {
        // (p + p0 gamma) / (gamma - 1) (stiffenedFluid.cpp:19-23)
        T p0, g;
        vdw_ab<T, NS>(P, c, c0, p0, g);
        return (p + p0 * g) / (g - T(1));
    }
    if constexpr (may_be_rk<GAS>() && std::is_floating_point_v<T>) {
        if (is_rk<GAS>(P)) {
            double cd[NS];
            to_f64<T, NS>(c, cd);
            return T(rk_U<NS>(P, cd, double(c0), double(p)));
        }
    }
    if constexpr (may_be_kataoka<GAS>() && std::is_floating_point_v<T>) {
        if (is_kataoka<GAS>(P)) {
            double cd[NS];
            to_f64<T, NS>(c, cd);
            return T(kat_U<NS>(P, cd, double(c0), double(p)));
        }
    }
    if (is_vdw<GAS>(P)) {
        T a, b;
        vdw_ab<T, NS>(P, c, c0, a, b);
        const T acc = a * (c0 * c0);
        return (p + acc) * (T(1) - c0 * b) / g1 - acc;
    }
    return p / g1;
}

// sqSonicSpeed = dpdrho + dpdUv (U + p) / rho (mixtureIdeal.cpp:495-501,
// mixtureVanDerWaals.cpp:779-785).  Ideal: dpdrho = 0, dpdUv = g1.  Van der
// Waals (vanDerWaalsFluid.cpp:61-83, mixtureVanDerWaals.cpp:719-762): with
// m = 1 - c0 b, ac = a c0: dpdrho = (g1 (2 ac / m - b (U + ac c0) / m^2)
// - 2 ac) / M_mix, M_mix = sum_k X_k M_k, dpdUv = g1 / m.  Redlich-Kwong
// rk_c2.
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ T eos_c2(const DevParams &P, const T *c, T c0, T g1, T rho, T U,
                                    T p)
{
    if (is_stiff<GAS>(P)) {
        // dpdrho = 0, dpdUv = gamma - 1 (mixtureStiffened.cpp:621-663)
        T p0, g;
        vdw_ab<T, NS>(P, c, c0, p0, g);
        return T(0) + (g - T(1)) * (U + p) / rho;
    }
    if constexpr (may_be_rk<GAS>() && std::is_floating_point_v<T>) {
        if (is_rk<GAS>(P)) {
            double cd[NS];
            to_f64<T, NS>(c, cd);
            return T(rk_c2<NS>(P, cd, double(c0), double(rho), double(U), double(p)));
        }
    }
    if constexpr (may_be_kataoka<GAS>() && std::is_floating_point_v<T>) {
        if (is_kataoka<GAS>(P)) {
            double cd[NS];
            to_f64<T, NS>(c, cd);
            return T(kat_c2<NS>(P, cd, double(c0), double(rho), double(U), double(p)));
        }
    }
    if (is_vdw<GAS>(P)) {
        T a, b, mm = T(0);
        vdw_ab<T, NS>(P, c, c0, a, b);
#pragma unroll
        for (int k = 0; k < NS; ++k)
            mm += (c[k] / c0) * mixM<T>(P, k);
        const T m = T(1) - c0 * b;
        const T ac = a * c0;
        const T dpdrho = (g1 * (T(2) * ac / m - b * (U + ac * c0) / (m * m)) - T(2) * ac) / mm;
        return dpdrho + g1 / m * (U + p) / rho;
    }
    return T(0) + g1 * (U + p) / rho;
}

// Primitive recovery from the conserved vector, as after the stage update
// (Advection2dOrder.cpp:578-602): c_k = rho_k / M_k, c_0 = sum c_k,
// v = m / rho, U = E - rho v^2 / 2 - 0, p = (R / Cv) U.
// u: [rho_1..rho_Ns, rho, m_x, m_y, m_z, E]; w: [c_1..c_Ns, v_x, v_y, v_z, p]
template <class T, int NS, int GAS = -1>
__device__ __forceinline__ void prims_from_cons(const DevParams &P, const T *u, T *w)
{
    T c0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        w[k] = u[k] / mixM<T>(P, k);
        c0 += w[k];
    }
    const T rho = u[NS];
    const T vx = u[NS + 1] / rho, vy = u[NS + 2] / rho, vz = u[NS + 3] / rho;
    w[NS] = vx;
    w[NS + 1] = vy;
    w[NS + 2] = vz;
    const T v2 = ((T(0) + vx * vx) + vy * vy) + vz * vz;
    const T U = (u[NS + 4] - rho * v2 * T(0.5)) - T(0);
    w[NS + 3] = eos_p<T, NS, GAS>(P, w, c0, gamma1<T, NS>(P, w, c0), U);
}

// Reconstructed face side closed with the EOS (Advection2dOrder.cpp:233-289
// owner side, 301-372 neighbour side).
template <class T, int NS> struct Side {
    T rho[NS];  // partial densities
    T r0;       // total density
    T v[3];
    T m[3];
    T p, U, E, cs;
};

template <class T, int NS, int GAS = -1>
__device__ __forceinline__ void close_side(const DevParams &P, const T *c, const T *v,
                                           T p, Side<T, NS> &s)
{
    T c0 = 0, r0 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        c0 += c[k];
        s.rho[k] = c[k] * mixM<T>(P, k);
        r0 += s.rho[k];
    }
    s.r0 = r0;
    const T g1 = gamma1<T, NS>(P, c, c0);
    s.p = p;
    s.U = eos_U<T, NS, GAS>(P, c, c0, g1, p);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.v[a] = v[a];
        s.m[a] = v[a] * r0;
    }
    const T v2 = ((T(0) + v[0] * v[0]) + v[1] * v[1]) + v[2] * v[2];
    s.E = (s.U + r0 * v2 * T(0.5)) + T(0);
    // sqSonicSpeed; the turbulent term (2/3) k (1 + dpdUv) is 0
    const T c2 = eos_c2<T, NS, GAS>(P, c, c0, g1, r0, s.U, p);
    s.cs = Real<T>::sqrt_(c2);
}

// Flux through one axis-aligned face: g = F & N for every conserved field
// (rho_1..rho_Ns, rho, m_x, m_y, m_z, E).  ax: axis, sg: sign of the normal
// (+1 for inner and high-side faces, -1 for the low-side boundary faces).
// O is the owner side, B the neighbour side.  Returns false when no wave
// pattern applies (NaN speeds), as HLLCSolver.cpp:385-388 / HLLSolver.cpp:297-300.
template <class T, int NS>
__device__ __forceinline__ bool hllc_flux(const Side<T, NS> &O, const Side<T, NS> &B,
                                          int ax, T sg, T *g)
{
    const T uO = sg * O.v[ax], uN = sg * B.v[ax];  // v & N
    const T SO = tmin(uO - O.cs, uN - B.cs);
    const T SN = tmax(uO + O.cs, uN + B.cs);
    const T rO = O.r0, rN = B.r0;
    const T Sc = (B.p - O.p + rO * uO * (SO - uO) - rN * uN * (SN - uN)) /
                 (rO * (SO - uO) - rN * (SN - uN));
    const T PLR = T(0.5) * (O.p + B.p + rO * (SO - uO) * (Sc - uO) +
                            rN * (SN - uN) * (Sc - uN));
    // Wave pattern (HLLCSolver.cpp:140-388): SO >= 0 -> F(O); Sc >= 0 -> F*(O);
    // SN > 0 -> F*(B); SN <= 0 -> F(B); otherwise NaN speeds.  The star flux
    // of side X is F(X) + S (U*_X - U_X), evaluated as that sum, so one code
    // path serves both sides: pick the side, form F(X), add the star term.
    const bool own = (SO >= T(0)) || (Sc >= T(0));
    const bool star = own ? !(SO >= T(0)) : (SN > T(0));
    if (!own && !(SN > T(0)) && !(SN <= T(0))) {
#pragma unroll
        for (int q = 0; q < NS + 5; ++q)
            g[q] = SO;  // NaN
        return false;
    }
    const T u = own ? uO : uN;
    const T S = own ? SO : SN;
    T rho[NS], m[3];
#pragma unroll
    for (int k = 0; k < NS; ++k)
        rho[k] = own ? O.rho[k] : B.rho[k];
    const T r0 = own ? O.r0 : B.r0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
        m[b] = own ? O.m[b] : B.m[b];
    const T p = own ? O.p : B.p;
    const T E = own ? O.E : B.E;
    // F(X) & N
#pragma unroll
    for (int k = 0; k < NS; ++k)
        g[k] = rho[k] * u;
    g[NS] = r0 * u;
#pragma unroll
    for (int b = 0; b < 3; ++b)
        g[NS + 1 + b] = (b == ax) ? m[b] * u + sg * p : m[b] * u;
    g[NS + 4] = (E + p) * u;
    if (star) {
        const T factor = S - u;
        const T den = T(1) / (S - Sc);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const T rs = rho[k] * factor * den;
            g[k] = g[k] + S * (rs - rho[k]);
        }
        {
            const T rs = r0 * factor * den;
            g[NS] = g[NS] + S * (rs - r0);
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const T ms = (b == ax) ? (m[b] * factor - sg * (p - PLR)) * den
                                   : (m[b] * factor) * den;
            g[NS + 1 + b] = g[NS + 1 + b] + (ms - m[b]) * S;
        }
        const T Es = (E * factor - p * u + PLR * Sc) * den;
        g[NS + 4] = g[NS + 4] + S * (Es - E);
    }
    return true;
}

template <class T, int NS>
__device__ __forceinline__ bool hll_flux(const Side<T, NS> &O, const Side<T, NS> &B,
                                         int ax, T sg, T *g)
{
    const T uO = sg * O.v[ax], uN = sg * B.v[ax];
    const T SO = tmin(uO - O.cs, uN - B.cs);
    const T SN = tmax(uO + O.cs, uN + B.cs);
    if (SO >= T(0)) {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            g[k] = uO * O.rho[k];
        g[NS] = uO * O.r0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            g[NS + 1 + b] = (b == ax) ? O.m[b] * uO + sg * O.p : O.m[b] * uO;
        g[NS + 4] = uO * (O.E + O.p);
    } else if (SN > T(0)) {
        const T den = T(1) / (SN - SO);
#pragma unroll
        for (int k = 0; k < NS; ++k)
            g[k] = (SN * uO * O.rho[k] - SO * uN * B.rho[k] +
                    SN * SO * (B.rho[k] - O.rho[k])) * den;
        g[NS] = (SN * uO * O.r0 - SO * uN * B.r0 + SN * SO * (B.r0 - O.r0)) * den;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (b == ax)
                g[NS + 1 + b] = (((O.m[b] * uO + sg * O.p) * SN -
                                  (B.m[b] * uN + sg * B.p) * SO) +
                                 (B.m[b] - O.m[b]) * (SN * SO)) * den;
            else
                g[NS + 1 + b] = (((O.m[b] * uO) * SN - (B.m[b] * uN) * SO) +
                                 (B.m[b] - O.m[b]) * (SN * SO)) * den;
        }
        g[NS + 4] = (SN * (uO * (O.E + O.p)) - SO * (uN * (B.E + B.p)) +
                     SN * SO * (B.E - O.E)) * den;
    } else if (SN <= T(0)) {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            g[k] = uN * B.rho[k];
        g[NS] = uN * B.r0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            g[NS + 1 + b] = (b == ax) ? B.m[b] * uN + sg * B.p : B.m[b] * uN;
        g[NS + 4] = uN * (B.E + B.p);
    } else {
#pragma unroll
        for (int q = 0; q < NS + 5; ++q)
            g[q] = SO;
        return false;
    }
    return true;
}

// Limiter function phi(r) with xi = 2 / (1 + r) evaluated where used
// (minmodLimiter.cpp:14-27, vanLeerLimiter.cpp:14-24, vanAlbadaLimiter.cpp:15-26;
// zeroLimiter.cpp returns 0).  lim is uniform across the launch.
template <class T> __device__ __forceinline__ T phi(int lim, T r)
{
    if (lim == LIM_MINMOD) {
        if (r <= T(0))
            return T(0);
        else if (r <= T(1))
            return r;
        return tmin(T(1), T(2) / (T(1) + r));
    } else if (lim == LIM_VANLEER) {
        if (r < T(0))
            return T(0);
        return tmin(T(2) * r / (T(1) + r), T(2) / (T(1) + r));
    } else if (lim == LIM_VANALBADA) {
        if (r <= T(0))
            return T(0);
        return tmin((r * r + r) / (T(1) + r * r), T(2) / (T(1) + r));
    }
    return T(0);
}

// Boundary values (boundaryConditionValue.hpp:33-276 surface, 278-511 cell;
// slipFunction.cpp:15-35).  comp: -1 for a scalar, else the velocity
// component, ax the axis of the face.  slip reflects only the normal
// velocity component; for scalars slip is the identity.
template <class T>
__device__ __forceinline__ T bc_surface(int type, T val, T q, int comp, int ax)
{
    switch (type) {
    case BC_SLIP:
        return (comp == ax) ? q - q : q;  // v - (v & N) N
    case BC_FIXED_CELL:
        return (q + val) * T(0.5);
    case BC_FIXED_SURF:
        return val;
    default:
        return q;
    }
}

template <class T>
__device__ __forceinline__ T bc_cell(int type, T val, T q, int comp, int ax)
{
    switch (type) {
    case BC_SLIP:
        return (comp == ax) ? (q - q) * T(2) - q : q * T(2) - q;
    case BC_FIXED_CELL:
        return val;
    default:
        return q;
    }
}

}  // namespace sg
