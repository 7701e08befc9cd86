// capi.hip -- C ABI of the MI355X explicit finite-volume step (schemi_gpu.h).
//
// A context owns the device state (two conserved buffers: U^n, updated in
// place by stage 2, and U^1), the per-axis geometry tables, the parameter
// block and a small block of device scalars (time step, time, Courant
// reduction), so that a run of many steps needs no host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/schemi_gpu.h"
#include "fv_launch.hpp"
#include "fv_turb.hpp"
#include "fv_turb_launch.hpp"
#include "fv_star_launch.hpp"
#include "geometry.hpp"

using namespace sg;

// z planes per workgroup of the 3-D stage kernels
#ifndef SG_KCHUNK
#define SG_KCHUNK 16
#endif

namespace {

// device scalar block

__global__ void finalize_step_kernel(double *sc, const DevParams *P, double const_dt,
                                     int flip_parity, int src_scaled,
                                     unsigned long long *epsmax)
{
    // Schemi.cpp:802-868, 933-935 for one step
    if (sc[SC_DONE] != 0.0)
        return;
    const double dt = sc[SC_DT];
    const double t = sc[SC_TIME] + dt;
    unsigned long long bits = reinterpret_cast<unsigned long long *>(sc)[SC_SIGMAX];
    double dtn;
    unsigned long long *srcmin = reinterpret_cast<unsigned long long *>(sc) + SC_SRCMIN;
    // the source step of this step's diffusion stage, veryBig when none set it
    // (the turbulence model's ratios carry the coefficient already,
    // shearModel.cpp:131-138)
    const double srcv = __longlong_as_double(static_cast<long long>(*srcmin));
    const double dts = (*srcmin == SRCMIN_NONE) ? 1e-3 * __DBL_MAX__
                                                : (src_scaled ? srcv : P->ts_coeff * srcv);
    if (const_dt > 0.0) {
        dtn = const_dt;
    } else {
        const double smax = __longlong_as_double(static_cast<long long>(bits));
        dtn = P->courant / smax;
        if (P->ts_source)
            dtn = (dts < dtn) ? dts : dtn;  // std::min(dt, timestepSource)
        dtn = (dtn < P->min_time) ? P->min_time : dtn;  // std::max(dt, minTime)
    }
    sc[SC_SRCDT] = dts;
    *srcmin = SRCMIN_NONE;
    sc[SC_DTNEXT] = dtn;
    if (t + dtn > sc[SC_TEND])
        dtn = sc[SC_TEND] - t;
    sc[SC_TIME] = t;
    sc[SC_DT] = dtn;
    sc[SC_STEPS] = sc[SC_STEPS] + 1.0;
    if (flip_parity)
        sc[SC_PARITY] = 1.0 - sc[SC_PARITY];
    reinterpret_cast<unsigned long long *>(sc)[SC_SIGMAX] = 0ull;
    if (epsmax)
        *epsmax = 0ull;
    if (!(t < sc[SC_TEND]))
        sc[SC_DONE] = 1.0;
}

__global__ void set_scalars_kernel(double *sc, double dt, double t, double tend)
{
    sc[SC_DT] = dt;
    sc[SC_TIME] = t;
    sc[SC_TEND] = tend;
    reinterpret_cast<unsigned long long *>(sc)[SC_SIGMAX] = 0ull;
    reinterpret_cast<unsigned long long *>(sc)[SC_SRCMIN] = SRCMIN_NONE;
    sc[SC_SRCDT] = 1e-3 * __DBL_MAX__;
    sc[SC_DTNEXT] = 0.0;
    sc[SC_STEPS] = 0.0;
    sc[SC_DONE] = 0.0;
    sc[SC_PARITY] = 0.0;
}

// Moves the current state back into the first buffer when the MUSCL-Hancock
// steps of a call ended on the second one.
__global__ void restore_state_kernel(const double *sc, double *U0, const double *U1,
                                     long long count)
{
    if (sc[SC_PARITY] == 0.0)
        return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        U0[i] = U1[i];
}

template <int NS, bool TURB = false>
__global__ void prims_kernel(const DevParams *__restrict__ Pp, const double *__restrict__ U,
                             double *__restrict__ W, long long n)
{
    const DevParams &P = *Pp;
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= n)
        return;
    constexpr int NF = TURB ? NS + 7 : NS + 5;
    double u[NF], w[NS + 6];
    for (int q = 0; q < NF; ++q)
        u[q] = U[q * n + id];
    if constexpr (TURB)
        prims_turb<double, NS, -1>(P, u, w);
    else
        prims_from_cons<double, NS>(P, u, w);
    double c0 = 0;
    for (int k = 0; k < NS; ++k) {
        W[k * n + id] = w[k];
        c0 += w[k];
    }
    W[NS * n + id] = c0;
    for (int b = 0; b < 3; ++b)
        W[(NS + 1 + b) * n + id] = w[NS + b];
    const double vx = w[NS], vy = w[NS + 1], vz = w[NS + 2], p = w[NS + 3];
    const double v2 = ((0. + vx * vx) + vy * vy) + vz * vz;
    const double Uint = (u[NS + 4] - u[NS] * v2 * 0.5) - (TURB ? u[NS + 5] : 0.);
    double cv = 0;
    for (int k = 0; k < NS; ++k)
        cv += (w[k] / c0) * P.Cv[k];
    W[(NS + 4) * n + id] = p;
    W[(NS + 5) * n + id] = Uint;
    // TFromUv: U / ((c0 + stab) Cv) (idealFluid.cpp:37-41), with U + a c0^2
    // for van der Waals (vanDerWaalsFluid.cpp:45-49)
    double Ut = Uint;
    if (P.gas_model == GAS_VDW) {
        double a, b;
        vdw_ab<double, NS>(P, w, c0, a, b);
        Ut = Uint + a * (c0 * c0);
    }
    // stiffened: U - p0 (stiffenedFluid.cpp:37-41)
    if (P.gas_model == GAS_STIFF) {
        double p0, g;
        vdw_ab<double, NS>(P, w, c0, p0, g);
        Ut = Uint - p0;
    }
    // Redlich-Kwong: T = s^2 from its cubic (RedlichKwongFluid.cpp:89-97);
    // Kataoka-van der Waals: Newton (KataokaVanDerWaalsFluid.cpp:277-294)
    W[(NS + 6) * n + id] = (P.gas_model == GAS_RK)        ? rk_T<NS>(P, w, c0, Uint)
                           : (P.gas_model == GAS_KATAOKA) ? kat_T<NS>(P, w, c0, Uint)
                                                          : Ut / ((c0 + STAB) * cv);
    const double c2 = eos_c2<double, NS>(P, w, c0, gamma1<double, NS>(P, w, c0), u[NS], Uint, p);
    W[(NS + 7) * n + id] = __builtin_sqrt(c2);
}

template <class T, int NS>
__global__ void face_flux_kernel(const DevParams *__restrict__ Pp, long long nf,
                                 const int *__restrict__ axis, const double *__restrict__ sgn,
                                 const double *__restrict__ Lp, const double *__restrict__ Rp,
                                 double *__restrict__ F, int *err)
{
    const DevParams &P = *Pp;
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    constexpr int NQ = NS + 4, NF = NS + 5;
    T cL[NS], vL[3], cR[NS], vR[3];
    for (int k = 0; k < NS; ++k) {
        cL[k] = T(Lp[f * NQ + k]);
        cR[k] = T(Rp[f * NQ + k]);
    }
    for (int b = 0; b < 3; ++b) {
        vL[b] = T(Lp[f * NQ + NS + b]);
        vR[b] = T(Rp[f * NQ + NS + b]);
    }
    Side<T, NS> O, B;
    close_side<T, NS>(P, cL, vL, T(Lp[f * NQ + NS + 3]), O);
    close_side<T, NS>(P, cR, vR, T(Rp[f * NQ + NS + 3]), B);
    T g[NF];
    const bool ok = (P.flux == FLUX_HLL) ? hll_flux<T, NS>(O, B, axis[f], T(sgn[f]), g)
                                         : hllc_flux<T, NS>(O, B, axis[f], T(sgn[f]), g);
    if (!ok)
        atomicOr(err, ERRF_RIEMANN);
    for (int q = 0; q < NF; ++q)
        F[f * NF + q] = double(g[q]);
}

}  // namespace

struct schemi_gpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevParams hp{};
    DevParams *dP = nullptr;
    double *dgeo = nullptr;
    double *U[2] = {nullptr, nullptr};  // U^n (in place), U^1
    double *sc = nullptr;
    int *derr = nullptr;
    int dim = 1, ns = 1, lin = 0, prec = 0;
    int mode = MODE_SLOPES;
    bool hancock = false;  // SCHEMI_STEP_SECOND_ORDER: one MUSCL-Hancock stage per step
    void *scratch = nullptr;  // MODE_GENERAL: primitives and gradients
    void *W = nullptr, *Lg = nullptr, *Hs = nullptr;
    long long n = 0;
    int kchunk = 16;
    double const_dt = -1.0;
    double time = 0.0;
    int pending_error = 0;  // reported by the next step (reference throws lazily)
    MeshTables tables;
    bool timing = false;
    std::vector<hipEvent_t> ev;  // start/stop pairs of the timed stage launches
    double *adj = nullptr;       // adjoint: lambda of U^1 (Ns + 5 rows), partial sums
    bool diffusion = false;      // explicit molecular diffusion stage after each step
    double mu = 0, D = 0, kappa = 0;
    double *Wd = nullptr;        // diffusion: (2 Ns + 5) x n primitives
    std::vector<hipEvent_t> evd; // timed diffusion stages: start, between passes, end
    int sle_iters = 0;           // Jacobi solver: iteration limit (0: explicit solver)
    int molmass = 0, enthalpy = 0;
    int nf = 0;                  // conserved fields: ns + 5, turbulence ns + 7, kEpsA ns + 10
    bool turb = false;           // k-epsilon turbulence
    bool afield = false;         // kEpsA, BHR: the state carries rho a
    bool bfield = false;         // BHR: the state carries rho b
    double *db2 = nullptr;       // BHR: the second stage's b; kEpsA: the diffusion stage's b (n)
    bool stage_turb = false;     // a turbulent diffusion stage has filled Wd's k, epsilon
    bool star = false;           // #Reconstruction star: star states per face
    double *dS = nullptr;        // star buffer, (5 + na) x hp.star_nft
    TurbParams tp{};
    TurbParams *dTP = nullptr;
    unsigned long long *epsmax = nullptr;  // largest epsilon of the state (bits)
    double *Mb = nullptr, *dmax = nullptr;
    int *done = nullptr;
    int done_host[16] = {};
};

namespace {

int hip_fail(hipError_t e)
{
    return e == hipSuccess ? SCHEMI_OK : SCHEMI_ERR_SYSTEM;
}

#define HIPCHK(x)                                \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess)                    \
            return SCHEMI_ERR_SYSTEM;            \
    } while (0)

int launch_stage_turb(schemi_gpu_ctx *c, const double *Uin, double *Uout, const double *Un,
                      int variant)
{
    TurbStageLaunch L;
    L.P = c->dP;
    L.TP = c->dTP;
    L.geo = c->dgeo;
    L.Uin = Uin;
    L.Uout = Uout;
    L.Un = Un;
    L.dt = c->sc;
    L.sigmax = reinterpret_cast<unsigned long long *>(c->sc + SC_SIGMAX);
    L.err = c->derr;
    L.dim = c->dim;
    L.ns = c->ns;
    L.variant = variant;
    L.gas_model = c->hp.gas_model;
    L.slopes = c->mode == MODE_SLOPES ? 1 : 0;
    L.na = c->bfield ? 4 : c->afield ? 3 : 0;
    L.b2 = c->db2;
    L.n = c->n;
    L.stream = c->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, c->stream));
    }
    const int r = c->prec == SCHEMI_PREC_MIXED ? launch_turb_stage_mixed(L)
                                                : launch_turb_stage_f64(L);
    if (c->timing) {
        HIPCHK(hipEventRecord(e1, c->stream));
        c->ev.push_back(e0);
        c->ev.push_back(e1);
    }
    if (r == -1)
        return SCHEMI_ERR_INIT;
    return r == 0 ? SCHEMI_OK : SCHEMI_ERR_SYSTEM;
}

int launch_stage(schemi_gpu_ctx *c, const double *Uin, double *Uout, const double *Un,
                 int variant)
{
    if (c->turb)
        return launch_stage_turb(c, Uin, Uout, Un, variant);
    StageLaunch L;
    L.P = c->dP;
    L.geo = c->dgeo;
    L.Uin = Uin;
    L.Uout = Uout;
    L.Un = Un;
    L.dt = c->sc;
    L.sigmax = reinterpret_cast<unsigned long long *>(c->sc + SC_SIGMAX);
    L.err = c->derr;
    L.dim = c->dim;
    L.ns = c->ns;
    L.lin = c->lin;
    L.variant = variant;
    L.mode = c->mode;
    L.gas_model = c->hp.gas_model;
    L.W = c->W;
    L.Lg = c->Lg;
    L.Hs = c->Hs;
    L.nx = c->hp.N[0];
    L.ny = c->hp.N[1];
    L.nz = c->hp.N[2];
    L.kchunk = c->kchunk;
    L.stream = c->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, c->stream));
    }
    const int r = c->mode == MODE_LEAN               ? launch_stage_lean(L)
                  : c->prec == SCHEMI_PREC_MIXED ? launch_stage_mixed(L)
                                                 : launch_stage_f64(L);
    if (c->timing) {
        HIPCHK(hipEventRecord(e1, c->stream));
        c->ev.push_back(e0);
        c->ev.push_back(e1);
    }
    if (r == -1)
        return SCHEMI_ERR_INIT;
    return r == 0 ? SCHEMI_OK : SCHEMI_ERR_SYSTEM;
}

// Star states of the faces of state U (first stage: avg 0, second: 1).
int launch_star_stage(schemi_gpu_ctx *c, const double *U, int avg)
{
    if (!c->star)
        return SCHEMI_OK;
    StarLaunch S;
    S.P = c->dP;
    S.TP = c->dTP;
    S.geo = c->dgeo;
    S.U = U;
    S.S = c->dS;
    S.sc = c->sc;
    S.avg = avg;
    S.dim = c->dim;
    S.ns = c->ns;
    S.gas_model = c->hp.gas_model;
    S.slopes = (c->dim == 1 && c->mode == MODE_SLOPES) ? 1 : 0;
    S.nt = c->turb ? 2 : 0;
    S.na = c->bfield ? 4 : c->afield ? 3 : 0;
    S.n = c->n;
    S.stream = c->stream;
    const int r = launch_star(S);
    if (r == -1)
        return SCHEMI_ERR_INIT;
    return r == 0 ? SCHEMI_OK : SCHEMI_ERR_SYSTEM;
}

int one_step(schemi_gpu_ctx *c)
{
    if (c->hancock) {
        // secondOrderStepSolver::calculateStep (secondOrderStepSolver.cpp:33-48):
        // U^{n+1} = L_H(U^n), one stage; the kernel reads the buffer the
        // scalar block's parity names and writes the other one
        const int r = launch_stage(c, c->U[0], c->U[1], nullptr, SV_HANCOCK);
        if (r)
            return r;
    } else {
        // secondOrderStepSolverRK::calculateStep (secondOrderStepSolverRK.cpp:48-71):
        // U1 = L(Un); U2 = L(U1); U^{n+1} = 0.5 U2 + 0.5 Un (fused in stage 2);
        // with star reconstruction the faces' star states of each stage's
        // input, averaged (secondOrderStepSolverRK.cpp:73-85)
        int r = launch_star_stage(c, c->U[0], 0);
        if (r)
            return r;
        r = launch_stage(c, c->U[0], c->U[1], nullptr, SV_RK1);
        if (r)
            return r;
        r = launch_star_stage(c, c->U[1], 1);
        if (r)
            return r;
        r = launch_stage(c, c->U[1], c->U[0], c->U[0], SV_RK2);
        if (r)
            return r;
    }
    if (c->diffusion) {
        // Diffusion (secondOrderStepSolverRK.cpp:102-109,
        // secondOrderStepSolver.cpp:72-82) on the new state
        DiffusionLaunch L;
        L.P = c->dP;
        L.geo = c->dgeo;
        L.U = c->U[0];
        L.Ualt = c->U[1];
        L.hk = c->hancock ? 1 : 0;
        L.W = c->Wd;
        L.sc = c->sc;
        L.err = c->derr;
        L.mu = c->mu;
        L.D = c->D;
        L.kappa = c->kappa;
        L.dim = c->dim;
        L.ns = c->ns;
        L.n = c->n;
        L.stream = c->stream;
        L.mid = nullptr;
        L.Mb = c->Mb;
        L.dmax = c->dmax;
        L.done = c->done;
        L.done_host = c->done_host;
        L.maxit = c->sle_iters;
        L.mm = c->molmass;
        L.en = c->enthalpy;
        L.star = c->star ? c->dS : nullptr;
        if (c->Mb) {
            HIPCHK(hipMemsetAsync(c->dmax, 0,
                                  sizeof(double) * (size_t)(c->ns + 4) * (size_t)(c->sle_iters + 2),
                                  c->stream));
            HIPCHK(hipMemsetAsync(c->done, 0, sizeof(int) * (size_t)(c->ns + 4), c->stream));
        }
        hipEvent_t e[3] = {nullptr, nullptr, nullptr};
        if (c->timing) {
            for (hipEvent_t &x : e)
                HIPCHK(hipEventCreate(&x));
            HIPCHK(hipEventRecord(e[0], c->stream));
            L.mid = e[1];
        }
        if (c->turb) {
            TurbDiffusion T;
            T.TP = c->dTP;
            T.epsmax = c->epsmax;
            T.amodel = c->bfield ? 2 : c->afield ? 1 : 0;
            T.b2 = c->db2;
            T.bk = c->afield && !c->bfield ? c->db2 : nullptr;
            if (launch_diffusion_turb(L, T))
                return SCHEMI_ERR_SYSTEM;
            c->stage_turb = true;
        } else if (launch_diffusion(L)) {
            return SCHEMI_ERR_SYSTEM;
        }
        if (c->timing) {
            HIPCHK(hipEventRecord(e[2], c->stream));
            c->evd.insert(c->evd.end(), e, e + 3);
        }
    }
    hipLaunchKernelGGL(finalize_step_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, c->dP,
                       c->const_dt, c->hancock ? 1 : 0, c->turb ? 1 : 0, c->epsmax);
    return hip_fail(hipGetLastError());
}

// After the steps of a call: the current state back in U[0].
int restore_state(schemi_gpu_ctx *c)
{
    if (!c->hancock)
        return SCHEMI_OK;
    const long long count = (long long)(c->ns + 5) * c->n;
    const long long blocks = std::min<long long>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(restore_state_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream,
                       c->sc, c->U[0], c->U[1], count);
    return hip_fail(hipGetLastError());
}

// Device buffers of one API call, released on every exit path (error
// returns included) by hipFreeAsync on the context's stream.
struct callBuffers {
    hipStream_t stream;
    std::vector<void *> held;
    explicit callBuffers(hipStream_t s) : stream(s) {}
    callBuffers(const callBuffers &) = delete;
    callBuffers &operator=(const callBuffers &) = delete;
    template <class T> hipError_t get(T **out, size_t bytes)
    {
        void *q = nullptr;
        const hipError_t e = hipMallocAsync(&q, bytes, stream);
        if (e == hipSuccess) {
            held.push_back(q);
            *out = static_cast<T *>(q);
        }
        return e;
    }
    ~callBuffers()
    {
        for (void *q : held)
            (void)hipFreeAsync(q, stream);
    }
};

int read_error(schemi_gpu_ctx *c)
{
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, c->derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (e & ERRF_RIEMANN)
        return SCHEMI_ERR_RIEMANN;
    if (e & ERRF_BOUNDARY)
        return SCHEMI_ERR_BOUNDARY;
    if (e & ERRF_NEG_T)
        return SCHEMI_ERR_NEGATIVE_TEMPERATURE;
    return SCHEMI_OK;
}

bool bc_ok(int t)
{
    return t == SCHEMI_BC_FREE || t == SCHEMI_BC_SLIP || t == SCHEMI_BC_FIXED_CELL ||
           t == SCHEMI_BC_FIXED_SURFACE || t == SCHEMI_BC_PERIODIC;
}

}  // namespace

namespace {

// Scratch of schemi_gpu_mixing_width_noalloc: the per-workgroup partial sums
// of the mixing-width reduction (one per 256 cells) and its result, in a
// fixed device array that every device holds for the life of the process.
// Host code never hands the array's name to a kernel or a copy: its address
// on each device is looked up once, with that device current, when the first
// context is created there (noalloc_setup), and kept in noalloc_ptr.
constexpr long long NOALLOC_DOUBLES = SCHEMI_NOALLOC_MAX_CELLS / 256 + 1;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_DEVICES = 64;
std::mutex noalloc_mu;
double *noalloc_ptr[NOALLOC_DEVICES] = {};

int noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_DEVICES)
        return SCHEMI_ERR_SYSTEM;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_ptr[device])
        return SCHEMI_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return SCHEMI_ERR_SYSTEM;
    noalloc_ptr[device] = static_cast<double *>(p);
    return SCHEMI_OK;
}

int ensure_adj(schemi_gpu_ctx *c)
{
    if (c->adj)
        return SCHEMI_OK;
    const size_t count = (size_t)(c->ns + 5) * (size_t)c->n + (size_t)(c->n + 255) / 256 + 1;
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&c->adj), count * sizeof(double)));
    return SCHEMI_OK;
}

int adjoint_stage(schemi_gpu_ctx *c, const double *U, const double *lam, double *out,
                  double dt, double a, const double *add, double b)
{
    AdjointLaunch L;
    L.P = c->dP;
    L.geo = c->dgeo;
    L.U = U;
    L.lam = lam;
    L.out = out;
    L.add = add;
    L.dt = dt;
    L.a = a;
    L.b = b;
    L.dim = c->dim;
    L.ns = c->ns;
    L.gas_model = c->hp.gas_model;
    L.slopes = (c->mode == MODE_SLOPES) ? 1 : 0;
    L.lin = c->lin;
    L.n = c->n;
    L.stream = c->stream;
    const int r = launch_adjoint_stage(L);
    if (r == -1)
        return SCHEMI_ERR_INIT;
    return r == 0 ? SCHEMI_OK : SCHEMI_ERR_SYSTEM;
}

}  // namespace

extern "C" {

const char *schemi_gpu_strerror(int code)
{
    switch (code) {
    case SCHEMI_OK:
        return "noErrors";
    case SCHEMI_ERR_RIEMANN:
        return "RiemannSolverError: Fatal error in Riemann solver.";
    case SCHEMI_ERR_NEGATIVE_TEMPERATURE:
        return "negativeTemperatureError: Negative temperature after diffusion.";
    case SCHEMI_ERR_SYSTEM:
        return "systemError: HIP runtime failure";
    case SCHEMI_ERR_BOUNDARY:
        return "boundaryConditionError";
    case SCHEMI_ERR_MESH:
        return "meshGenerationError";
    case SCHEMI_ERR_INIT:
        return "initialisationError: configuration not supported by the GPU step";
    case SCHEMI_ERR_SCRATCH:
        return "scratchCapacityError: more cells than the no-allocation scratch holds "
               "(SCHEMI_NOALLOC_MAX_CELLS)";
    default:
        return "unknown error";
    }
}

int schemi_gpu_create(const schemi_mesh_desc *mesh, const schemi_mix_desc *mix,
                      const schemi_scheme_desc *scheme, schemi_gpu_ctx **out)
{
    if (!mesh || !mix || !scheme || !out)
        return SCHEMI_ERR_INIT;
    *out = nullptr;
    const int dim = mesh->dim;
    if (dim < 1 || dim > 3)
        return SCHEMI_ERR_MESH;
    // at least two cells along every axis of a 2-D/3-D mesh (the first-order
    // kernel relies on a neighbour across each other axis)
    for (int a = 0; a < dim; ++a)
        if (mesh->n[a] < 1 || (dim >= 2 && mesh->n[a] < 2) || mesh->n[a] > (1 << 30))
            return SCHEMI_ERR_MESH;
    if (mix->ns < 1 || mix->ns > 3 ||
        (mix->gas_model != SCHEMI_GAS_IDEAL && mix->gas_model != SCHEMI_GAS_VAN_DER_WAALS &&
         mix->gas_model != SCHEMI_GAS_REDLICH_KWONG && mix->gas_model != SCHEMI_GAS_STIFFENED &&
         mix->gas_model != SCHEMI_GAS_KATAOKA))
        return SCHEMI_ERR_INIT;
    if (mix->gas_model == SCHEMI_GAS_KATAOKA)
        for (int k = 0; k < mix->ns; ++k)
            if (!(mix->sigmaLJ[k] > 0.0) || !(mix->Tcrit[k] > 0.0) || !(mix->Pcrit[k] > 0.0))
                return SCHEMI_ERR_INIT;
    if (mix->gas_model == SCHEMI_GAS_STIFFENED)
        for (int k = 0; k < mix->ns; ++k)
            if (!(mix->gamma[k] > 1.0) || !(mix->p0[k] >= 0.0))
                return SCHEMI_ERR_INIT;
    if (mix->gas_model == SCHEMI_GAS_VAN_DER_WAALS || mix->gas_model == SCHEMI_GAS_REDLICH_KWONG)
        for (int k = 0; k < mix->ns; ++k)
            if (!(mix->Tcrit[k] > 0.0) || !(mix->Pcrit[k] > 0.0))
                return SCHEMI_ERR_INIT;
    if (scheme->flux != SCHEMI_FLUX_HLL && scheme->flux != SCHEMI_FLUX_HLLC)
        return SCHEMI_ERR_INIT;
    const int lim = scheme->limiter;
    if (lim != SCHEMI_LIM_ZERO && lim != SCHEMI_LIM_MINMOD && lim != SCHEMI_LIM_VANLEER &&
        lim != SCHEMI_LIM_LINEAR && lim != SCHEMI_LIM_VANALBADA)
        return SCHEMI_ERR_INIT;
    if (scheme->precision != SCHEMI_PREC_FP64 && scheme->precision != SCHEMI_PREC_MIXED)
        return SCHEMI_ERR_INIT;
    if (scheme->step_solver != SCHEMI_STEP_SECOND_ORDER &&
        scheme->step_solver != SCHEMI_STEP_SECOND_ORDER_RK)
        return SCHEMI_ERR_INIT;
    if (scheme->diffusion &&
        ((mix->gas_model != SCHEMI_GAS_IDEAL && mix->gas_model != SCHEMI_GAS_VAN_DER_WAALS) ||
         (mix->gas_model != SCHEMI_GAS_IDEAL && scheme->diffusion_enthalpy_flow) ||
         !(scheme->transport_mu >= 0.0) ||
         !(scheme->transport_D >= 0.0) || !(scheme->transport_kappa >= 0.0) ||
         (scheme->diffusion_solver != SCHEMI_SLE_EXPLICIT &&
          scheme->diffusion_solver != SCHEMI_SLE_JACOBI &&
          scheme->diffusion_solver != SCHEMI_SLE_GAUSS_SEIDEL) ||
         (scheme->diffusion_solver != SCHEMI_SLE_EXPLICIT && scheme->diffusion_iterations < 1) ||
         (scheme->diffusion_enthalpy_flow && scheme->diffusion_solver != SCHEMI_SLE_EXPLICIT)))
        return SCHEMI_ERR_INIT;
    if (scheme->reconstruction != SCHEMI_RECONSTRUCTION_LINEAR &&
        scheme->reconstruction != SCHEMI_RECONSTRUCTION_STAR)
        return SCHEMI_ERR_INIT;
    if ((scheme->timestep_source != 0 && scheme->timestep_source != 1) ||
        !(scheme->source_coeff >= 0.0))
        return SCHEMI_ERR_INIT;
    // turbulence: the shear, zero and kEpsA k-epsilon models with SSP-RK2,
    // ideal or van der Waals mixture, slopes that vanish or 1-D TVD slopes,
    // the explicit diffusion stage without the species corrections
    if (scheme->turbulence != 0 && scheme->turbulence != 1)
        return SCHEMI_ERR_INIT;
    if (scheme->turbulence &&
        ((scheme->turbulence_model != SCHEMI_TURB_SHEAR &&
          scheme->turbulence_model != SCHEMI_TURB_ZERO &&
          scheme->turbulence_model != SCHEMI_TURB_KEPSA &&
          scheme->turbulence_model != SCHEMI_TURB_BHR) ||
         scheme->step_solver != SCHEMI_STEP_SECOND_ORDER_RK ||
         (mix->gas_model != SCHEMI_GAS_IDEAL && mix->gas_model != SCHEMI_GAS_VAN_DER_WAALS) ||
         lim == SCHEMI_LIM_LINEAR ||
         (scheme->diffusion && (scheme->diffusion_solver != SCHEMI_SLE_EXPLICIT ||
                                scheme->diffusion_molar_mass || scheme->diffusion_enthalpy_flow)) ||
         !(scheme->turb_Cmu > 0.0) || !(scheme->turb_min_k > 0.0) ||
         !(scheme->turb_min_eps > 0.0)))
        return SCHEMI_ERR_INIT;

    schemi_gpu_ctx *c = new (std::nothrow) schemi_gpu_ctx();
    if (!c)
        return SCHEMI_ERR_SYSTEM;
    c->dim = dim;
    c->ns = mix->ns;
    c->lin = (lim == SCHEMI_LIM_LINEAR) ? 1 : 0;
    c->prec = scheme->precision;
    c->const_dt = scheme->const_dt;
    c->hancock = scheme->step_solver == SCHEMI_STEP_SECOND_ORDER;
    c->diffusion = scheme->diffusion != 0;
    c->mu = scheme->transport_mu;
    c->D = scheme->transport_D;
    c->kappa = scheme->transport_kappa;
    c->molmass = (c->diffusion && scheme->diffusion_molar_mass) ? 1 : 0;
    c->turb = scheme->turbulence != 0;
    c->afield = c->turb && (scheme->turbulence_model == SCHEMI_TURB_KEPSA ||
                            scheme->turbulence_model == SCHEMI_TURB_BHR);
    c->bfield = c->turb && scheme->turbulence_model == SCHEMI_TURB_BHR;
    c->nf = mix->ns + 5 + (c->turb ? 2 : 0) + (c->afield ? 3 : 0) + (c->bfield ? 1 : 0);
    if (c->turb) {
        TurbParams &T = c->tp;
        T.model = scheme->turbulence_model;
        T.Cmu = scheme->turb_Cmu;
        T.sigmaSc = scheme->turb_sigma_sc;
        T.sigmaT = scheme->turb_sigma_t;
        T.sigmaE = scheme->turb_sigma_e;
        T.sigmaK = scheme->turb_sigma_k;
        T.sigmaEps = scheme->turb_sigma_eps;
        T.CMS_R = scheme->turb_CMS_R;
        T.CMS_D = scheme->turb_CMS_D;
        T.mink = scheme->turb_min_k;
        T.mineps = scheme->turb_min_eps;
        T.C1 = scheme->turb_C[0];
        T.C2 = scheme->turb_C[1];
        T.C3 = scheme->turb_C[2];
        if (c->bfield) {
            // BHRModel.cpp:86-137: C0, C1, C2, C3, C4, Ca, Cb1, Cb2, alpha2,
            // alpha3, alpha4, CMS_M
            const double *K = scheme->turb_bhr_C;
            T.C0 = K[0];
            T.C1 = K[1];
            T.C2 = K[2];
            T.C3 = K[3];
            T.C4 = K[4];
            T.Ca = K[5];
            T.Cb1 = K[6];
            T.Cb2 = K[7];
            T.alpha2 = K[8];
            T.alpha3 = K[9];
            T.alpha4 = K[10];
            T.CMSM = K[11];
            T.sigmaa = scheme->turb_sigma_a;
            T.sigmab = scheme->turb_sigma_b;
            T.minb = scheme->turb_min_b;
            for (int p = 0; p < 6; ++p) {
                T.bcta[p] = scheme->bc_a[p].type;
                for (int b = 0; b < 3; ++b)
                    T.bcva[p][b] = scheme->bc_a[p].value[b];
                T.bctb[p] = scheme->bc_b[p].type;
                T.bcvb[p] = scheme->bc_b[p].value[0];
            }
        } else if (c->afield) {
            // kEpsAModel.cpp:86-121: C0, C1, C2, C3, Ca, CaMax, CMSA, chi
            const double *K = scheme->turb_kepsa_C;
            T.C0 = K[0];
            T.C1 = K[1];
            T.C2 = K[2];
            T.C3 = K[3];
            T.Ca = K[4];
            T.CaMax = K[5];
            T.CMSA = K[6];
            T.chi = K[7];
            T.sigmaa = scheme->turb_sigma_a;
            for (int p = 0; p < 6; ++p) {
                T.bcta[p] = scheme->bc_a[p].type;
                for (int b = 0; b < 3; ++b)
                    T.bcva[p][b] = scheme->bc_a[p].value[b];
            }
        }
        T.ts_coeff = scheme->source_coeff;
        for (int p = 0; p < 6; ++p) {
            T.bct[p][0] = scheme->bc_k[p].type;
            T.bcv[p][0] = scheme->bc_k[p].value[0];
            T.bct[p][1] = scheme->bc_eps[p].type;
            T.bcv[p][1] = scheme->bc_eps[p].value[0];
        }
    }
    c->enthalpy = (c->diffusion && scheme->diffusion_enthalpy_flow) ? 1 : 0;
    // Jacobi sweeps serve the Jacobi and the Gauss-Seidel solvers
    c->sle_iters = (c->diffusion && scheme->diffusion_solver != SCHEMI_SLE_EXPLICIT)
                       ? scheme->diffusion_iterations
                       : 0;
    DevParams &P = c->hp;
    std::memset(&P, 0, sizeof P);
    P.Ns = mix->ns;
    P.dim = dim;
    for (int a = 0; a < 3; ++a)
        P.N[a] = (a < dim) ? static_cast<int>(mesh->n[a]) : 1;
    P.n = (long long)P.N[0] * P.N[1] * P.N[2];
    c->n = P.n;
    P.stride[0] = 1;
    P.stride[1] = P.N[0];
    P.stride[2] = (long long)P.N[0] * P.N[1];
    P.limiter = lim;
    P.flux = scheme->flux;
    for (int k = 0; k < mix->ns; ++k) {
        P.M[k] = mix->M[k];
        P.Cv[k] = mix->Cv[k];
        P.Mf[k] = static_cast<float>(mix->M[k]);
        P.Cvf[k] = static_cast<float>(mix->Cv[k]);
        P.iMf[k] = static_cast<float>(1.0 / mix->M[k]);
        P.CvMf[k] = static_cast<float>(mix->Cv[k] / mix->M[k]);
    }
    P.R = mix->R;
    P.Rf = static_cast<float>(mix->R);
    P.gas_model = mix->gas_model;
    if (mix->gas_model == SCHEMI_GAS_VAN_DER_WAALS) {
        // mixtureVanDerWaals.cpp:19-58: b_kk = Tc R / (8 Pc), a_kk = Pc b b 27,
        // b_kl = (0.5 (cbrt b_kk + cbrt b_ll))^3, a_kl = sqrt(a_kk a_ll)
        for (int k = 0; k < mix->ns; ++k) {
            const double b = mix->Tcrit[k] * mix->R / (8 * mix->Pcrit[k]);
            P.bkl[k][k] = b;
            P.akl[k][k] = mix->Pcrit[k] * b * b * 27;
        }
        for (int k = 0; k < mix->ns; ++k)
            for (int l = 0; l < mix->ns; ++l)
                if (k != l) {
                    P.bkl[k][l] = std::pow(0.5 * (std::cbrt(P.bkl[k][k]) + std::cbrt(P.bkl[l][l])),
                                           3.0);
                    P.akl[k][l] = std::sqrt(P.akl[k][k] * P.akl[l][l]);
                }
    }
    if (mix->gas_model == SCHEMI_GAS_REDLICH_KWONG) {
        // mixtureRedlichKwong.cpp:19-67: chi = cbrt(2) - 1 (a constant
        // expression in the reference, folded to the correctly rounded
        // 0x1.428a2f98d728bp+0), Sigma_a = 1 / (9 chi), Sigma_b = chi / 3,
        // b_kk = Sigma_b R Tc / Pc, a_kk = Sigma_a R R Tc Tc sqrt(Tc) / Pc,
        // b_kl = (0.5 (cbrt b_kk + cbrt b_ll))^3, a_kl = sqrt(a_kk a_ll)
        const double chi = 0x1.428a2f98d728bp+0 - 1.;
        const double Sa = 1. / (9. * chi), Sb = chi / 3.;
        for (int k = 0; k < mix->ns; ++k) {
            const double R = mix->R, Tc = mix->Tcrit[k], Pc = mix->Pcrit[k];
            P.bkl[k][k] = Sb * R * Tc / Pc;
            P.akl[k][k] = Sa * R * R * Tc * Tc * std::sqrt(Tc) / Pc;
        }
        for (int k = 0; k < mix->ns; ++k)
            for (int l = 0; l < mix->ns; ++l)
                if (k != l) {
                    P.bkl[k][l] = std::pow(0.5 * (std::cbrt(P.bkl[k][k]) + std::cbrt(P.bkl[l][l])),
                                           3.0);
                    P.akl[k][l] = std::sqrt(P.akl[k][k] * P.akl[l][l]);
                }
    }
    if (mix->gas_model == SCHEMI_GAS_KATAOKA) {
        // mixtureKataokaVanDerWaals.cpp:19-77: b_kk from the critical point or
        // sigma, V0_kk = sigma^3 / sqrt(2) N_A, eps_kk = eps; across species
        // (0.5 (cbrt x_kk + cbrt x_ll))^3 for b and V0, sqrt(eps_kk eps_ll)
        const double NA = 6.02214076E23;
        for (int k = 0; k < mix->ns; ++k) {
            const double s3 = mix->sigmaLJ[k] * mix->sigmaLJ[k] * mix->sigmaLJ[k];
            P.bkl[k][k] = mix->kataoka_b_critical
                              ? mix->kataoka_b_coeff * mix->R * mix->Tcrit[k] / (8 * mix->Pcrit[k])
                              : mix->kataoka_b_coeff * s3 * NA;
            P.akl[k][k] = s3 / std::sqrt(2.0) * NA;
            P.ekl[k][k] = mix->epsLJ[k];
        }
        for (int k = 0; k < mix->ns; ++k)
            for (int l = 0; l < mix->ns; ++l)
                if (k != l) {
                    P.bkl[k][l] = std::pow(0.5 * (std::cbrt(P.bkl[k][k]) + std::cbrt(P.bkl[l][l])),
                                           3.0);
                    P.akl[k][l] = std::pow(0.5 * (std::cbrt(P.akl[k][k]) + std::cbrt(P.akl[l][l])),
                                           3.0);
                    P.ekl[k][l] = std::sqrt(P.ekl[k][k] * P.ekl[l][l]);
                }
    }
    if (mix->gas_model == SCHEMI_GAS_STIFFENED) {
        // mixtureStiffened.cpp:19-43: p0_kl and gamma_kl, geometric means off
        // the diagonal; kept in the a and b matrices
        for (int k = 0; k < mix->ns; ++k) {
            P.akl[k][k] = mix->p0[k];
            P.bkl[k][k] = mix->gamma[k];
        }
        for (int k = 0; k < mix->ns; ++k)
            for (int l = 0; l < mix->ns; ++l)
                if (k != l) {
                    P.akl[k][l] = std::sqrt(P.akl[k][k] * P.akl[l][l]);
                    P.bkl[k][l] = std::sqrt(P.bkl[k][k] * P.bkl[l][l]);
                }
    }
    for (int k = 0; k < mix->ns; ++k)
        for (int l = 0; l < mix->ns; ++l) {
            P.aklf[k][l] = static_cast<float>(P.akl[k][l]);
            P.bklf[k][l] = static_cast<float>(P.bkl[k][l]);
        }
    P.gravity = scheme->gravity ? 1 : 0;
    for (int b = 0; b < 3; ++b)
        P.g[b] = scheme->g[b];
    P.courant = scheme->courant;
    P.min_time = scheme->min_time;
    P.ts_source = scheme->timestep_source;
    P.ts_coeff = scheme->source_coeff;

    // boundary conditions; periodic must pair both patches of an axis for
    // every field
    bool periodic[3] = {false, false, false};
    for (int p = 0; p < 6; ++p) {
        for (int k = 0; k < mix->ns; ++k) {
            P.bct[p][k] = scheme->bc_c[k][p].type;
            P.bcv[p][k][0] = scheme->bc_c[k][p].value[0];
        }
        P.bct[p][mix->ns] = scheme->bc_v[p].type;
        for (int b = 0; b < 3; ++b)
            P.bcv[p][mix->ns][b] = scheme->bc_v[p].value[b];
        P.bct[p][mix->ns + 1] = scheme->bc_p[p].type;
        P.bcv[p][mix->ns + 1][0] = scheme->bc_p[p].value[0];
    }
    for (int a = 0; a < dim; ++a) {
        const int plo = patch_of(a, 0), phi = patch_of(a, 1);
        int nper = 0, nfield = 0;
        for (int q = 0; q < mix->ns + 2; ++q) {
            for (int p : {plo, phi}) {
                const int t = P.bct[p][q];
                if (!bc_ok(t)) {
                    delete c;
                    return SCHEMI_ERR_BOUNDARY;
                }
                if (t == SCHEMI_BC_FIXED_SURFACE)
                    c->pending_error = SCHEMI_ERR_BOUNDARY;
                nper += (t == SCHEMI_BC_PERIODIC);
                ++nfield;
            }
        }
        if (c->turb)
            for (int f = 0; f < (c->bfield ? 4 : c->afield ? 3 : 2); ++f)
                for (int p : {plo, phi}) {
                    const int t = f < 2 ? c->tp.bct[p][f] : f == 2 ? c->tp.bcta[p] : c->tp.bctb[p];
                    if (!bc_ok(t)) {
                        delete c;
                        return SCHEMI_ERR_BOUNDARY;
                    }
                    if (t == SCHEMI_BC_FIXED_SURFACE)
                        c->pending_error = SCHEMI_ERR_BOUNDARY;
                    nper += (t == SCHEMI_BC_PERIODIC);
                    ++nfield;
                }
        if (nper != 0 && nper != nfield) {
            delete c;
            return SCHEMI_ERR_BOUNDARY;
        }
        periodic[a] = nper != 0;
        P.periodic[a] = periodic[a] ? 1 : 0;
    }
    for (int a = 0; a < 3; ++a)
        P.zero_entry[a] = (dim >= 2) ? 1 : 0;

    c->tables = build_tables(dim, mesh->n, mesh->extent, mesh->origin, periodic);
    const bool separable = c->tables.separable;
    // orientation of the high boundary faces: the reference points a face
    // normal away from its owner's centre (mesh.cpp:34-53); the high plane
    // x_f(N) = o + L can lie below the last cell centre along z in 3-D, where
    // the vertices are spaced by Ly / Nz.  A face through the centre (no
    // orientation) and a periodic axis with a reversed face are declined.
    for (int a = 0; a < dim; ++a) {
        const AxisTables &T = c->tables.ax[a];
        const double d = T.xf[T.N] - T.xc[T.N - 1];
        if (d == 0.0 || (d < 0.0 && periodic[a])) {
            delete c;
            return SCHEMI_ERR_MESH;
        }
        P.hi_flip[a] = d < 0.0 ? 1 : 0;
    }
    // Kernel family.  In 2-D and 3-D the TVD ratio along an axis has an exact
    // zero entry from the neighbours across another axis (their centres share
    // the coordinate; every axis has at least two cells, checked above), so
    // phi = 0 and the slopes vanish: first-order kernel, unless the full
    // limiter evaluation is requested.  The linear limiter keeps its slopes;
    // on non-separable meshes its face offsets have transverse components.
    const char *full = std::getenv("SCHEMI_GPU_FULL_LIMITER");
    const bool force_slopes = full && full[0] == '1' && !c->hancock;
    if (lim == SCHEMI_LIM_LINEAR)
        c->mode = (dim >= 2 && (!separable || c->hancock)) ? MODE_GENERAL : MODE_SLOPES;
    else if (lim == SCHEMI_LIM_ZERO || dim >= 2)
        c->mode = force_slopes ? MODE_SLOPES : MODE_FIRST_ORDER;
    else
        c->mode = MODE_SLOPES;
    // #Reconstruction star (diffusion stage): star_kernel serves the meshes
    // whose face values are the cell values and 1-D TVD slopes.  After a
    // MUSCL-Hancock step under gravity the reference replaces the star
    // velocity by the interpolated one, which is the linear stage of a laminar
    // state; the other Hancock cases are not built.
    if (c->diffusion && scheme->reconstruction == SCHEMI_RECONSTRUCTION_STAR) {
        if (c->hancock) {
            if (!scheme->gravity || c->turb) {
                delete c;
                return SCHEMI_ERR_INIT;
            }
        } else if (lim == SCHEMI_LIM_LINEAR || c->mode == MODE_GENERAL) {
            delete c;
            return SCHEMI_ERR_INIT;
        } else {
            c->star = true;
        }
    }
    if (c->star) {
        long long off = 0;
        for (int a = 0; a < 3; ++a) {
            P.star_off[a] = off;
            if (a < dim)
                off += (long long)(P.N[a] + 1) * (P.n / P.N[a]);
        }
        P.star_nft = off;
    }
    // Mixed path of the ideal mixture with SSP-RK2 where the slopes vanish:
    // the lean kernel, when the mesh splits into whole 64 x SG_LEAN_TY tiles
    // and every wall is free or slip (SCHEMI_GPU_LEAN=0 keeps stage0_kernel)
    {
        const char *lean = std::getenv("SCHEMI_GPU_LEAN");
        bool walls = true;
        for (int a = 0; a < dim; ++a)
            for (int side = 0; side < 2 && !periodic[a]; ++side)
                for (int q = 0; q < mix->ns + 2; ++q) {
                    const int t = P.bct[patch_of(a, side)][q];
                    walls = walls && (t == BC_FREE || t == BC_SLIP);
                }
        if (c->mode == MODE_FIRST_ORDER && c->prec == SCHEMI_PREC_MIXED && !c->turb &&
            mix->gas_model == SCHEMI_GAS_IDEAL && !c->hancock && dim >= 2 && walls &&
            P.N[0] % 64 == 0 && P.N[1] % SG_LEAN_TY == 0 && !(lean && lean[0] == '0'))
            c->mode = MODE_LEAN;
    }
    // MUSCL-Hancock: the slope kernel in 1-D, the first-order kernel where
    // the slopes vanish, the general-geometry kernels for the linear limiter
    // in 2-D/3-D (every slot's predicted side is stored)
    // periodic axes need whole tiles so the wrapped neighbours are staged
    {
        const bool fo = c->mode == MODE_FIRST_ORDER;
        const int tile[2] = {dim == 1 ? TileShape<1>::TX
                                      : (fo ? TileShape<2, true>::TX : TileShape<2>::TX),
                             fo ? TileShape<2, true>::TY : TileShape<2>::TY};
        for (int a = 0; a < 2 && a < dim; ++a)
            if (c->mode != MODE_GENERAL && !c->turb && periodic[a] && (P.N[a] % tile[a]) != 0) {
                delete c;
                return SCHEMI_ERR_INIT;
            }
    }
    // flatten the per-axis tables
    std::vector<double> geo;
    auto put = [&](const std::vector<double> &v) {
        const int off = static_cast<int>(geo.size());
        geo.insert(geo.end(), v.begin(), v.end());
        return off;
    };
    for (int a = 0; a < 3; ++a) {
        const AxisTables &T = c->tables.ax[a];
        P.off_xc[a] = put(T.xc);
        P.off_e[a] = put(T.e);
        P.off_xf[a] = put(T.xf);
        P.off_wo[a] = put(T.wo);
        P.off_wn[a] = put(T.wn);
        P.off_dn[a] = put(T.dn);
        P.off_dlo[a] = put(T.dlo);
        P.off_dhi[a] = put(T.dhi);
        P.ghost_lo[a] = T.ghost_lo;
        P.ghost_hi[a] = T.ghost_hi;
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int cls = 0; cls < 2; ++cls)
                P.off_td[a][b][cls] = (a == b) ? 0 : put(c->tables.td[a][b][cls]);
    c->kchunk = SG_KCHUNK;
    if (const char *kc = std::getenv("SCHEMI_GPU_KCHUNK")) {
        // z planes per workgroup of the 3-D kernels (launch shape only)
        const int v = std::atoi(kc);
        if (v >= 1 && v <= 4096)
            c->kchunk = v;
    }

    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return SCHEMI_ERR_SYSTEM;
    }
    if (noalloc_setup(c->device) != SCHEMI_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return SCHEMI_ERR_SYSTEM;
    }
    const size_t state_bytes = sizeof(double) * (size_t)c->nf * (size_t)c->n;
    bool ok = hipMalloc(&c->dP, sizeof(DevParams)) == hipSuccess &&
              hipMalloc(&c->dgeo, geo.size() * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->U[0], state_bytes) == hipSuccess &&
              hipMalloc(&c->U[1], state_bytes) == hipSuccess &&
              hipMalloc(&c->sc, SC_N * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->derr, sizeof(int)) == hipSuccess;
    if (ok && c->diffusion)
        ok = hipMalloc(&c->Wd, sizeof(double) *
                                   (size_t)(2 * mix->ns + (c->bfield   ? 16
                                                           : c->afield ? 13
                                                           : c->turb   ? 9
                                                                       : 5)) *
                                   (size_t)c->n) == hipSuccess;
    if (ok && c->afield)
        ok = hipMalloc(&c->db2, sizeof(double) * (size_t)c->n) == hipSuccess;
    if (ok && c->star)
        ok = hipMalloc(&c->dS, sizeof(double) *
                                   (size_t)(5 + (c->bfield ? 4 : c->afield ? 3 : 0)) *
                                   (size_t)P.star_nft) == hipSuccess;
    if (ok && (c->turb || c->star))
        ok = hipMalloc(&c->dTP, sizeof(TurbParams)) == hipSuccess &&
             hipMalloc(&c->epsmax, sizeof(unsigned long long)) == hipSuccess &&
             hipMemcpyAsync(c->dTP, &c->tp, sizeof(TurbParams), hipMemcpyHostToDevice,
                            c->stream) == hipSuccess &&
             hipMemsetAsync(c->epsmax, 0, sizeof(unsigned long long), c->stream) == hipSuccess;
    if (ok && c->sle_iters > 0)
        ok = hipMalloc(&c->Mb, sizeof(double) * diffusion_matrix_doubles(mix->ns) *
                                   (size_t)c->n) == hipSuccess &&
             hipMalloc(&c->dmax, sizeof(double) * (size_t)(mix->ns + 4) *
                                     (size_t)(c->sle_iters + 2)) == hipSuccess &&
             hipMalloc(&c->done, sizeof(int) * (size_t)(mix->ns + 4)) == hipSuccess;
    if (ok && c->mode == MODE_GENERAL) {
        const size_t tsz = (c->prec == SCHEMI_PREC_MIXED) ? sizeof(float) : sizeof(double);
        const size_t nq = (size_t)(mix->ns + 4) * (size_t)c->n;
        const size_t nh = c->hancock ? (size_t)(2 * dim) * (size_t)(mix->ns + 10) * (size_t)c->n
                                     : 0;
        ok = hipMalloc(&c->scratch, tsz * (nq * (size_t)(1 + dim) + nh)) == hipSuccess;
        if (ok) {
            c->W = c->scratch;
            c->Lg = static_cast<char *>(c->scratch) + tsz * nq;
            c->Hs = static_cast<char *>(c->scratch) + tsz * nq * (size_t)(1 + dim);
        }
    }
    ok = ok &&
         hipMemcpyAsync(c->dP, &P, sizeof P, hipMemcpyHostToDevice, c->stream) == hipSuccess &&
         hipMemcpyAsync(c->dgeo, geo.data(), geo.size() * sizeof(double),
                        hipMemcpyHostToDevice, c->stream) == hipSuccess &&
         hipMemsetAsync(c->derr, 0, sizeof(int), c->stream) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, 0.0, 0.0,
                           0.0);
        ok = hipStreamSynchronize(c->stream) == hipSuccess;
    }
    if (!ok) {
        schemi_gpu_destroy(c);
        return SCHEMI_ERR_SYSTEM;
    }
    *out = c;
    return SCHEMI_OK;
}

void schemi_gpu_destroy(schemi_gpu_ctx *c)
{
    if (!c)
        return;
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev)
        (void)hipEventDestroy(e);
    for (hipEvent_t e : c->evd)
        (void)hipEventDestroy(e);
    (void)hipFree(c->dP);
    (void)hipFree(c->dgeo);
    (void)hipFree(c->U[0]);
    (void)hipFree(c->U[1]);
    (void)hipFree(c->sc);
    (void)hipFree(c->derr);
    (void)hipFree(c->scratch);
    (void)hipFree(c->adj);
    (void)hipFree(c->Wd);
    (void)hipFree(c->Mb);
    (void)hipFree(c->dmax);
    (void)hipFree(c->done);
    (void)hipFree(c->dTP);
    (void)hipFree(c->epsmax);
    (void)hipFree(c->dS);
    (void)hipFree(c->db2);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

int schemi_gpu_set_state(schemi_gpu_ctx *c, const double *cons, int64_t ncells,
                         int32_t numbering)
{
    if (!c || !cons || ncells != c->n || numbering != SCHEMI_NUMBERING_REFERENCE)
        return SCHEMI_ERR_INIT;
    const size_t bytes = sizeof(double) * (size_t)c->nf * (size_t)c->n;
    HIPCHK(hipMemcpyAsync(c->U[0], cons, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->derr, 0, sizeof(int), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->time = 0.0;
    c->stage_turb = false;
    return SCHEMI_OK;
}

int schemi_gpu_get_state(schemi_gpu_ctx *c, double *cons, double *prim, int32_t numbering)
{
    if (!c || numbering != SCHEMI_NUMBERING_REFERENCE)
        return SCHEMI_ERR_INIT;
    const size_t bytes = sizeof(double) * (size_t)c->nf * (size_t)c->n;
    if (cons)
        HIPCHK(hipMemcpyAsync(cons, c->U[0], bytes, hipMemcpyDeviceToHost, c->stream));
    callBuffers buf(c->stream);
    if (prim) {
        double *dW = nullptr;
        const size_t pbytes = sizeof(double) * (size_t)(c->ns + 8) * (size_t)c->n;
        HIPCHK(buf.get(&dW, pbytes));
        const int bs = 256;
        const dim3 g((unsigned)((c->n + bs - 1) / bs));
#define SG_PRIMS(NS)                                                                      \
    do {                                                                                  \
        if (c->turb)                                                                      \
            hipLaunchKernelGGL((prims_kernel<NS, true>), g, dim3(bs), 0, c->stream, c->dP, \
                               c->U[0], dW, c->n);                                        \
        else                                                                              \
            hipLaunchKernelGGL((prims_kernel<NS, false>), g, dim3(bs), 0, c->stream, c->dP, \
                               c->U[0], dW, c->n);                                        \
    } while (0)
        switch (c->ns) {
        case 1:
            SG_PRIMS(1);
            break;
        case 2:
            SG_PRIMS(2);
            break;
        default:
            SG_PRIMS(3);
            break;
        }
#undef SG_PRIMS
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(prim, dW, pbytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCHEMI_OK;
}

int schemi_gpu_get_turbulence_stage(schemi_gpu_ctx *c, double *out, int32_t numbering)
{
    if (!c || !out || numbering != SCHEMI_NUMBERING_REFERENCE || !c->turb || !c->stage_turb)
        return SCHEMI_ERR_INIT;
    // Wd rows k and epsilon of the stage's primitives (DW<NS>::K, EPS)
    const size_t row = sizeof(double) * (size_t)c->n;
    const size_t k = (size_t)(2 * c->ns + 5) * (size_t)c->n;
    HIPCHK(hipMemcpyAsync(out, c->Wd + k, 2 * row, hipMemcpyDeviceToHost, c->stream));
    if (c->afield && !c->bfield)
        HIPCHK(hipMemcpyAsync(out + 2 * c->n, c->db2, row, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCHEMI_OK;
}

int schemi_gpu_step(schemi_gpu_ctx *c, double dt, double *dt_next)
{
    if (!c)
        return SCHEMI_ERR_INIT;
    if (c->pending_error)
        return c->pending_error;
    hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, dt, c->time,
                       HUGE_VAL);
    HIPCHK(hipGetLastError());
    int r = one_step(c);
    if (!r)
        r = restore_state(c);
    if (r)
        return r;
    double sc[SC_N];
    HIPCHK(hipMemcpyAsync(sc, c->sc, sizeof sc, hipMemcpyDeviceToHost, c->stream));
    r = read_error(c);
    c->time = sc[SC_TIME];
    if (dt_next)
        *dt_next = sc[SC_DTNEXT];
    return r;
}

int schemi_gpu_run(schemi_gpu_ctx *c, int64_t nsteps, double dt0, double t_end, double *t_out,
                   int64_t *steps_out, double *dt_out)
{
    if (!c || nsteps < 0)
        return SCHEMI_ERR_INIT;
    if (c->pending_error)
        return c->pending_error;
    // first step: min(dt0, t_end - t) (Schemi.cpp:665-670)
    double dt = dt0;
    if (c->time + dt > t_end)
        dt = t_end - c->time;
    hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, dt, c->time,
                       t_end);
    HIPCHK(hipGetLastError());
    for (int64_t s = 0; s < nsteps; ++s) {
        const int r = one_step(c);
        if (r)
            return r;
    }
    if (const int r = restore_state(c))
        return r;
    double sc[SC_N];
    HIPCHK(hipMemcpyAsync(sc, c->sc, sizeof sc, hipMemcpyDeviceToHost, c->stream));
    const int r = read_error(c);
    c->time = sc[SC_TIME];
    if (t_out)
        *t_out = sc[SC_TIME];
    if (steps_out)
        *steps_out = static_cast<int64_t>(sc[SC_STEPS]);
    if (dt_out)
        *dt_out = sc[SC_DT];
    return r;
}

int schemi_gpu_stage(schemi_gpu_ctx *c, double dt)
{
    if (!c)
        return SCHEMI_ERR_INIT;
    if (c->pending_error)
        return c->pending_error;
    hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, dt, c->time,
                       HUGE_VAL);
    HIPCHK(hipGetLastError());
    int r = launch_stage(c, c->U[0], c->U[1], nullptr, c->hancock ? SV_HANCOCK : SV_RK1);
    if (r)
        return r;
    const size_t bytes = sizeof(double) * (size_t)c->nf * (size_t)c->n;
    HIPCHK(hipMemcpyAsync(c->U[0], c->U[1], bytes, hipMemcpyDeviceToDevice, c->stream));
    return read_error(c);
}

int schemi_gpu_face_flux(schemi_gpu_ctx *c, int64_t nfaces, const int32_t *axis,
                         const double *sgn, const double *left, const double *right,
                         double *flux)
{
    if (!c || nfaces < 0 || (nfaces > 0 && (!axis || !sgn || !left || !right || !flux)))
        return SCHEMI_ERR_INIT;
    if (nfaces == 0)
        return SCHEMI_OK;
    for (int64_t f = 0; f < nfaces; ++f)
        if (axis[f] < 0 || axis[f] > 2)
            return SCHEMI_ERR_INIT;
    const int NQ = c->ns + 4, NF = c->ns + 5;
    int *dax = nullptr;
    double *dsg = nullptr, *dL = nullptr, *dR = nullptr, *dF = nullptr;
    callBuffers buf(c->stream);
    HIPCHK(buf.get(&dax, nfaces * sizeof(int)));
    HIPCHK(buf.get(&dsg, nfaces * sizeof(double)));
    HIPCHK(buf.get(&dL, nfaces * NQ * sizeof(double)));
    HIPCHK(buf.get(&dR, nfaces * NQ * sizeof(double)));
    HIPCHK(buf.get(&dF, nfaces * NF * sizeof(double)));
    HIPCHK(hipMemcpyAsync(dax, axis, nfaces * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dsg, sgn, nfaces * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dL, left, nfaces * NQ * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dR, right, nfaces * NQ * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->derr, 0, sizeof(int), c->stream));
    const int bs = 256;
    const dim3 g((unsigned)((nfaces + bs - 1) / bs));
#define SG_FF(T, NS)                                                                    \
    hipLaunchKernelGGL((face_flux_kernel<T, NS>), g, dim3(bs), 0, c->stream, c->dP,    \
                       (long long)nfaces, dax, dsg, dL, dR, dF, c->derr)
    if (c->prec == SCHEMI_PREC_MIXED) {
        if (c->ns == 1) SG_FF(float, 1); else if (c->ns == 2) SG_FF(float, 2); else SG_FF(float, 3);
    } else {
        if (c->ns == 1) SG_FF(double, 1); else if (c->ns == 2) SG_FF(double, 2); else SG_FF(double, 3);
    }
#undef SG_FF
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(flux, dF, nfaces * NF * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return read_error(c);
}

int schemi_gpu_device_state(schemi_gpu_ctx *c, double **dev_ptr, int64_t *ncells)
{
    if (!c)
        return SCHEMI_ERR_INIT;
    if (dev_ptr)
        *dev_ptr = c->U[0];
    if (ncells)
        *ncells = c->n;
    return SCHEMI_OK;
}

int schemi_gpu_state_device(schemi_gpu_ctx *c, const double *src, double *dst)
{
    if (!c)
        return SCHEMI_ERR_INIT;
    const size_t bytes = sizeof(double) * (size_t)c->nf * (size_t)c->n;
    if (src)
        HIPCHK(hipMemcpyAsync(c->U[0], src, bytes, hipMemcpyDeviceToDevice, c->stream));
    if (dst)
        HIPCHK(hipMemcpyAsync(dst, c->U[0], bytes, hipMemcpyDeviceToDevice, c->stream));
    return read_error(c);
}

int schemi_gpu_adjoint_step(schemi_gpu_ctx *c, double dt, const double *lam_in, double *lam_out,
                            int32_t what)
{
    if (!c || !lam_in || !lam_out || (what != SCHEMI_ADJ_STEP && what != SCHEMI_ADJ_STAGE))
        return SCHEMI_ERR_INIT;
    const bool kernel = c->mode == MODE_FIRST_ORDER || c->mode == MODE_LEAN ||
                        (c->mode == MODE_SLOPES && c->dim == 1);
    if (!kernel || c->hancock || c->diffusion || c->turb || c->hp.gas_model == GAS_RK ||
        c->hp.gas_model == GAS_KATAOKA)
        return SCHEMI_ERR_INIT;
    if (c->pending_error)
        return c->pending_error;
    if (const int r = ensure_adj(c))
        return r;
    if (what == SCHEMI_ADJ_STAGE) {
        // lam_out = (dL/dU)^T lam_in at U^n
        if (const int r = adjoint_stage(c, c->U[0], lam_in, c->adj, dt, 1.0, nullptr, 0.0))
            return r;
        const size_t bytes = sizeof(double) * (size_t)(c->ns + 5) * (size_t)c->n;
        HIPCHK(hipMemcpyAsync(lam_out, c->adj, bytes, hipMemcpyDeviceToDevice, c->stream));
        return read_error(c);
    }
    // U^1 = L(U^n) recomputed into U[1]; U^{n+1} = (L(U^1) + U^n) / 2, so
    // lam^n = (1/2) (dL/dU|_{U^n})^T (dL/dU|_{U^1})^T lam^{n+1} + (1/2) lam^{n+1}
    hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, dt, c->time,
                       HUGE_VAL);
    HIPCHK(hipGetLastError());
    if (const int r = launch_stage(c, c->U[0], c->U[1], nullptr, SV_RK1))
        return r;
    if (const int r = adjoint_stage(c, c->U[1], lam_in, c->adj, dt, 1.0, nullptr, 0.0))
        return r;
    if (const int r = adjoint_stage(c, c->U[0], c->adj, lam_out, dt, 0.5, lam_in, 0.5))
        return r;
    return read_error(c);
}

int schemi_gpu_mixing_width(schemi_gpu_ctx *c, int32_t species, double *J, double *dJdU)
{
    if (!c || !J || species < 0 || species >= c->ns)
        return SCHEMI_ERR_INIT;
    if (const int r = ensure_adj(c))
        return r;
    double *partial = c->adj + (size_t)(c->ns + 5) * (size_t)c->n;
    double *dJ = partial + (c->n + 255) / 256;
    const int r = launch_mixing_width(c->dP, c->dgeo, c->U[0], c->ns, c->n, species, partial, dJ,
                                      dJdU, c->stream);
    if (r)
        return r == -1 ? SCHEMI_ERR_INIT : SCHEMI_ERR_SYSTEM;
    HIPCHK(hipMemcpyAsync(J, dJ, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return read_error(c);
}

int schemi_gpu_mixing_width_noalloc(schemi_gpu_ctx *c, int32_t species, double *J,
                                    double *dJdU)
{
    if (!c || !J || species < 0 || species >= c->ns)
        return SCHEMI_ERR_INIT;
    if (c->n > SCHEMI_NOALLOC_MAX_CELLS)
        return SCHEMI_ERR_SCRATCH;
    double *base = (c->device >= 0 && c->device < NOALLOC_DEVICES) ? noalloc_ptr[c->device]
                                                                   : nullptr;
    if (!base)
        return SCHEMI_ERR_SYSTEM;
    // every partial slot and the result are written before they are read:
    // nothing carried over from an earlier call is used
    double *partial = base;
    double *dJ = base + (c->n + 255) / 256;
    const int r = launch_mixing_width(c->dP, c->dgeo, c->U[0], c->ns, c->n, species, partial, dJ,
                                      dJdU, c->stream);
    if (r)
        return r == -1 ? SCHEMI_ERR_INIT : SCHEMI_ERR_SYSTEM;
    HIPCHK(hipMemcpyAsync(J, dJ, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return read_error(c);
}

int schemi_gpu_source_timestep(schemi_gpu_ctx *c, double *dt_source)
{
    if (!c || !dt_source)
        return SCHEMI_ERR_INIT;
    HIPCHK(hipMemcpyAsync(dt_source, c->sc + SC_SRCDT, sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCHEMI_OK;
}

int schemi_gpu_sync(schemi_gpu_ctx *c)
{
    if (!c)
        return SCHEMI_ERR_INIT;
    return read_error(c);
}

int schemi_gpu_timing(schemi_gpu_ctx *c, int32_t enable, double *avg_ms, int64_t *launches)
{
    if (!c)
        return SCHEMI_ERR_INIT;
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0;
    int64_t cnt = 0;
    for (size_t q = 0; q + 1 < c->ev.size(); q += 2) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev[q], c->ev[q + 1]) == hipSuccess) {
            tot += ms;
            ++cnt;
        }
    }
    if (avg_ms)
        *avg_ms = cnt ? tot / cnt : 0.0;
    if (launches)
        *launches = cnt;
    for (hipEvent_t e : c->ev)
        (void)hipEventDestroy(e);
    c->ev.clear();
    c->timing = enable != 0;
    return SCHEMI_OK;
}

int schemi_gpu_diffusion_timing(schemi_gpu_ctx *c, double *prims_ms, double *update_ms,
                                int64_t *launches)
{
    if (!c)
        return SCHEMI_ERR_INIT;
    HIPCHK(hipStreamSynchronize(c->stream));
    double t0 = 0, t1 = 0;
    int64_t cnt = 0;
    for (size_t q = 0; q + 2 < c->evd.size(); q += 3) {
        float a = 0, b = 0;
        if (hipEventElapsedTime(&a, c->evd[q], c->evd[q + 1]) == hipSuccess &&
            hipEventElapsedTime(&b, c->evd[q + 1], c->evd[q + 2]) == hipSuccess) {
            t0 += a;
            t1 += b;
            ++cnt;
        }
    }
    if (prims_ms)
        *prims_ms = cnt ? t0 / cnt : 0.0;
    if (update_ms)
        *update_ms = cnt ? t1 / cnt : 0.0;
    if (launches)
        *launches = cnt;
    for (hipEvent_t e : c->evd)
        (void)hipEventDestroy(e);
    c->evd.clear();
    return SCHEMI_OK;
}

int schemi_gpu_face_permutation(const schemi_mesh_desc *mesh, int64_t *ref_to_gpu,
                                int64_t nfaces)
{
    if (!mesh || !ref_to_gpu || mesh->dim < 1 || mesh->dim > 3)
        return SCHEMI_ERR_INIT;
    const std::vector<int64_t> p = face_permutation(mesh->dim, mesh->n);
    if ((int64_t)p.size() != nfaces)
        return SCHEMI_ERR_INIT;
    std::memcpy(ref_to_gpu, p.data(), p.size() * sizeof(int64_t));
    return SCHEMI_OK;
}

int schemi_gpu_axis_table(const schemi_mesh_desc *mesh, int32_t periodic_mask, int32_t axis,
                          int32_t which, double *out, int64_t *len)
{
    if (!mesh || !len || mesh->dim < 1 || mesh->dim > 3 || axis < 0 || axis > 2 ||
        which < 0 || which > 15)
        return SCHEMI_ERR_INIT;
    const bool per[3] = {(periodic_mask & 1) != 0, (periodic_mask & 2) != 0,
                         (periodic_mask & 4) != 0};
    const MeshTables M = build_tables(mesh->dim, mesh->n, mesh->extent, mesh->origin, per);
    const AxisTables &T = M.ax[axis];
    std::vector<double> v;
    switch (which) {
    case 0: v = T.xc; break;
    case 1: v = T.e; break;
    case 2: v = T.xf; break;
    case 3: v = T.wo; break;
    case 4: v = T.wn; break;
    case 5: v = T.dn; break;
    case 6: v = T.dlo; break;
    case 7: v = T.dhi; break;
    case 8: v = {T.ghost_lo, T.ghost_hi}; break;
    case 9: v = {M.separable ? 1.0 : 0.0}; break;
    default: {
        const int b = (which - 10) / 2, cls = (which - 10) % 2;
        if (b != axis)
            v = M.td[axis][b][cls];
        break;
    }
    }
    *len = static_cast<int64_t>(v.size());
    if (out)
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    return SCHEMI_OK;
}

}  // extern "C"
