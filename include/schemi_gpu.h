/*
 * schemi_gpu.h -- C ABI of the MI355X explicit finite-volume step.
 *
 * The library replaces the reference solver's explicit SSP-RK2 step
 * (secondOrderStepSolverRK::calculateStep,
 *  Schemi/src/solverStages/secondOrderStepSolverRK.cpp:48-120, created by
 *  abstractStepSolver::createStepSolver, Schemi/src/interfaces/abstractStepSolver.cpp:56-146)
 * together with the pieces it drives: Advection2dOrder
 * (Schemi/src/Advections/Advection2dOrder.cpp:16-636), the HLL/HLLC flux
 * (Schemi/src/RiemannSolvers/HLLSolver.cpp, HLLCSolver.cpp), the TVD limiter
 * (Schemi/src/limiters/TVDLimiter.cpp), the ideal-mixture closure
 * (Schemi/src/thermodynamics/mixtureIdeal.cpp) and the Courant time step of the
 * main loop (Schemi/main/Schemi.cpp:807-868).
 *
 * Conventions
 *  - Plain C types only.  Host arrays are borrowed for the duration of a call.
 *  - Conserved state is structure-of-arrays, Ns + 5 fields of ncells doubles:
 *      rho_1 .. rho_Ns, rho, m_x, m_y, m_z, E
 *    (partial densities, total density evolved separately, momentum, total
 *    energy; Schemi/templates/bunchOfFields.hpp:28-45).
 *  - Primitive output is Ns + 1 + 3 + 4 fields:
 *      c_1 .. c_Ns, c, v_x, v_y, v_z, p, U, T, sound speed
 *  - Cell order: SCHEMI_NUMBERING_REFERENCE is the reference's
 *    id = i + Nx (j + Ny k) (Schemi/src/dataStructures/mesh.cpp:1223-1230), which
 *    is also the device layout (x fastest: coalesced owner/neighbour gathers
 *    along x, constant-stride gathers along y and z).
 *  - Every function returns an errors code with the reference's values
 *    (Schemi/enumerations/errorsEnum.hpp:15-30): 0 = noErrors.
 *  - One context owns its device buffers and one HIP stream; calls on one
 *    context are not re-entrant.
 */
#ifndef SCHEMI_GPU_H
#define SCHEMI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCHEMI_MAX_SPECIES 9

/* errors (Schemi/enumerations/errorsEnum.hpp:15-30) */
#define SCHEMI_OK 0
#define SCHEMI_ERR_RIEMANN 1
#define SCHEMI_ERR_NEGATIVE_TEMPERATURE 2
#define SCHEMI_ERR_SYSTEM 4
#define SCHEMI_ERR_BOUNDARY 5
#define SCHEMI_ERR_MESH 10
#define SCHEMI_ERR_INIT 11
/* no counterpart in the reference: the context has more cells than the
 * no-allocation scratch of schemi_gpu_mixing_width_noalloc holds */
#define SCHEMI_ERR_SCRATCH 100

/* boundaryConditionType (boundaryConditionTypesEnum.hpp:15-34); PERIODIC has
 * no counterpart in the reference */
#define SCHEMI_BC_FREE 1
#define SCHEMI_BC_SLIP 2
#define SCHEMI_BC_FIXED_CELL 3
#define SCHEMI_BC_FIXED_SURFACE 4
#define SCHEMI_BC_PERIODIC 32

/* flowSolver (flowSolverEnum.hpp:15-18) */
#define SCHEMI_FLUX_HLL 0
#define SCHEMI_FLUX_HLLC 2

/* typeOfTVDLimiter (typeOfTVDLimiterEnum.hpp:15-32) */
#define SCHEMI_LIM_ZERO 0
#define SCHEMI_LIM_MINMOD 1
#define SCHEMI_LIM_VANLEER 2
#define SCHEMI_LIM_LINEAR 3
#define SCHEMI_LIM_VANALBADA 4

/* precision of the reconstruction / limiter / Riemann solve */
#define SCHEMI_PREC_FP64 0  /* every operation in FP64, reference order */
#define SCHEMI_PREC_MIXED 1 /* FP32 flux path, FP64 accumulation and update */

#define SCHEMI_NUMBERING_REFERENCE 0

/* Patch order: 0 tail (x=0), 1 point (x=L), 2 bottom (z=0), 3 right (y=0),
 * 4 left (y=L), 5 top (z=L) (Schemi/main/Schemi.cpp:575-592). */
typedef struct {
    int32_t type;
    double value[3]; /* fixed value; scalars use value[0] */
} schemi_bc;

/* Mesh: the reference's parallelepiped generators
 * (mesh::oneD/twoD/threeDParallelepiped, mesh.cpp:296-3018). */
typedef struct {
    int32_t dim;       /* 1, 2, 3 */
    int64_t n[3];      /* cells per direction (unused ones = 1) */
    double extent[3];  /* radius-vector of the parallelepiped */
    double origin[3];  /* zero point */
} schemi_mesh_desc;

/* gasModel (Schemi/enumerations/gasModelEnum.hpp:15-18) */
#define SCHEMI_GAS_IDEAL 0
#define SCHEMI_GAS_VAN_DER_WAALS 1
#define SCHEMI_GAS_REDLICH_KWONG 2
#define SCHEMI_GAS_STIFFENED 3
#define SCHEMI_GAS_KATAOKA 4

/* Mixture (abstractMixtureThermodynamics::createThermodynamics,
 * Schemi/src/interfaces/abstractMixtureThermodynamics.cpp:32-63): molar
 * masses, molar Cv, and for the van der Waals (mixtureVanDerWaals.cpp:19-58)
 * and Redlich-Kwong (mixtureRedlichKwong.cpp:19-67) mixtures the critical
 * temperature and pressure of every species (ignored by the ideal mixture).
 * The Redlich-Kwong temperature is the root of a cubic in sqrt(T), solved
 * in FP64 by Newton's method on both precision paths (the reference solves
 * it in long double; results agree to a few ulps, not bit for bit). */
typedef struct {
    int32_t ns; /* 1 <= ns <= 3: the kernels are instantiated for up to three
                   species (SCHEMI_ERR_INIT otherwise); the arrays below are
                   sized for SCHEMI_MAX_SPECIES */
    double M[SCHEMI_MAX_SPECIES];
    double Cv[SCHEMI_MAX_SPECIES];
    double R;
    int32_t gas_model; /* SCHEMI_GAS_* */
    double Tcrit[SCHEMI_MAX_SPECIES];
    double Pcrit[SCHEMI_MAX_SPECIES];
    /* stiffened mixture (mixtureStiffened.cpp:19-43, read by the reference
     * from set/stiffenedFluidData.txt): p0 and gamma of every species */
    double p0[SCHEMI_MAX_SPECIES];
    double gamma[SCHEMI_MAX_SPECIES];
    /* Kataoka-van der Waals mixture (mixtureKataokaVanDerWaals.cpp:19-77,
     * read by the reference from set/KataokaVanDerWaalsFluidData.txt):
     * Lennard-Jones epsilon and sigma per species, and b from the critical
     * point (critical = 1: coeff R Tc / (8 Pc)) or from sigma (0: coeff
     * sigma^3 N_A).  Its temperature is found by the reference's Newton
     * iteration with the device's pow and exp (results a few ulps from the
     * reference's, not bit for bit). */
    double epsLJ[SCHEMI_MAX_SPECIES];
    double sigmaLJ[SCHEMI_MAX_SPECIES];
    int32_t kataoka_b_critical;
    double kataoka_b_coeff;
} schemi_mix_desc;

/* typeOfSolverEnum (Schemi/enumerations/typeOfSolverEnum.hpp:15):
 * SecondOrder is the MUSCL-Hancock step (secondOrderStepSolver.cpp:33-48),
 * SecondOrderRK the SSP-RK2 step (secondOrderStepSolverRK.cpp:48-71).
 * Both run with every limiter, flux, EOS and mesh here. */
#define SCHEMI_STEP_SECOND_ORDER 0
#define SCHEMI_STEP_SECOND_ORDER_RK 1

/* Scheme: flux, limiter, precision, gravity, time-step control, one
 * boundary condition per field (c_k, v, p) and patch, and the step solver. */
typedef struct {
    int32_t flux;
    int32_t limiter;
    int32_t precision;
    int32_t gravity;
    double g[3];
    double courant;
    double min_time;   /* #Minimal_timestep */
    double const_dt;   /* #Constant_timestep; <= 0 selects the Courant step */
    schemi_bc bc_c[SCHEMI_MAX_SPECIES][6];
    schemi_bc bc_v[6];
    schemi_bc bc_p[6];
    int32_t step_solver; /* SCHEMI_STEP_* */
    /* Molecular diffusion stage after each step (Diffusion.cpp:20-1061,
     * called at secondOrderStepSolverRK.cpp:102-109 and
     * secondOrderStepSolver.cpp:72-82; #Diffusion_flag on) with linear
     * reconstruction (#Reconstruction linear) and the constant transport
     * model of set/transportCoefficients.txt (viscosity, diffusion and heat
     * conduction coefficients); the matrix solver, molar-mass correction and
     * enthalpy flow below; no turbulence.  Ideal or van der Waals mixture
     * (on a periodic axis the wrapped
     * face is an inner face spanning the centre distance through the
     * boundary; no reference counterpart); otherwise
     * SCHEMI_ERR_INIT.  A negative temperature after the stage returns
     * SCHEMI_ERR_NEGATIVE_TEMPERATURE (negativeTemperatureError). */
    int32_t diffusion;
    double transport_mu;
    double transport_D;
    double transport_kappa;
    /* matrix solver of the stage (matrixSolverEnum.hpp:15-25 values):
     * SCHEMI_SLE_EXPLICIT, SCHEMI_SLE_JACOBI (JacobiSolver.cpp: implicit
     * Laplacians solved to the reference's tolerance, at most
     * diffusion_iterations iterations, #Number_of_iterations_of_SLE_matrix_solver;
     * the host waits for the solves once per batch of iterations), or
     * SCHEMI_SLE_GAUSS_SEIDEL: the same systems solved by the parallel Jacobi
     * sweeps to the same stop test (GaussSeidelSolver.cpp's sequential sweeps
     * have no parallel form with the same rounding; both converge to the
     * system's solution, so the results agree to the solver tolerance, not
     * bit for bit). */
    int32_t diffusion_solver;
    int32_t diffusion_iterations;
    /* #Flag_for_molar_mass_diffusion_correction on (1): the species fluxes
     * corrected to the mole-fraction gradients (caclulateDFluxes,
     * transportCoefficients.hpp:285-395, Diffusion.cpp:326-377);
     * #Flag_for_enthalpy_flow_solution explicit (1): the enthalpy carried by
     * the species fluxes (Diffusion.cpp:379-421, 605-635; explicit matrix
     * solver and ideal mixture only, else SCHEMI_ERR_INIT). */
    int32_t diffusion_molar_mass;
    int32_t diffusion_enthalpy_flow;
    /* #Source_timestep_correction of main.txt (Schemi.cpp:190-245):
     * timestep_source 0 "Courant", 1 "Courant-Source" (the Courant step is
     * capped by the source step of the previous diffusion stage,
     * Schemi.cpp:828-833, which the explicit enthalpy flow sets to
     * source_coeff * min(C Cv T / |div enthalpy flow + stab|),
     * Diffusion.cpp:696-700); "Courant-Source-Diffusion" is not built
     * (SCHEMI_ERR_INIT for values other than 0 and 1). */
    int32_t timestep_source;
    double source_coeff;
    /* k-epsilon turbulence (#Turbulence_flag on, #Type_of_source_flag; the
     * models of abstractTurbulenceModel::createTurbulenceModel,
     * Schemi/src/interfaces/abstractTurbulenceModel.cpp:235-320): the state
     * carries rho k and rho epsilon after E (Ns + 7 fields, see
     * schemi_gpu_set_state); the flux carries the turbulent pressure
     * (2/3) rho k (HLLCSolver.cpp, HLLSolver.cpp) and, with the diffusion
     * stage, the turbulent transport coefficients, the k and epsilon
     * Laplacians and the model's sources apply (Diffusion.cpp:129-933).
     * Models: SCHEMI_TURB_SHEAR (shearModel.cpp, its source step capping a
     * "Courant-Source" step), SCHEMI_TURB_ZERO (zeroModel.cpp) and
     * SCHEMI_TURB_KEPSA (kEpsAModel.cpp: the state also carries rho a after
     * rho epsilon, Ns + 10 fields; the algebraic b is recomputed from the
     * mixture in every diffusion stage and is not part of the state; the
     * source step caps "Courant-Source" as the shear model's).  The
     * parameters are those of set/turbulentParameters.txt in the models'
     * meaning (kEpsModels.cpp, shearModel.cpp:35-59): Cmu, turbulent Schmidt
     * and Prandtl numbers (T, e), sigma_k, sigma_epsilon, CMS_R, CMS_D, the
     * minimal k and epsilon, and the shear model's C1, C2, C3; bc_k, bc_eps
     * the boundary conditions of set/k.txt and set/epsilon.txt; for kEpsA
     * also sigma_a, the eight constants of kEpsAModel.cpp:73-80 (C0, C1, C2,
     * C3, Ca, CaMax, CMSA, chi; turb_C is not read) and the vector boundary
     * conditions of set/a.txt.  Supported
     * with SecondOrderRK, the ideal or van der Waals mixture, and limiters
     * whose slopes vanish (2-D/3-D TVD limiters, the zero limiter) or 1-D
     * meshes with a TVD limiter; the diffusion stage (if on) with the
     * explicit solver and neither the molar-mass correction nor the enthalpy
     * flow.  Otherwise SCHEMI_ERR_INIT. */
    int32_t turbulence;
    int32_t turbulence_model;
    double turb_Cmu;
    double turb_sigma_sc;
    double turb_sigma_t;
    double turb_sigma_e;
    double turb_sigma_k;
    double turb_sigma_eps;
    double turb_CMS_R;
    double turb_CMS_D;
    double turb_min_k;
    double turb_min_eps;
    double turb_C[3];
    schemi_bc bc_k[6];
    schemi_bc bc_eps[6];
    double turb_sigma_a;
    double turb_kepsa_C[8];
    schemi_bc bc_a[6];
    /* #Reconstruction of main.txt (Schemi.cpp:152-171) for the diffusion
     * stage: SCHEMI_RECONSTRUCTION_LINEAR (0; grad v, div v and, with
     * turbulence, grad rho, grad p, grad a from the interpolated cell values)
     * or SCHEMI_RECONSTRUCTION_STAR (1; from the faces' Riemann star states,
     * averaged over the two SSP-RK2 stages: Diffusion.cpp:176-193, 529-537,
     * secondOrderStepSolverRK.cpp:73-85).  Star needs limiters whose slopes
     * vanish or a 1-D TVD limiter (not the linear limiter); with
     * MUSCL-Hancock it is built only under gravity, where the reference
     * replaces the star velocity by the interpolated one
     * (secondOrderStepSolver.cpp:74-76) and a laminar stage therefore equals
     * the linear one.  Otherwise SCHEMI_ERR_INIT.  Ignored without the
     * diffusion stage. */
    int32_t reconstruction;
    /* SCHEMI_TURB_BHR (BHRModel.cpp with the Goncharov initialisation model
     * off, as in every reference deck; the state carries rho a and rho b
     * after rho epsilon, Ns + 11 fields): sigma_a in turb_sigma_a, sigma_b,
     * the minimal b, the twelve constants of BHRModel.cpp:73-84 (C0, C1, C2,
     * C3, C4, Ca, Cb1, Cb2, alpha2, alpha3, alpha4, CMS_M), and the boundary
     * conditions of set/a.txt (bc_a) and set/b.txt.  With the explicit solver
     * b has no sources (the reference's explicit branch leaves them out). */
    double turb_sigma_b;
    double turb_min_b;
    double turb_bhr_C[12];
    schemi_bc bc_b[6];
} schemi_scheme_desc;

#define SCHEMI_RECONSTRUCTION_LINEAR 0
#define SCHEMI_RECONSTRUCTION_STAR 1

/* turbulenceModel (Schemi/enumerations/turbulenceModelEnum.hpp) values */
#define SCHEMI_TURB_ZERO 0
#define SCHEMI_TURB_SHEAR 2
#define SCHEMI_TURB_KEPSA 6
#define SCHEMI_TURB_BHR 7

#define SCHEMI_SLE_GAUSS_SEIDEL 0
#define SCHEMI_SLE_EXPLICIT 3
#define SCHEMI_SLE_JACOBI 6

typedef struct schemi_gpu_ctx schemi_gpu_ctx;

/* Creates a context on the current HIP device: validates the description
 * (mesh::instance / phaseInitialization checks), builds the per-axis
 * geometry tables and allocates the device state. */
int schemi_gpu_create(const schemi_mesh_desc *mesh, const schemi_mix_desc *mix,
                      const schemi_scheme_desc *scheme, schemi_gpu_ctx **out);

void schemi_gpu_destroy(schemi_gpu_ctx *ctx);

/* Uploads the conserved state (Ns + 5 fields x ncells doubles; with
 * turbulence Ns + 7: rho k and rho epsilon after E; with the kEpsA model
 * Ns + 10: rho a_x, rho a_y, rho a_z after rho epsilon; with BHR Ns + 11:
 * rho b after rho a). */
int schemi_gpu_set_state(schemi_gpu_ctx *ctx, const double *cons,
                         int64_t ncells, int32_t numbering);

/* Downloads the conserved state (the fields of schemi_gpu_set_state) and,
 * when prim is not NULL, the primitives recovered as in
 * Advection2dOrder.cpp:578-602 (Ns + 8 fields; with turbulence the internal
 * energy is E - rho v^2 / 2 - rho k). */
int schemi_gpu_get_state(schemi_gpu_ctx *ctx, double *cons, double *prim,
                         int32_t numbering);

/* One SSP-RK2 step with time step dt (secondOrderStepSolverRK::calculateStep).
 * On return *dt_next (if not NULL) holds the Courant step computed from the
 * new state (Schemi.cpp:807-833), capped by the step's source step when
 * timestep_source is 1. */
int schemi_gpu_step(schemi_gpu_ctx *ctx, double dt, double *dt_next);

/* Source time step of the last finished step (mesh::timestepSource(),
 * Schemi/headers/mesh.hpp:155-157, set by the diffusion stage at
 * Diffusion.cpp:696-700 and read by the main loop's "Courant-Source" step at
 * Schemi.cpp:828-833): source_coeff times the stage's minimum ratio, or the
 * reference's veryBig (1e-3 DBL_MAX) when no stage set one.  A drop-in caller
 * whose own loop forms the time step writes it to mesh::timestepSourceRef(). */
int schemi_gpu_source_timestep(schemi_gpu_ctx *ctx, double *dt_source);

/* The turbulent fields that the last step's diffusion stage leaves in the
 * reference's phase besides the state: the k and epsilon the stage started
 * from (rows 0 and 1; the phase's turbulent transport coefficients,
 * nu_t = Cmu k^2 / epsilon and nu_t / sigma for the others, are formed from
 * them at Diffusion.cpp:129-135 with transportCoefficients.hpp:65-83) and,
 * with the kEpsA model, the stage's algebraic b, which the phase keeps as its
 * b (row 2; kEpsAModel.cpp:232-257, Diffusion.cpp:525-527, 1006).  out:
 * 2 (kEpsA: 3) x ncells doubles.  SCHEMI_ERR_INIT without turbulence or
 * before a step with the diffusion stage since the last set_state.
 * Synchronous. */
int schemi_gpu_get_turbulence_stage(schemi_gpu_ctx *ctx, double *out, int32_t numbering);

/* nsteps SSP-RK2 steps with the time step kept on the device: the first
 * step uses dt0, every later one the Courant step of the previous state,
 * clamped so that the time does not pass t_end (Schemi.cpp:933-935).  Stops
 * early when t_end is reached.  Outputs may be NULL. */
int schemi_gpu_run(schemi_gpu_ctx *ctx, int64_t nsteps, double dt0,
                   double t_end, double *t_out, int64_t *steps_out,
                   double *dt_out);

/* One forward-Euler stage (Advection2dOrder) with time step dt. */
int schemi_gpu_stage(schemi_gpu_ctx *ctx, double dt);

/* Numerical flux of nfaces independent faces (unit-test entry of
 * abstractFlowSolver::calculateFlows, abstractFlowSolver.hpp:31-33).
 * Face f has unit normal along axis[f] with sign sgn[f] (+1/-1); left and
 * right states are primitives [c_1..c_Ns, v_x, v_y, v_z, p] (Ns + 4 doubles
 * per face, face-major).  Output: Ns + 5 scalar fluxes per face
 * (rho_1..rho_Ns, rho, m_x, m_y, m_z, E), i.e. F & N. */
int schemi_gpu_face_flux(schemi_gpu_ctx *ctx, int64_t nfaces,
                         const int32_t *axis, const double *sgn,
                         const double *left, const double *right,
                         double *flux);

/* Device pointer of the conserved state (Ns + 5 fields x ncells), for
 * callers that keep the state resident and pass it to other device code. */
int schemi_gpu_device_state(schemi_gpu_ctx *ctx, double **dev_ptr,
                            int64_t *ncells);

/* Blocks until the context's stream is idle; returns any pending error. */
int schemi_gpu_sync(schemi_gpu_ctx *ctx);

/* Device-to-device state copies (SoA, Ns + 5 rows of ncells doubles, cells
 * in the reference's order): src (nullable) becomes the current state, then
 * the current state is copied to dst (nullable).  For checkpoints kept in
 * device memory by the caller; synchronous. */
int schemi_gpu_state_device(schemi_gpu_ctx *ctx, const double *src, double *dst);

/* Adjoint (no counterpart in the reference: gradients of functionals of the
 * solution with respect to the initial state).  With the current state U^n
 * and device arrays in the state layout, what = SCHEMI_ADJ_STEP writes
 *   lambda_out = (dU^{n+1}/dU^n)^T lambda_in
 * for one SSP-RK2 step of size dt; SCHEMI_ADJ_STAGE the same for one stage
 * (schemi_gpu_stage).  The step is linearised in FP64 at the forward states
 * (U^1 is recomputed in the context's precision), with dt held fixed and the
 * flux's branches (HLLC wave pattern, wave-speed min/max) frozen at the
 * forward switch state (and the limiter's: min/max of the ratio entries, the
 * branches of phi).  The current state is left unchanged; lambda_out may
 * equal lambda_in.  Supported with SecondOrderRK for every limiter on 1-D
 * meshes and, on 2-D/3-D meshes, where the face values are the cell values
 * (zero limiter, TVD limiters); ideal, van der Waals and stiffened mixtures.
 * Otherwise SCHEMI_ERR_INIT.  Synchronous. */
#define SCHEMI_ADJ_STEP 0
#define SCHEMI_ADJ_STAGE 1
int schemi_gpu_adjoint_step(schemi_gpu_ctx *ctx, double dt, const double *lambda_in,
                            double *lambda_out, int32_t what);

/* Mixing width J = sum_i V_i Y_i (1 - Y_i), Y = rho_species / rho, of the
 * current state (no counterpart in the reference, whose mixedZoneWidth1D,
 * output.cpp:310-391, is a threshold width); dJ/dU into dJdU (device, state
 * layout, nullable; the Ns + 5 rows up to E are written, the turbulent rows
 * of a turbulent state are not).  Deterministic; synchronous. */
int schemi_gpu_mixing_width(schemi_gpu_ctx *ctx, int32_t species, double *J, double *dJdU);

/* No-allocation variant of schemi_gpu_mixing_width.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and the same result, bit for bit, but the reduction's partial
 * sums live in a fixed library-owned device array of
 * SCHEMI_NOALLOC_MAX_CELLS / 256 + 1 doubles (512 KiB; every device the
 * library runs on holds its own copy until the process ends), so nothing is
 * allocated or freed.  Contexts with more than SCHEMI_NOALLOC_MAX_CELLS cells
 * get SCHEMI_ERR_SCRATCH.  Every call on one device shares that array: calls
 * that overlap on a device (from several host threads or contexts at once)
 * race and must be serialised by the caller.  Synchronous. */
#define SCHEMI_NOALLOC_MAX_CELLS 16777216
int schemi_gpu_mixing_width_noalloc(schemi_gpu_ctx *ctx, int32_t species, double *J,
                                    double *dJdU);

/* Average duration (ms) of the last stage-kernel launches timed with HIP
 * events on the context's stream (0 when none were timed); enables timing
 * when enable != 0. */
/* Average durations (ms) of the two passes of the diffusion stages run while
 * timing was enabled (schemi_gpu_timing), measured with HIP events on the
 * context's stream; clears them.  No counterpart in the reference (its
 * Diffusion stage keeps a wall-clock total, Diffusion.cpp:1056-1059). */
int schemi_gpu_diffusion_timing(schemi_gpu_ctx *ctx, double *prims_ms, double *update_ms,
                                int64_t *launches);

int schemi_gpu_timing(schemi_gpu_ctx *ctx, int32_t enable, double *avg_ms,
                      int64_t *launches);

const char *schemi_gpu_strerror(int code);

/* Reference face id -> structured face (axis, i, j, k) and back; cells are
 * not permuted.  The face permutation is a bijection between the
 * reference's face blocks (mesh.cpp:1301-1790) and axis-major face arrays. */
int schemi_gpu_face_permutation(const schemi_mesh_desc *mesh, int64_t *ref_to_gpu,
                                int64_t nfaces);

/* Per-axis geometry table of a mesh (host only, no device needed), as used
 * by the kernels: which = 0 cell centres, 1 edge lengths, 2 face coordinates,
 * 3 owner weights, 4 neighbour weights, 5 centre distances across faces,
 * 6 x_f(i) - x_c(i), 7 x_f(i+1) - x_c(i), 8 limiter ghost distances (low,
 * high), 9 separability flag (1 when every face centre shares the cells'
 * transverse coordinates), 10 + 2 b + cls the offsets of the centres of the
 * faces normal to `axis` from their cells' centres along transverse axis b,
 * per cell index along b (cls 0 inner faces, 1 boundary-plane faces; empty
 * for b == axis).  periodic_mask bit a marks periodic axes.  With
 * out == NULL only *len is set.  Replaces the reference's mesh accessors
 * (mesh.hpp:39-176: cells(), surfaces(), surfOwnW(), surfNeiW()). */
int schemi_gpu_axis_table(const schemi_mesh_desc *mesh, int32_t periodic_mask,
                          int32_t axis, int32_t which, double *out, int64_t *len);

#ifdef __cplusplus
}
#endif

#endif
