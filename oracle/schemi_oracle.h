/*
 * schemi_oracle.h -- CPU oracle for the explicit finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * solver's SSP-RK2 step (Schemi/src/solverStages/secondOrderStepSolverRK.cpp),
 * used by the tests, by the smoke check and by the CPU-baseline leg of the
 * benchmark as the checker. No product code path may call it.
 *
 * The mesh is built exactly as the reference builds it (reference cell and face
 * numbering, per-cell face order, owner/neighbour lists, normals, weights),
 * and every arithmetic operation follows the reference's operation order, so
 * that with -ffp-contract=off the oracle reproduces the reference bit for bit.
 */
#ifndef SCHEMI_ORACLE_H
#define SCHEMI_ORACLE_H

#include <stddef.h>

#define OR_MAXS 9

/* Reference enum values (Schemi/enumerations/ headers). */
enum { OR_BC_BLANK = 0, OR_BC_FREE = 1, OR_BC_SLIP = 2, OR_BC_FIXED_CELL = 3,
       OR_BC_FIXED_SURF = 4, OR_BC_INNER = 5,
       OR_BC_PERIODIC = 32 /* no reference counterpart */ };
enum { OR_FLUX_HLL = 0, OR_FLUX_HLLC = 2 };
/* gasModel (gasModelEnum.hpp:15-18) */
enum { OR_GAS_IDEAL = 0, OR_GAS_VDW = 1, OR_GAS_RK = 2, OR_GAS_STIFFENED = 3, OR_GAS_KATAOKA = 4 };
enum { OR_LIM_ZERO = 0, OR_LIM_MINMOD = 1, OR_LIM_VANLEER = 2, OR_LIM_LINEAR = 3,
       OR_LIM_VANALBADA = 4 };
/* errors enum: noErrors=0, RiemannSolverError=1, boundaryConditionError=5,
 * initialisationError=11 */

typedef struct {
    int type;
    double val[3];
} or_bc;

typedef struct {
    int dim;            /* 1, 2 or 3 */
    long N[3];
    double L[3];        /* radius-vector of the parallelepiped */
    double origin[3];   /* zero point */
    int Ns;
    double M[OR_MAXS];
    double Cv[OR_MAXS];
    double R;
    int flux;
    int limiter;
    int gravity;
    double g[3];
    /* boundary conditions per patch: 0 tail, 1 point, 2 bottom, 3 right,
     * 4 left, 5 top */
    or_bc bc_c[OR_MAXS][6];
    or_bc bc_v[6];
    or_bc bc_p[6];
    /* equation of state: OR_GAS_IDEAL, OR_GAS_VDW or OR_GAS_RK (Redlich-Kwong)
     * with the species' critical temperature and pressure */
    int gas_model;
    double Tc[OR_MAXS];
    double Pc[OR_MAXS];
    /* mixing matrices (row-major), filled by or_create, inputs ignored:
     * van der Waals / Redlich-Kwong a_kl, b_kl; stiffened p0_kl in a_kl and
     * gamma_kl in b_kl */
    double a_kl[OR_MAXS * OR_MAXS];
    double b_kl[OR_MAXS * OR_MAXS];
    /* stiffened fluid (OR_GAS_STIFFENED): p0 and gamma of every species */
    double p0[OR_MAXS];
    double gam[OR_MAXS];
    /* Kataoka-van der Waals (OR_GAS_KATAOKA): Lennard-Jones epsilon and
     * sigma of every species, b from the critical point (kb_critical = 1:
     * kb_coeff R Tc / (8 Pc)) or from sigma (0: kb_coeff sigma^3 N_A);
     * matrices V0_kl in a_kl, b_kl in b_kl, eps_kl in e_kl */
    double epsLJ[OR_MAXS];
    double sigLJ[OR_MAXS];
    int kb_critical;
    double kb_coeff;
    double e_kl[OR_MAXS * OR_MAXS];
    /* explicit molecular diffusion stage after each step (Diffusion.cpp):
     * constant transport coefficients (abstractTransportModel.cpp:14-34:
     * viscosity mu, diffusion coefficient D, heat conduction kappa), explicit
     * matrix solver, linear reconstruction, no turbulence, no enthalpy flow,
     * no molar-mass correction; ideal mixture.  On a periodic axis (no
     * reference counterpart) the wrapped face is an inner face whose centre
     * distance runs through the boundary. */
    int diffusion;
    double mu, D, kappa;
    /* matrix solver of the stage: OR_SLE_EXPLICIT, OR_SLE_JACOBI
     * (JacobiSolver.cpp) or OR_SLE_GAUSS_SEIDEL (GaussSeidelSolver.cpp) with
     * at most sle_iters iterations */
    int sle_solver;
    long sle_iters;
    /* enthalpy flow solved explicitly (1) or not (0), molar-mass diffusion
     * correction on (1) or off (0) (set/transportCoefficients.txt) */
    int enth_flow;
    int molmass;
    /* time-step flag of main.txt: 0 "Courant", 1 "Courant-Source" (the
     * Courant step is capped by the source step of the previous diffusion
     * stage, Schemi.cpp:828-833); ts_coeff is the first source coefficient
     * of the same entry (Schemi.cpp:205-225) */
    int ts_source;
    double ts_coeff;
    /* k-epsilon turbulence (#Turbulence_flag on, #Type_of_source_flag):
     * turb_model OR_TURB_SHEAR (shearModel.cpp), OR_TURB_ZERO
     * (zeroModel.cpp: transport and diffusion of k and epsilon, no
     * sources), OR_TURB_KEPSA (kEpsAModel.cpp: with the turbulent mass
     * flux a, transported as rho a, and the algebraic b of calculate_b) or
     * OR_TURB_BHR (BHRModel.cpp: rho a and a transported rho b);
     * the parameters of set/turbulentParameters.txt (kEpsModels.cpp,
     * abstractTurbulenceModel.hpp:26-57; sigmaa for a); tC the model
     * constants (shear: C1, C2, C3; kEpsA: C0, C1, C2, C3, Ca, CaMax, CMSA,
     * chi); bc_k, bc_eps and bc_a the boundary conditions of set/k.txt,
     * set/epsilon.txt and set/a.txt.  Turbulence needs the SSP-RK2 step;
     * with the diffusion stage, the explicit matrix solver and no molar-mass
     * correction or enthalpy flow. */
    int turb;
    int turb_model;
    double Cmu, sigmaSc, sigmaT, sigmaE, sigmaK, sigmaEps, CMS_R, CMS_D, mink, mineps;
    double tC[12];
    or_bc bc_k[6];
    or_bc bc_eps[6];
    double sigmaa;
    or_bc bc_a[6];
    /* BHR (BHRModel.cpp, the Goncharov initialisation model off as in every
     * reference deck): the transported b as rho b with sigma_b, its minimal
     * value and set/b.txt; tC holds C0, C1, C2, C3, C4, Ca, Cb1, Cb2,
     * alpha2, alpha3, alpha4, CMS_M */
    double sigmab, minb;
    or_bc bc_b[6];
    /* #Reconstruction of main.txt for the diffusion stage: 0 linear, 1 star
     * (the cell gradients of v, rho, p and a from the step's star states,
     * Diffusion.cpp:176-193, 529-537) */
    int star_rec;
} or_config;

/* turbulenceModel (turbulenceModelEnum.hpp) values of the built models */
enum { OR_TURB_ZERO = 0, OR_TURB_SHEAR = 2, OR_TURB_KEPSA = 6, OR_TURB_BHR = 7 };

enum { OR_SLE_EXPLICIT = 0, OR_SLE_JACOBI = 1, OR_SLE_GAUSS_SEIDEL = 2 };

typedef struct or_solver or_solver;

/* Cell state in the reference layout.  Arrays are owned by the solver. */
typedef struct {
    long n;
    int Ns;
    double *c[OR_MAXS + 1];    /* molar concentrations, c[0] = sum */
    double *rho[OR_MAXS + 1];  /* partial densities, rho[0] = total */
    double *v[3];
    double *p;
    double *U;
    double *T;
    double *m[3];
    double *E;
    /* turbulence: k, epsilon and the conserved rho k, rho epsilon (zero
     * without turbulence); kEpsA: a, rho a and b (zero otherwise) */
    double *k, *eps, *rk, *reps;
    double *a[3], *ra[3], *b;
    double *rb; /* BHR: rho b (zero otherwise) */
} or_state;

or_solver *or_create(const or_config *cfg);
void or_destroy(or_solver *s);
long or_ncells(const or_solver *s);
long or_nfaces(const or_solver *s);
or_state *or_get_state(or_solver *s);

/* Fills the state from primitives (c_1..c_Ns, v, p) the way the reference's
 * initialisation does (phaseInitialization.cpp:1827-1883). */
void or_init_primitive(or_solver *s, const double *c, const double *v,
                       const double *p);
/* Turbulent fields for or_init_primitive (call it first): k and epsilon,
 * clipped to the minimal values and multiplied by the density as the
 * reference's initialisation does (phaseInitialization.cpp:1814-1848, 1873-1877). */
void or_init_turbulence(or_solver *s, const double *k, const double *eps);
/* kEpsA: the initial a (3 n, component-major; multiplied by the density by
 * or_init_primitive, phaseInitialization.cpp:1850) and the conserved rho a
 * for or_set_conserved (call either before those). */
void or_init_turbulence_a(or_solver *s, const double *a);
void or_set_turbulence_a_conserved(or_solver *s, const double *ra);
/* BHR: the initial b (n; rho b = rho b, phaseInitialization.cpp:1852) and
 * the conserved rho b for or_set_conserved (call either before those). */
void or_init_turbulence_b(or_solver *s, const double *b);
void or_set_turbulence_b_conserved(or_solver *s, const double *rb);
/* Conserved turbulent fields rho k, rho epsilon for or_set_conserved (call
 * it first). */
void or_set_turbulence_conserved(or_solver *s, const double *rk, const double *reps);
/* Sets the full reference state (as written by the reference's output). */
void or_set_full(or_solver *s, const double *c0, const double *c,
                 const double *rho0, const double *rho, const double *v,
                 const double *p, const double *U, const double *T,
                 const double *m, const double *E);
/* Sets the state from the conserved vector only (rho_1..rho_Ns, rho, m, E),
 * deriving primitives as the stage update does. */
void or_set_conserved(or_solver *s, const double *rho_k, const double *rho,
                      const double *m, const double *E);

/* One SSP-RK2 step with time step dt.  mode 0: literal reference state
 * (concentrations averaged independently); mode 1: the state is re-derived
 * from the conserved variables at the start of the step.  Returns an errors
 * code (0 = noErrors). */
int or_step(or_solver *s, double dt, int mode);
/* One forward-Euler stage (Advection2dOrder). */
int or_stage(or_solver *s, double dt);
/* One MUSCL-Hancock step (secondOrderStepSolver: Advection with the Hancock
 * half-step predictor of the face states); mode as for or_step. */
int or_step_hancock(or_solver *s, double dt, int mode);
/* Courant time step of the reference main loop (Schemi.cpp:807-833). */
double or_courant_dt(or_solver *s, double courant, double min_time);

/* Geometry accessors (reference numbering). */
void or_cell_geometry(const or_solver *s, double *rc, double *V);
void or_face_geometry(const or_solver *s, double *rc, double *N, double *S,
                      long *owner, long *neigh, double *wo, double *wn);
void or_cell_faces(const or_solver *s, long *faces /* 6 per cell, -1 pad */);

/* Point kernels exposed for unit tests. */
double or_hypot3(double x, double y, double z);
/* Face flux between two face states (HLL or HLLC).  state layout:
 * [rho_0..rho_Ns, v0,v1,v2, p, U(ignored, recomputed), ...]; see the C file.
 * Outputs the scalar fluxes (F & N) for rho_k, m (3) and E. */
int or_face_flux(const or_solver *s, const double *N, const double *cL,
                 const double *vL, double pL, const double *cR,
                 const double *vR, double pR, double *Frho, double *Fm,
                 double *FE);
double or_limiter_phi(int limiter, double r);

#endif
