"""Case description: mesh, mixture, scheme and boundary conditions.

A :class:`Case` carries what the reference reads from its ``set/*.txt`` deck
for the explicit step (Schemi/main/Schemi.cpp:84-163 and
Schemi/src/thermodynamics/phaseInitialization.cpp:17-660): the parallelepiped
(``#Radius-vector_of_system's_parallelepiped``, ``#Number_of_cells_in_direction``),
the species (molar mass, molar Cv, critical temperature and pressure per
``sub_k.txt``), the universal gas constant, the equation of state, the flux
scheme and limiter, gravity, and one boundary condition per field and patch.

Cells are numbered as in the reference: ``id = i + Nx (j + Ny k)``, x fastest
(Schemi/src/dataStructures/mesh.cpp:1223-1230).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

from .enums import boundaryConditionType as BCT
from .enums import flowSolver, gasModel, turbulenceModel, typeOfTVDLimiter

Vec3 = Tuple[float, float, float]


@dataclass
class BC:
    """One boundary condition: type plus fixed value (scalar in ``value[0]``)."""

    type: BCT = BCT.freeBoundary
    value: Vec3 = (0.0, 0.0, 0.0)


def _free6() -> List[BC]:
    return [BC() for _ in range(6)]


@dataclass
class Turbulence:
    """k-epsilon turbulence (``#Turbulence_flag on``): the model of
    ``#Type_of_source_flag`` (``shear``: shearModel.cpp; ``zero``:
    zeroModel.cpp; ``kEpsA``: kEpsAModel.cpp, with the turbulent mass flux a
    transported as rho a and the algebraic b), the parameters of
    ``set/turbulentParameters.txt`` in the order the k-epsilon models read
    them (kEpsModels.cpp, shearModel.cpp:35-59, kEpsAModel.cpp:48-80), and
    the boundary conditions of ``set/k.txt``, ``set/epsilon.txt`` and (kEpsA)
    ``set/a.txt``.  ``C`` holds the model constants: C1, C2, C3 for the shear
    model; C0, C1, C2, C3, Ca, CaMax, CMSA, chi for kEpsA (:func:`kEpsA`)."""

    model: turbulenceModel = turbulenceModel.shearModel
    Cmu: float = 0.09
    sigmaSc: float = 0.7      # Turbulent_Schmidt
    sigmaT: float = 0.7       # Turbulent_Prandtl_T
    sigmaE: float = 1e6       # Turbulent_Prandtl_e
    sigmaK: float = 0.7       # sigma_k
    sigmaEps: float = 0.6     # sigma_epsilon
    CMS_R: float = 0.1
    CMS_D: float = 1.0
    mink: float = 1e-15       # Minimal_k
    mineps: float = 1e-15     # Minimal_epsilon
    C: Tuple[float, ...] = (1.44, 1.92, 2.0)  # shear: C1, C2, C3
    bc_k: List[BC] = field(default_factory=_free6)
    bc_eps: List[BC] = field(default_factory=_free6)
    sigmaa: float = 1.4       # sigma_a (kEpsA, BHR)
    bc_a: List[BC] = field(default_factory=_free6)
    sigmab: float = 0.7       # sigma_b (BHR)
    minb: float = 1e-15       # Minimal_b (BHR)
    bc_b: List[BC] = field(default_factory=_free6)

    def params(self):
        return (self.Cmu, self.sigmaSc, self.sigmaT, self.sigmaE, self.sigmaK,
                self.sigmaEps, self.CMS_R, self.CMS_D, self.mink, self.mineps)

    @property
    def a_field(self) -> bool:
        """The model carries the turbulent mass flux a (kEpsA, BHR)."""
        return self.model in (turbulenceModel.kEpsAModel, turbulenceModel.BHRModel)

    @property
    def b_field(self) -> bool:
        """The model transports b (BHR)."""
        return self.model == turbulenceModel.BHRModel

    @classmethod
    def BHR(cls, C0=0.92, C1=1.44, C2=1.92, C3=2.0, C4=1.1, Ca=2.5, Cb1=2.0, Cb2=5.0,
            alpha2=0.15, alpha3=0.2, alpha4=0.1, CMSM=1.0, **kw) -> "Turbulence":
        """The BHR model with its twelve constants (BHRModel.cpp:86-137; the
        defaults of the reference's decks) and the Goncharov initialisation
        model off."""
        return cls(model=turbulenceModel.BHRModel,
                   C=(C0, C1, C2, C3, C4, Ca, Cb1, Cb2, alpha2, alpha3, alpha4, CMSM), **kw)

    @classmethod
    def kEpsA(cls, C0=0.92, C1=1.44, C2=1.92, C3=2.0, Ca=2.5, CaMax=5.0, CMSA=1.5,
              chi=1.0, **kw) -> "Turbulence":
        """The kEpsA model with its eight constants (kEpsAModel.cpp:86-121);
        CMSA >= 1e15 switches the thetaA limiter off."""
        return cls(model=turbulenceModel.kEpsAModel,
                   C=(C0, C1, C2, C3, Ca, CaMax, CMSA, chi), **kw)


@dataclass
class Case:
    dim: int
    N: Tuple[int, int, int]
    L: Vec3
    M: Sequence[float]
    Cv: Sequence[float]
    origin: Vec3 = (0.0, 0.0, 0.0)
    R: float = 1.0
    flux: flowSolver = flowSolver.HLLC
    limiter: typeOfTVDLimiter = typeOfTVDLimiter.minmod
    gravity: bool = False
    g: Vec3 = (0.0, 0.0, 0.0)
    courant: float = 0.4
    min_time: float = 1e-15
    bc_c: List[List[BC]] = field(default_factory=list)
    bc_v: List[BC] = field(default_factory=_free6)
    bc_p: List[BC] = field(default_factory=_free6)
    # equation of state (gasModel) and the species' critical temperature and
    # pressure (used by the van der Waals model)
    gas_model: gasModel = gasModel.ideal
    Tc: Sequence[float] = ()
    Pc: Sequence[float] = ()
    # stiffened fluid (./set/stiffenedFluidData.txt): p0 and gamma per species
    p0: Sequence[float] = ()
    gammas: Sequence[float] = ()
    # Kataoka-van der Waals (./set/KataokaVanDerWaalsFluidData.txt):
    # Lennard-Jones epsilon and sigma per species, and how b is formed
    # ("critical": coeff R Tc / (8 Pc); "Kataoka": coeff sigma^3 N_A)
    epsLJ: Sequence[float] = ()
    sigmaLJ: Sequence[float] = ()
    kataoka_b: tuple = ("critical", 1.0)
    # step solver (#Approximation_order): 1 SecondOrderRK (SSP-RK2),
    # 0 SecondOrder (MUSCL-Hancock)
    step_solver: int = 1
    # molecular diffusion stage after each step (#Diffusion_flag on, explicit
    # matrix solver, linear reconstruction, constant transport coefficients
    # of ./set/transportCoefficients.txt: viscosity, diffusion coefficient,
    # heat conduction coefficient)
    diffusion: bool = False
    transport: tuple = (2e-5, 2e-5, 2e-2)
    # its matrix solver (#Type_of_SLE_solver_flag): "explicit" or "Jacobi",
    # and the iteration limit (#Number_of_iterations_of_SLE_matrix_solver)
    sle_solver: str = "explicit"
    sle_iters: int = 2000
    # #Flag_for_enthalpy_flow_solution explicit (True) or no, and
    # #Flag_for_molar_mass_diffusion_correction on (True) or off
    enthalpy_flow: bool = False
    molar_mass_correction: bool = False
    # #Reconstruction of the stage's cell gradients of v, rho, p (and a):
    # "linear" (interpolated cell values) or "star" (the step's Riemann star
    # states on the faces, averaged over the two SSP-RK2 stages)
    reconstruction: str = "linear"
    # #Source_timestep_correction: "Courant-Source" (True) or "Courant", and
    # its two coefficients (source, diffusion)
    timestep_source: bool = False
    source_coeffs: tuple = (0.25, 0.5)
    # k-epsilon turbulence (None: #Turbulence_flag off)
    turbulence: Optional[Turbulence] = None

    def __post_init__(self):
        if not self.bc_c:
            self.bc_c = [_free6() for _ in self.M]
        if not self.Tc:
            self.Tc = [1.0] * len(self.M)
        if not self.Pc:
            self.Pc = [1.0] * len(self.M)
        if not self.p0:
            self.p0 = [0.0] * len(self.M)
        if not self.gammas:
            self.gammas = [1.4] * len(self.M)
        if not self.epsLJ:
            self.epsLJ = [1.0] * len(self.M)
        if not self.sigmaLJ:
            self.sigmaLJ = [1e-8] * len(self.M)
        n = list(self.N)
        if self.dim < 3:
            n[2] = 1
        if self.dim < 2:
            n[1] = 1
        self.N = tuple(int(x) for x in n)

    @property
    def Ns(self) -> int:
        return len(self.M)

    @property
    def nfields(self) -> int:
        """Conserved fields per cell: rho_1..rho_Ns, rho, m (3), E, and with
        turbulence rho k, rho epsilon, with the kEpsA and BHR models rho a (3),
        with BHR rho b."""
        t = self.turbulence
        if t is None:
            return self.Ns + 5
        return self.Ns + 7 + (3 if t.a_field else 0) + (1 if t.b_field else 0)

    @property
    def ncells(self) -> int:
        return self.N[0] * self.N[1] * self.N[2]

    def set_bc(self, patch: int, type_: BCT, field_: str = "all", value=None):
        """Sets one patch's boundary condition on ``c``, ``v``, ``p`` or all."""
        if field_ in ("all", "c"):
            for k in range(self.Ns):
                self.bc_c[k][patch] = BC(type_, tuple(value or (0.0, 0.0, 0.0)))
        if field_ in ("all", "v"):
            self.bc_v[patch] = BC(type_, tuple(value or (0.0, 0.0, 0.0)))
        if field_ in ("all", "p"):
            self.bc_p[patch] = BC(type_, tuple(value or (0.0, 0.0, 0.0)))

    def patch_used(self, patch: int) -> bool:
        # tail/point always; right/left from 2D; bottom/top in 3D
        return patch in (0, 1) or (patch in (3, 4) and self.dim >= 2) or (
            patch in (2, 5) and self.dim >= 3)
