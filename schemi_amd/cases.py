"""Initial-condition generators for the benchmark workloads.

The reference sets initial fields by zones (Schemi/src/interfaces/zone.cpp:25-126)
or reads them back through its restart reader
(Schemi/src/thermodynamics/phaseInitialization.cpp:1382-1744); these
generators produce the same primitive fields (molar concentrations, velocity,
pressure) in reference cell order and convert them to the conserved vector with
the reference's initialisation arithmetic.  Random perturbations use an
in-repo SplitMix64 stream so that the fields are reproducible across languages.

Gas constant R = 1 (``#Defines_universal_gas_constant 1``), molar Cv_k =
R / (gamma_k - 1).
"""
from __future__ import annotations

import math

import numpy as np

from .case import BC, Case
from .enums import boundaryConditionType as BCT
from .enums import flowSolver, gasModel, typeOfTVDLimiter
from .gpu import conserved_from_primitive

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator; uniform doubles as (x >> 11) * 2**-53."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def cv_of(gamma: float, R: float = 1.0) -> float:
    return R / (gamma - 1.0)


def _cell_centres(case: Case):
    """Cell-centre coordinates in reference order (x fastest)."""
    Nx, Ny, Nz = case.N
    x = case.origin[0] + (np.arange(Nx) + 0.5) * (case.L[0] / Nx)
    y = (case.origin[1] + (np.arange(Ny) + 0.5) * (case.L[1] / Ny)) if case.dim >= 2 \
        else np.array([0.5 * case.L[1]])
    zlen = case.L[1] if case.dim == 3 else case.L[2]
    z = (case.origin[2] + (np.arange(Nz) + 0.5) * (zlen / Nz)) if case.dim == 3 \
        else np.array([0.5 * case.L[2]])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X.reshape(-1), Y.reshape(-1), Z.reshape(-1)


def shock_tube_1d(n: int = 1000):
    """1-D two-component shock tube: left species A (gamma 1.4, M 29) rho 1,
    p 1; right species B (gamma 1.67, M 4) rho 0.125, p 0.1; diaphragm at 0.5;
    HLLC, minmod, transmissive ends."""
    case = Case(dim=1, N=(n, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0],
                Cv=[cv_of(1.4), cv_of(1.67)], flux=flowSolver.HLLC,
                limiter=typeOfTVDLimiter.minmod, courant=0.4, min_time=1e-15)
    x, _, _ = _cell_centres(case)
    left = x < 0.5
    c = np.zeros((2, n))
    c[0] = np.where(left, 1.0 / 29.0, 0.0)
    c[1] = np.where(left, 0.0, 0.125 / 4.0)
    v = np.zeros((3, n))
    p = np.where(left, 1.0, 0.1)
    return case, (c, v, p)


def rankine_hugoniot(mach: float, gamma: float, rho1: float, p1: float):
    """Post-shock state of a normal shock moving into gas at rest."""
    m2 = mach * mach
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0))
    w = mach * math.sqrt(gamma * p1 / rho1)
    u2 = w * (1.0 - rho1 / rho2)
    return rho2, p2, u2


def richtmyer_meshkov_3d(nx: int = 256, ny: int = 128, nz: int = 128, seed: int = 2024,
                         x_shock: float = 0.4, x_iface: float = 0.6, modes: int = 32,
                         periodic_yz: bool = True):
    """3-D Richtmyer-Meshkov: Mach 1.2 shock in light gas (gamma 1.4, M 29,
    rho 1, p 1) hitting a heavy layer (gamma 1.09, M 146, rho 5.03, p 1) at
    x = 0.6 + eta(y, z), eta a sum of 32 seeded Fourier modes with integer
    wavevectors 2 <= |k| <= 16, amplitudes U[0, 0.005], phases U[0, 2 pi).
    Domain [0,2]x[0,1]x[0,1]; HLLC + van Albada; transmissive (free) x,
    periodic y/z (periodic_yz=False: slip y/z walls, the surrogate the
    reference itself can run, as it has no periodic boundary)."""
    case = Case(dim=3, N=(nx, ny, nz), L=(2.0, 1.0, 1.0), M=[29.0, 146.0],
                Cv=[cv_of(1.4), cv_of(1.09)], flux=flowSolver.HLLC,
                limiter=typeOfTVDLimiter.vanAlbada, courant=0.4, min_time=1e-15)
    for pt in (2, 3, 4, 5):
        case.set_bc(pt, BCT.periodic if periodic_yz else BCT.slip)
    rng = SplitMix64(seed)
    ks = []
    while len(ks) < modes:
        ky = int(rng.uniform() * 33) - 16
        kz = int(rng.uniform() * 33) - 16
        kk = math.hypot(ky, kz)
        if 2.0 <= kk <= 16.0:
            ks.append((ky, kz, 0.005 * rng.uniform(), 2.0 * math.pi * rng.uniform()))
    x, y, z = _cell_centres(case)
    eta = np.zeros_like(x)
    for ky, kz, amp, ph in ks:
        eta += amp * np.cos(2.0 * math.pi * (ky * y + kz * z) + ph)
    dx = case.L[0] / nx
    # heavy mass fraction: smooth over one cell around the perturbed interface
    yh = 0.5 * (1.0 + np.tanh((x - (x_iface + eta)) / dx))
    rho2, p2, u2 = rankine_hugoniot(1.2, 1.4, 1.0, 1.0)
    rho_l = np.where(x < x_shock, rho2, 1.0)
    p = np.where(x < x_shock, p2, 1.0)
    u = np.where(x < x_shock, u2, 0.0)
    rho_h = 5.03
    rho = 1.0 / (yh / rho_h + (1.0 - yh) / rho_l)
    c = np.stack([(1.0 - yh) * rho / 29.0, yh * rho / 146.0])
    v = np.stack([u, np.zeros_like(u), np.zeros_like(u)])
    return case, conserved_from_primitive(case, c, v, p)


def richtmyer_meshkov_bhr_3d(nx: int = 256, ny: int = 128, nz: int = 128, seed: int = 2024):
    """richtmyer_meshkov_3d with the reference's turbulent deck settings: the
    diffusion stage (constant mu, D, kappa of 1e-4; explicit solver) with the
    BHR model (its default constants, Goncharov initialisation off) and
    #Reconstruction star.  Initial turbulence: k = 1e-4 in a Gaussian of width
    0.02 around the unperturbed interface over a 1e-8 background,
    eps = Cmu^(3/4) k^(3/2) / 0.01, a = 0, b = 4e-3 Y (1 - Y) + 1e-15 with Y
    the heavy mass fraction; free x, periodic y/z for every field."""
    from .case import Turbulence

    base, U = richtmyer_meshkov_3d(nx, ny, nz, seed)
    case = base
    case.diffusion, case.transport = True, (1e-4, 1e-4, 1e-4)
    case.reconstruction = "star"
    case.turbulence = Turbulence.BHR()
    t = case.turbulence
    for pt in (2, 3, 4, 5):
        t.bc_k[pt] = BC(BCT.periodic)
        t.bc_eps[pt] = BC(BCT.periodic)
        t.bc_a[pt] = BC(BCT.periodic)
        t.bc_b[pt] = BC(BCT.periodic)
    x, _, _ = _cell_centres(case)
    Ns, n = case.Ns, case.ncells
    rho = U[Ns]
    yh = U[1] / rho
    kt = 1e-4 * np.exp(-((x - 0.6) / 0.02) ** 2) + 1e-8
    et = 0.09 ** 0.75 * kt ** 1.5 / 0.01
    bt = 4e-3 * yh * (1.0 - yh) + 1e-15
    extra = np.stack([kt * rho, et * rho, np.zeros(n), np.zeros(n), np.zeros(n), bt * rho])
    # E gains rho k (the turbulent energy is part of the total energy)
    U = U.copy()
    U[Ns + 4] = U[Ns + 4] + extra[0]
    return case, np.concatenate([U, extra])


def rayleigh_taylor_2d(nx: int = 256, ny: int = 1024, seed: int = 7, a0: float = 0.01,
                       modes=None, periodic_x: bool = True):
    """2-D Rayleigh-Taylor: heavy (rho 3, M 87, gamma 1.4) over light (rho 1,
    M 29, gamma 1.4) on [0,0.25]x[0,1], gravity -1 in y, hydrostatic pressure
    with p = 2.5 at y = 0.5, interface 0.5 + a0 cos(8 pi x) + per-column noise
    of +-1 % of a0 (or a sum of single modes a_k cos(2 pi k x / 0.25)), tanh
    mass-fraction profile of width 2 cells; HLLC + van Leer; reflecting (slip)
    walls in y, periodic in x (periodic_x=False: slip x walls, the reference's
    surrogate)."""
    case = Case(dim=2, N=(nx, ny, 1), L=(0.25, 1.0, 1.0 / ny), M=[87.0, 29.0],
                Cv=[cv_of(1.4), cv_of(1.4)], flux=flowSolver.HLLC,
                limiter=typeOfTVDLimiter.vanLeer, gravity=True, g=(0.0, -1.0, 0.0),
                courant=0.4, min_time=1e-15)
    for pt in (0, 1):
        case.set_bc(pt, BCT.periodic if periodic_x else BCT.slip)
    for pt in (3, 4):
        case.set_bc(pt, BCT.slip)
    rng = SplitMix64(seed)
    xs = (np.arange(nx) + 0.5) * (0.25 / nx)
    if modes is None:
        eta = 0.5 + a0 * np.cos(8.0 * math.pi * xs) + np.array(
            [0.01 * a0 * (2.0 * rng.uniform() - 1.0) for _ in range(nx)])
    else:
        eta = 0.5 + sum(a * np.cos(2.0 * math.pi * (k + 1) * xs / 0.25)
                        for k, a in enumerate(modes))
    dy = 1.0 / ny
    ys = (np.arange(ny) + 0.5) * dy
    w = 2.0 * dy
    Y, _ = np.meshgrid(ys, xs, indexing="ij")
    yh = 0.5 * (1.0 + np.tanh((Y - eta[None, :]) / w))
    rho = 1.0 / (yh / 3.0 + (1.0 - yh) / 1.0)
    # hydrostatic p per column: dp/dy = -rho, p = 2.5 at y = 0.5
    p = np.zeros_like(rho)
    j0 = int(np.floor(0.5 / dy))
    p[j0] = 2.5 - 0.5 * rho[j0] * (ys[j0] - 0.5) * 2.0
    for j in range(j0 + 1, ny):
        p[j] = p[j - 1] - 0.5 * (rho[j - 1] + rho[j]) * dy
    for j in range(j0 - 1, -1, -1):
        p[j] = p[j + 1] + 0.5 * (rho[j + 1] + rho[j]) * dy
    c = np.stack([(yh * rho / 87.0).reshape(-1), ((1.0 - yh) * rho / 29.0).reshape(-1)])
    v = np.zeros((3, nx * ny))
    return case, conserved_from_primitive(case, c, v, p.reshape(-1))


def _rk_concentration(case: Case, X, p, T):
    """Total molar concentration c of the Redlich-Kwong mixture with mole
    fractions X at pressure p and temperature T: the root of
    c R T / (1 - c b) - a c^2 / (sqrt(T) (1 + c b)) = p (RedlichKwongFluid.cpp
    pressure) with a, b mixed by the mole fractions (mixtureRedlichKwong.cpp),
    by Newton's method from the ideal-gas value."""
    from .gpu import rk_matrices

    am, bm = rk_matrices(case)
    a = sum(X[k] * X[l] * am[k, l] for k in range(case.Ns) for l in range(case.Ns))
    b = sum(X[k] * X[l] * bm[k, l] for k in range(case.Ns) for l in range(case.Ns))
    R, sT = float(case.R), np.sqrt(T)
    c = p / (R * T)
    for _ in range(100):
        f = c * R * T / (1 - c * b) - a * c * c / (sT * (1 + c * b)) - p
        df = R * T / (1 - c * b) ** 2 - a * c * (2 + c * b) / (sT * (1 + c * b) ** 2)
        dc = f / df
        c = c - dc
        if np.all(np.abs(dc) <= 1e-15 * np.abs(c)):
            break
    return c


def kelvin_helmholtz_2d(n: int = 512, seed: int = 11, T0: float = 300.0, p0: float = 5.0e6):
    """2-D Kelvin-Helmholtz shear layer u = +-0.5 tanh((y - 0.5) / 0.01) in a
    periodic box [0,1]^2 (SI units, R = 8.314462618 J/(mol K)) with two
    species, CO2-like (M 0.044 kg/mol, Cv 28.5 J/(mol K), Tc 304.13 K,
    Pc 7.3773 MPa) below the layer and N2-like (M 0.028, Cv 20.8, Tc 126.19 K,
    Pc 3.3958 MPa) above it (mass-fraction tanh profile of the same width), at
    T = 300 K and p = 5 MPa through the Redlich-Kwong cubic mixture (its
    temperature and pressure found per cell by the root of a cubic in sqrt(T),
    solved in FP64 on the device), a v perturbation of amplitude up to 1e-3
    over modes 1-4 (seeded amplitudes and phases, Gaussian envelope of width
    0.05 around the layer); HLL + minmod."""
    case = Case(dim=2, N=(n, n, 1), L=(1.0, 1.0, 1.0 / n), M=[0.044, 0.028],
                Cv=[28.5, 20.8], R=8.314462618, flux=flowSolver.HLL,
                limiter=typeOfTVDLimiter.minmod, courant=0.4, min_time=1e-15,
                gas_model=gasModel.RedlichKwong, Tc=[304.13, 126.19], Pc=[7.3773e6, 3.3958e6])
    for pt in (0, 1, 3, 4):
        case.set_bc(pt, BCT.periodic)
    rng = SplitMix64(seed)
    x, y, _ = _cell_centres(case)
    u = 0.5 * np.tanh((y - 0.5) / 0.01)
    vv = np.zeros_like(x)
    for k in range(1, 5):
        amp, ph = 1e-3 * rng.uniform(), 2.0 * math.pi * rng.uniform()
        vv += amp * np.sin(2.0 * math.pi * k * x + ph) * np.exp(-((y - 0.5) / 0.05) ** 2)
    Y1 = 0.5 * (1.0 - np.tanh((y - 0.5) / 0.01))  # CO2 below the layer
    n1, n2 = Y1 / case.M[0], (1.0 - Y1) / case.M[1]
    X = [n1 / (n1 + n2), n2 / (n1 + n2)]
    p = np.full_like(x, p0)
    c0 = _rk_concentration(case, X, p, np.full_like(x, T0))
    c = np.stack([X[0] * c0, X[1] * c0])
    v = np.stack([u, vv, np.zeros_like(u)])
    return case, conserved_from_primitive(case, c, v, p)


def _vdw_concentration(case: Case, X, p, T):
    """Total molar concentration c of the van der Waals mixture with mole
    fractions X at pressure p and temperature T: the gas root of
    c R T / (1 - c b) - a c^2 = p (vanDerWaalsFluid.cpp:14-22 with U from T)
    with a, b mixed by the mole fractions (mixtureVanDerWaals.cpp:104-109),
    by Newton's method from the ideal-gas value."""
    from .gpu import vdw_matrices

    am, bm = vdw_matrices(case)
    a = sum(X[k] * X[l] * am[k, l] for k in range(case.Ns) for l in range(case.Ns))
    b = sum(X[k] * X[l] * bm[k, l] for k in range(case.Ns) for l in range(case.Ns))
    R = float(case.R)
    c = p / (R * T)
    for _ in range(100):
        f = c * R * T / (1 - c * b) - a * c * c - p
        df = R * T / (1 - c * b) ** 2 - 2 * a * c
        dc = f / df
        c = c - dc
        if np.all(np.abs(dc) <= 1e-15 * np.abs(c)):
            break
    return c


def kelvin_helmholtz_turbulent_2d(n: int = 512, seed: int = 11, T0: float = 300.0,
                                  p0: float = 5.0e6):
    """The 2-D Kelvin-Helmholtz shear layer with k-epsilon turbulence: the
    layer, box, species and perturbation of kelvin_helmholtz_2d through the
    van der Waals cubic mixture (a, b from Tc, Pc), the diffusion stage with
    the molecular coefficients of CO2/N2 near 5 MPa (viscosity 1.8e-5 Pa s,
    diffusion 3e-7 m^2/s, heat conduction 0.03 W/(m K); explicit solver) and
    the shear k-epsilon model with the reference decks' parameters
    (Cmu 0.09, turbulent Schmidt / Prandtl 0.7, sigma_k 0.7, sigma_eps 0.6,
    C1 1.44, C2 1.92, C3 2.0); k = 1.5e-4 m^2/s^2 (2 % intensity of the
    0.5 m/s stream) in a Gaussian of width 0.02 around the layer over a 1e-8
    background, eps = Cmu^(3/4) k^(3/2) / 0.01 m; HLL + minmod, periodic box."""
    from .case import Turbulence

    case = Case(dim=2, N=(n, n, 1), L=(1.0, 1.0, 1.0 / n), M=[0.044, 0.028],
                Cv=[28.5, 20.8], R=8.314462618, flux=flowSolver.HLL,
                limiter=typeOfTVDLimiter.minmod, courant=0.4, min_time=1e-15,
                gas_model=gasModel.vanDerWaals, Tc=[304.13, 126.19], Pc=[7.3773e6, 3.3958e6],
                diffusion=True, transport=(1.8e-5, 3e-7, 0.03), turbulence=Turbulence())
    for pt in (0, 1, 3, 4):
        case.set_bc(pt, BCT.periodic)
        case.turbulence.bc_k[pt] = BC(BCT.periodic)
        case.turbulence.bc_eps[pt] = BC(BCT.periodic)
    rng = SplitMix64(seed)
    x, y, _ = _cell_centres(case)
    u = 0.5 * np.tanh((y - 0.5) / 0.01)
    vv = np.zeros_like(x)
    for k in range(1, 5):
        amp, ph = 1e-3 * rng.uniform(), 2.0 * math.pi * rng.uniform()
        vv += amp * np.sin(2.0 * math.pi * k * x + ph) * np.exp(-((y - 0.5) / 0.05) ** 2)
    Y1 = 0.5 * (1.0 - np.tanh((y - 0.5) / 0.01))
    n1, n2 = Y1 / case.M[0], (1.0 - Y1) / case.M[1]
    X = [n1 / (n1 + n2), n2 / (n1 + n2)]
    p = np.full_like(x, p0)
    c0 = _vdw_concentration(case, X, p, np.full_like(x, T0))
    c = np.stack([X[0] * c0, X[1] * c0])
    v = np.stack([u, vv, np.zeros_like(u)])
    kt = 1.5e-4 * np.exp(-((y - 0.5) / 0.02) ** 2) + 1e-8
    et = 0.09 ** 0.75 * kt ** 1.5 / 0.01
    return case, conserved_from_primitive(case, c, v, p, kt, et)
