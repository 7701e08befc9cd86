"""ctypes binding of the C ABI in ``include/schemi_gpu.h``.

This is the Python side of the drop-in boundary: every call goes through
``schemi_amd/libschemi_gpu.so``.  There is no CPU fallback: when the library
is missing or no HIP device is present, constructing a solver raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .case import Case
from .enums import boundaryConditionType as BCT
from .enums import errors

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libschemi_gpu.so")
MAXS = 9

PREC_FP64 = 0
PREC_MIXED = 1


class SchemiGpuError(RuntimeError):
    """Raised with the reference's error code (errorsEnum.hpp:15-30)."""

    def __init__(self, code: int, where: str = ""):
        self.code = int(code)
        msg = _lib().schemi_gpu_strerror(self.code).decode()
        super().__init__(f"{where}: {msg} (errors::{_name(self.code)})")


def _name(code):
    try:
        return errors(code).name
    except ValueError:
        return str(code)


class _BC(C.Structure):
    _fields_ = [("type", C.c_int32), ("value", C.c_double * 3)]


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n", C.c_int64 * 3), ("extent", C.c_double * 3),
                ("origin", C.c_double * 3)]


class MixDesc(C.Structure):
    _fields_ = [("ns", C.c_int32), ("M", C.c_double * MAXS), ("Cv", C.c_double * MAXS),
                ("R", C.c_double), ("gas_model", C.c_int32), ("Tcrit", C.c_double * MAXS),
                ("Pcrit", C.c_double * MAXS), ("p0", C.c_double * MAXS),
                ("gamma", C.c_double * MAXS), ("epsLJ", C.c_double * MAXS),
                ("sigmaLJ", C.c_double * MAXS), ("kataoka_b_critical", C.c_int32),
                ("kataoka_b_coeff", C.c_double)]


class SchemeDesc(C.Structure):
    _fields_ = [("flux", C.c_int32), ("limiter", C.c_int32), ("precision", C.c_int32),
                ("gravity", C.c_int32), ("g", C.c_double * 3), ("courant", C.c_double),
                ("min_time", C.c_double), ("const_dt", C.c_double),
                ("bc_c", (_BC * 6) * MAXS), ("bc_v", _BC * 6), ("bc_p", _BC * 6),
                ("step_solver", C.c_int32), ("diffusion", C.c_int32),
                ("transport_mu", C.c_double), ("transport_D", C.c_double),
                ("transport_kappa", C.c_double), ("diffusion_solver", C.c_int32),
                ("diffusion_iterations", C.c_int32), ("diffusion_molar_mass", C.c_int32),
                ("diffusion_enthalpy_flow", C.c_int32), ("timestep_source", C.c_int32),
                ("source_coeff", C.c_double), ("turbulence", C.c_int32),
                ("turbulence_model", C.c_int32), ("turb_Cmu", C.c_double),
                ("turb_sigma_sc", C.c_double), ("turb_sigma_t", C.c_double),
                ("turb_sigma_e", C.c_double), ("turb_sigma_k", C.c_double),
                ("turb_sigma_eps", C.c_double), ("turb_CMS_R", C.c_double),
                ("turb_CMS_D", C.c_double), ("turb_min_k", C.c_double),
                ("turb_min_eps", C.c_double), ("turb_C", C.c_double * 3),
                ("bc_k", _BC * 6), ("bc_eps", _BC * 6), ("turb_sigma_a", C.c_double),
                ("turb_kepsa_C", C.c_double * 8), ("bc_a", _BC * 6),
                ("reconstruction", C.c_int32), ("turb_sigma_b", C.c_double),
                ("turb_min_b", C.c_double), ("turb_bhr_C", C.c_double * 12),
                ("bc_b", _BC * 6)]


_LIB = None
_SYMBOLS = ("schemi_gpu_create", "schemi_gpu_destroy", "schemi_gpu_set_state",
            "schemi_gpu_get_state", "schemi_gpu_step", "schemi_gpu_run",
            "schemi_gpu_stage", "schemi_gpu_face_flux", "schemi_gpu_device_state",
            "schemi_gpu_sync", "schemi_gpu_timing", "schemi_gpu_strerror",
            "schemi_gpu_face_permutation", "schemi_gpu_axis_table", "schemi_gpu_state_device",
            "schemi_gpu_adjoint_step", "schemi_gpu_mixing_width",
            "schemi_gpu_mixing_width_noalloc", "schemi_gpu_diffusion_timing",
            "schemi_gpu_source_timestep", "schemi_gpu_get_turbulence_stage")


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "schemi_amd/libschemi_gpu.so is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, dp, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.schemi_gpu_create.restype = C.c_int
        L.schemi_gpu_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(MixDesc),
                                        C.POINTER(SchemeDesc), C.POINTER(vp)]
        L.schemi_gpu_destroy.restype = None
        L.schemi_gpu_destroy.argtypes = [vp]
        L.schemi_gpu_set_state.restype = C.c_int
        L.schemi_gpu_set_state.argtypes = [vp, dp, C.c_int64, C.c_int32]
        L.schemi_gpu_get_state.restype = C.c_int
        L.schemi_gpu_get_state.argtypes = [vp, dp, dp, C.c_int32]
        L.schemi_gpu_step.restype = C.c_int
        L.schemi_gpu_step.argtypes = [vp, C.c_double, dp]
        L.schemi_gpu_run.restype = C.c_int
        L.schemi_gpu_run.argtypes = [vp, C.c_int64, C.c_double, C.c_double, dp, lp, dp]
        L.schemi_gpu_stage.restype = C.c_int
        L.schemi_gpu_stage.argtypes = [vp, C.c_double]
        L.schemi_gpu_face_flux.restype = C.c_int
        L.schemi_gpu_face_flux.argtypes = [vp, C.c_int64, C.POINTER(C.c_int32), dp, dp, dp,
                                           dp]
        L.schemi_gpu_device_state.restype = C.c_int
        L.schemi_gpu_device_state.argtypes = [vp, C.POINTER(dp), lp]
        L.schemi_gpu_sync.restype = C.c_int
        L.schemi_gpu_sync.argtypes = [vp]
        L.schemi_gpu_state_device.restype = C.c_int
        L.schemi_gpu_state_device.argtypes = [vp, vp, vp]
        L.schemi_gpu_adjoint_step.restype = C.c_int
        L.schemi_gpu_adjoint_step.argtypes = [vp, C.c_double, vp, vp, C.c_int32]
        L.schemi_gpu_mixing_width.restype = C.c_int
        L.schemi_gpu_mixing_width.argtypes = [vp, C.c_int32, dp, vp]
        L.schemi_gpu_source_timestep.restype = C.c_int
        L.schemi_gpu_source_timestep.argtypes = [vp, dp]
        L.schemi_gpu_get_turbulence_stage.restype = C.c_int
        L.schemi_gpu_get_turbulence_stage.argtypes = [vp, dp, C.c_int32]
        L.schemi_gpu_mixing_width_noalloc.restype = C.c_int
        L.schemi_gpu_mixing_width_noalloc.argtypes = [vp, C.c_int32, dp, vp]
        L.schemi_gpu_timing.restype = C.c_int
        L.schemi_gpu_timing.argtypes = [vp, C.c_int32, dp, lp]
        L.schemi_gpu_diffusion_timing.restype = C.c_int
        L.schemi_gpu_diffusion_timing.argtypes = [vp, dp, dp, lp]
        L.schemi_gpu_strerror.restype = C.c_char_p
        L.schemi_gpu_strerror.argtypes = [C.c_int]
        L.schemi_gpu_face_permutation.restype = C.c_int
        L.schemi_gpu_face_permutation.argtypes = [C.POINTER(MeshDesc), lp, C.c_int64]
        L.schemi_gpu_axis_table.restype = C.c_int
        L.schemi_gpu_axis_table.argtypes = [C.POINTER(MeshDesc), C.c_int32, C.c_int32,
                                            C.c_int32, dp, lp]
        _LIB = L
    return _LIB


def library():
    """The loaded C-ABI library (raises when it is not built)."""
    return _lib()


def exported_symbols():
    return _SYMBOLS


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def mesh_desc(case: Case) -> MeshDesc:
    m = MeshDesc()
    m.dim = case.dim
    for a in range(3):
        m.n[a] = int(case.N[a])
        m.extent[a] = float(case.L[a])
        m.origin[a] = float(case.origin[a])
    return m


def mix_desc(case: Case) -> MixDesc:
    x = MixDesc()
    x.ns = case.Ns
    for k in range(case.Ns):
        x.M[k] = float(case.M[k])
        x.Cv[k] = float(case.Cv[k])
    x.R = float(case.R)
    x.gas_model = int(case.gas_model)
    for k in range(case.Ns):
        x.Tcrit[k] = float(case.Tc[k])
        x.Pcrit[k] = float(case.Pc[k])
        x.p0[k] = float(case.p0[k])
        x.gamma[k] = float(case.gammas[k])
        x.epsLJ[k] = float(case.epsLJ[k])
        x.sigmaLJ[k] = float(case.sigmaLJ[k])
    x.kataoka_b_critical = 1 if case.kataoka_b[0] == "critical" else 0
    x.kataoka_b_coeff = float(case.kataoka_b[1])
    return x


def scheme_desc(case: Case, precision: int = PREC_FP64, const_dt: float = -1.0) -> SchemeDesc:
    s = SchemeDesc()
    s.flux = int(case.flux)
    s.limiter = int(case.limiter)
    s.precision = int(precision)
    s.gravity = 1 if case.gravity else 0
    for a in range(3):
        s.g[a] = float(case.g[a])
    s.courant = float(case.courant)
    s.min_time = float(case.min_time)
    s.const_dt = float(const_dt)
    for p in range(6):
        for k in range(case.Ns):
            s.bc_c[k][p].type = int(case.bc_c[k][p].type)
            for a in range(3):
                s.bc_c[k][p].value[a] = float(case.bc_c[k][p].value[a])
        s.bc_v[p].type = int(case.bc_v[p].type)
        s.bc_p[p].type = int(case.bc_p[p].type)
        for a in range(3):
            s.bc_v[p].value[a] = float(case.bc_v[p].value[a])
            s.bc_p[p].value[a] = float(case.bc_p[p].value[a])
    s.step_solver = int(case.step_solver)
    s.diffusion = 1 if case.diffusion else 0
    s.transport_mu, s.transport_D, s.transport_kappa = (float(x) for x in case.transport)
    # matrixSolver enum values (matrixSolverEnum.hpp:15-25)
    s.diffusion_solver = {"GaussSeidel": 0, "explicit": 3, "Jacobi": 6}.get(case.sle_solver, -1)
    s.diffusion_iterations = int(case.sle_iters)
    s.diffusion_molar_mass = 1 if case.molar_mass_correction else 0
    s.diffusion_enthalpy_flow = 1 if case.enthalpy_flow else 0
    s.timestep_source = 1 if case.timestep_source else 0
    s.source_coeff = float(case.source_coeffs[0])
    s.reconstruction = 1 if case.reconstruction == "star" else 0
    t = case.turbulence
    if t is not None:
        s.turbulence = 1
        s.turbulence_model = int(t.model)
        (s.turb_Cmu, s.turb_sigma_sc, s.turb_sigma_t, s.turb_sigma_e, s.turb_sigma_k,
         s.turb_sigma_eps, s.turb_CMS_R, s.turb_CMS_D, s.turb_min_k,
         s.turb_min_eps) = (float(x) for x in t.params())
        if t.b_field:
            for i, x in enumerate(t.C[:12]):
                s.turb_bhr_C[i] = float(x)
            s.turb_sigma_a = float(t.sigmaa)
            s.turb_sigma_b = float(t.sigmab)
            s.turb_min_b = float(t.minb)
        elif t.a_field:
            for i, x in enumerate(t.C[:8]):
                s.turb_kepsa_C[i] = float(x)
            s.turb_sigma_a = float(t.sigmaa)
        else:
            for i, x in enumerate(t.C[:3]):
                s.turb_C[i] = float(x)
        for p in range(6):
            for dst, src in ((s.bc_k, t.bc_k), (s.bc_eps, t.bc_eps), (s.bc_a, t.bc_a),
                             (s.bc_b, t.bc_b)):
                dst[p].type = int(src[p].type)
                for a in range(3):
                    dst[p].value[a] = float(src[p].value[a])
    return s


def face_permutation(case: Case) -> np.ndarray:
    """Reference face id -> structured (axis-major) face id."""
    nf = _nfaces(case)
    out = np.zeros(nf, dtype=np.int64)
    m = mesh_desc(case)
    r = _lib().schemi_gpu_face_permutation(C.byref(m),
                                           out.ctypes.data_as(C.POINTER(C.c_int64)), nf)
    if r:
        raise SchemiGpuError(r, "face_permutation")
    return out


TABLES = ("xc", "e", "xf", "wo", "wn", "dn", "dlo", "dhi", "ghost", "separable",
          "td_x_inner", "td_x_bnd", "td_y_inner", "td_y_bnd", "td_z_inner", "td_z_bnd")


def axis_table(case: Case, axis: int, which: str, periodic_mask: int = 0) -> np.ndarray:
    """One per-axis geometry table of ``case``'s mesh (host computation)."""
    m = mesh_desc(case)
    w = TABLES.index(which)
    n = C.c_int64()
    r = _lib().schemi_gpu_axis_table(C.byref(m), periodic_mask, axis, w, None, C.byref(n))
    if r:
        raise SchemiGpuError(r, "axis_table")
    out = np.zeros(n.value)
    r = _lib().schemi_gpu_axis_table(C.byref(m), periodic_mask, axis, w, _dp(out),
                                     C.byref(n))
    if r:
        raise SchemiGpuError(r, "axis_table")
    return out


def _nfaces(case: Case) -> int:
    Nx, Ny, Nz = case.N
    if case.dim == 1:
        return Nx + 1
    if case.dim == 2:
        return Ny * (Nx + 1) + Nx * (Ny + 1)
    return Ny * Nz * (Nx + 1) + Nx * Nz * (Ny + 1) + Nx * Ny * (Nz + 1)


class GpuSolver:
    """One device context on the current GPU: the step of ``case``, SSP-RK2
    or MUSCL-Hancock as ``case.step_solver`` selects."""

    def __init__(self, case: Case, precision: int = PREC_FP64, const_dt: float = -1.0):
        self.case = case
        self.precision = precision
        self._m = mesh_desc(case)
        self._x = mix_desc(case)
        self._s = scheme_desc(case, precision, const_dt)
        self._ctx = C.c_void_p()
        r = _lib().schemi_gpu_create(C.byref(self._m), C.byref(self._x), C.byref(self._s),
                                     C.byref(self._ctx))
        if r:
            self._ctx = None
            raise SchemiGpuError(r, "schemi_gpu_create")
        self.n = case.ncells
        self.nf = case.nfields

    def close(self):
        if getattr(self, "_ctx", None):
            _lib().schemi_gpu_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    def _check(self, r, where):
        if r:
            raise SchemiGpuError(r, where)

    def set_conserved(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nf, self.n)
        self._check(_lib().schemi_gpu_set_state(self._ctx, _dp(U), self.n, 0), "set_state")

    def conserved(self) -> np.ndarray:
        U = np.zeros((self.nf, self.n))
        self._check(_lib().schemi_gpu_get_state(self._ctx, _dp(U), None, 0), "get_state")
        return U

    def primitives(self) -> dict:
        Ns, n = self.case.Ns, self.n
        W = np.zeros((Ns + 8, n))
        self._check(_lib().schemi_gpu_get_state(self._ctx, None, _dp(W), 0), "get_state")
        return {"c": W[:Ns], "c0": W[Ns], "v": W[Ns + 1:Ns + 4], "p": W[Ns + 4],
                "U": W[Ns + 5], "T": W[Ns + 6], "cs": W[Ns + 7]}

    def step(self, dt: float) -> float:
        nxt = C.c_double(0.0)
        self._check(_lib().schemi_gpu_step(self._ctx, float(dt), C.byref(nxt)), "step")
        return nxt.value

    def stage(self, dt: float):
        self._check(_lib().schemi_gpu_stage(self._ctx, float(dt)), "stage")

    def run(self, nsteps: int, dt0: float, t_end: float = float("inf")):
        t, s, d = C.c_double(), C.c_int64(), C.c_double()
        self._check(_lib().schemi_gpu_run(self._ctx, int(nsteps), float(dt0), float(t_end),
                                          C.byref(t), C.byref(s), C.byref(d)), "run")
        return t.value, s.value, d.value

    def face_flux(self, axis, sgn, left, right) -> np.ndarray:
        axis = np.ascontiguousarray(axis, dtype=np.int32)
        sgn = np.ascontiguousarray(sgn, dtype=np.float64)
        left = np.ascontiguousarray(left, dtype=np.float64)
        right = np.ascontiguousarray(right, dtype=np.float64)
        nf = axis.shape[0]
        F = np.zeros((nf, self.case.Ns + 5))
        self._check(_lib().schemi_gpu_face_flux(
            self._ctx, nf, axis.ctypes.data_as(C.POINTER(C.c_int32)), _dp(sgn), _dp(left),
            _dp(right), _dp(F)), "face_flux")
        return F

    def device_state(self):
        p = C.POINTER(C.c_double)()
        n = C.c_int64()
        self._check(_lib().schemi_gpu_device_state(self._ctx, C.byref(p), C.byref(n)),
                    "device_state")
        return C.cast(p, C.c_void_p).value, n.value

    def sync(self):
        self._check(_lib().schemi_gpu_sync(self._ctx), "sync")

    # -- device-resident state, adjoint and mixing width (device pointers) ----
    def state_device(self, src: int = 0, dst: int = 0):
        """src (device address or 0) becomes the state; then the state is
        copied to dst (device address or 0)."""
        self._check(_lib().schemi_gpu_state_device(self._ctx, src or None, dst or None),
                    "state_device")

    def adjoint_step(self, dt: float, lam_in: int, lam_out: int, what: int = 0):
        """lam_out = (dU^{n+1}/dU^n)^T lam_in at the current state (device
        addresses of (Ns + 5, n) FP64 arrays); what 1: one stage."""
        self._check(_lib().schemi_gpu_adjoint_step(self._ctx, float(dt), lam_in, lam_out,
                                                   int(what)), "adjoint_step")

    def mixing_width(self, species: int = 0, grad: int = 0) -> float:
        """sum_i V_i Y (1 - Y), Y = rho_species / rho; dJ/dU into grad (device
        address or 0)."""
        J = C.c_double()
        self._check(_lib().schemi_gpu_mixing_width(self._ctx, int(species), C.byref(J),
                                                   grad or None), "mixing_width")
        return J.value

    def source_timestep(self) -> float:
        """Source time step of the last finished step (veryBig when none)."""
        v = C.c_double()
        self._check(_lib().schemi_gpu_source_timestep(self._ctx, C.byref(v)), "source_timestep")
        return v.value

    def turbulence_stage(self) -> dict:
        """k and epsilon at the start of the last diffusion stage and, with the
        kEpsA model, that stage's algebraic b."""
        t = self.case.turbulence
        kepsa = t is not None and t.a_field and not t.b_field
        out = np.zeros((3 if kepsa else 2, self.n))
        self._check(_lib().schemi_gpu_get_turbulence_stage(self._ctx, _dp(out), 0),
                    "get_turbulence_stage")
        d = {"k": out[0], "eps": out[1]}
        if kepsa:
            d["b"] = out[2]
        return d

    def mixing_width_noalloc(self, species: int = 0, grad: int = 0) -> float:
        """mixing_width through the no-allocation entry point (partial sums in
        the library's fixed device scratch)."""
        J = C.c_double()
        self._check(_lib().schemi_gpu_mixing_width_noalloc(self._ctx, int(species),
                                                           C.byref(J), grad or None),
                    "mixing_width_noalloc")
        return J.value

    def timing(self, enable: bool):
        avg, cnt = C.c_double(), C.c_int64()
        self._check(_lib().schemi_gpu_timing(self._ctx, 1 if enable else 0, C.byref(avg),
                                             C.byref(cnt)), "timing")
        return avg.value, cnt.value

    def diffusion_timing(self):
        """Average (prims pass ms, update pass ms, stages) of the diffusion
        stages timed while timing was on."""
        a, b, cnt = C.c_double(), C.c_double(), C.c_int64()
        self._check(_lib().schemi_gpu_diffusion_timing(self._ctx, C.byref(a), C.byref(b),
                                                       C.byref(cnt)), "diffusion_timing")
        return a.value, b.value, cnt.value


_LIBM = None


def _libm():
    global _LIBM
    if _LIBM is None:
        _LIBM = C.CDLL("libm.so.6")
        for f in (_LIBM.cbrt, _LIBM.pow, _LIBM.sqrt):
            f.restype = C.c_double
        _LIBM.cbrt.argtypes = [C.c_double]
        _LIBM.sqrt.argtypes = [C.c_double]
        _LIBM.pow.argtypes = [C.c_double, C.c_double]
    return _LIBM


def rk_matrices(case: Case):
    """Redlich-Kwong mixing matrices a_kl, b_kl (mixtureRedlichKwong.cpp:19-67):
    chi = cbrt(2) - 1 with the correctly rounded cube root (the reference's
    constant expression), Sigma_a = 1 / (9 chi), Sigma_b = chi / 3."""
    L = _libm()
    chi = float.fromhex("0x1.428a2f98d728bp+0") - 1.0
    Sa, Sb = 1.0 / (9.0 * chi), chi / 3.0
    n = case.Ns
    a = np.zeros((n, n))
    b = np.zeros((n, n))
    R = float(case.R)
    for k in range(n):
        Tc, Pc = float(case.Tc[k]), float(case.Pc[k])
        b[k, k] = Sb * R * Tc / Pc
        a[k, k] = Sa * R * R * Tc * Tc * L.sqrt(Tc) / Pc
    for k in range(n):
        for l in range(n):
            if k != l:
                b[k, l] = L.pow(0.5 * (L.cbrt(b[k, k]) + L.cbrt(b[l, l])), 3.0)
                a[k, l] = L.sqrt(a[k, k] * a[l, l])
    return a, b


def _rk_root(A, C, D):
    """Positive root of A s^3 + C s + D (A > 0, D < 0) by Newton's method from
    the root without D, elementwise (the device's rk_root)."""
    s = np.where(C < 0, np.sqrt(np.maximum(-C / A, 0.0)), np.cbrt(-D / A))
    for _ in range(60):
        s2 = s * s
        ds = ((A * s2 + C) * s + D) / (3.0 * A * s2 + C)
        s = s - ds
        if np.all(np.abs(ds) <= 1e-13 * s):
            break
    return s


def kataoka_matrices(case: Case):
    """V0_kl, eps_kl, b_kl of the Kataoka-van der Waals mixture
    (mixtureKataokaVanDerWaals.cpp:19-77)."""
    L = _libm()
    NA = 6.02214076E23
    n = case.Ns
    V0, eps, b = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    crit, coeff = case.kataoka_b[0] == "critical", float(case.kataoka_b[1])
    for k in range(n):
        s3 = float(case.sigmaLJ[k]) * float(case.sigmaLJ[k]) * float(case.sigmaLJ[k])
        b[k, k] = coeff * float(case.R) * float(case.Tc[k]) / (8 * float(case.Pc[k])) if crit \
            else coeff * s3 * NA
        V0[k, k] = s3 / L.sqrt(2.0) * NA
        eps[k, k] = float(case.epsLJ[k])
    for k in range(n):
        for l in range(n):
            if k != l:
                b[k, l] = L.pow(0.5 * (L.cbrt(b[k, k]) + L.cbrt(b[l, l])), 3.0)
                V0[k, l] = L.pow(0.5 * (L.cbrt(V0[k, k]) + L.cbrt(V0[l, l])), 3.0)
                eps[k, l] = L.sqrt(eps[k, k] * eps[l, l])
    return V0, eps, b


def _kataoka_U(case: Case, c, c0, p):
    """UvFromp of the Kataoka-van der Waals mixture (KataokaVanDerWaalsFluid.cpp:
    17-62, 242-261): T from c R T / (1 - c b) + pA(T) = p by the reference's
    Newton iteration, then U = c Cv T + UvA(T)."""
    V0m, epsm, bm = kataoka_matrices(case)
    Ns = case.Ns
    X = [c[k] / c0 for k in range(Ns)]
    cv = sum(X[k] * case.Cv[k] for k in range(Ns))
    V0 = np.zeros_like(c0)
    eps = np.zeros_like(c0)
    b = np.zeros_like(c0)
    for k in range(Ns):
        for l in range(Ns):
            V0 = V0 + X[k] * X[l] * V0m[k, l]
            eps = eps + X[k] * X[l] * epsm[k, l]
            b = b + X[k] * X[l] * bm[k, l]
    R = float(case.R)

    def series(T):
        V0c, epsRT = V0 * c0, 0.5 * eps / (R * T)
        qa, su, sp = np.ones_like(T), np.zeros_like(T), np.zeros_like(T)
        for J in range(2, 14):
            x = np.power(V0c, (J - 1.0) / 3.0) * np.exp((J - 1.0) * epsRT)
            qa, su, sp = qa + x, su - (J - 1.0) * x, sp - (J - 1.0) / 3.0 * x
        return qa, su, sp

    def f(T):
        qa, su, sp = series(T)
        return c0 * R * T / (1 - c0 * b) + c0 * R * T * sp / qa - p

    df = c0 * R / (1 - c0 * b)
    x1 = p / ((c0 + 5 * np.finfo(float).eps) * R)
    x2 = x1 - f(x1) / df
    for _ in range(100):
        if np.all(np.abs(x2 - x1) <= 1e-12):
            break
        x1 = x2
        x2 = x1 - f(x1) / df
    qa, su, sp = series(x2)
    return c0 * cv * x2 + 0.5 * c0 * eps * su / qa


def vdw_matrices(case: Case):
    """van der Waals mixing matrices a_kl, b_kl (mixtureVanDerWaals.cpp:19-58),
    with glibc's cbrt and pow as the reference's build uses them."""
    global _LIBM
    if _LIBM is None:
        _LIBM = C.CDLL("libm.so.6")
        for f in (_LIBM.cbrt, _LIBM.pow, _LIBM.sqrt):
            f.restype = C.c_double
        _LIBM.cbrt.argtypes = [C.c_double]
        _LIBM.sqrt.argtypes = [C.c_double]
        _LIBM.pow.argtypes = [C.c_double, C.c_double]
    n = case.Ns
    a = np.zeros((n, n))
    b = np.zeros((n, n))
    for k in range(n):
        b[k, k] = float(case.Tc[k]) * float(case.R) / (8 * float(case.Pc[k]))
        a[k, k] = float(case.Pc[k]) * b[k, k] * b[k, k] * 27
    for k in range(n):
        for l in range(n):
            if k != l:
                b[k, l] = _LIBM.pow(0.5 * (_LIBM.cbrt(b[k, k]) + _LIBM.cbrt(b[l, l])), 3.0)
                a[k, l] = _LIBM.sqrt(a[k, k] * a[l, l])
    return a, b


def conserved_from_primitive(case: Case, c, v, p, kt=None, et=None, at=None,
                             bt=None) -> np.ndarray:
    """Conserved vector from primitives (c_k, v, p), with the reference's
    initialisation arithmetic (phaseInitialization.cpp:1827-1883) and its
    UvFromp (idealFluid.cpp:19-23, vanDerWaalsFluid.cpp:24-31; Redlich-Kwong
    RedlichKwongFluid.cpp:59-70 with its cubic solved by Newton's method in
    FP64, a few ulps from the reference's long-double root).  With turbulence
    k and epsilon are clipped to the minimal values, rho k = k rho and
    rho eps = rho eps are appended and rho k is added to E
    (phaseInitialization.cpp:1814-1848, 1873-1877); with the kEpsA model the
    turbulent mass flux at (3, n) follows as rho a = a rho (:1850).  Pure host
    arithmetic on numpy arrays; used to prepare initial states."""
    c = np.asarray(c, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    Ns = case.Ns
    n = c.shape[1]
    c0 = np.zeros(n)
    for k in range(Ns):
        c0 = c0 + c[k]
    rho_k = np.stack([c[k] * case.M[k] for k in range(Ns)])
    rho = np.zeros(n)
    for k in range(Ns):
        rho = rho + rho_k[k]
    m = v * rho
    cv = np.zeros(n)
    for k in range(Ns):
        cv = cv + (c[k] / c0) * case.Cv[k]
    if int(case.gas_model) == 1:
        am, bm = vdw_matrices(case)
        X = [c[k] / c0 for k in range(Ns)]
        a = np.zeros(n)
        b = np.zeros(n)
        for k in range(Ns):
            for l in range(Ns):
                a = a + X[k] * X[l] * am[k, l]
                b = b + X[k] * X[l] * bm[k, l]
        acc = a * (c0 * c0)
        U = (p + acc) * (1 - c0 * b) / (case.R / cv) - acc
    elif int(case.gas_model) == 3:
        # stiffened: (p + p0 gamma) / (gamma - 1), p0 and gamma mixed by
        # geometric means (stiffenedFluid.cpp:19-23, mixtureStiffened.cpp:19-43)
        X = [c[k] / c0 for k in range(Ns)]
        p0m = np.zeros(n)
        gm = np.zeros(n)
        for k in range(Ns):
            for l in range(Ns):
                p0kl = case.p0[k] if k == l else _libm().sqrt(case.p0[k] * case.p0[l])
                gkl = case.gammas[k] if k == l else _libm().sqrt(case.gammas[k] * case.gammas[l])
                p0m = p0m + X[k] * X[l] * p0kl
                gm = gm + X[k] * X[l] * gkl
        U = (p + p0m * gm) / (gm - 1.0)
    elif int(case.gas_model) == 4:
        U = _kataoka_U(case, c, c0, p)
    elif int(case.gas_model) == 2:
        am, bm = rk_matrices(case)
        X = [c[k] / c0 for k in range(Ns)]
        a = np.zeros(n)
        b = np.zeros(n)
        for k in range(Ns):
            for l in range(Ns):
                a = a + X[k] * X[l] * am[k, l]
                b = b + X[k] * X[l] * bm[k, l]
        s = _rk_root(c0 * case.R, -p * (1 - c0 * b), -a * (c0 * c0) * (1 - c0 * b) / (1 + c0 * b))
        U = c0 * cv * (s * s) - 1.5 * a * c0 / (b * s) * np.log1p(c0 * b)
    else:
        U = p / (case.R / cv)
    v2 = ((0.0 + v[0] * v[0]) + v[1] * v[1]) + v[2] * v[2]
    t = case.turbulence
    if t is None:
        E = (U + rho * v2 * 0.5) + 0.0
        return np.concatenate([rho_k, rho[None], m, E[None]])
    kt = np.maximum(np.asarray(kt, dtype=np.float64), t.mink)
    et = np.maximum(np.asarray(et, dtype=np.float64), t.mineps)
    rk = kt * rho
    E = (U + rho * v2 * 0.5) + rk
    rows = [rho_k, rho[None], m, E[None], rk[None], (rho * et)[None]]
    if t.a_field:
        rows.append(np.asarray(at, dtype=np.float64) * rho)
    if t.b_field:
        rows.append((rho * np.asarray(bt, dtype=np.float64))[None])
    return np.concatenate(rows)


def periodic(case: Case, axis: int):
    """Marks both patches of ``axis`` periodic for every field."""
    lo, hi = [(0, 1), (3, 4), (2, 5)][axis]
    for p in (lo, hi):
        case.set_bc(p, BCT.periodic)
    return case
