"""ctypes wrapper of the C oracle (test infrastructure only).

Only the tests, the smoke check and the benchmark's CPU-baseline leg use this
module, as the checker. It loads ``oracle/liboracle.so`` built by
``oracle/Makefile`` (``make -C oracle``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXS = 9


class _BC(C.Structure):
    _fields_ = [("type", C.c_int), ("val", C.c_double * 3)]


class _Config(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("N", C.c_long * 3),
        ("L", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("Ns", C.c_int),
        ("M", C.c_double * MAXS),
        ("Cv", C.c_double * MAXS),
        ("R", C.c_double),
        ("flux", C.c_int),
        ("limiter", C.c_int),
        ("gravity", C.c_int),
        ("g", C.c_double * 3),
        ("bc_c", (_BC * 6) * MAXS),
        ("bc_v", _BC * 6),
        ("bc_p", _BC * 6),
        ("gas_model", C.c_int),
        ("Tc", C.c_double * MAXS),
        ("Pc", C.c_double * MAXS),
        ("a_kl", C.c_double * (MAXS * MAXS)),
        ("b_kl", C.c_double * (MAXS * MAXS)),
        ("p0", C.c_double * MAXS),
        ("gam", C.c_double * MAXS),
        ("epsLJ", C.c_double * MAXS),
        ("sigLJ", C.c_double * MAXS),
        ("kb_critical", C.c_int),
        ("kb_coeff", C.c_double),
        ("e_kl", C.c_double * (MAXS * MAXS)),
        ("diffusion", C.c_int),
        ("mu", C.c_double),
        ("D", C.c_double),
        ("kappa", C.c_double),
        ("sle_solver", C.c_int),
        ("sle_iters", C.c_long),
        ("enth_flow", C.c_int),
        ("molmass", C.c_int),
        ("ts_source", C.c_int),
        ("ts_coeff", C.c_double),
        ("turb", C.c_int),
        ("turb_model", C.c_int),
        ("Cmu", C.c_double),
        ("sigmaSc", C.c_double),
        ("sigmaT", C.c_double),
        ("sigmaE", C.c_double),
        ("sigmaK", C.c_double),
        ("sigmaEps", C.c_double),
        ("CMS_R", C.c_double),
        ("CMS_D", C.c_double),
        ("mink", C.c_double),
        ("mineps", C.c_double),
        ("tC", C.c_double * 12),
        ("bc_k", _BC * 6),
        ("bc_eps", _BC * 6),
        ("sigmaa", C.c_double),
        ("bc_a", _BC * 6),
        ("sigmab", C.c_double),
        ("minb", C.c_double),
        ("bc_b", _BC * 6),
        ("star_rec", C.c_int),
    ]


class _State(C.Structure):
    _fields_ = [
        ("n", C.c_long),
        ("Ns", C.c_int),
        ("c", C.POINTER(C.c_double) * (MAXS + 1)),
        ("rho", C.POINTER(C.c_double) * (MAXS + 1)),
        ("v", C.POINTER(C.c_double) * 3),
        ("p", C.POINTER(C.c_double)),
        ("U", C.POINTER(C.c_double)),
        ("T", C.POINTER(C.c_double)),
        ("m", C.POINTER(C.c_double) * 3),
        ("E", C.POINTER(C.c_double)),
        ("k", C.POINTER(C.c_double)),
        ("eps", C.POINTER(C.c_double)),
        ("rk", C.POINTER(C.c_double)),
        ("reps", C.POINTER(C.c_double)),
        ("a", C.POINTER(C.c_double) * 3),
        ("ra", C.POINTER(C.c_double) * 3),
        ("b", C.POINTER(C.c_double)),
        ("rb", C.POINTER(C.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(_Config)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_ncells.restype = C.c_long
        L.or_ncells.argtypes = [C.c_void_p]
        L.or_nfaces.restype = C.c_long
        L.or_nfaces.argtypes = [C.c_void_p]
        L.or_get_state.restype = C.POINTER(_State)
        L.or_get_state.argtypes = [C.c_void_p]
        L.or_init_primitive.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_set_full.argtypes = [C.c_void_p] + [dp] * 10
        L.or_set_conserved.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_init_turbulence.argtypes = [C.c_void_p, dp, dp]
        L.or_set_turbulence_conserved.argtypes = [C.c_void_p, dp, dp]
        L.or_init_turbulence_a.argtypes = [C.c_void_p, dp]
        L.or_set_turbulence_a_conserved.argtypes = [C.c_void_p, dp]
        L.or_init_turbulence_b.argtypes = [C.c_void_p, dp]
        L.or_set_turbulence_b_conserved.argtypes = [C.c_void_p, dp]
        L.or_step.restype = C.c_int
        L.or_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.or_step_hancock.restype = C.c_int
        L.or_step_hancock.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.or_stage.restype = C.c_int
        L.or_stage.argtypes = [C.c_void_p, C.c_double]
        L.or_courant_dt.restype = C.c_double
        L.or_courant_dt.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.or_cell_geometry.argtypes = [C.c_void_p, dp, dp]
        lp = C.POINTER(C.c_long)
        L.or_face_geometry.argtypes = [C.c_void_p, dp, dp, dp, lp, lp, dp, dp]
        L.or_cell_faces.argtypes = [C.c_void_p, lp]
        L.or_hypot3.restype = C.c_double
        L.or_hypot3.argtypes = [C.c_double] * 3
        L.or_face_flux.restype = C.c_int
        L.or_face_flux.argtypes = [C.c_void_p, dp, dp, dp, C.c_double, dp, dp,
                                   C.c_double, dp, dp, dp]
        L.or_limiter_phi.restype = C.c_double
        L.or_limiter_phi.argtypes = [C.c_int, C.c_double]
        _lib = L
    return _lib


def _need(a, size, name):
    """The C side reads exactly size doubles: refuse any other length."""
    if a.size != size or a.dtype != np.float64:
        raise ValueError(f"oracle: {name} needs {size} float64 values, got {a.size} {a.dtype}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(case) -> _Config:
    cfg = _Config()
    cfg.dim = case.dim
    for a in range(3):
        cfg.N[a] = int(case.N[a])
        cfg.L[a] = float(case.L[a])
        cfg.origin[a] = float(case.origin[a])
        cfg.g[a] = float(case.g[a])
    cfg.Ns = case.Ns
    for k in range(case.Ns):
        cfg.M[k] = float(case.M[k])
        cfg.Cv[k] = float(case.Cv[k])
    cfg.R = float(case.R)
    cfg.gas_model = int(getattr(case, "gas_model", 0))
    for k in range(case.Ns):
        cfg.Tc[k] = float(case.Tc[k])
        cfg.Pc[k] = float(case.Pc[k])
        cfg.p0[k] = float(case.p0[k])
        cfg.gam[k] = float(case.gammas[k])
        cfg.epsLJ[k] = float(case.epsLJ[k])
        cfg.sigLJ[k] = float(case.sigmaLJ[k])
    cfg.kb_critical = 1 if case.kataoka_b[0] == "critical" else 0
    cfg.kb_coeff = float(case.kataoka_b[1])
    cfg.diffusion = 1 if getattr(case, "diffusion", False) else 0
    cfg.mu, cfg.D, cfg.kappa = (float(x) for x in getattr(case, "transport", (0, 0, 0)))
    cfg.sle_solver = {"explicit": 0, "Jacobi": 1, "GaussSeidel": 2}[
        getattr(case, "sle_solver", "explicit")]
    cfg.sle_iters = int(getattr(case, "sle_iters", 2000))
    cfg.enth_flow = 1 if getattr(case, "enthalpy_flow", False) else 0
    cfg.molmass = 1 if getattr(case, "molar_mass_correction", False) else 0
    cfg.ts_source = 1 if getattr(case, "timestep_source", False) else 0
    cfg.ts_coeff = float(getattr(case, "source_coeffs", (0.25, 0.5))[0])
    cfg.star_rec = 1 if getattr(case, "reconstruction", "linear") == "star" else 0
    cfg.flux = int(case.flux)
    cfg.limiter = int(case.limiter)
    cfg.gravity = 1 if case.gravity else 0
    tb = getattr(case, "turbulence", None)
    if tb is not None:
        cfg.turb = 1
        cfg.turb_model = int(tb.model)
        (cfg.Cmu, cfg.sigmaSc, cfg.sigmaT, cfg.sigmaE, cfg.sigmaK, cfg.sigmaEps, cfg.CMS_R,
         cfg.CMS_D, cfg.mink, cfg.mineps) = (float(x) for x in tb.params())
        for i, x in enumerate(tb.C):
            cfg.tC[i] = float(x)
        cfg.sigmaa = float(tb.sigmaa)
        cfg.sigmab = float(tb.sigmab)
        cfg.minb = float(tb.minb)
        for p in range(6):
            for dst, src in ((cfg.bc_k, tb.bc_k), (cfg.bc_eps, tb.bc_eps), (cfg.bc_a, tb.bc_a),
                             (cfg.bc_b, tb.bc_b)):
                dst[p].type = int(src[p].type)
                for a in range(3):
                    dst[p].val[a] = float(src[p].value[a])
    for p in range(6):
        for k in range(case.Ns):
            cfg.bc_c[k][p].type = int(case.bc_c[k][p].type)
            for a in range(3):
                cfg.bc_c[k][p].val[a] = float(case.bc_c[k][p].value[a])
        cfg.bc_v[p].type = int(case.bc_v[p].type)
        cfg.bc_p[p].type = int(case.bc_p[p].type)
        for a in range(3):
            cfg.bc_v[p].val[a] = float(case.bc_v[p].value[a])
            cfg.bc_p[p].val[a] = float(case.bc_p[p].value[a])
    return cfg


class Oracle:
    """The reference's SSP-RK2 step restated on the CPU (checker only)."""

    def __init__(self, case):
        self.case = case
        self._cfg = make_config(case)
        self._s = lib().or_create(C.byref(self._cfg))
        if not self._s:
            raise ValueError("oracle: bad configuration")
        self.n = lib().or_ncells(self._s)
        self.nf = lib().or_nfaces(self._s)

    def __del__(self):
        if getattr(self, "_s", None):
            lib().or_destroy(self._s)
            self._s = None

    # -- state ---------------------------------------------------------------
    @property
    def a_field(self) -> bool:
        return bool(self._cfg.turb) and self._cfg.turb_model in (6, 7)

    @property
    def b_field(self) -> bool:
        return bool(self._cfg.turb) and self._cfg.turb_model == 7

    def init_primitive(self, c, v, p, k=None, eps=None, a=None, b=None):
        """Initial state from primitives; with turbulence also k and epsilon
        (clipped to the minimal values as the reference's initialisation) and
        with the kEpsA model a (3, n)."""
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        n = self.n
        _need(c, self.case.Ns * n, "c")
        _need(v, 3 * n, "v")
        _need(p, n, "p")
        if self._cfg.turb:
            k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1)
            eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1)
            _need(k, n, "k")
            _need(eps, n, "eps")
            lib().or_init_turbulence(self._s, _dp(k), _dp(eps))
        if self.a_field:
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            _need(a, 3 * n, "a")
            lib().or_init_turbulence_a(self._s, _dp(a))
        if self.b_field:
            b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
            _need(b, n, "b")
            lib().or_init_turbulence_b(self._s, _dp(b))
        lib().or_init_primitive(self._s, _dp(c), _dp(v), _dp(p))

    def set_full(self, st: dict):
        a = {k: np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
             for k, v in st.items()}
        lib().or_set_full(self._s, _dp(a["c0"]), _dp(a["c"]), _dp(a["rho0"]),
                          _dp(a["rho"]), _dp(a["v"]), _dp(a["p"]), _dp(a["U"]),
                          _dp(a["T"]), _dp(a["m"]), _dp(a["E"]))

    def set_conserved(self, rho_k, rho, m, E, rk=None, reps=None, ra=None, rb=None):
        """State from the conserved vector; with turbulence also rho k and
        rho epsilon, and with the kEpsA model rho a (3, n)."""
        n = self.n
        if self._cfg.turb:
            t = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (rk, reps)]
            _need(t[0], n, "rk")
            _need(t[1], n, "reps")
            lib().or_set_turbulence_conserved(self._s, _dp(t[0]), _dp(t[1]))
        if self.a_field:
            ra = np.ascontiguousarray(ra, dtype=np.float64).reshape(-1)
            _need(ra, 3 * n, "ra")
            lib().or_set_turbulence_a_conserved(self._s, _dp(ra))
        if self.b_field:
            rb = np.ascontiguousarray(rb, dtype=np.float64).reshape(-1)
            _need(rb, n, "rb")
            lib().or_set_turbulence_b_conserved(self._s, _dp(rb))
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
             for x in (rho_k, rho, m, E)]
        for x, size, name in zip(a, (self.case.Ns * n, n, 3 * n, n), ("rho_k", "rho", "m", "E")):
            _need(x, size, name)
        lib().or_set_conserved(self._s, *[_dp(x) for x in a])

    def set_conserved_all(self, U):
        """State from a (nfields, n) conserved array in the library's layout:
        rho_1..rho_Ns, rho, m_x, m_y, m_z, E[, rho k, rho epsilon[, rho a]]."""
        Ns = self.case.Ns
        if self.b_field:
            self.set_conserved(U[:Ns], U[Ns], U[Ns + 1:Ns + 4], U[Ns + 4], U[Ns + 5], U[Ns + 6],
                               U[Ns + 7:Ns + 10], U[Ns + 10])
        elif self.a_field:
            self.set_conserved(U[:Ns], U[Ns], U[Ns + 1:Ns + 4], U[Ns + 4], U[Ns + 5], U[Ns + 6],
                               U[Ns + 7:Ns + 10])
        elif self._cfg.turb:
            self.set_conserved(U[:Ns], U[Ns], U[Ns + 1:Ns + 4], U[Ns + 4], U[Ns + 5], U[Ns + 6])
        else:
            self.set_conserved(U[:Ns], U[Ns], U[Ns + 1:Ns + 4], U[Ns + 4])

    def state(self) -> dict:
        st = lib().or_get_state(self._s).contents
        n, Ns = self.n, self.case.Ns

        def arr(ptr):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

        return {
            "c0": arr(st.c[0]),
            "c": np.stack([arr(st.c[k]) for k in range(1, Ns + 1)]),
            "rho0": arr(st.rho[0]),
            "rho": np.stack([arr(st.rho[k]) for k in range(1, Ns + 1)]),
            "v": np.stack([arr(st.v[a]) for a in range(3)]),
            "p": arr(st.p),
            "U": arr(st.U),
            "T": arr(st.T),
            "m": np.stack([arr(st.m[a]) for a in range(3)]),
            "E": arr(st.E),
            "k": arr(st.k),
            "eps": arr(st.eps),
            "rk": arr(st.rk),
            "reps": arr(st.reps),
            "a": np.stack([arr(st.a[c]) for c in range(3)]),
            "ra": np.stack([arr(st.ra[c]) for c in range(3)]),
            "b": arr(st.b),
            "rb": arr(st.rb),
        }

    def conserved(self) -> np.ndarray:
        """(Ns + 5, n) array: rho_1..rho_Ns, rho, m_x, m_y, m_z, E; with
        turbulence (Ns + 7, n), rho k and rho epsilon appended; with the kEpsA
        model (Ns + 10, n), rho a appended."""
        s = self.state()
        rows = [s["rho"], s["rho0"][None], s["m"], s["E"][None]]
        if self._cfg.turb:
            rows += [s["rk"][None], s["reps"][None]]
        if self.a_field:
            rows += [s["ra"]]
        if self.b_field:
            rows += [s["rb"][None]]
        return np.concatenate(rows)

    # -- stepping ------------------------------------------------------------
    def step(self, dt: float, mode: int = 0) -> int:
        """One step of the case's step solver (SSP-RK2 or MUSCL-Hancock)."""
        if int(getattr(self.case, "step_solver", 1)) == 0:
            return lib().or_step_hancock(self._s, float(dt), int(mode))
        return lib().or_step(self._s, float(dt), int(mode))

    def stage(self, dt: float) -> int:
        return lib().or_stage(self._s, float(dt))

    def courant_dt(self, courant: float, min_time: float) -> float:
        return lib().or_courant_dt(self._s, float(courant), float(min_time))

    # -- geometry --------------------------------------------------------------
    def cell_geometry(self):
        rc = np.zeros(3 * self.n)
        V = np.zeros(self.n)
        lib().or_cell_geometry(self._s, _dp(rc), _dp(V))
        return rc.reshape(self.n, 3), V

    def face_geometry(self):
        nf = self.nf
        rc, N, S = np.zeros(3 * nf), np.zeros(3 * nf), np.zeros(nf)
        own, nei = np.zeros(nf, dtype=np.int64), np.zeros(nf, dtype=np.int64)
        wo, wn = np.zeros(nf), np.zeros(nf)
        lp = C.POINTER(C.c_long)
        lib().or_face_geometry(self._s, _dp(rc), _dp(N), _dp(S),
                               own.ctypes.data_as(lp), nei.ctypes.data_as(lp),
                               _dp(wo), _dp(wn))
        return dict(rc=rc.reshape(nf, 3), N=N.reshape(nf, 3), S=S, owner=own,
                    neighbour=nei, wo=wo, wn=wn)

    def cell_faces(self):
        f = np.zeros(6 * self.n, dtype=np.int64)
        lib().or_cell_faces(self._s, f.ctypes.data_as(C.POINTER(C.c_long)))
        return f.reshape(self.n, 6)

    def face_flux(self, N, cL, vL, pL, cR, vR, pR):
        Ns = self.case.Ns
        Fr, Fm, FE = np.zeros(Ns + 1), np.zeros(3), np.zeros(1)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (N, cL, vL, cR, vR)]
        err = lib().or_face_flux(self._s, _dp(a[0]), _dp(a[1]), _dp(a[2]),
                                 float(pL), _dp(a[3]), _dp(a[4]), float(pR),
                                 _dp(Fr), _dp(Fm), _dp(FE))
        return err, Fr, Fm, FE[0]


def hypot3(x, y, z):
    return lib().or_hypot3(x, y, z)


def limiter_phi(lim, r):
    return lib().or_limiter_phi(int(lim), float(r))
