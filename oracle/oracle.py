"""ctypes front end of the CPU restatement (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker or the timed CPU baseline; the product never does.

Mesh arrays are the reference's own (as dumped by oracle/_ref/joe_ref_dump or
exported by libjoemi): a dict with COUNTS, CVOFA, FAOCV_I/V, NBOCV_I/V, X_FA,
FA_NORMAL, X_CV, CV_VOLUME, NBOCV_LSG, BFA_LSG, ZONES, ZONE_NAMES.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from . import jmdump

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BC = {"CBC": 1, "CBC_SUBSONIC_INLET": 2, "CBC_SUBSONIC_OUTLET": 3, "SYMMETRY": 4, "NEUMANN": 5, "WALL": 6,
      "FARFIELD": 7, "PROFILES": 8}


def read_profiles(path):
    """A boundary-profile file as the reference reads it (JoeWithModels.cpp:272-284):
    "n=<rows>\td=<values>", then rows of values (y, rho, u, v, p, scalars)."""
    text = open(path).read()
    mm = re.match(r"n=(\d+)\s*d=(\d+)", text)
    n, d = int(mm.group(1)), int(mm.group(2))
    vals = np.array([float(t) for t in text[mm.end():].split()[:n * d]], np.float64)
    return vals.reshape(n, d)
SCHEME = {"HLLC": 0, "ROE": 1, "AUSM": 2}


class _Mesh(C.Structure):
    _fields_ = [("ncv", C.c_int), ("nfa", C.c_int), ("nfa_b", C.c_int), ("nnz", C.c_int),
                ("cvofa", C.c_void_p), ("faocv_i", C.c_void_p), ("faocv_v", C.c_void_p),
                ("nbocv_i", C.c_void_p), ("nbocv_v", C.c_void_p),
                ("x_fa", C.c_void_p), ("fa_normal", C.c_void_p), ("x_cv", C.c_void_p), ("cv_volume", C.c_void_p),
                ("nbocv_lsg", C.c_void_p), ("bfa_lsg", C.c_void_p),
                ("nzone", C.c_int), ("zone_bc", C.c_void_p), ("zone_ifa", C.c_void_p), ("zone_par", C.c_void_p),
                ("zone_scal", C.c_void_p), ("zone_scal2", C.c_void_p), ("profiles", C.c_void_p),
                ("nprof", C.c_int), ("nval_prof", C.c_int)]


class _Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("R_gas", C.c_double), ("mu_ref", C.c_double), ("gravity", C.c_double * 3),
                ("scheme", C.c_int), ("second_order", C.c_int), ("grad_scheme", C.c_int), ("limiter", C.c_int),
                ("symm_bc", C.c_int), ("eps_limiter", C.c_double), ("cfl", C.c_double), ("time_step_min", C.c_int),
                ("under_relax", C.c_double), ("ls_maxiter", C.c_int), ("ls_abs", C.c_double), ("ls_rel", C.c_double),
                ("cfl_after", C.c_int), ("cfl_interval", C.c_int), ("cfl_factor", C.c_double),
                ("cfl_max", C.c_double), ("nsteps", C.c_int), ("resid_energ_th", C.c_double),
                ("visc_mode", C.c_int), ("mu_power_law", C.c_double), ("sl_tref", C.c_double),
                ("sl_sref", C.c_double), ("T_ref", C.c_double), ("pr", C.c_double), ("pr_turb", C.c_double),
                ("visc_limiter", C.c_int), ("sa", C.c_int), ("sa_init", C.c_double), ("relax_scal", C.c_double),
                ("time_step_sum", C.c_int), ("coupled", C.c_int), ("sst", C.c_int), ("sst_init", C.c_double * 2),
                ("sst_limit_pk", C.c_int), ("sst_limiter_shear", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle oracle)")
        L = C.CDLL(path)
        L.or_ctx_create.restype = C.c_void_p
        L.or_ctx_create.argtypes = [C.POINTER(_Mesh), C.POINTER(_Params)]
        L.or_ctx_destroy.argtypes = [C.c_void_p]
        L.or_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_set_scalar.argtypes = [C.c_void_p, C.c_void_p]
        L.or_set_step.argtypes = [C.c_void_p, C.c_int]
        L.or_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_residual.argtypes = [C.c_void_p, C.c_int]
        L.or_step.argtypes = [C.c_void_p]
        L.or_step.restype = C.c_int
        L.or_run.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.or_run.restype = C.c_int
        L.or_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.or_get.restype = C.c_longlong
        L.or_cfl.argtypes = [C.c_void_p]
        L.or_cfl.restype = C.c_double
        L.or_linear_solve.argtypes = [C.c_void_p]
        L.or_wall_values.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.or_linear_solve.restype = C.c_int
        _LIB = L
    return _LIB


def parse_joe_in(text: str) -> dict:
    """Joe.in as ParamMap reads it (Param.h:60-63, Param.cpp:247-298): tokens
    split on ' =\\t', '#' starts a comment, repeated keys append; lookup uses
    the first entry."""
    params: dict = {}
    for line in text.splitlines():
        s = line.split("#", 1)[0]
        toks = [t for t in s.replace("=", " ").replace("\t", " ").split(" ") if t]
        if toks:
            params.setdefault(toks[0], []).append(toks)
    return params


def _after(tokens, key, default=None):
    if key in tokens[1:-1]:
        return tokens[tokens.index(key, 1) + 1]
    if default is None:
        raise KeyError(key)
    return default


def make_params(joe_in: str) -> tuple[_Params, dict]:
    P = parse_joe_in(joe_in)

    def get(k, default=None):
        if k in P:
            return P[k][0][1] if len(P[k][0]) > 1 else ""
        if default is None:
            raise KeyError(k)
        return default

    p = _Params()
    p.gamma = float(get("GAMMA", "1.4"))
    p.R_gas = float(get("P_REF")) / float(get("RHO_REF")) / float(get("T_REF"))
    p.mu_ref = float(get("MU_REF", "0.0"))
    # viscous keywords (UgpWithCvCompFlow.h:155-186)
    mode = get("MU_MODE", "POWERLAW")
    p.visc_mode = {"POWERLAW": 0, "SUTHERLAND": 1}[mode]
    p.mu_power_law = float(get("MU_POWER_LAW", "0.76"))
    p.sl_tref = float(get("SL_TREF", "273.15"))
    p.sl_sref = float(get("SL_SREF", "110.4"))
    p.T_ref = float(get("T_REF"))
    p.pr = float(get("PR", "0.72"))
    p.pr_turb = float(get("PRTURB", "0.9"))
    p.visc_limiter = 1 if get("VISC_LIMITER", "YES") == "YES" else 0
    for i, k in enumerate(("GX", "GY", "GZ")):
        p.gravity[i] = float(get(k, "0.0"))
    p.scheme = SCHEME[get("SPACE_INTEGRATION", "HLLC")]
    p.second_order = 1 if "SPATIAL_SECOND_ORDER" in P else 0
    gr = P.get("GRAD_RECONSTRUCTION", [["GRAD_RECONSTRUCTION", "SCHEME", "GREENGAUSS", "LIMITER", "YES", "EPS",
                                         "0.001"]])[0]
    sch = _after(gr, "SCHEME")
    p.grad_scheme = {"GREENGAUSS": 0, "LEASTSQUARES": 1}[sch]
    p.limiter = 1 if _after(gr, "LIMITER") in ("YES", "BARTH_JESPERSEN_MOD") else 0
    p.eps_limiter = float(_after(gr, "EPS"))
    ti = get("TIME_INTEGRATION", "BACKWARD_EULER")
    p.coupled = 1 if ti == "BACKWARD_EULER_COUPLED" else 0
    if p.coupled and get("SPACE_INTEGRATION", "HLLC") != "HLLC":
        raise ValueError("the coupled restatement covers HLLC only")
    # calcFluxCoupled's own default is ANALYTICAL (JoeWithModels.cpp:4615), but
    # init has already registered ORIGINAL_NUMERIC by reading the key with that
    # default while printing it (UgpWithCvCompFlow.h:479; Param.cpp:139-149)
    p.symm_bc = {"ORIGINAL_NUMERIC": 0, "NEW_NUMERIC": 1, "ANALYTICAL": 2}[get("SYMM_BC", "ORIGINAL_NUMERIC")]
    if "CFL" in P:
        p.cfl, p.time_step_min = float(get("CFL")), 0
    else:
        p.cfl, p.time_step_min = float(get("CFL_MIN")), 1
    p.time_step_sum = {"MAX_BASED": 0, "SUM_BASED": 1}[get("TIME_STEP", "MAX_BASED")]
    p.under_relax = float(get("UNDER_RELAXATION", "0.3"))
    ls = P.get("LINEAR_SOLVER_NS_TRESHOLDS", [["X", "MAX_ITER", "30", "ABS_RESID", "1.0e-8", "REL_RESID", "1.0e-2"]])[0]
    p.ls_maxiter = int(_after(ls, "MAX_ITER"))
    p.ls_abs = float(_after(ls, "ABS_RESID"))
    p.ls_rel = float(_after(ls, "REL_RESID"))
    cr = P.get("CFL_RAMP", [["X", "AFTER_ITER", "100", "INTERVAL_ITER", "10", "FACTOR_CFL", "1.0", "MAX_CFL",
                             "100.0"]])[0]
    p.cfl_after = int(_after(cr, "AFTER_ITER"))
    p.cfl_interval = int(_after(cr, "INTERVAL_ITER"))
    p.cfl_factor = float(_after(cr, "FACTOR_CFL"))
    p.cfl_max = float(_after(cr, "MAX_CFL"))
    p.nsteps = int(get("NSTEPS", "100"))
    p.resid_energ_th = float(get("RESID_ENERG_TH", "0.0"))
    # Spalart-Allmaras (the reference's RansTurbSA mix-in, chosen here by TURB_MODEL)
    p.sa = 1 if get("TURB_MODEL", "NONE") == "SA" else 0
    if p.sa:
        p.sa_init = float(P["INITIAL_CONDITION_TURB"][0][1])
        p.relax_scal = float(get("UNDER_RELAXATION_SCALARS", "0.2"))
    # Menter SST k-omega (RansTurbKOmSST): kine, omega initial values and options
    p.sst = 1 if get("TURB_MODEL", "NONE") == "SST" else 0
    if p.sst:
        p.sst_init[0] = float(P["INITIAL_CONDITION_TURB"][0][1])
        p.sst_init[1] = float(P["INITIAL_CONDITION_TURB"][0][2])
        p.relax_scal = float(get("UNDER_RELAXATION_SCALARS", "0.2"))
        p.sst_limit_pk = int(get("SST_LIMIT_PK", "1"))
        p.sst_limiter_shear = int(get("SST_LIMITER_SHEARSTRESS", "0"))
    return p, P


class Oracle:
    """A CPU solve context over reference-layout mesh arrays."""

    def __init__(self, mesh: dict, joe_in: str):
        L = lib()
        self.params, P = make_params(joe_in)
        counts = mesh["COUNTS"]
        self.nno, self.nfa, self.nfa_b, self.ncv, _ncvg, self.nnz = (int(x) for x in counts[:6])
        self._keep = {}

        def arr(name, dt):
            a = np.ascontiguousarray(mesh[name], dtype=dt)
            self._keep[name] = a
            return a.ctypes.data

        zl = [z for z in jmdump.zones(mesh) if z[1] == 1]
        prof_file = None
        zbc = np.zeros(len(zl), np.int32)
        zifa = np.zeros((len(zl), 2), np.int32)
        zpar = np.zeros((len(zl), 8), np.float64)
        for i, (name, _k, f, l) in enumerate(zl):
            toks = P[name][0]
            kw = toks[1]
            zbc[i] = BC[kw]
            zifa[i] = (f, l)
            if kw == "WALL":
                zpar[i, 0] = float(toks[toks.index("TEMP") + 1]) if "TEMP" in toks else 0.0
            elif kw == "PROFILES":
                prof_file = toks[2]  # the last PROFILES zone's file serves all of them (:249-256)
            else:
                vals = [float(t) for t in toks[2:]]
                zpar[i, :len(vals)] = vals
        # "<zone>.<scalar> = v" (scalarZoneIsDirichlet) or "= HOOK" (scalarZoneIsHook)
        names = ["kine", "omega"] if self.params.sst else ["nuSA"]
        zs = [np.zeros((len(zl), 2), np.float64) for _ in range(2)]
        for s, sname in enumerate(names):
            for i, (name, _k, f, l) in enumerate(zl):
                key = name + "." + sname
                if key in P:
                    v = P[key][0][1]
                    zs[s][i] = ((2.0, 0.0) if v in ("HOOK", "Hook", "hook") else
                                (3.0, 0.0) if v == "PROFILES" else (1.0, float(v)))
        zscal, zscal2 = zs
        self._keep.update(zbc=zbc, zifa=zifa, zpar=zpar, zscal=zscal, zscal2=zscal2)
        self.zone_names = [z[0] for z in zl]
        self.nscal = 1 if self.params.sa else (2 if self.params.sst else 0)
        m = _Mesh()
        m.ncv, m.nfa, m.nfa_b, m.nnz = self.ncv, self.nfa, self.nfa_b, self.nnz
        m.cvofa = arr("CVOFA", np.int32)
        m.faocv_i = arr("FAOCV_I", np.int32)
        m.faocv_v = arr("FAOCV_V", np.int32)
        m.nbocv_i = arr("NBOCV_I", np.int32)
        m.nbocv_v = arr("NBOCV_V", np.int32)
        m.x_fa = arr("X_FA", np.float64)
        m.fa_normal = arr("FA_NORMAL", np.float64)
        m.x_cv = arr("X_CV", np.float64)
        m.cv_volume = arr("CV_VOLUME", np.float64)
        m.nbocv_lsg = arr("NBOCV_LSG", np.float64)
        m.bfa_lsg = arr("BFA_LSG", np.float64)
        m.nzone = len(zl)
        m.zone_bc = zbc.ctypes.data
        m.zone_ifa = zifa.ctypes.data
        m.zone_par = zpar.ctypes.data
        m.zone_scal = zscal.ctypes.data
        m.zone_scal2 = zscal2.ctypes.data
        if prof_file is not None:
            prof = np.ascontiguousarray(read_profiles(prof_file))
            if prof.shape[1] - 5 != self.nscal:
                raise ValueError(f"{prof_file} does not have the required number of values")
            self._keep["prof"] = prof
            m.profiles, m.nprof, m.nval_prof = prof.ctypes.data, prof.shape[0], prof.shape[1]
        self._mesh = m
        self.ctx = L.or_ctx_create(C.byref(m), C.byref(self.params))

    def __del__(self):
        if getattr(self, "ctx", None) and _LIB is not None:
            _LIB.or_ctx_destroy(self.ctx)
            self.ctx = None

    def set_state(self, rho, rhou, rhoE):
        rho = np.ascontiguousarray(rho, np.float64)
        rhou = np.ascontiguousarray(rhou, np.float64)
        rhoE = np.ascontiguousarray(rhoE, np.float64)
        lib().or_set_state(self.ctx, rho.ctypes.data, rhou.ctypes.data, rhoE.ctypes.data)

    def state(self):
        rho = np.empty(self.ncv)
        rhou = np.empty((self.ncv, 3))
        rhoE = np.empty(self.ncv)
        lib().or_get_state(self.ctx, rho.ctypes.data, rhou.ctypes.data, rhoE.ctypes.data)
        return rho, rhou, rhoE

    def set_scalar(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        lib().or_set_scalar(self.ctx, phi.ctypes.data)

    def residual(self, implicit: bool = True):
        lib().or_residual(self.ctx, 1 if implicit else 0)

    def linear_solve(self) -> int:
        return lib().or_linear_solve(self.ctx)

    def step(self) -> int:
        return lib().or_step(self.ctx)

    def set_step(self, step: int):
        """The step count of a solve resumed from a restart file."""
        lib().or_set_step(self.ctx, int(step))

    def run(self, max_steps: int):
        hist = np.zeros((max_steps, 5 + self.nscal))
        n = lib().or_run(self.ctx, max_steps, hist.ctypes.data)
        return n, hist[:n]

    def wall_values(self, zone: str) -> np.ndarray:
        """writeWallValues of a boundary zone: [n][16] (see or_wall_values)."""
        k = self.zone_names.index(zone)
        n = lib().or_wall_values(self.ctx, k, None)
        out = np.empty((n, 16))
        lib().or_wall_values(self.ctx, k, out.ctypes.data)
        return out

    def get(self, name: str) -> np.ndarray:
        n = lib().or_get(self.ctx, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n)
        lib().or_get(self.ctx, name.encode(), out.ctypes.data)
        return out

    @property
    def cfl(self) -> float:
        return lib().or_cfl(self.ctx)


def chemtable_lookup(x1, x2, x3, data, zm, zv, pc, vars):
    """The oracle's Lookup (ChemtableCartesianLinear_single.cpp:374-429) of
    variables vars at the points: [len(vars)][n] float64."""
    L = lib()
    f = L.or_chemtable_lookup
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_int64] + [C.c_void_p] * 3 + \
        [C.c_int, C.c_void_p, C.c_void_p]
    f.restype = None
    x1, x2, x3 = (np.ascontiguousarray(x, np.float64) for x in (x1, x2, x3))
    data = np.ascontiguousarray(data, np.float32)
    zm, zv, pc = (np.ascontiguousarray(x, np.float64) for x in (zm, zv, pc))
    v = np.ascontiguousarray(vars, np.int32)
    n1, n2, n3, nvar = data.shape
    out = np.empty((len(v), zm.size))
    f(n1, n2, n3, nvar, x1.ctypes.data, x2.ctypes.data, x3.ctypes.data, data.ctypes.data, zm.size, zm.ctypes.data,
      zv.ctypes.data, pc.ctypes.data, len(v), v.ctypes.data, out.ctypes.data)
    return out

