"""ctypes front end of liboracle.so, the CPU restatement of the CitcomCU hot
path.  TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
CPU-baseline leg of bench.py, as the checker of the HIP path."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def build(quiet: bool = True) -> None:
    subprocess.run(["make", "-C", HERE, "cpu"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


class OrcParams(C.Structure):
    _fields_ = [
        ("levels", C.c_int), ("mgunitx", C.c_int), ("mgunity", C.c_int), ("mgunitz", C.c_int),
        ("topvbc", C.c_int), ("botvbc", C.c_int), ("toptbc", C.c_int), ("bottbc", C.c_int),
        ("toptbcval", C.c_float), ("bottbcval", C.c_float),
        ("accuracy", C.c_double), ("p_iterations", C.c_int),
        ("v_steps_low", C.c_int), ("v_steps_high", C.c_int), ("down_heavy", C.c_int),
        ("up_heavy", C.c_int), ("mg_cycle", C.c_int),
        ("aug_lagr", C.c_int), ("aug_number", C.c_double), ("precond", C.c_int),
        ("rheol", C.c_int), ("tdepv", C.c_int), ("num_mat", C.c_int),
        ("visc0", C.c_float * 4), ("viscE", C.c_float * 4), ("viscT", C.c_float * 4),
        ("z_layer", C.c_float * 4),
        ("vmin_on", C.c_int), ("vmax_on", C.c_int), ("visc_min", C.c_float), ("visc_max", C.c_float),
        ("rayleigh", C.c_float), ("finetunedt", C.c_float), ("temp_iterations", C.c_int),
        ("gamma", C.c_float), ("Q0", C.c_double), ("fixed_timestep", C.c_float),
        ("smoother", C.c_int),
    ]


def params_from_problem(p, smoother: int = 0) -> OrcParams:
    q = OrcParams()
    q.levels = p.levels
    q.mgunitx, q.mgunity, q.mgunitz = p.mgunit
    q.topvbc, q.botvbc, q.toptbc, q.bottbc = p.topvbc, p.botvbc, p.toptbc, p.bottbc
    q.toptbcval, q.bottbcval = p.toptbcval, p.bottbcval
    q.accuracy = p.accuracy
    q.p_iterations = p.p_iterations
    q.v_steps_low, q.v_steps_high = p.v_steps_low, p.v_steps_high
    q.down_heavy, q.up_heavy, q.mg_cycle = p.down_heavy, p.up_heavy, p.mg_cycle
    q.aug_lagr, q.aug_number, q.precond = p.aug_lagr, p.aug_number, p.precond
    q.rheol, q.tdepv, q.num_mat = p.rheol, p.tdepv, p.num_mat
    for i in range(4):
        q.visc0[i], q.viscE[i], q.viscT[i], q.z_layer[i] = p.visc0[i], p.viscE[i], p.viscT[i], p.z_layer[i]
    q.vmin_on, q.vmax_on, q.visc_min, q.visc_max = p.vmin_on, p.vmax_on, p.visc_min, p.visc_max
    q.rayleigh, q.finetunedt, q.temp_iterations = p.rayleigh, p.finetunedt, p.temp_iterations
    q.gamma, q.Q0, q.fixed_timestep = p.gamma, float(np.float32(p.Q0)), p.fixed_timestep
    q.smoother = smoother
    return q


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.POINTER(OrcParams), fp, fp, fp]
        L.orc_free.argtypes = [P]
        L.orc_get.restype = P
        L.orc_get.argtypes = [P, C.c_char_p, C.c_int, C.POINTER(C.c_long)]
        L.orc_set_viscosity.argtypes = [P, fp]
        L.orc_build_operators.argtypes = [P]
        L.orc_matvec.argtypes = [P, C.c_int, dp, dp, C.c_int]
        L.orc_gauss_seidel.argtypes = [P, C.c_int, dp, dp, dp, C.c_int, C.c_int]
        L.orc_coarse_solve.argtypes = [P, dp, dp]
        L.orc_affine_solve.argtypes = [P, P, dp, dp]
        L.orc_project.argtypes = [P, C.c_int, dp, dp]
        L.orc_interp.argtypes = [P, C.c_int, dp, dp]
        L.orc_div.argtypes = [P, C.c_int, dp, dp]
        L.orc_grad.argtypes = [P, C.c_int, dp, dp]
        L.orc_strip.argtypes = [P, C.c_int, dp]
        L.orc_multi_grid.restype = C.c_double
        L.orc_multi_grid.argtypes = [P, dp, dp, P, C.c_double, C.c_int]
        L.orc_solve_del2_u.restype = C.c_int
        L.orc_solve_del2_u.argtypes = [P, dp, dp, C.c_double, C.c_int, C.POINTER(C.c_int)]
        L.orc_solve_stokes.restype = C.c_int
        L.orc_solve_stokes.argtypes = [P, dp, dp, dp, dp]
        L.orc_thermal_buoyancy.argtypes = [P, fp, fp]
        L.orc_assemble_forces.argtypes = [P, fp, dp]
        L.orc_visc_from_T.argtypes = [P, fp, fp]
        L.orc_pg_timestep.restype = C.c_float
        L.orc_pg_timestep.argtypes = [P, fp, fp, fp, fp]
        L.orc_heat_flux.argtypes = [P, fp, fp, fp]
        L.orc_vrms.argtypes = [P, fp]
        L.orc_vrms.restype = C.c_double
        _lib = L
    return _lib


_DT = {"X": np.float32, "Y": np.float32, "Z": np.float32, "NODE": np.uint32, "GNX": np.float32,
       "GDA": np.float32, "ECO": np.float32, "TWW": np.float32, "MASS": np.float32,
       "EVI": np.float32, "EG": np.float32, "K": np.float32, "BI": np.float64,
       "BPI": np.float64, "TB": np.float32, "mat": np.int32}


class Oracle:
    """One problem instance of the restated hot path (all MG levels)."""

    def __init__(self, problem, smoother: int = 0):
        self.p = problem
        self.params = params_from_problem(problem, smoother)
        xc, yc, zc = (np.ascontiguousarray(c, dtype=np.float32) for c in problem.coords)
        self.h = lib().orc_create(C.byref(self.params), xc, yc, zc)
        self.levmax = problem.levels - 1

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def get(self, name: str, lev: int | None = None) -> np.ndarray:
        lev = self.levmax if lev is None else lev
        n = C.c_long()
        ptr = lib().orc_get(self.h, name.encode(), lev, C.byref(n))
        dt = np.dtype(_DT[name])
        buf = (C.c_char * (n.value * dt.itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dt).copy()

    def nno(self, lev=None):
        nox, noy, noz = self.p.dims(self.levmax if lev is None else lev)
        return nox * noy * noz

    def nel(self, lev=None):
        nox, noy, noz = self.p.dims(self.levmax if lev is None else lev)
        return (nox - 1) * (noy - 1) * (noz - 1)

    # ---- operators
    def set_viscosity(self, evi: np.ndarray) -> None:
        lib().orc_set_viscosity(self.h, np.ascontiguousarray(evi, dtype=np.float32))

    def build_operators(self) -> None:
        lib().orc_build_operators(self.h)

    def matvec(self, lev, u, strip=True):
        out = np.zeros_like(u)
        lib().orc_matvec(self.h, lev, np.ascontiguousarray(u), out, int(strip))
        return out

    def gauss_seidel(self, lev, d0, F, cycles, guess):
        d = np.ascontiguousarray(d0, dtype=np.float64).copy()
        Ad = np.zeros_like(d)
        lib().orc_gauss_seidel(self.h, lev, d, np.ascontiguousarray(F), Ad, cycles, int(guess))
        return d, Ad

    def coarse_solve(self, F):
        """smoother 1: the coarsest-level smoother of multi_grid as a matrix"""
        d = np.zeros(3 * self.nno(0))
        lib().orc_coarse_solve(self.h, np.ascontiguousarray(F, dtype=np.float64), d)
        return d

    def affine_solve(self, din, F):
        """smoother 1: the level above the coarsest, Ma din + Mb F"""
        out = np.zeros(3 * self.nno(1))
        a = None if din is None else np.ascontiguousarray(din, dtype=np.float64)
        lib().orc_affine_solve(self.h, None if a is None else a.ctypes.data_as(C.c_void_p),
                               np.ascontiguousarray(F, dtype=np.float64), out)
        return out

    def project(self, lev, fine):
        out = np.zeros(3 * self.nno(lev - 1))
        lib().orc_project(self.h, lev, np.ascontiguousarray(fine), out)
        return out

    def interp(self, lev, coarse):
        out = np.zeros(3 * self.nno(lev))
        lib().orc_interp(self.h, lev, np.ascontiguousarray(coarse), out)
        return out

    def div(self, lev, U):
        out = np.zeros(self.nel(lev))
        lib().orc_div(self.h, lev, np.ascontiguousarray(U), out)
        return out

    def grad(self, lev, P):
        out = np.zeros(3 * self.nno(lev))
        lib().orc_grad(self.h, lev, np.ascontiguousarray(P), out)
        return out

    def strip(self, lev, v):
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        lib().orc_strip(self.h, lev, v)
        return v

    # ---- solvers
    def solve_del2_u(self, F, acc):
        F = np.ascontiguousarray(F, dtype=np.float64).copy()
        d0 = np.zeros_like(F)
        cnt = C.c_int()
        valid = lib().orc_solve_del2_u(self.h, d0, F, acc, self.levmax, C.byref(cnt))
        return d0, valid, cnt.value

    def solve_stokes(self, U, P, F):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        P = np.ascontiguousarray(P, dtype=np.float64).copy()
        stats = np.zeros(8)
        lib().orc_solve_stokes(self.h, U, P, np.ascontiguousarray(F, dtype=np.float64), stats)
        return U, P, stats

    # ---- energy
    def thermal_buoyancy(self, T):
        out = np.zeros(self.nno(), dtype=np.float32)
        lib().orc_thermal_buoyancy(self.h, np.ascontiguousarray(T, dtype=np.float32), out)
        return out

    def assemble_forces(self, buoy):
        out = np.zeros(3 * self.nno())
        lib().orc_assemble_forces(self.h, np.ascontiguousarray(buoy, dtype=np.float32), out)
        return out

    def visc_from_T(self, T):
        out = np.zeros(8 * self.nel(), dtype=np.float32)
        lib().orc_visc_from_T(self.h, np.ascontiguousarray(T, dtype=np.float32), out)
        return out

    def pg_timestep(self, T, Tdot, V):
        T = np.ascontiguousarray(T, dtype=np.float32).copy()
        Tdot = np.ascontiguousarray(Tdot, dtype=np.float32).copy()
        stats = np.zeros(4, dtype=np.float32)
        dt = lib().orc_pg_timestep(self.h, T, Tdot, np.ascontiguousarray(V, dtype=np.float32), stats)
        return T, Tdot, float(dt), stats

    def vrms(self, V):
        return float(lib().orc_vrms(self.h, np.ascontiguousarray(V, dtype=np.float32)))

    def heat_flux(self, T, V):
        nu = np.zeros(2, dtype=np.float32)
        lib().orc_heat_flux(self.h, np.ascontiguousarray(T, dtype=np.float32),
                            np.ascontiguousarray(V, dtype=np.float32), nu)
        return float(nu[0]), float(nu[1])


def v_from_vector(U: np.ndarray, nno: int) -> np.ndarray:
    """V[d][node] = U[3 node + d] as float (Stokes_flow_Incomp.c:530-552;
    every prescribed velocity of these workloads is zero)."""
    return U.reshape(nno, 3).T.astype(np.float32).copy()
