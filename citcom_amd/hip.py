"""Python mirror of the reference's hot-path interface over libcitcom_hip.so.

``Citcom`` holds the device-resident state of one problem (the part of
``struct All_variables`` the timestep touches, global_defs.h:931-1102) and
exposes the reference's entry points by their names:

* ``solve_constrained_flow_iterative``  (Stokes_flow_Incomp.c:50-90)
* ``PG_timestep``                       (Advection_diffusion.c:251-345)
* ``general_stokes_solver``             (Drive_solvers.c:45-165)
* ``get_system_viscosity``              (Viscosity_structures.c:369-443)
* ``construct_stiffness_B_matrix``      (Construct_arrays.c:834-889)
* ``thermal_buoyancy``, ``assemble_forces``, ``heat_flux``

Everything runs through the C ABI of include/citcom_hip.h.  There is no CPU
fallback: without the built library or a GPU the constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import deck as _deck

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcitcom_hip.so")


class ChipProblem(C.Structure):
    _fields_ = [
        ("levels", C.c_int), ("mgunitx", C.c_int), ("mgunity", C.c_int), ("mgunitz", C.c_int),
        ("xc", C.POINTER(C.c_float)), ("yc", C.POINTER(C.c_float)), ("zc", C.POINTER(C.c_float)),
        ("topvbc", C.c_int), ("botvbc", C.c_int), ("toptbc", C.c_int), ("bottbc", C.c_int),
        ("toptbcval", C.c_float), ("bottbcval", C.c_float),
        ("accuracy", C.c_double), ("p_iterations", C.c_int),
        ("v_steps_low", C.c_int), ("v_steps_high", C.c_int), ("down_heavy", C.c_int),
        ("up_heavy", C.c_int), ("mg_cycle", C.c_int),
        ("aug_lagr", C.c_int), ("aug_number", C.c_double), ("precond", C.c_int),
        ("rheol", C.c_int), ("tdepv", C.c_int), ("num_mat", C.c_int), ("visc_update", C.c_int),
        ("visc0", C.c_float * 4), ("viscE", C.c_float * 4), ("viscT", C.c_float * 4), ("z_layer", C.c_float * 4),
        ("vmin_on", C.c_int), ("vmax_on", C.c_int), ("visc_min", C.c_float), ("visc_max", C.c_float),
        ("rayleigh", C.c_float), ("finetunedt", C.c_float), ("temp_iterations", C.c_int),
        ("gamma", C.c_float), ("Q0", C.c_double), ("fixed_timestep", C.c_float), ("mixed", C.c_int),
    ]


class StokesStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("mg_cycles", C.c_int), ("inner_solves", C.c_int),
                ("vdotv", C.c_double), ("pdotp", C.c_double), ("incompressibility", C.c_double),
                ("dvelocity", C.c_double), ("dpressure", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PGStats(C.Structure):
    _fields_ = [("timestep", C.c_float), ("T_interior", C.c_float), ("sub_iterations", C.c_int),
                ("elapsed_time", C.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ChipError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChipError(f"{LIB_PATH} is missing: run citcom_amd/build.py (or __graft_entry__.build())")
        # torch bundles a HIP runtime with the same soname (libamdhip64.so.7)
        # as /opt/rocm's.  Whichever loads first serves the whole process; when
        # it is /opt/rocm's, torch later finds no device.  Loading torch first
        # keeps torch tensors (the learned network) and this library usable
        # together in one process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.chip_last_error.restype = C.c_char_p
        L.chip_create.argtypes = [C.POINTER(ChipProblem), C.POINTER(P)]
        L.chip_destroy.argtypes = [P]
        L.chip_set_stream.argtypes = [P, P]
        L.chip_get_stream.restype = P
        L.chip_get_stream.argtypes = [P]
        L.chip_upload.argtypes = [P, C.c_char_p, C.c_int, P, C.c_long]
        L.chip_download.argtypes = [P, C.c_char_p, C.c_int, P, C.c_long]
        L.chip_get_system_viscosity.argtypes = [P]
        L.chip_construct_stiffness_B_matrix.argtypes = [P]
        L.chip_apply.argtypes = [P, C.c_char_p, C.c_int, P, P, P, P, C.c_int]
        L.chip_solve_constrained_flow_iterative.argtypes = [P, C.POINTER(StokesStats)]
        L.chip_solve_del2_u.argtypes = [P, P, P, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.chip_PG_timestep.argtypes = [P, C.POINTER(PGStats)]
        L.chip_thermal_buoyancy.argtypes = [P]
        L.chip_assemble_forces.argtypes = [P]
        L.chip_heat_flux.argtypes = [P, C.POINTER(C.c_float)]
        L.chip_vrms.argtypes = [P, C.POINTER(C.c_double)]
        L.chip_timestep.argtypes = [P, C.POINTER(PGStats), C.POINTER(StokesStats), C.POINTER(C.c_float)]
        L.chip_general_stokes_solver.argtypes = [P, C.POINTER(StokesStats)]
        L.chip_profile_enable.argtypes = [P, C.c_char_p, C.c_int]
        L.chip_profile_read.argtypes = [P, C.POINTER(C.c_long), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.chip_set_option.argtypes = [P, C.c_char_p, C.c_int]
        L.chip_set_coarse_net.argtypes = [P, C.c_int, P, P, P]
        L.chip_coarse_net_state.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.chip_averages.argtypes = [P, P, P, P]
        L.chip_set_warm_net.argtypes = [P, P, C.c_long, P, C.c_long]
        L.chip_warm_record.argtypes = [P, P, P]
        L.chip_warm_scales.argtypes = [P, P]
        L.chip_device_bytes.restype = C.c_long
        L.chip_device_bytes.argtypes = [P]
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ChipError(f"{what}: {lib().chip_last_error().decode()}")


def problem_struct(p: _deck.Problem, mixed: int = 0):
    s = ChipProblem()
    keep = [np.ascontiguousarray(c, dtype=np.float32) for c in p.coords]
    s.levels = p.levels
    s.mgunitx, s.mgunity, s.mgunitz = p.mgunit
    s.xc, s.yc, s.zc = (k.ctypes.data_as(C.POINTER(C.c_float)) for k in keep)
    s.topvbc, s.botvbc, s.toptbc, s.bottbc = p.topvbc, p.botvbc, p.toptbc, p.bottbc
    s.toptbcval, s.bottbcval = p.toptbcval, p.bottbcval
    s.accuracy, s.p_iterations = p.accuracy, p.p_iterations
    s.v_steps_low, s.v_steps_high, s.down_heavy, s.up_heavy, s.mg_cycle = (
        p.v_steps_low, p.v_steps_high, p.down_heavy, p.up_heavy, p.mg_cycle)
    s.aug_lagr, s.aug_number, s.precond = p.aug_lagr, p.aug_number, p.precond
    s.rheol, s.tdepv, s.num_mat, s.visc_update = p.rheol, p.tdepv, p.num_mat, p.visc_update
    for i in range(4):
        s.visc0[i], s.viscE[i], s.viscT[i], s.z_layer[i] = p.visc0[i], p.viscE[i], p.viscT[i], p.z_layer[i]
    s.vmin_on, s.vmax_on, s.visc_min, s.visc_max = p.vmin_on, p.vmax_on, p.visc_min, p.visc_max
    s.rayleigh, s.finetunedt, s.temp_iterations = p.rayleigh, p.finetunedt, p.temp_iterations
    s.gamma, s.Q0, s.fixed_timestep = p.gamma, float(np.float32(p.Q0)), p.fixed_timestep
    s.mixed = int(mixed)
    return s, keep


@dataclass
class Shapes:
    nno: int
    nel: int


class Citcom:
    """Device-resident hot-path state of one CitcomCU problem."""

    def __init__(self, problem: _deck.Problem, mixed: int = 0):
        """mixed: 0 the reference's arithmetic, 1 FP32 multigrid inside FP64
        refinement, 2 the same with the fine levels' stiffness in FP16"""
        self.p = problem
        self.levmax = problem.levels - 1
        s, keep = problem_struct(problem, mixed)
        h = C.c_void_p()
        _check(lib().chip_create(C.byref(s), C.byref(h)), "chip_create")
        self.h = h
        nox, noy, noz = problem.dims()
        self.shape = Shapes(nox * noy * noz, (nox - 1) * (noy - 1) * (noz - 1))

    @classmethod
    def from_deck(cls, path_or_keys, **kw):
        keys = _deck.read_deck(path_or_keys) if isinstance(path_or_keys, str) else {
            k: str(v) for k, v in path_or_keys.items()}
        return cls(_deck.problem_from_deck(keys), **kw)

    def close(self):
        if getattr(self, "h", None):
            lib().chip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- sizes
    def nno(self, lev=None):
        nox, noy, noz = self.p.dims(self.levmax if lev is None else lev)
        return nox * noy * noz

    def nel(self, lev=None):
        nox, noy, noz = self.p.dims(self.levmax if lev is None else lev)
        return (nox - 1) * (noy - 1) * (noz - 1)

    # ---- state
    _DT = {"T": np.float32, "Tdot": np.float32, "U": np.float64, "P": np.float64, "F": np.float64,
           "V": np.float32, "buoyancy": np.float32, "EVI": np.float32, "K": np.float32, "BI": np.float64,
           "BPI": np.float64, "EG": np.float32, "MASS": np.float32, "NODE": np.uint32, "GNX": np.float32,
           "GDA": np.float32, "ECO": np.float32, "TWW": np.float32}

    def _count(self, which, lev):
        lev = self.levmax if lev < 0 else lev
        n, e = self.nno(lev), self.nel(lev)
        return {"T": n, "Tdot": n, "U": 3 * n, "P": e, "F": 3 * n, "V": 3 * n, "buoyancy": n, "EVI": 8 * e,
                "K": 243 * n, "BI": 3 * n, "BPI": e, "EG": 24 * e, "MASS": n, "NODE": n, "GNX": 216 * e,
                "GDA": 11 * e, "ECO": 13 * e, "TWW": 9 * e}[which]

    def upload(self, which: str, arr, lev: int = -1) -> None:
        a = np.ascontiguousarray(arr, dtype=self._DT[which])
        n = self._count(which, lev)
        if a.size != n:
            raise ValueError(f"{which}: expected {n} values, got {a.size}")
        _check(lib().chip_upload(self.h, which.encode(), lev, a.ctypes.data, n), f"upload {which}")

    def download(self, which: str, lev: int = -1) -> np.ndarray:
        n = self._count(which, lev)
        out = np.empty(n, dtype=self._DT[which])
        _check(lib().chip_download(self.h, which.encode(), lev, out.ctypes.data, n), f"download {which}")
        return out

    # ---- reference entry points
    def get_system_viscosity(self):
        _check(lib().chip_get_system_viscosity(self.h), "get_system_viscosity")

    def construct_stiffness_B_matrix(self):
        _check(lib().chip_construct_stiffness_B_matrix(self.h), "construct_stiffness_B_matrix")

    def thermal_buoyancy(self):
        _check(lib().chip_thermal_buoyancy(self.h), "thermal_buoyancy")

    def assemble_forces(self):
        _check(lib().chip_assemble_forces(self.h), "assemble_forces")

    def solve_constrained_flow_iterative(self) -> dict:
        st = StokesStats()
        _check(lib().chip_solve_constrained_flow_iterative(self.h, C.byref(st)), "solve_constrained_flow_iterative")
        return st.as_dict()

    def general_stokes_solver(self) -> dict:
        st = StokesStats()
        _check(lib().chip_general_stokes_solver(self.h, C.byref(st)), "general_stokes_solver")
        return st.as_dict()

    def PG_timestep(self) -> dict:
        st = PGStats()
        _check(lib().chip_PG_timestep(self.h, C.byref(st)), "PG_timestep")
        return st.as_dict()

    def vrms(self) -> float:
        """volume-RMS velocity of the current V (no reference routine)"""
        v = C.c_double()
        _check(lib().chip_vrms(self.h, C.byref(v)), "vrms")
        return v.value

    def averages(self) -> dict:
        """layer profiles of averages (Process_velocity.c:179-224): Have.T,
        Have.vrms, Have.Vi, each [noz] (chip_averages)"""
        noz = self.p.dims()[2]
        out = {k: np.zeros(noz, np.float32) for k in ("T", "vrms", "Vi")}
        _check(lib().chip_averages(self.h, out["T"].ctypes.data, out["vrms"].ctypes.data, out["Vi"].ctypes.data),
               "averages")
        return out

    def heat_flux(self) -> tuple[float, float]:
        nu = (C.c_float * 2)()
        _check(lib().chip_heat_flux(self.h, nu), "heat_flux")
        return float(nu[0]), float(nu[1])

    def timestep(self, with_nu: bool = True) -> dict:
        pg, st = PGStats(), StokesStats()
        nu = (C.c_float * 2)()
        _check(lib().chip_timestep(self.h, C.byref(pg), C.byref(st), nu if with_nu else None), "timestep")
        return {"pg": pg.as_dict(), "stokes": st.as_dict(), "Nu_top": float(nu[0]), "Nu_bot": float(nu[1])}

    def solve_del2_u(self, F: np.ndarray, acc: float):
        F = np.ascontiguousarray(F, dtype=np.float64)
        d0 = np.zeros_like(F)
        valid, cyc = C.c_int(), C.c_int()
        _check(lib().chip_solve_del2_u(self.h, F.ctypes.data, d0.ctypes.data, acc, C.byref(valid), C.byref(cyc)),
               "solve_del2_u")
        return d0, valid.value, cyc.value

    def apply(self, op: str, lev: int, x=None, x2=None, sweeps: int = 0):
        """Single operator on host arrays (see chip_apply in include/citcom_hip.h)."""
        n3, ne = 3 * self.nno(lev), self.nel(lev)
        out_n = {"matvec": n3, "gs0": n3, "gs1": n3, "box0": n3, "box1": n3, "coarse": n3, "direct": n3,
                 "affine0": n3, "affine1": n3,
                 "div": ne, "grad": n3,
                 "strip": n3, "cnet": n3,
                 "project": 3 * self.nno(lev - 1) if lev > 0 else 0, "interp": n3}[op]
        # what chip_apply reads from x (in) and x2 (in2): it copies these many
        # doubles from the host pointers, so short or missing arrays are refused
        n_in = {"matvec": n3, "gs1": n3, "box1": n3, "affine1": n3, "div": n3, "grad": ne, "strip": n3,
                "project": n3, "interp": 3 * self.nno(lev - 1) if lev > 0 else 0}.get(op, 0)
        n_in2 = {"gs0": n3, "gs1": n3, "box0": n3, "box1": n3, "coarse": n3, "direct": n3, "affine0": n3, "affine1": n3,
                 "cnet": n3}.get(op, 0)
        for name, arr, need in (("x", x, n_in), ("x2", x2, n_in2)):
            if need and (arr is None or np.size(arr) != need):
                raise ChipError(f"apply {op}: {name} needs {need} values, got "
                                f"{'none' if arr is None else np.size(arr)}")
        out = np.zeros(out_n)
        out2 = np.zeros(n3)
        a = np.ascontiguousarray(x, dtype=np.float64) if x is not None else np.zeros(1)
        b = np.ascontiguousarray(x2, dtype=np.float64) if x2 is not None else np.zeros(1)
        _check(lib().chip_apply(self.h, op.encode(), lev, a.ctypes.data, b.ctypes.data, out.ctypes.data,
                                out2.ctypes.data, sweeps), f"apply {op}")
        return (out, out2) if op in ("gs0", "gs1", "box0", "box1") else out

    # ---- measurement
    def profile(self, kernel: str | None):
        _check(lib().chip_profile_enable(self.h, (kernel or "").encode(), int(kernel is not None)), "profile")

    def profile_read(self):
        n, ms, b = C.c_long(), C.c_double(), C.c_double()
        _check(lib().chip_profile_read(self.h, C.byref(n), C.byref(ms), C.byref(b)), "profile_read")
        return n.value, ms.value, b.value

    def stream_handle(self) -> int:
        return lib().chip_get_stream(self.h) or 0

    def set_stream(self, handle: int) -> None:
        _check(lib().chip_set_stream(self.h, C.c_void_p(handle)), "set_stream")

    def set_option(self, name: str, value: int) -> None:
        _check(lib().chip_set_option(self.h, name.encode(), int(value)), f"set_option {name}")

    def coarse_net_state(self) -> tuple[bool, int]:
        """(learned coarse correction still in use, inner solves that fell back
        to the plain multigrid)"""
        on, fb = C.c_int(), C.c_int()
        _check(lib().chip_coarse_net_state(self.h, C.byref(on), C.byref(fb)), "coarse_net_state")
        return bool(on.value), fb.value

    def set_coarse_net(self, weights: np.ndarray, bias: np.ndarray, out_scale: np.ndarray) -> None:
        """learned coarse-level correction (chip_set_coarse_net): weights as BF16
        bits [layers][16][27][16], bias [layers][16], out_scale [3]"""
        w = np.ascontiguousarray(weights, dtype=np.uint16)
        b = np.ascontiguousarray(bias, dtype=np.float32)
        ys = np.ascontiguousarray(out_scale, dtype=np.float32)
        if w.ndim != 4 or w.shape[1:] != (16, 27, 16) or b.shape != (w.shape[0], 16) or ys.shape != (3,):
            raise ChipError("set_coarse_net: weights [layers][16][27][16], bias [layers][16], out_scale [3]")
        _check(lib().chip_set_coarse_net(self.h, w.shape[0], w.ctypes.data, b.ctypes.data, ys.ctypes.data),
               "set_coarse_net")

    def set_warm_net(self, weights_bf16: np.ndarray, bias: np.ndarray) -> None:
        """the learned warm start's U-Net (chip_set_warm_net): BF16 bits of the
        weights in layer order, FP32 biases"""
        w = np.ascontiguousarray(weights_bf16, dtype=np.uint16)
        b = np.ascontiguousarray(bias, dtype=np.float32)
        _check(lib().chip_set_warm_net(self.h, w.ctypes.data, w.size, b.ctypes.data, b.size), "set_warm_net")

    def warm_record(self, features_ptr: int | None, targets_ptr: int | None) -> None:
        """record training samples into device buffers (chip_warm_record)"""
        _check(lib().chip_warm_record(self.h, C.c_void_p(features_ptr or None), C.c_void_p(targets_ptr or None)),
               "warm_record")

    def warm_scales(self) -> np.ndarray:
        out = np.zeros(8)
        _check(lib().chip_warm_scales(self.h, out.ctypes.data), "warm_scales")
        return out

    def device_bytes(self) -> int:
        return lib().chip_device_bytes(self.h)

    # ---- convenience: initial state as read_instructions leaves it
    def initialize(self, T: np.ndarray) -> None:
        """Upload T, zero U, P, Tdot and form the buoyancy (convection_initial_temperature,
        Convection.c:308-461, ends with thermal_buoyancy)."""
        self.upload("T", T)
        self.upload("Tdot", np.zeros(self.nno(), np.float32))
        self.upload("U", np.zeros(3 * self.nno()))
        self.upload("P", np.zeros(self.nel()))
        self.thermal_buoyancy()
