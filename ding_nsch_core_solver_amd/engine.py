"""ctypes binding of the C ABI in include/pp_solver.h (libppsolve.so, built in-tree).

The library must be built (``python -m ding_nsch_core_solver_amd.build``); there is no fallback
path: every call goes to the HIP engine and fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

import numpy as np

from .build import LIB_PATH

PP_ICPCG, PP_ILDLT_PCG, PP_SPLIT_ICPCG, PP_SPLIT_ILDLT_PCG = 1, 2, 3, 4
PP_FP64, PP_MIXED = 0, 1
# pp_status (include/pp_solver.h)
PP_OK, PP_ERR_FACTOR, PP_ERR_PAP, PP_ERR_RZ0, PP_ERR_RZ, PP_ERR_NODIAG, PP_ERR_SWEEP = 0, 1, 2, 3, 4, 5, 6
PP_ERR_BICG_RHO, PP_ERR_BICG_RHAT_V, PP_ERR_BICG_TT, PP_ERR_BICG_OMEGA = 7, 8, 9, 10
PP_ERR_ARG, PP_ERR_HIP = 100, 101

# New scheme strings (the reference's names prefixed with hip_) and their codes. The split
# aliases mirror the reference dispatch (pressure.cpp:1048-1058).
SCHEME_CODES = {
    "hip_icpcg": PP_ICPCG,
    "hip_ildlt_pcg": PP_ILDLT_PCG,
    "hip_split_icpcg": PP_SPLIT_ICPCG,
    "hip_liu_split_icpcg": PP_SPLIT_ICPCG,
    "hip_paper_split_icpcg": PP_SPLIT_ICPCG,
    "hip_split_ildlt_pcg": PP_SPLIT_ILDLT_PCG,
    "hip_liu_split_ildlt_pcg": PP_SPLIT_ILDLT_PCG,
    "hip_paper_split_ildlt_pcg": PP_SPLIT_ILDLT_PCG,
}


PHASES = ("setup", "spmv", "update", "rnorm", "forward", "backward", "zmean", "rz", "pupdate", "output")


class PPProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dt", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int),
        ("bc_type", ctypes.c_int * 4), ("bc_value", ctypes.c_double * 4),
        ("scheme", ctypes.c_int), ("precision", ctypes.c_int),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
        ("density_ratio", ctypes.c_double), ("step_index", ctypes.c_int),
        ("inner_max_iterations", ctypes.c_int), ("inner_tolerance", ctypes.c_double),
        ("ring_depth", ctypes.c_int),
    ]


class PPVelocityBC(ctypes.Structure):
    """pp_velocity_bc: velocity boundary conditions of the correction step."""
    _fields_ = [
        ("u_type", ctypes.c_int * 4), ("u_value", ctypes.c_double * 4),
        ("v_type", ctypes.c_int * 4), ("v_value", ctypes.c_double * 4),
        ("bottom_wall_velocity_x", ctypes.c_double), ("top_wall_velocity_x", ctypes.c_double),
    ]


def make_velocity_bc(vbc) -> PPVelocityBC:
    out = PPVelocityBC()
    for k in range(4):
        out.u_type[k], out.u_value[k] = int(vbc.u_type[k]), float(vbc.u_value[k])
        out.v_type[k], out.v_value[k] = int(vbc.v_type[k]), float(vbc.v_value[k])
    out.bottom_wall_velocity_x = float(vbc.bottom_wall_velocity_x)
    out.top_wall_velocity_x = float(vbc.top_wall_velocity_x)
    return out


class PPMomentumProblem(ctypes.Structure):
    """pp_momentum_problem: grid, time step, Reynolds number, body force, periodicity, step index
    and the BiCGSTAB controls of one momentum predictor step."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dt", ctypes.c_double), ("re", ctypes.c_double),
        ("body_force_x", ctypes.c_double), ("body_force_y", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int), ("step_index", ctypes.c_int),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
    ]


def make_momentum_problem(mc) -> PPMomentumProblem:
    mp = PPMomentumProblem()
    mp.nx, mp.ny = mc.nx, mc.ny
    mp.dx, mp.dy, mp.dt, mp.re = mc.lx / mc.nx, mc.ly / mc.ny, mc.dt, mc.re
    mp.body_force_x, mp.body_force_y = mc.body_force_x, mc.body_force_y
    mp.periodic_x, mp.periodic_y = int(mc.periodic_x), int(mc.periodic_y)
    mp.step_index, mp.max_iterations, mp.tolerance = int(mc.step_index), int(mc.max_iterations), float(mc.tolerance)
    return mp


class PPChProblem(ctypes.Structure):
    """pp_ch_problem: grid, alpha0 and the stabilization coefficients, ch_inner_iterations and
    ch_tolerance of one Cahn-Hilliard linear solve."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int),
        ("alpha0", ctypes.c_double), ("beta1", ctypes.c_double), ("beta2", ctypes.c_double),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
    ]


class PPChReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int), ("preconditioner", ctypes.c_int),
        ("iterate_residual", ctypes.c_double), ("absolute_residual", ctypes.c_double),
        ("relative_residual", ctypes.c_double), ("equation_residual", ctypes.c_double),
        ("kernel_ms", ctypes.c_double), ("levels", ctypes.c_int), ("status", ctypes.c_int),
        ("message", ctypes.c_char * 256),
    ]


PP_CH_DIAGONAL, PP_CH_ILDLT, PP_CH_DIAGONAL_FALLBACK = 0, 1, 2
CH_SOLVER_NAMES = {PP_CH_DIAGONAL: "SparsePCG[Diagonal]", PP_CH_ILDLT: "SparsePCG[ILDLT]",
                   PP_CH_DIAGONAL_FALLBACK: "SparsePCG[DiagonalFallback]"}


def make_ch_problem(cc) -> PPChProblem:
    cp = PPChProblem()
    cp.nx, cp.ny = cc.nx, cc.ny
    cp.dx, cp.dy = cc.lx / cc.nx, cc.ly / cc.ny
    cp.periodic_x, cp.periodic_y = int(cc.periodic_x), int(cc.periodic_y)
    cp.alpha0, cp.beta1, cp.beta2 = float(cc.alpha0), float(cc.beta1), float(cc.beta2)
    cp.max_iterations, cp.tolerance = int(cc.max_iterations), float(cc.tolerance)
    return cp


def _ch_report_dict(rep: PPChReport) -> dict:
    return {"iterations": rep.iterations, "preconditioner": rep.preconditioner,
            "solver_name": CH_SOLVER_NAMES.get(rep.preconditioner, "?"), "iterate_residual": rep.iterate_residual,
            "absolute_residual": rep.absolute_residual, "relative_residual": rep.relative_residual,
            "equation_residual": rep.equation_residual, "kernel_ms": rep.kernel_ms, "levels": rep.levels,
            "status": rep.status, "message": rep.message.decode(errors="replace")}


class PPReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int), ("outer_cycles", ctypes.c_int),
        ("relative_residual", ctypes.c_double), ("absolute_residual", ctypes.c_double),
        ("iterate_residual", ctypes.c_double), ("kernel_ms", ctypes.c_double),
        ("phase_ms", ctypes.c_double * 10), ("status", ctypes.c_int),
        ("message", ctypes.c_char * 256),
    ]


class EngineError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


_lib = None

# Every entry point declared in include/pp_solver.h.
EXPORTS = ("pp_create", "pp_destroy", "pp_last_error", "pp_last_kernel_ms", "pp_solve", "pp_solve_device",
           "pp_solve_coeffs", "pp_solve_coeffs_device", "pp_adjoint", "pp_adjoint_device", "pp_solve_batched_device",
           "pp_set_tuning", "pp_velocity_correction", "pp_velocity_correction_device", "pp_solve_monitored",
           "pp_solve_device_monitored", "pp_solve_device_field", "pp_pressure_step", "pp_bicgstab",
           "pp_bicgstab_device", "pp_set_wait_limit", "pp_solve_device_noalloc", "pp_momentum_predictor",
           "pp_momentum_predictor_device", "pp_ch_solve", "pp_ch_solve_device", "pp_set_exact_mean")

# One PCG monitor record (pp_monitor_record): the scalar fields of one CGMonitor call.
MONITOR_DTYPE = np.dtype([("iteration", "<i4"), ("pad", "<i4"), ("alpha", "<f8"), ("beta", "<f8"),
                          ("absolute_residual", "<f8"), ("relative_residual", "<f8"), ("iterate_residual", "<f8"),
                          ("preconditioned_residual", "<f8")])


def library_available() -> bool:
    return os.path.exists(LIB_PATH)


def _bind_shared_hip_runtime():
    """Make the engine use the same HIP runtime as PyTorch in this process.

    PyTorch-ROCm wheels bundle their own libamdhip64 (soname libamdhip64.so.7) and load it by
    the file name libamdhip64.so. If the engine were loaded first it would map the system
    runtime under that soname, and PyTorch would then map a second runtime that sees no GPU
    (and whose stream handles the engine could not use). Loading PyTorch's copy first, by path,
    makes the engine's libamdhip64.so.7 dependency resolve to it. Without PyTorch installed the
    engine links the system runtime as usual."""
    spec = importlib.util.find_spec("torch")
    for root in (spec.submodule_search_locations or []) if spec else []:
        path = os.path.join(root, "lib", "libamdhip64.so")
        if os.path.exists(path):
            ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
            return path
    return None


def lib():
    global _lib
    if _lib is None:
        if not library_available():
            raise FileNotFoundError(f"HIP engine not built: {LIB_PATH} (python -m ding_nsch_core_solver_amd.build)")
        _bind_shared_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.pp_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.pp_create.restype = ctypes.c_int
        L.pp_destroy.argtypes = [vp]
        L.pp_destroy.restype = None
        L.pp_last_error.argtypes = [vp]
        L.pp_last_error.restype = ctypes.c_char_p
        L.pp_last_kernel_ms.argtypes = [vp]
        L.pp_last_kernel_ms.restype = ctypes.c_double
        L.pp_solve.argtypes = [vp, ctypes.POINTER(PPProblem), dp, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int,
                               dp, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.POINTER(PPReport)]
        L.pp_solve.restype = ctypes.c_int
        L.pp_solve_device.argtypes = [vp, ctypes.POINTER(PPProblem), vp, vp, vp, vp, vp, vp,
                                      ctypes.POINTER(PPReport), vp]
        L.pp_solve_device.restype = ctypes.c_int
        ip = ctypes.POINTER(ctypes.c_int)
        L.pp_solve_monitored.argtypes = [vp, ctypes.POINTER(PPProblem), dp, ctypes.c_int, dp, ctypes.c_int, dp,
                                         ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, vp,
                                         ctypes.c_int, ip, ctypes.POINTER(PPReport)]
        L.pp_solve_monitored.restype = ctypes.c_int
        L.pp_solve_device_monitored.argtypes = [vp, ctypes.POINTER(PPProblem), vp, vp, vp, vp, vp, vp, vp,
                                                ctypes.c_int, ip, ctypes.POINTER(PPReport), vp]
        L.pp_solve_device_monitored.restype = ctypes.c_int
        L.pp_solve_device_field.argtypes = [vp, ctypes.POINTER(PPProblem), ctypes.c_int, vp, vp, vp, vp, vp, vp,
                                            ctypes.POINTER(PPReport), vp]
        L.pp_solve_device_field.restype = ctypes.c_int
        L.pp_pressure_step.argtypes = [vp, ctypes.POINTER(PPProblem), ctypes.POINTER(PPVelocityBC), ctypes.c_int, dp,
                                       dp, dp, dp, dp, dp, dp, dp, dp, ctypes.POINTER(PPReport)]
        L.pp_pressure_step.restype = ctypes.c_int
        L.pp_solve_coeffs.argtypes = [vp, ctypes.POINTER(PPProblem), dp, dp, dp, dp, ctypes.POINTER(PPReport)]
        L.pp_solve_coeffs.restype = ctypes.c_int
        L.pp_solve_coeffs_device.argtypes = [vp, ctypes.POINTER(PPProblem), vp, vp, vp, vp, ctypes.POINTER(PPReport), vp]
        L.pp_solve_coeffs_device.restype = ctypes.c_int
        L.pp_adjoint.argtypes = [vp, ctypes.POINTER(PPProblem), dp, dp, dp, dp, dp, dp, dp, ctypes.POINTER(PPReport)]
        L.pp_adjoint.restype = ctypes.c_int
        L.pp_adjoint_device.argtypes = [vp, ctypes.POINTER(PPProblem), vp, vp, vp, vp, vp, vp, vp,
                                        ctypes.POINTER(PPReport), vp]
        L.pp_adjoint_device.restype = ctypes.c_int
        pv = ctypes.POINTER(ctypes.c_void_p)
        L.pp_solve_batched_device.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PPProblem), pv, pv, pv, pv, pv, pv,
                                              ctypes.POINTER(PPReport), vp]
        L.pp_solve_batched_device.restype = ctypes.c_int
        ip = ctypes.POINTER(ctypes.c_int)
        L.pp_set_tuning.argtypes = [vp, ctypes.c_int, ip, ip, ip]
        L.pp_set_tuning.restype = ctypes.c_int
        L.pp_ch_solve_device.argtypes = [vp, ctypes.POINTER(PPChProblem), vp, vp, ctypes.POINTER(PPChReport), vp]
        L.pp_ch_solve_device.restype = ctypes.c_int
        L.pp_ch_solve.argtypes = [vp, ctypes.POINTER(PPChProblem), dp, dp, ctypes.POINTER(PPChReport)]
        L.pp_ch_solve.restype = ctypes.c_int
        mpp = ctypes.POINTER(PPMomentumProblem)
        L.pp_momentum_predictor_device.argtypes = [vp, mpp, ctypes.POINTER(PPVelocityBC), ctypes.c_int] + [vp] * 11 + [
            ctypes.POINTER(PPReport), vp]
        L.pp_momentum_predictor_device.restype = ctypes.c_int
        L.pp_momentum_predictor.argtypes = [vp, mpp, ctypes.POINTER(PPVelocityBC), ctypes.c_int] + [dp] * 11 + [
            ctypes.POINTER(PPReport)]
        L.pp_momentum_predictor.restype = ctypes.c_int
        L.pp_solve_device_noalloc.argtypes = [ctypes.POINTER(PPProblem), vp, vp, vp, vp, vp, vp,
                                              ctypes.POINTER(PPReport), vp]
        L.pp_solve_device_noalloc.restype = ctypes.c_int
        L.pp_set_wait_limit.argtypes = [vp, ctypes.c_int]
        L.pp_set_wait_limit.restype = ctypes.c_int
        L.pp_set_exact_mean.argtypes = [vp, ctypes.c_int]
        L.pp_set_exact_mean.restype = ctypes.c_int
        vb = ctypes.POINTER(PPVelocityBC)
        L.pp_velocity_correction_device.argtypes = [vp, ctypes.POINTER(PPProblem), vb, ctypes.c_int, vp, vp, vp, vp,
                                                    vp, vp, vp, vp, dp, vp]
        L.pp_velocity_correction_device.restype = ctypes.c_int
        L.pp_velocity_correction.argtypes = [vp, ctypes.POINTER(PPProblem), vb, ctypes.c_int, dp, dp, dp, dp, dp, dp,
                                             dp, dp, dp]
        L.pp_velocity_correction.restype = ctypes.c_int
        L.pp_bicgstab.argtypes = [vp, ctypes.c_int, ip, ip, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.POINTER(PPReport)]
        L.pp_bicgstab.restype = ctypes.c_int
        L.pp_bicgstab_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.POINTER(PPReport), vp]
        L.pp_bicgstab_device.restype = ctypes.c_int
        _lib = L
    return _lib


def make_problem(case, scheme: int = PP_ICPCG, precision: int = PP_FP64, inner_max_iterations: int = 0,
                 inner_tolerance: float = 0.0, ring_depth: int = 0) -> PPProblem:
    pb = PPProblem()
    pb.nx, pb.ny = case.nx, case.ny
    pb.dx, pb.dy, pb.dt = case.dx, case.dy, case.dt
    pb.periodic_x, pb.periodic_y = int(case.periodic_x), int(case.periodic_y)
    for s in range(4):
        pb.bc_type[s] = int(case.bc_type[s])
        pb.bc_value[s] = float(case.bc_value[s])
    pb.scheme, pb.precision = int(scheme), int(precision)
    pb.max_iterations = int(case.max_iterations)
    pb.tolerance = float(case.tolerance)
    pb.density_ratio = float(case.density_ratio)
    pb.step_index = int(case.step_index)
    pb.inner_max_iterations = int(inner_max_iterations)
    pb.inner_tolerance = float(inner_tolerance)
    pb.ring_depth = int(ring_depth)
    return pb


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _report_dict(rep: PPReport) -> dict:
    return {
        "iterations": rep.iterations, "outer_cycles": rep.outer_cycles,
        "relative_residual": rep.relative_residual, "absolute_residual": rep.absolute_residual,
        "iterate_residual": rep.iterate_residual, "kernel_ms": rep.kernel_ms,
        "phase_ms": dict(zip(PHASES, list(rep.phase_ms))), "status": rep.status,
        "message": rep.message.decode(errors="replace"),
    }


TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")


def load_tuning(path: str = TUNING_FILE) -> dict:
    """{(nx, ny): ring_depth} from the autotuner's table (empty if there is none)."""
    import json
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    return {(int(e["nx"]), int(e["ny"])): int(e["ring_depth"]) for e in data.get("entries", [])}


class Engine:
    """One engine context on one device (pp_create / pp_destroy). The autotuner's table
    (tuning_gfx950.json beside this module) is installed when present, unless tuning=False."""

    def __init__(self, device: int = 0, tuning: bool = True):
        self._lib = lib()
        h = ctypes.c_void_p()
        rc = self._lib.pp_create(ctypes.byref(h), device)
        if rc != 0:
            raise EngineError(rc, f"pp_create failed ({rc})")
        self.h = h
        self.device = device
        if tuning:
            self.set_tuning(load_tuning())

    def set_tuning(self, table: dict):
        """Install {(nx, ny): ring_depth} (replaces the previous table)."""
        keys = sorted(table)
        arr = lambda vals: (ctypes.c_int * max(len(vals), 1))(*vals)  # noqa: E731
        rc = self._lib.pp_set_tuning(self.h, len(keys), arr([k[0] for k in keys]), arr([k[1] for k in keys]),
                                     arr([table[k] for k in keys]))
        if rc != 0:
            raise EngineError(rc, self.last_error())

    def set_wait_limit(self, sleeps: int = 0):
        """Sleeps a sweep hand-off wait may spend before the solve fails with PP_ERR_SWEEP
        (0: the built-in budget)."""
        rc = self._lib.pp_set_wait_limit(self.h, int(sleeps))
        if rc != 0:
            raise EngineError(rc, self.last_error())

    def set_exact_mean(self, exact: bool = True):
        """Velocity correction's pressure mean in the reference's sequential order (exact=True, its
        bits) or the tree order (exact=False, the default: within rounding, microseconds)."""
        rc = self._lib.pp_set_exact_mean(self.h, int(bool(exact)))
        if rc != 0:
            raise EngineError(rc, self.last_error())

    def close(self):
        if getattr(self, "h", None):
            self._lib.pp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def solve(self, case, scheme: int = PP_ICPCG, precision: int = PP_FP64, **kw):
        """Host-array solve. Returns (p interior as (nx*ny,) float64, report dict)."""
        pb = make_problem(case, scheme, precision, **kw)
        nx, ny = case.nx, case.ny
        rho = np.ascontiguousarray(case.rho, dtype=np.float64).reshape(nx, ny)
        u = np.ascontiguousarray(case.u, dtype=np.float64).reshape(nx + 1, ny)
        v = np.ascontiguousarray(case.v, dtype=np.float64).reshape(nx, ny + 1)
        pn = None if case.p_n is None else np.ascontiguousarray(case.p_n, dtype=np.float64).reshape(nx, ny)
        pm = None if case.p_nm1 is None else np.ascontiguousarray(case.p_nm1, dtype=np.float64).reshape(nx, ny)
        out = np.zeros((nx, ny))
        rep = PPReport()
        rc = self._lib.pp_solve(self.h, ctypes.byref(pb), _dp(rho), ny, _dp(u), ny, _dp(v), ny + 1,
                                _dp(pn), ny, _dp(pm), ny, _dp(out), ny, ctypes.byref(rep))
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return out.reshape(-1), _report_dict(rep)

    def solve_monitored(self, case, scheme: int = PP_ICPCG, capacity: int | None = None, precision: int = PP_FP64,
                        **kw):
        """Host-array solve with the PCG monitor log (FP64 only; the library rejects PP_MIXED).
        Returns (p, report dict, records): one MONITOR_DTYPE record per monitor call, up to
        `capacity` (default max_iterations + 1); the report's "monitor_calls" is the number of calls."""
        pb = make_problem(case, scheme, precision, **kw)
        nx, ny = case.nx, case.ny
        rho = np.ascontiguousarray(case.rho, dtype=np.float64).reshape(nx, ny)
        u = np.ascontiguousarray(case.u, dtype=np.float64).reshape(nx + 1, ny)
        v = np.ascontiguousarray(case.v, dtype=np.float64).reshape(nx, ny + 1)
        pn = None if case.p_n is None else np.ascontiguousarray(case.p_n, dtype=np.float64).reshape(nx, ny)
        pm = None if case.p_nm1 is None else np.ascontiguousarray(case.p_nm1, dtype=np.float64).reshape(nx, ny)
        cap = pb.max_iterations + 1 if capacity is None else int(capacity)
        records = np.zeros(max(cap, 1), dtype=MONITOR_DTYPE)
        count = ctypes.c_int(0)
        out = np.zeros((nx, ny))
        rep = PPReport()
        rc = self._lib.pp_solve_monitored(self.h, ctypes.byref(pb), _dp(rho), ny, _dp(u), ny, _dp(v), ny + 1,
                                          _dp(pn), ny, _dp(pm), ny, _dp(out), ny, records.ctypes.data, cap,
                                          ctypes.byref(count), ctypes.byref(rep))
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        report = _report_dict(rep)
        report["monitor_calls"] = count.value
        return out.reshape(-1), report, records[:min(count.value, cap)]

    def solve_device_monitored(self, case, d_rho, d_u, d_v, d_pn, d_pnm1, d_out, d_records, capacity: int,
                               scheme: int = PP_ICPCG, stream=None, **kw):
        """Device-pointer FP64 solve with the PCG monitor log written to d_records (device memory
        for `capacity` records). Returns the report dict with "monitor_calls"."""
        pb = make_problem(case, scheme, PP_FP64, **kw)
        rep = PPReport()
        count = ctypes.c_int(0)
        rc = self._lib.pp_solve_device_monitored(self.h, ctypes.byref(pb), d_rho, d_u, d_v, d_pn, d_pnm1, d_out,
                                                 d_records, int(capacity), ctypes.byref(count), ctypes.byref(rep),
                                                 stream)
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        report = _report_dict(rep)
        report["monitor_calls"] = count.value
        return report

    def solve_device_field(self, case, d_rho_mid, d_u_star, d_v_star, d_pressure, d_pressure_prev, d_p_corr,
                           scheme: int = PP_ICPCG, precision: int = PP_FP64, ghost: int = 3, stream=None, **kw):
        """The backend contract on ghosted Field2D device buffers (see pp_solve_device_field):
        d_p_corr receives pressure_correction_ with its ghosts. Returns the report dict."""
        pb = make_problem(case, scheme, precision, **kw)
        rep = PPReport()
        rc = self._lib.pp_solve_device_field(self.h, ctypes.byref(pb), int(ghost), d_rho_mid, d_u_star, d_v_star,
                                             d_pressure, d_pressure_prev, d_p_corr, ctypes.byref(rep), stream)
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return _report_dict(rep)

    def solve_device(self, case, d_rho, d_u, d_v, d_pn, d_pnm1, d_out, scheme: int = PP_ICPCG,
                     precision: int = PP_FP64, stream=None, **kw):
        """Solve on device pointers (ints, e.g. torch tensor.data_ptr()). Returns the report dict."""
        pb = make_problem(case, scheme, precision, **kw)
        rep = PPReport()
        rc = self._lib.pp_solve_device(self.h, ctypes.byref(pb), d_rho, d_u, d_v, d_pn, d_pnm1, d_out,
                                       ctypes.byref(rep), stream)
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return _report_dict(rep)

    def _check(self, rc, rep):
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return _report_dict(rep)

    def solve_coeffs(self, case, ku, kv, src, scheme: int = PP_ICPCG, precision: int = PP_FP64, **kw):
        """Coefficient-level solve on host arrays (k_u (nx+1)*ny, k_v nx*(ny+1), src nx*ny).
        Returns (p (nx*ny,), report dict)."""
        pb = make_problem(case, scheme, precision, **kw)
        ku, kv, src = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (ku, kv, src))
        assert ku.size == (case.nx + 1) * case.ny and kv.size == case.nx * (case.ny + 1) and src.size == case.n
        out = np.zeros(case.n)
        rep = PPReport()
        rc = self._lib.pp_solve_coeffs(self.h, ctypes.byref(pb), _dp(ku), _dp(kv), _dp(src), _dp(out), ctypes.byref(rep))
        return out, self._check(rc, rep)

    def adjoint(self, case, ku, kv, p, dLdp, scheme: int = PP_ICPCG, precision: int = PP_FP64, **kw):
        """Adjoint of solve_coeffs on host arrays. Returns (dL/dsrc, dL/dk_u, dL/dk_v, report dict)."""
        pb = make_problem(case, scheme, precision, **kw)
        ku, kv, p, dLdp = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (ku, kv, p, dLdp))
        lam = np.zeros(case.n)
        dku = np.zeros((case.nx + 1) * case.ny)
        dkv = np.zeros(case.nx * (case.ny + 1))
        rep = PPReport()
        rc = self._lib.pp_adjoint(self.h, ctypes.byref(pb), _dp(ku), _dp(kv), _dp(p), _dp(dLdp), _dp(lam), _dp(dku),
                                  _dp(dkv), ctypes.byref(rep))
        return lam, dku, dkv, self._check(rc, rep)

    def solve_coeffs_device(self, case, d_ku, d_kv, d_src, d_out, scheme: int = PP_ICPCG, stream=None,
                            precision: int = PP_FP64, **kw):
        pb = make_problem(case, scheme, precision, **kw)
        rep = PPReport()
        rc = self._lib.pp_solve_coeffs_device(self.h, ctypes.byref(pb), d_ku, d_kv, d_src, d_out, ctypes.byref(rep),
                                              stream)
        return self._check(rc, rep)

    def adjoint_device(self, case, d_ku, d_kv, d_p, d_g, d_lam, d_dku, d_dkv, scheme: int = PP_ICPCG, stream=None,
                       precision: int = PP_FP64, **kw):
        pb = make_problem(case, scheme, precision, **kw)
        rep = PPReport()
        rc = self._lib.pp_adjoint_device(self.h, ctypes.byref(pb), d_ku, d_kv, d_p, d_g, d_lam, d_dku, d_dkv,
                                         ctypes.byref(rep), stream)
        return self._check(rc, rep)

    def solve_batched_device(self, cases, d_rho, d_u, d_v, d_out, d_pn=None, d_pnm1=None, scheme: int = PP_ICPCG,
                             stream=None, precision=PP_FP64, **kw):
        """Batched solve; d_* are lists of device pointers (ints); scheme and precision are one value
        or one per system. Returns the report dicts. Raises EngineError for the first failing system
        (reports still describe every system)."""
        count = len(cases)
        schemes = scheme if isinstance(scheme, (list, tuple)) else [scheme] * count
        precs = precision if isinstance(precision, (list, tuple)) else [precision] * count
        pbs = (PPProblem * count)(*[make_problem(c, s, pr, **kw) for c, s, pr in zip(cases, schemes, precs)])
        P = ctypes.c_void_p * count
        ptrs = lambda xs: None if xs is None else P(*[x if x else None for x in xs])  # noqa: E731
        reps = (PPReport * count)()
        rc = self._lib.pp_solve_batched_device(self.h, count, pbs, ptrs(d_rho), ptrs(d_u), ptrs(d_v), ptrs(d_pn),
                                               ptrs(d_pnm1), ptrs(d_out), reps, stream)
        out = [_report_dict(r) for r in reps]
        if rc != 0:
            bad = next((k for k, r in enumerate(out) if r["status"] != 0), None)
            msg = out[bad]["message"] if bad is not None else self.last_error()
            err = EngineError(rc, msg)
            err.reports = out
            err.index = bad
            raise err
        return out

    @staticmethod
    def _check_field_sizes(case, ghost, **fields):
        """Host Field2D buffers must have exactly their ghosted sizes: the library copies whole
        buffers of the sizes the grid implies."""
        nx, ny, g = case.nx, case.ny, int(ghost)
        if not 1 <= g <= 8:
            return  # the library rejects the ghost count itself
        shapes = {"u_star": (nx + 1, ny), "u": (nx + 1, ny), "v_star": (nx, ny + 1), "v": (nx, ny + 1)}
        for name, a in fields.items():
            rows, cols = shapes.get(name, (nx, ny))
            want = (rows + 2 * g) * (cols + 2 * g)
            if np.asarray(a).size != want:
                raise ValueError(f"{name}: {np.asarray(a).size} values, expected {want} "
                                 f"({rows} x {cols} with {g} ghost layers)")

    def velocity_correction(self, case, vbc, rho_mid, u_star, v_star, p_corr, u, v, pressure, pressure_prev,
                            scheme: int = PP_ICPCG, ghost: int = 3):
        """Velocity correction and pressure-history shift on host Field2D buffers (ghost layers
        included). Returns new (u, v, pressure, pressure_prev, (speed_pre, speed_post)); the inputs
        are not modified."""
        self._check_field_sizes(case, ghost, rho_mid=rho_mid, u_star=u_star, v_star=v_star, p_corr=p_corr, u=u, v=v,
                                pressure=pressure, pressure_prev=pressure_prev)
        pb = make_problem(case, scheme, PP_FP64)
        ins = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (rho_mid, u_star, v_star, p_corr)]
        outs = [np.array(a, dtype=np.float64, copy=True).reshape(-1) for a in (u, v, pressure, pressure_prev)]
        speeds = np.zeros(2)
        rc = self._lib.pp_velocity_correction(self.h, ctypes.byref(pb), ctypes.byref(make_velocity_bc(vbc)), ghost,
                                              *[_dp(a) for a in ins], *[_dp(a) for a in outs], _dp(speeds))
        if rc != 0:
            raise EngineError(rc, self.last_error())
        return outs[0], outs[1], outs[2], outs[3], (float(speeds[0]), float(speeds[1]))

    def pressure_step(self, case, vbc, rho_mid, u_star, v_star, u, v, pressure, pressure_prev,
                      scheme: int = PP_ICPCG, precision: int = PP_FP64, ghost: int = 3, **kw):
        """The pressure part of a timestep on host Field2D buffers (pp_pressure_step). Returns
        (p_corr, u, v, pressure, pressure_prev, (speed_pre, speed_post), report); inputs unchanged."""
        self._check_field_sizes(case, ghost, rho_mid=rho_mid, u_star=u_star, v_star=v_star, u=u, v=v,
                                pressure=pressure, pressure_prev=pressure_prev)
        pb = make_problem(case, scheme, precision, **kw)
        f64 = lambda a: np.array(a, dtype=np.float64, copy=True).reshape(-1)  # noqa: E731
        rho_mid, u_star, v_star = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                                   for a in (rho_mid, u_star, v_star))
        u, v, pressure, prev = f64(u), f64(v), f64(pressure), f64(pressure_prev)
        p_corr = np.zeros_like(pressure)
        speeds = np.zeros(2)
        rep = PPReport()
        rc = self._lib.pp_pressure_step(self.h, ctypes.byref(pb), ctypes.byref(make_velocity_bc(vbc)), int(ghost),
                                        _dp(rho_mid), _dp(u_star), _dp(v_star), _dp(p_corr), _dp(u), _dp(v),
                                        _dp(pressure), _dp(prev), _dp(speeds), ctypes.byref(rep))
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return p_corr, u, v, pressure, prev, (float(speeds[0]), float(speeds[1])), _report_dict(rep)

    def velocity_correction_device(self, case, vbc, d_rho_mid, d_u_star, d_v_star, d_p_corr, d_u, d_v, d_pressure,
                                   d_pressure_prev, scheme: int = PP_ICPCG, ghost: int = 3, stream=None):
        """Device-pointer form; returns (speed_pre, speed_post)."""
        pb = make_problem(case, scheme, PP_FP64)
        speeds = np.zeros(2)
        rc = self._lib.pp_velocity_correction_device(self.h, ctypes.byref(pb), ctypes.byref(make_velocity_bc(vbc)),
                                                     ghost, d_rho_mid, d_u_star, d_v_star, d_p_corr, d_u, d_v,
                                                     d_pressure, d_pressure_prev, _dp(speeds), stream)
        if rc != 0:
            raise EngineError(rc, self.last_error())
        return float(speeds[0]), float(speeds[1])

    def bicgstab(self, system, x0=None, max_iterations: int = 50, tolerance: float = 1.0e-8,
                 use_relative: bool = False):
        """The momentum predictor's solve (pp_bicgstab) on a CSR system with row_ptr, col_idx (int32),
        values and rhs (float64). Returns (x, report); defaults are the reference's
        momentum_iterations, momentum_tolerance and absolute test (momentum.cpp:469-472)."""
        rp = np.ascontiguousarray(system.row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(system.col_idx, dtype=np.int32)
        val = np.ascontiguousarray(system.values, dtype=np.float64)
        rhs = np.ascontiguousarray(system.rhs, dtype=np.float64)
        n = len(rp) - 1
        x = np.zeros(max(n, 0)) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
        if n > 0 and (len(x) != n or len(rhs) != n or len(ci) < rp[n] or len(val) < rp[n]):
            raise ValueError("CSR, right-hand side and x sizes do not match")
        ipp = ctypes.POINTER(ctypes.c_int)
        rep = PPReport()
        rc = self._lib.pp_bicgstab(self.h, n, rp.ctypes.data_as(ipp), ci.ctypes.data_as(ipp), _dp(val), _dp(rhs),
                                   _dp(x) if n > 0 else None, int(max_iterations), float(tolerance),
                                   int(bool(use_relative)), ctypes.byref(rep))
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return x, _report_dict(rep)

    def bicgstab_device(self, n, d_row_ptr, d_col_idx, d_values, d_rhs, d_x, max_iterations: int = 50,
                        tolerance: float = 1.0e-8, use_relative: bool = False, stream=None) -> dict:
        """Device-pointer form (pp_bicgstab_device); d_x is solved in place. Returns the report."""
        rep = PPReport()
        rc = self._lib.pp_bicgstab_device(self.h, int(n), d_row_ptr, d_col_idx, d_values, d_rhs, d_x,
                                          int(max_iterations), float(tolerance), int(bool(use_relative)),
                                          ctypes.byref(rep), stream)
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return _report_dict(rep)

    def ch_solve(self, cc, rhs=None, c0=None):
        """The Cahn-Hilliard linear solve (pp_ch_solve) of a workloads.ChCase (or its rhs / c0
        overridden). Returns (c, report); the context keeps the operator for the next solve."""
        rhs = np.ascontiguousarray(cc.rhs if rhs is None else rhs, dtype=np.float64).reshape(-1)
        c = np.array(cc.c0 if c0 is None else c0, dtype=np.float64, copy=True).reshape(-1)
        rep = PPChReport()
        rc = self._lib.pp_ch_solve(self.h, ctypes.byref(make_ch_problem(cc)), _dp(rhs), _dp(c), ctypes.byref(rep))
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return c, _ch_report_dict(rep)

    def ch_solve_device(self, cc, d_rhs, d_c, stream=None) -> dict:
        """Device-pointer form (pp_ch_solve_device); d_c is solved in place."""
        rep = PPChReport()
        rc = self._lib.pp_ch_solve_device(self.h, ctypes.byref(make_ch_problem(cc)), d_rhs, d_c, ctypes.byref(rep),
                                          stream)
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return _ch_report_dict(rep)

    MOMENTUM_INPUTS = ("u", "v", "rho_mid", "eta", "eta_prev", "surface_fx_u", "surface_fy_v")

    def momentum_predictor(self, mc):
        """The momentum predictor (pp_momentum_predictor) on a workloads.MomentumCase's host Field2D
        buffers. Returns (u_star, v_star, u_rhs_prev, v_rhs_prev, report): the predictor's fields and
        this step's convection fields, all ghosted and flattened; the case is not modified."""
        ins = [np.ascontiguousarray(getattr(mc, k), dtype=np.float64).reshape(-1) for k in self.MOMENTUM_INPUTS]
        u_prev = np.array(mc.u_prev, dtype=np.float64, copy=True).reshape(-1)
        v_prev = np.array(mc.v_prev, dtype=np.float64, copy=True).reshape(-1)
        u_star, v_star = np.zeros_like(u_prev), np.zeros_like(v_prev)
        rep = PPReport()
        rc = self._lib.pp_momentum_predictor(self.h, ctypes.byref(make_momentum_problem(mc)),
                                             ctypes.byref(make_velocity_bc(mc.vbc)), int(mc.ghost),
                                             *[_dp(a) for a in ins], _dp(u_prev), _dp(v_prev), _dp(u_star),
                                             _dp(v_star), ctypes.byref(rep))
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return u_star, v_star, u_prev, v_prev, _report_dict(rep)

    def momentum_predictor_device(self, mc, d_u, d_v, d_rho_mid, d_eta, d_eta_prev, d_sfx, d_sfy, d_u_prev, d_v_prev,
                                  d_u_star, d_v_star, stream=None) -> dict:
        """Device-pointer form (pp_momentum_predictor_device); d_u_prev / d_v_prev are updated in place."""
        rep = PPReport()
        rc = self._lib.pp_momentum_predictor_device(self.h, ctypes.byref(make_momentum_problem(mc)),
                                                    ctypes.byref(make_velocity_bc(mc.vbc)), int(mc.ghost), d_u, d_v,
                                                    d_rho_mid, d_eta, d_eta_prev, d_sfx, d_sfy, d_u_prev, d_v_prev,
                                                    d_u_star, d_v_star, ctypes.byref(rep), stream)
        if rc != 0:
            raise EngineError(rc, rep.message.decode(errors="replace") or self.last_error())
        return _report_dict(rep)

    def last_error(self) -> str:
        return self._lib.pp_last_error(self.h).decode(errors="replace")

    def last_kernel_ms(self) -> float:
        """Device time of the kernels the last call ran (HIP events on that call's stream)."""
        return float(self._lib.pp_last_kernel_ms(self.h))
