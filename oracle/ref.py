"""ctypes access to the reference solver compiled from its own sources (oracle/_ref/libdingref.so).

Checker only: imported by tests/, the smoke check and bench.py's CPU-baseline leg. The library is
built by ``make -C oracle ref`` where the reference sources exist; callers must handle its absence
(``available()``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libdingref.so")
_lib = None


class RefProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("dt", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int),
        ("bc_type", ctypes.c_int * 4), ("bc_value", ctypes.c_double * 4),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
        ("density_ratio", ctypes.c_double), ("step_index", ctypes.c_int),
    ]


class RefVelocityBc(ctypes.Structure):
    _fields_ = [
        ("u_type", ctypes.c_int * 4), ("u_value", ctypes.c_double * 4),
        ("v_type", ctypes.c_int * 4), ("v_value", ctypes.c_double * 4),
        ("bottom_wall_velocity_x", ctypes.c_double), ("top_wall_velocity_x", ctypes.c_double),
    ]


class RefChConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("lx", ctypes.c_double), ("ly", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int), ("cn", ctypes.c_double), ("pe", ctypes.c_double),
        ("ch_inner_iterations", ctypes.c_int), ("ch_tolerance", ctypes.c_double),
    ]


class RefMomentumProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("dt", ctypes.c_double), ("re", ctypes.c_double),
        ("body_force_x", ctypes.c_double), ("body_force_y", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int), ("step_index", ctypes.c_int),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
    ]


RefChSolveFn = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


class RefRunConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("dt", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int),
        ("re", ctypes.c_double), ("ca", ctypes.c_double), ("pe", ctypes.c_double), ("cn", ctypes.c_double),
        ("density_ratio", ctypes.c_double), ("viscosity_ratio", ctypes.c_double),
        ("surface_tension_multiplier", ctypes.c_double),
        ("body_force_x", ctypes.c_double), ("body_force_y", ctypes.c_double),
        ("center_x", ctypes.c_double), ("center_y", ctypes.c_double), ("radius", ctypes.c_double),
        ("width", ctypes.c_double),
        ("ch_inner_iterations", ctypes.c_int), ("ch_tolerance", ctypes.c_double),
        ("momentum_iterations", ctypes.c_int), ("momentum_tolerance", ctypes.c_double),
        ("poisson_iterations", ctypes.c_int), ("pressure_tolerance", ctypes.c_double),
    ]


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"reference checker not built: {LIB_PATH}")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ref_pressure_solve.argtypes = [ctypes.c_char_p, ctypes.POINTER(RefProblem), dp, dp, dp, dp, dp, dp, dp,
                                         ip, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_pressure_solve.restype = ctypes.c_int
        L.ref_apply_pressure_bc.argtypes = [ctypes.POINTER(RefProblem), dp]
        L.ref_sys_build.argtypes = [ctypes.POINTER(RefProblem), ctypes.c_int, dp, dp, dp, dp, dp,
                                    ctypes.c_char_p, ctypes.c_int]
        L.ref_sys_build.restype = ctypes.c_void_p
        L.ref_sys_free.argtypes = [ctypes.c_void_p]
        L.ref_sys_sizes.argtypes = [ctypes.c_void_p, ip, ip, ip]
        L.ref_sys_csr.argtypes = [ctypes.c_void_p, ip, ip, dp, dp]
        L.ref_sys_factor.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
        L.ref_sys_factor.restype = ctypes.c_int
        L.ref_sys_apply_matrix.argtypes = [ctypes.c_void_p, dp, dp]
        L.ref_sys_apply_precond.argtypes = [ctypes.c_void_p, dp, dp]
        L.ref_sys_pcg.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp, ip, dp, dp,
                                  ctypes.c_char_p, ctypes.c_int]
        L.ref_sys_pcg.restype = ctypes.c_int
        L.ref_sys_pcg_monitored.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp, ip, dp, dp,
                                            ctypes.c_void_p, ctypes.c_int, ip, ctypes.c_char_p, ctypes.c_int]
        L.ref_sys_pcg_monitored.restype = ctypes.c_int
        L.ref_pressure_ghosts.argtypes = [ctypes.POINTER(RefProblem), ctypes.c_int, dp]
        L.ref_pressure_ghosts.restype = ctypes.c_int
        L.ref_velocity_correction.argtypes = [ctypes.c_char_p, ctypes.POINTER(RefProblem), ctypes.POINTER(RefVelocityBc),
                                              ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, dp, dp, ctypes.c_int, dp,
                                              ctypes.c_char_p, ctypes.c_int]
        L.ref_velocity_correction.restype = ctypes.c_int
        L.ref_bicgstab.argtypes = [ctypes.c_int, ip, ip, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ip, dp, dp, dp, ctypes.c_int, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_bicgstab.restype = ctypes.c_int
        vp = ctypes.c_void_p
        L.ref_ch_create.argtypes = [ctypes.POINTER(RefChConfig), ctypes.c_char_p, ctypes.c_int]
        L.ref_ch_create.restype = vp
        L.ref_ch_destroy.argtypes = [vp]
        L.ref_ch_destroy.restype = None
        L.ref_ch_solve.argtypes = [vp, ctypes.c_double, dp, dp, ip, dp, dp, ctypes.c_char_p, ctypes.c_int,
                                   ctypes.c_int, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_ch_solve.restype = ctypes.c_int
        L.ref_ch_operator.argtypes = [vp, ip, ip, ip, ip, dp, ip, dp, dp]
        L.ref_ch_operator.restype = None
        L.ref_momentum_predictor.argtypes = [ctypes.POINTER(RefMomentumProblem), ctypes.POINTER(RefVelocityBc),
                                             ctypes.c_int] + [dp] * 11 + [ip, dp, ctypes.c_int, dp, ctypes.c_char_p,
                                                                          ctypes.c_int]
        L.ref_momentum_predictor.restype = ctypes.c_int
        L.ref_run_create.argtypes = [ctypes.POINTER(RefRunConfig), ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.c_int]
        L.ref_run_create.restype = vp
        L.ref_run_destroy.argtypes = [vp]
        L.ref_run_destroy.restype = None
        L.ref_run_step_index.argtypes = [vp]
        L.ref_run_step_index.restype = ctypes.c_int
        L.ref_run_phase_momentum.argtypes = [vp, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_run_phase_momentum.restype = ctypes.c_int
        L.ref_run_phase.argtypes = [vp, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_run_phase.restype = ctypes.c_int
        L.ref_run_phase_external_ch.argtypes = [vp, RefChSolveFn, vp, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_run_phase_external_ch.restype = ctypes.c_int
        L.ref_run_momentum.argtypes = [vp, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_run_momentum.restype = ctypes.c_int
        L.ref_run_set_momentum_report.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.ref_run_set_momentum_report.restype = None
        L.ref_run_pressure.argtypes = [vp, ip, dp, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_run_pressure.restype = ctypes.c_int
        L.ref_run_finish.argtypes = [vp, dp, ctypes.c_char_p, ctypes.c_int]
        L.ref_run_finish.restype = ctypes.c_int
        L.ref_run_field.argtypes = [vp, ctypes.c_char_p, ip]
        L.ref_run_field.restype = dp
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def problem(case) -> RefProblem:
    pb = RefProblem()
    pb.nx, pb.ny = case.nx, case.ny
    pb.lx, pb.ly, pb.dt = case.lx, case.ly, case.dt
    pb.periodic_x, pb.periodic_y = int(case.periodic_x), int(case.periodic_y)
    for s in range(4):
        pb.bc_type[s] = int(case.bc_type[s])
        pb.bc_value[s] = float(case.bc_value[s])
    pb.max_iterations = case.max_iterations
    pb.tolerance = case.tolerance
    pb.density_ratio = case.density_ratio
    pb.step_index = case.step_index
    return pb


def pressure_solve(scheme: str, case, ghosted: bool = False):
    """Solver::solve_pressure_correction() with cfg_.pressure_scheme = scheme.

    Returns (p interior (nx*ny,), iterations, relative residual[, ghosted Field2D buffer])."""
    L = lib()
    pb = problem(case)
    p = np.zeros(case.n)
    g = np.zeros((case.nx + 6) * (case.ny + 6)) if ghosted else None
    it = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = L.ref_pressure_solve(scheme.encode(), ctypes.byref(pb), _dp(case.rho), _dp(case.u), _dp(case.v),
                              _dp(case.p_n), _dp(case.p_nm1), _dp(p), _dp(g), ctypes.byref(it), ctypes.byref(res),
                              err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    if ghosted:
        return p, it.value, res.value, g
    return p, it.value, res.value


class RefSystem:
    """The reference's assembled CSR system for a case, with its factor and operator applications."""

    def __init__(self, case, use_split=False):
        L = lib()
        self.case = case
        err = ctypes.create_string_buffer(512)
        self.h = L.ref_sys_build(ctypes.byref(problem(case)), int(use_split), _dp(case.rho), _dp(case.u),
                                 _dp(case.v), _dp(case.p_n), _dp(case.p_nm1), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        n, nnz, hd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.ref_sys_sizes(self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(hd))
        self.n, self.nnz, self.has_dirichlet = n.value, nnz.value, bool(hd.value)
        self.row_ptr = np.zeros(self.n + 1, dtype=np.int32)
        self.col_idx = np.zeros(self.nnz, dtype=np.int32)
        self.values = np.zeros(self.nnz)
        self.rhs = np.zeros(self.n)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ref_sys_csr(self.h, self.row_ptr.ctypes.data_as(ip), self.col_idx.ctypes.data_as(ip), _dp(self.values),
                      _dp(self.rhs))

    def factor(self, kind=1):
        """kind 1 = ICC(0), 2 = ILDLT(0). Returns (ok, lower (nnz,), diag (n,), floor)."""
        lower = np.zeros(self.nnz)
        diag = np.zeros(self.n)
        fl = ctypes.c_double(0.0)
        ok = lib().ref_sys_factor(self.h, kind, _dp(lower), _dp(diag), ctypes.byref(fl))
        return bool(ok), lower, diag, fl.value

    def apply_matrix(self, x):
        y = np.zeros(self.n)
        lib().ref_sys_apply_matrix(self.h, _dp(np.ascontiguousarray(x, dtype=np.float64)), _dp(y))
        return y

    def apply_precond(self, r):
        z = np.zeros(self.n)
        lib().ref_sys_apply_precond(self.h, _dp(np.ascontiguousarray(r, dtype=np.float64)), _dp(z))
        return z

    def pcg(self, max_iterations, tolerance):
        x = np.zeros(self.n)
        it, rel, ab = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        err = ctypes.create_string_buffer(512)
        rc = lib().ref_sys_pcg(self.h, max_iterations, tolerance, _dp(x), ctypes.byref(it), ctypes.byref(rel),
                               ctypes.byref(ab), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return x, it.value, rel.value, ab.value

    def pcg_monitored(self, max_iterations, tolerance, capacity=None):
        """The reference's PCG with a CGMonitor recording every call; returns (x, iterations, rel,
        abs, records, error text or None) like the oracle's pcg_monitored."""
        from oracle.cpu import MONITOR_DTYPE
        cap = max_iterations + 1 if capacity is None else capacity
        records = np.zeros(max(cap, 1), dtype=MONITOR_DTYPE)
        x = np.zeros(self.n)
        it, rel, ab, count = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        err = ctypes.create_string_buffer(512)
        rc = lib().ref_sys_pcg_monitored(self.h, max_iterations, tolerance, _dp(x), ctypes.byref(it),
                                         ctypes.byref(rel), ctypes.byref(ab), records.ctypes.data, cap,
                                         ctypes.byref(count), err, 512)
        return (x, it.value, rel.value, ab.value, records[:min(count.value, cap)],
                err.value.decode() if rc != 0 else None)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_sys_free(self.h)
            self.h = None


def velocity_correction(scheme: str, case, vbc, ghost: int, rho_mid, u_star, v_star, p_corr, u, v, pressure,
                        pressure_prev, repeats: int = 1, timing: list | None = None):
    """The reference's apply_pressure_velocity_correction on ghosted Field2D buffers.
    Returns the new (u, v, pressure, pressure_prev, (speed_pre, speed_post)). With repeats > 1 the
    call runs that many times on the same state (outputs of the last call) and, when `timing` is a
    list, the median seconds of one call (set-up excluded) is appended to it."""
    b = RefVelocityBc()
    for k in range(4):
        b.u_type[k], b.u_value[k] = int(vbc.u_type[k]), float(vbc.u_value[k])
        b.v_type[k], b.v_value[k] = int(vbc.v_type[k]), float(vbc.v_value[k])
    b.bottom_wall_velocity_x = float(vbc.bottom_wall_velocity_x)
    b.top_wall_velocity_x = float(vbc.top_wall_velocity_x)
    u, v, pressure, prev = (np.array(a, dtype=np.float64, copy=True) for a in (u, v, pressure, pressure_prev))
    speeds = np.zeros(2)
    seconds = np.zeros(1)
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho_mid, u_star, v_star, p_corr)]
    err = ctypes.create_string_buffer(512)
    rc = lib().ref_velocity_correction(scheme.encode(), ctypes.byref(problem(case)), ctypes.byref(b), int(ghost),
                                       *[_dp(a) for a in ins], _dp(u), _dp(v), _dp(pressure), _dp(prev), _dp(speeds),
                                       int(repeats), _dp(seconds), err, 512)
    if rc:
        raise RuntimeError(err.value.decode())
    if timing is not None:
        timing.append(float(seconds[0]))
    return u, v, pressure, prev, (float(speeds[0]), float(speeds[1]))


def pressure_ghosts(case, field_ghosted: np.ndarray, ghost: int) -> np.ndarray:
    """The reference's apply_pressure_bc on a Field2D buffer with `ghost` layers, in place."""
    assert field_ghosted.size == (case.nx + 2 * ghost) * (case.ny + 2 * ghost)
    lib().ref_pressure_ghosts(ctypes.byref(problem(case)), int(ghost), _dp(field_ghosted))
    return field_ghosted


def bicgstab(system, x0=None, max_iterations: int = 200, tolerance: float = 1e-8, use_relative: bool = False,
             repeats: int = 1, timing: list | None = None):
    """The reference's build_diagonal_preconditioner + solve_preconditioned_bicgstab on a CSR system
    (row_ptr, col_idx, values, rhs), after its finalize_row_maps. Returns (x, iterations, absolute,
    relative, iterate residual); raises RuntimeError with the reference's text. With a timing list,
    the mean seconds of one solve over `repeats` is appended to it."""
    L = lib()
    rp, ci, val, rhs = system.row_ptr, system.col_idx, system.values, system.rhs
    assert rp.dtype == np.int32 and ci.dtype == np.int32
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = ctypes.c_int(0)
    ab, rel, itr = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    seconds = np.zeros(1)
    err = ctypes.create_string_buffer(512)
    ipp = ctypes.POINTER(ctypes.c_int)
    rc = L.ref_bicgstab(n, rp.ctypes.data_as(ipp), ci.ctypes.data_as(ipp), _dp(val), _dp(rhs), _dp(x),
                        int(max_iterations), float(tolerance), int(bool(use_relative)), ctypes.byref(it),
                        ctypes.byref(ab), ctypes.byref(rel), ctypes.byref(itr), int(repeats), _dp(seconds), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    if timing is not None:
        timing.append(float(seconds[0]))
    return x, it.value, ab.value, rel.value, itr.value


class RefRun:
    """A full run of the reference Solver (solver.cpp:224-283), stepped in three parts so that the
    pressure step between the momentum predictor and the rest of the timestep can be the reference's
    own or another one working on the same Field2D buffers:

        run.phase_momentum()   # history shifts, CH step, momentum fluxes and predictor (u_star_, v_star_)
        run.pressure()         # solve_pressure_correction + apply_pressure_velocity_correction
        run.finish()           # chemical potential, diagnostics; the step counter advances

    field(name) is a writable (nx + 2g, ny + 2g) view of a Solver Field2D buffer (u is (nx+1) x ny,
    v nx x (ny+1) before the ghosts)."""

    def __init__(self, spec: dict, scheme: str = "icpcg", restart_file: str | None = None):
        rc = RefRunConfig()
        for name, _ in RefRunConfig._fields_:
            setattr(rc, name, spec[name])
        err = ctypes.create_string_buffer(512)
        self.h = lib().ref_run_create(ctypes.byref(rc), scheme.encode(),
                                      restart_file.encode() if restart_file else None, err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.spec = dict(spec)

    @property
    def step(self) -> int:
        return lib().ref_run_step_index(self.h)

    def _call(self, fn, *args):
        err = ctypes.create_string_buffer(512)
        if fn(self.h, *args, err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def phase_momentum(self) -> float:
        sec = ctypes.c_double(0.0)
        self._call(lib().ref_run_phase_momentum, ctypes.byref(sec))
        return sec.value

    def phase(self) -> float:
        """History shifts and the CH stage of the step (advance_phase_state)."""
        sec = ctypes.c_double(0.0)
        self._call(lib().ref_run_phase, ctypes.byref(sec))
        return sec.value

    def phase_external_ch(self, solve) -> float:
        """phase() with the CH linear solve done by solve(alpha0, rhs, c) -> (iterations, iterate
        residual, equation residual), which updates c (natural order) in place."""
        def cb(user, alpha0, rhs_p, c_p, n, it_p, itr_p, eq_p):
            try:
                rhs = np.ctypeslib.as_array(rhs_p, shape=(n,))
                c = np.ctypeslib.as_array(c_p, shape=(n,))
                it, itr, eq = solve(alpha0, rhs, c)
                it_p[0], itr_p[0], eq_p[0] = int(it), float(itr), float(eq)
                return 0
            except Exception as e:  # noqa: BLE001 - reported through the driver's error text
                self._cb_error = e
                return 1
        fn = RefChSolveFn(cb)
        sec = ctypes.c_double(0.0)
        self._cb_error = None
        try:
            self._call(lib().ref_run_phase_external_ch, fn, None, ctypes.byref(sec))
        except RuntimeError:
            if self._cb_error is not None:
                raise self._cb_error
            raise
        return sec.value

    def momentum(self) -> float:
        """The reference's momentum predictor of the step (after phase())."""
        sec = ctypes.c_double(0.0)
        self._call(lib().ref_run_momentum, ctypes.byref(sec))
        return sec.value

    def set_momentum_report(self, residual: float, iterations: int):
        """Record an external predictor's residual and iterations for the step's diagnostics."""
        lib().ref_run_set_momentum_report(self.h, float(residual), int(iterations))

    def momentum_case(self):
        """The momentum predictor's inputs of the current step as a workloads.MomentumCase (copies
        of the Solver's ghosted buffers; call after phase())."""
        from ding_nsch_core_solver_amd import workloads as W
        sp = self.spec
        f = lambda k: self.field(k).reshape(-1).copy()  # noqa: E731
        return W.MomentumCase(nx=sp["nx"], ny=sp["ny"], lx=sp["lx"], ly=sp["ly"], dt=sp["dt"], re=sp["re"],
                              body_force_x=sp["body_force_x"], body_force_y=sp["body_force_y"],
                              periodic_x=bool(sp["periodic_x"]), periodic_y=bool(sp["periodic_y"]),
                              step_index=self.step, max_iterations=sp["momentum_iterations"],
                              tolerance=sp["momentum_tolerance"], ghost=3, u=f("u"), v=f("v"), rho_mid=f("rho_mid"),
                              eta=f("eta"), eta_prev=f("eta_previous_step"), surface_fx_u=f("surface_fx_u"),
                              surface_fy_v=f("surface_fy_v"), u_prev=f("momentum_u_rhs_prev"),
                              v_prev=f("momentum_v_rhs_prev"))

    def pressure(self):
        """The reference's pressure step. Returns (iterations, relative residual, seconds)."""
        it, res, sec = ctypes.c_int(0), ctypes.c_double(0.0), ctypes.c_double(0.0)
        self._call(lib().ref_run_pressure, ctypes.byref(it), ctypes.byref(res), ctypes.byref(sec))
        return it.value, res.value, sec.value

    def finish(self) -> dict:
        out = np.zeros(6)
        self._call(lib().ref_run_finish, _dp(out))
        return {"ch_residual": out[0], "momentum_residual": out[1], "momentum_iterations": int(out[2]),
                "ch_iterations": int(out[3]), "mass": out[4], "max_velocity": out[5]}

    def field(self, name: str) -> np.ndarray:
        dims = (ctypes.c_int * 3)()
        ptr = lib().ref_run_field(self.h, name.encode(), dims)
        if not ptr:
            raise KeyError(name)
        nx, ny, g = dims[0], dims[1], dims[2]
        return np.ctypeslib.as_array(ptr, shape=((nx + 2 * g) * (ny + 2 * g),)).reshape(nx + 2 * g, ny + 2 * g)

    def close(self):
        if getattr(self, "h", None):
            lib().ref_run_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def momentum_predictor(mc, repeats: int = 1, timing: list | None = None):
    """The reference's compute_momentum_fluxes + solve_momentum_predictor on a workloads.MomentumCase.
    Returns (u_star, v_star, u_prev, v_prev, iterations, absolute residual) like the oracle's; with a
    timing list, the median seconds of one call over `repeats` is appended."""
    b = RefVelocityBc()
    vbc = mc.vbc
    for k in range(4):
        b.u_type[k], b.u_value[k] = int(vbc.u_type[k]), float(vbc.u_value[k])
        b.v_type[k], b.v_value[k] = int(vbc.v_type[k]), float(vbc.v_value[k])
    b.bottom_wall_velocity_x = float(vbc.bottom_wall_velocity_x)
    b.top_wall_velocity_x = float(vbc.top_wall_velocity_x)
    mp = RefMomentumProblem()
    mp.nx, mp.ny = mc.nx, mc.ny
    mp.lx, mp.ly, mp.dt, mp.re = mc.lx, mc.ly, mc.dt, mc.re
    mp.body_force_x, mp.body_force_y = mc.body_force_x, mc.body_force_y
    mp.periodic_x, mp.periodic_y = int(mc.periodic_x), int(mc.periodic_y)
    mp.step_index, mp.max_iterations, mp.tolerance = mc.step_index, mc.max_iterations, mc.tolerance
    ins = [np.ascontiguousarray(getattr(mc, k), dtype=np.float64)
           for k in ("u", "v", "rho_mid", "eta", "eta_prev", "surface_fx_u", "surface_fy_v")]
    u_prev, v_prev = np.array(mc.u_prev, dtype=np.float64, copy=True), np.array(mc.v_prev, dtype=np.float64, copy=True)
    u_star, v_star = np.zeros_like(u_prev), np.zeros_like(v_prev)
    it, ab = ctypes.c_int(0), ctypes.c_double(0.0)
    seconds = np.zeros(1)
    err = ctypes.create_string_buffer(512)
    rc = lib().ref_momentum_predictor(ctypes.byref(mp), ctypes.byref(b), int(mc.ghost), *[_dp(a) for a in ins],
                                      _dp(u_prev), _dp(v_prev), _dp(u_star), _dp(v_star), ctypes.byref(it),
                                      ctypes.byref(ab), int(repeats), _dp(seconds), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    if timing is not None:
        timing.append(float(seconds[0]))
    return u_star, v_star, u_prev, v_prev, it.value, ab.value


class RefCh:
    """A reference Solver used for its CH linear solve (solve_phase_linear_system_eq25), keeping its
    cached operator and preconditioner across solve() calls as a run does."""

    def __init__(self, cc):
        rc = RefChConfig()
        rc.nx, rc.ny, rc.lx, rc.ly = cc.nx, cc.ny, cc.lx, cc.ly
        rc.periodic_x, rc.periodic_y = int(cc.periodic_x), int(cc.periodic_y)
        rc.cn, rc.pe = cc.cn, cc.pe
        rc.ch_inner_iterations, rc.ch_tolerance = cc.max_iterations, cc.tolerance
        err = ctypes.create_string_buffer(512)
        self.h = lib().ref_ch_create(ctypes.byref(rc), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())

    def solve(self, alpha0, rhs, c0, repeats: int = 1, timing: list | None = None):
        """Returns (c, iterations, iterate residual, equation residual, solver name)."""
        c = np.array(c0, dtype=np.float64, copy=True).reshape(-1)
        it, itr, eq = ctypes.c_int(0), ctypes.c_double(0.0), ctypes.c_double(0.0)
        name = ctypes.create_string_buffer(128)
        sec = np.zeros(1)
        err = ctypes.create_string_buffer(512)
        rc = lib().ref_ch_solve(self.h, float(alpha0), _dp(np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)),
                                _dp(c), ctypes.byref(it), ctypes.byref(itr), ctypes.byref(eq), name, 128, int(repeats),
                                _dp(sec), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        if timing is not None:
            timing.append(float(sec[0]))
        return c, it.value, itr.value, eq.value, name.value.decode()

    def operator(self):
        """(row_ptr, col_idx, values, kind (1 ILDLT, 0 diagonal), lower, D) of the cached operator."""
        L = lib()
        n, nnz, kind = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ref_ch_operator(self.h, ctypes.byref(n), ctypes.byref(nnz), None, None, None, ctypes.byref(kind), None, None)
        rp, ci = np.zeros(n.value + 1, dtype=np.int32), np.zeros(nnz.value, dtype=np.int32)
        val, lower, D = np.zeros(nnz.value), np.zeros(nnz.value), np.zeros(n.value)
        L.ref_ch_operator(self.h, ctypes.byref(n), ctypes.byref(nnz), rp.ctypes.data_as(ip), ci.ctypes.data_as(ip),
                          _dp(val), ctypes.byref(kind), _dp(lower), _dp(D))
        return rp, ci, val, kind.value, lower, D

    def close(self):
        if getattr(self, "h", None):
            lib().ref_ch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
