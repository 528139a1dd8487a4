"""ctypes access to the CPU oracle (oracle/liboracle.so, built from oracle/pp_oracle.cpp).

Checker only: imported by tests/, the smoke check and bench.py's CPU-baseline leg.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

SCHEMES = {"icpcg": 1, "ildlt_pcg": 2, "split_icpcg": 3, "split_ildlt_pcg": 4}
SEQUENTIAL, PAIRWISE, CANONICAL = 0, 1, 2
CHUNKED = 3  # BiCGSTAB dot products in the engine's order (512-row chunk trees, then a tree over the chunks)

# One PCG monitor record (the scalar fields of CGMonitorData, ch_sparse_krylov.hpp:47-58).
MONITOR_DTYPE = np.dtype([("iteration", "<i4"), ("pad", "<i4"), ("alpha", "<f8"), ("beta", "<f8"),
                          ("absolute_residual", "<f8"), ("relative_residual", "<f8"), ("iterate_residual", "<f8"),
                          ("preconditioned_residual", "<f8")])


class OrProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("dt", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int),
        ("bc_type", ctypes.c_int * 4), ("bc_value", ctypes.c_double * 4),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
        ("density_ratio", ctypes.c_double), ("step_index", ctypes.c_int),
    ]


class OrVelocityBc(ctypes.Structure):
    _fields_ = [
        ("u_type", ctypes.c_int * 4), ("u_value", ctypes.c_double * 4),
        ("v_type", ctypes.c_int * 4), ("v_value", ctypes.c_double * 4),
        ("bottom_wall_velocity_x", ctypes.c_double), ("top_wall_velocity_x", ctypes.c_double),
    ]


class OrChProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("lx", ctypes.c_double), ("ly", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int),
        ("alpha0", ctypes.c_double), ("beta1", ctypes.c_double), ("beta2", ctypes.c_double),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
    ]


# CH solve reductions in the engine's orders: the ILDLT PCG's 512-thread tree over natural rows, the
# diagonal CG's 512-row chunk trees (BiCGSTAB's order), each attempt by its preconditioner
CH_ENGINE = 2


def ch_problem(cc) -> OrChProblem:
    cp = OrChProblem()
    cp.nx, cp.ny, cp.lx, cp.ly = cc.nx, cc.ny, cc.lx, cc.ly
    cp.periodic_x, cp.periodic_y = int(cc.periodic_x), int(cc.periodic_y)
    cp.alpha0, cp.beta1, cp.beta2 = cc.alpha0, cc.beta1, cc.beta2
    cp.max_iterations, cp.tolerance = cc.max_iterations, cc.tolerance
    return cp


class OrMomentumProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("dt", ctypes.c_double), ("re", ctypes.c_double),
        ("body_force_x", ctypes.c_double), ("body_force_y", ctypes.c_double),
        ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int), ("step_index", ctypes.c_int),
        ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double),
    ]


def momentum_problem(mc, struct=None):
    mp = (struct or OrMomentumProblem)()
    mp.nx, mp.ny = mc.nx, mc.ny
    mp.lx, mp.ly, mp.dt, mp.re = mc.lx, mc.ly, mc.dt, mc.re
    mp.body_force_x, mp.body_force_y = mc.body_force_x, mc.body_force_y
    mp.periodic_x, mp.periodic_y = int(mc.periodic_x), int(mc.periodic_y)
    mp.step_index, mp.max_iterations, mp.tolerance = mc.step_index, mc.max_iterations, mc.tolerance
    return mp


def available() -> bool:
    return os.path.exists(LIB_PATH)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"oracle not built: {LIB_PATH} (make -C oracle oracle)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_pressure_solve.argtypes = [ctypes.c_int, ctypes.POINTER(OrProblem), dp, dp, dp, dp, dp, ctypes.c_int,
                                        dp, ip, dp, ctypes.c_char_p, ctypes.c_int]
        L.or_pressure_solve.restype = ctypes.c_int
        L.or_pressure_solve_monitored.argtypes = [ctypes.c_int, ctypes.POINTER(OrProblem), dp, dp, dp, dp, dp,
                                                  ctypes.c_int, dp, ip, dp, ctypes.c_void_p, ctypes.c_int, ip,
                                                  ctypes.c_char_p, ctypes.c_int]
        L.or_pressure_solve_monitored.restype = ctypes.c_int
        L.or_sys_pcg_monitored.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp, ip, dp, dp,
                                           ctypes.c_void_p, ctypes.c_int, ip, ctypes.c_char_p, ctypes.c_int]
        L.or_sys_pcg_monitored.restype = ctypes.c_int
        L.or_sys_build.argtypes = [ctypes.POINTER(OrProblem), ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_int,
                                   ctypes.c_char_p, ctypes.c_int]
        L.or_sys_build.restype = ctypes.c_void_p
        L.or_sys_free.argtypes = [ctypes.c_void_p]
        L.or_sys_sizes.argtypes = [ctypes.c_void_p, ip, ip, ip]
        L.or_sys_csr.argtypes = [ctypes.c_void_p, ip, ip, dp, dp]
        L.or_sys_factor.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
        L.or_sys_factor.restype = ctypes.c_int
        L.or_sys_apply_matrix.argtypes = [ctypes.c_void_p, dp, dp]
        L.or_sys_apply_precond.argtypes = [ctypes.c_void_p, dp, dp]
        L.or_sys_pcg.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp, ip, dp, dp,
                                 ctypes.c_char_p, ctypes.c_int]
        L.or_sys_pcg.restype = ctypes.c_int
        L.or_apply_pressure_bc.argtypes = [ctypes.POINTER(OrProblem), dp]
        L.or_pressure_ghosts.argtypes = [ctypes.POINTER(OrProblem), ctypes.c_int, dp]
        L.or_solve_coeffs.argtypes = [ctypes.c_int, ctypes.POINTER(OrProblem), dp, dp, dp, ctypes.c_int, dp, ip, dp,
                                      ctypes.c_char_p, ctypes.c_int]
        L.or_solve_coeffs.restype = ctypes.c_int
        L.or_adjoint.argtypes = [ctypes.c_int, ctypes.POINTER(OrProblem), dp, dp, dp, dp, ctypes.c_int, dp, dp, dp,
                                 ip, dp, ctypes.c_char_p, ctypes.c_int]
        L.or_adjoint.restype = ctypes.c_int
        L.or_velocity_correction.argtypes = [ctypes.c_int, ctypes.POINTER(OrProblem), ctypes.POINTER(OrVelocityBc),
                                             ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_velocity_correction.restype = None
        L.or_velocity_correction_order.argtypes = [ctypes.c_int, ctypes.POINTER(OrProblem),
                                                   ctypes.POINTER(OrVelocityBc), ctypes.c_int, dp, dp, dp, dp, dp, dp,
                                                   dp, dp, dp, ctypes.c_int]
        L.or_velocity_correction_order.restype = None
        L.or_bicgstab.argtypes = [ctypes.c_int, ip, ip, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_int, ip, dp, dp, dp, ctypes.c_char_p, ctypes.c_int]
        L.or_bicgstab.restype = ctypes.c_int
        chp = ctypes.POINTER(OrChProblem)
        L.or_ch_operator.argtypes = [chp, ip, ip, ip, ip, dp, ctypes.c_char_p, ctypes.c_int]
        L.or_ch_operator.restype = ctypes.c_int
        L.or_ch_ildlt.argtypes = [chp, dp, dp, dp]
        L.or_ch_ildlt.restype = ctypes.c_int
        L.or_ch_solve.argtypes = [chp, dp, dp, ctypes.c_int, ctypes.c_int, ip, dp, dp, dp, ip, ctypes.c_char_p,
                                  ctypes.c_int]
        L.or_ch_solve.restype = ctypes.c_int
        mp_ = ctypes.POINTER(OrMomentumProblem)
        vb_ = ctypes.POINTER(OrVelocityBc)
        L.or_momentum_system.argtypes = [mp_, vb_, ctypes.c_int] + [dp] * 9 + [ip, ip, ip, ip, dp, dp, dp,
                                                                            ctypes.c_char_p, ctypes.c_int]
        L.or_momentum_system.restype = ctypes.c_int
        L.or_momentum_predictor.argtypes = [mp_, vb_, ctypes.c_int] + [dp] * 11 + [ctypes.c_int, ip, dp,
                                                                                 ctypes.c_char_p, ctypes.c_int]
        L.or_momentum_predictor.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def problem(case) -> OrProblem:
    pb = OrProblem()
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


def pressure_solve(scheme: str, case, reduction: int = SEQUENTIAL):
    """Oracle pressure correction. Returns (p interior (nx*ny,), iterations, relative residual)."""
    L = lib()
    p = np.zeros(case.n)
    it = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = L.or_pressure_solve(SCHEMES[scheme], ctypes.byref(problem(case)), _dp(case.rho), _dp(case.u), _dp(case.v),
                             _dp(case.p_n), _dp(case.p_nm1), reduction, _dp(p), ctypes.byref(it), ctypes.byref(res),
                             err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return p, it.value, res.value


def pressure_solve_monitored(scheme: str, case, reduction: int = SEQUENTIAL, capacity: int | None = None):
    """pressure_solve with the PCG monitor log. Returns (p, iterations, relative residual, records)
    where records holds one MONITOR_DTYPE entry per monitor call (up to capacity)."""
    cap = case.max_iterations + 1 if capacity is None else capacity
    records = np.zeros(max(cap, 1), dtype=MONITOR_DTYPE)
    p = np.zeros(case.n)
    it, res, count = ctypes.c_int(0), ctypes.c_double(0.0), ctypes.c_int(0)
    err = ctypes.create_string_buffer(512)
    rc = lib().or_pressure_solve_monitored(SCHEMES[scheme], ctypes.byref(problem(case)), _dp(case.rho), _dp(case.u),
                                           _dp(case.v), _dp(case.p_n), _dp(case.p_nm1), reduction, _dp(p),
                                           ctypes.byref(it), ctypes.byref(res), records.ctypes.data, cap,
                                           ctypes.byref(count), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return p, it.value, res.value, records[:min(count.value, cap)]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def solve_coeffs(scheme: str, case, ku, kv, src, reduction: int = SEQUENTIAL):
    """Coefficient-level solve: c_f = k_f / h^2 on faces, right-hand side src (+ boundary terms).
    Returns (p (nx*ny,), iterations, relative residual)."""
    ku, kv, src = _f64(ku), _f64(kv), _f64(src)
    assert ku.size == (case.nx + 1) * case.ny and kv.size == case.nx * (case.ny + 1) and src.size == case.n
    p = np.zeros(case.n)
    it, res = ctypes.c_int(0), ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = lib().or_solve_coeffs(SCHEMES[scheme], ctypes.byref(problem(case)), _dp(ku), _dp(kv), _dp(src), reduction,
                               _dp(p), ctypes.byref(it), ctypes.byref(res), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return p, it.value, res.value


def adjoint(scheme: str, case, ku, kv, p, dLdp, reduction: int = SEQUENTIAL):
    """FP64 adjoint of solve_coeffs. Returns (dL/dsrc, dL/dk_u, dL/dk_v, iterations, residual)."""
    ku, kv, p, dLdp = _f64(ku), _f64(kv), _f64(p), _f64(dLdp)
    lam = np.zeros(case.n)
    dku = np.zeros((case.nx + 1) * case.ny)
    dkv = np.zeros(case.nx * (case.ny + 1))
    it, res = ctypes.c_int(0), ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = lib().or_adjoint(SCHEMES[scheme], ctypes.byref(problem(case)), _dp(ku), _dp(kv), _dp(p), _dp(dLdp), reduction,
                          _dp(lam), _dp(dku), _dp(dkv), ctypes.byref(it), ctypes.byref(res), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return lam, dku, dkv, it.value, res.value


def apply_pressure_bc(case, field_ghosted: np.ndarray) -> np.ndarray:
    """Ghost fill (3 layers) of a pressure-like Field2D buffer, in place; returns it."""
    assert field_ghosted.size == (case.nx + 6) * (case.ny + 6)
    lib().or_apply_pressure_bc(ctypes.byref(problem(case)), _dp(field_ghosted))
    return field_ghosted


def pressure_ghosts(case, field_ghosted: np.ndarray, ghost: int) -> np.ndarray:
    """Ghost fill of a pressure-like Field2D buffer with `ghost` layers, in place; returns it."""
    assert field_ghosted.size == (case.nx + 2 * ghost) * (case.ny + 2 * ghost)
    lib().or_pressure_ghosts(ctypes.byref(problem(case)), int(ghost), _dp(field_ghosted))
    return field_ghosted


class OracleSystem:
    """The oracle's assembled CSR system for a case, with factor and operator applications."""

    def __init__(self, case, use_split=False, reduction=SEQUENTIAL):
        L = lib()
        err = ctypes.create_string_buffer(512)
        self.h = L.or_sys_build(ctypes.byref(problem(case)), int(use_split), _dp(case.rho), _dp(case.u), _dp(case.v),
                                _dp(case.p_n), _dp(case.p_nm1), reduction, err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        n, nnz, hd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.or_sys_sizes(self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(hd))
        self.n, self.nnz, self.has_dirichlet = n.value, nnz.value, bool(hd.value)
        self.row_ptr = np.zeros(self.n + 1, dtype=np.int32)
        self.col_idx = np.zeros(self.nnz, dtype=np.int32)
        self.values = np.zeros(self.nnz)
        self.rhs = np.zeros(self.n)
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_sys_csr(self.h, self.row_ptr.ctypes.data_as(ip), self.col_idx.ctypes.data_as(ip), _dp(self.values),
                     _dp(self.rhs))

    def dense(self) -> np.ndarray:
        A = np.zeros((self.n, self.n))
        for r in range(self.n):
            for q in range(self.row_ptr[r], self.row_ptr[r + 1]):
                A[r, self.col_idx[q]] = self.values[q]
        return A

    def factor(self, kind=1):
        lower = np.zeros(self.nnz)
        diag = np.zeros(self.n)
        fl = ctypes.c_double(0.0)
        ok = lib().or_sys_factor(self.h, kind, _dp(lower), _dp(diag), ctypes.byref(fl))
        return bool(ok), lower, diag, fl.value

    def apply_matrix(self, x):
        y = np.zeros(self.n)
        lib().or_sys_apply_matrix(self.h, _dp(np.ascontiguousarray(x, dtype=np.float64)), _dp(y))
        return y

    def apply_precond(self, r):
        z = np.zeros(self.n)
        lib().or_sys_apply_precond(self.h, _dp(np.ascontiguousarray(r, dtype=np.float64)), _dp(z))
        return z

    def pcg(self, max_iterations, tolerance):
        x = np.zeros(self.n)
        it, rel, ab = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        err = ctypes.create_string_buffer(512)
        rc = lib().or_sys_pcg(self.h, max_iterations, tolerance, _dp(x), ctypes.byref(it), ctypes.byref(rel),
                              ctypes.byref(ab), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return x, it.value, rel.value, ab.value

    def pcg_monitored(self, max_iterations, tolerance, capacity=None):
        """pcg with the monitor log; returns (x, iterations, rel, abs, records, error text or None):
        a breakdown is returned as its message with the records logged before it."""
        cap = max_iterations + 1 if capacity is None else capacity
        records = np.zeros(max(cap, 1), dtype=MONITOR_DTYPE)
        x = np.zeros(self.n)
        it, rel, ab, count = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        err = ctypes.create_string_buffer(512)
        rc = lib().or_sys_pcg_monitored(self.h, max_iterations, tolerance, _dp(x), ctypes.byref(it), ctypes.byref(rel),
                                        ctypes.byref(ab), records.ctypes.data, cap, ctypes.byref(count), err, 512)
        return (x, it.value, rel.value, ab.value, records[:min(count.value, cap)],
                err.value.decode() if rc != 0 else None)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sys_free(self.h)
            self.h = None


def velocity_bc(vbc) -> OrVelocityBc:
    out = OrVelocityBc()
    for k in range(4):
        out.u_type[k], out.u_value[k] = int(vbc.u_type[k]), float(vbc.u_value[k])
        out.v_type[k], out.v_value[k] = int(vbc.v_type[k]), float(vbc.v_value[k])
    out.bottom_wall_velocity_x = float(vbc.bottom_wall_velocity_x)
    out.top_wall_velocity_x = float(vbc.top_wall_velocity_x)
    return out


def velocity_correction(case, vbc, split: bool, ghost: int, rho_mid, u_star, v_star, p_corr, u, v, pressure,
                        pressure_prev, tree_mean: bool = False):
    """Solver::apply_pressure_velocity_correction on ghosted Field2D buffers (see pp_oracle.cpp).
    Returns the new (u, v, pressure, pressure_prev, (speed_pre, speed_post)); the inputs are not
    modified. tree_mean: the pressure mean in the engine's tree order instead of the reference's
    sequential sum."""
    u, v, pressure, prev = (np.array(a, dtype=np.float64, copy=True) for a in (u, v, pressure, pressure_prev))
    speeds = np.zeros(2)
    args = [_f64(a) for a in (rho_mid, u_star, v_star, p_corr)]
    lib().or_velocity_correction_order(int(split), ctypes.byref(problem(case)), ctypes.byref(velocity_bc(vbc)),
                                       int(ghost), *[_dp(a) for a in args], _dp(u), _dp(v), _dp(pressure), _dp(prev),
                                       _dp(speeds), int(bool(tree_mean)))
    return u, v, pressure, prev, (float(speeds[0]), float(speeds[1]))


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def bicgstab(system, x0=None, max_iterations: int = 200, tolerance: float = 1e-8, use_relative: bool = False,
             reduction: int = SEQUENTIAL):
    """BiCGSTAB with the diagonal preconditioner (ch_sparse_krylov.hpp:421-431, 839-963) on a CSR
    system (row_ptr, col_idx, values, rhs). Returns (x, iterations, absolute, relative, iterate
    residual); raises RuntimeError with the reference's text on a breakdown."""
    L = lib()
    rp, ci, val, rhs = system.row_ptr, system.col_idx, system.values, system.rhs
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = ctypes.c_int(0)
    ab, rel, itr = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = L.or_bicgstab(n, _ip(rp), _ip(ci), _dp(val), _dp(rhs), _dp(x), int(max_iterations), float(tolerance),
                       int(bool(use_relative)), int(reduction), ctypes.byref(it), ctypes.byref(ab), ctypes.byref(rel),
                       ctypes.byref(itr), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return x, it.value, ab.value, rel.value, itr.value


def _momentum_inputs(mc):
    return [_f64(getattr(mc, k)) for k in ("u", "v", "rho_mid", "eta", "eta_prev", "surface_fx_u", "surface_fy_v")]


def momentum_system(mc):
    """The momentum predictor's assembled system (momentum.cpp:157-468, see pp_oracle.cpp):
    a workloads.CsrSystem-like object with row_ptr, col_idx, values, rhs and x0 (the initial guess)."""
    import types
    L = lib()
    ins = _momentum_inputs(mc)
    prev = [_f64(mc.u_prev), _f64(mc.v_prev)]
    n, nnz = ctypes.c_int(0), ctypes.c_int(0)
    err = ctypes.create_string_buffer(512)
    args = [ctypes.byref(momentum_problem(mc)), ctypes.byref(velocity_bc(mc.vbc)), int(mc.ghost)]
    args += [_dp(a) for a in ins + prev]
    rc = L.or_momentum_system(*args, ctypes.byref(n), ctypes.byref(nnz), None, None, None, None, None, err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    rp = np.zeros(n.value + 1, dtype=np.int32)
    ci = np.zeros(nnz.value, dtype=np.int32)
    val, rhs, x0 = np.zeros(nnz.value), np.zeros(n.value), np.zeros(n.value)
    rc = L.or_momentum_system(*args, ctypes.byref(n), ctypes.byref(nnz), _ip(rp), _ip(ci), _dp(val), _dp(rhs),
                              _dp(x0), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return types.SimpleNamespace(n=n.value, row_ptr=rp, col_idx=ci, values=val, rhs=rhs, x0=x0)


def momentum_predictor(mc, reduction: int = SEQUENTIAL):
    """The whole predictor step. Returns (u_star, v_star, u_prev, v_prev, iterations, absolute
    residual): the ghosted predictor fields and this step's convection fields."""
    L = lib()
    u_prev, v_prev = np.array(mc.u_prev, dtype=np.float64, copy=True), np.array(mc.v_prev, dtype=np.float64, copy=True)
    u_star, v_star = np.zeros_like(u_prev), np.zeros_like(v_prev)
    it, ab = ctypes.c_int(0), ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = L.or_momentum_predictor(ctypes.byref(momentum_problem(mc)), ctypes.byref(velocity_bc(mc.vbc)), int(mc.ghost),
                                 *[_dp(a) for a in _momentum_inputs(mc)], _dp(u_prev), _dp(v_prev), _dp(u_star),
                                 _dp(v_star), int(reduction), ctypes.byref(it), ctypes.byref(ab), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return u_star, v_star, u_prev, v_prev, it.value, ab.value


def ch_operator(cc):
    """The CH operator alpha0 I - beta1 Lap + beta2 Lap^2 as CSR (row_ptr, col_idx, values)."""
    import types
    L = lib()
    n, nnz = ctypes.c_int(0), ctypes.c_int(0)
    err = ctypes.create_string_buffer(512)
    cp = ch_problem(cc)
    if L.or_ch_operator(ctypes.byref(cp), ctypes.byref(n), ctypes.byref(nnz), None, None, None, err, 512):
        raise RuntimeError(err.value.decode())
    rp, ci, val = np.zeros(n.value + 1, dtype=np.int32), np.zeros(nnz.value, dtype=np.int32), np.zeros(nnz.value)
    if L.or_ch_operator(ctypes.byref(cp), ctypes.byref(n), ctypes.byref(nnz), _ip(rp), _ip(ci), _dp(val), err, 512):
        raise RuntimeError(err.value.decode())
    return types.SimpleNamespace(n=n.value, row_ptr=rp, col_idx=ci, values=val)


def ch_ildlt(cc):
    """(ok, lower (nnz, aligned with the operator's CSR), D (n), floor) of the ILDLT(0) factor."""
    A = ch_operator(cc)
    lower, D, fl = np.zeros(len(A.values)), np.zeros(A.n), ctypes.c_double(0.0)
    ok = lib().or_ch_ildlt(ctypes.byref(ch_problem(cc)), _dp(lower), _dp(D), ctypes.byref(fl))
    return bool(ok), lower, D, fl.value


def ch_solve(cc, rhs, c0, reduction: int = SEQUENTIAL, precond: int = -1):
    """The CH linear solve (phase_field.cpp:437-503). Returns (c, iterations, absolute, relative,
    iterate residual, preconditioner used: 1 ILDLT, 0 diagonal, 2 diagonal after a fallback)."""
    L = lib()
    c = np.array(c0, dtype=np.float64, copy=True).reshape(-1)
    it, pre = ctypes.c_int(0), ctypes.c_int(0)
    ab, rel, itr = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = L.or_ch_solve(ctypes.byref(ch_problem(cc)), _dp(_f64(rhs)), _dp(c), int(precond), int(reduction),
                       ctypes.byref(it), ctypes.byref(ab), ctypes.byref(rel), ctypes.byref(itr), ctypes.byref(pre),
                       err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return c, it.value, ab.value, rel.value, itr.value, pre.value
