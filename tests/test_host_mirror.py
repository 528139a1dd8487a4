"""The C++ host mirror of the reference backends (ding_nsch_core_solver_amd/host/pp_backend.*):
scheme mapping and ghost filling on the CPU (ghosts bit-for-bit against the compiled reference's
apply_pressure_bc, or the oracle where the reference is absent), a full solve through
PressureBackend on the GPU (bit-for-bit against the Python binding of the same engine), and the
velocity correction through PressureBackend::correct_velocity (bit-for-bit against the oracle)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "ding_nsch_core_solver_amd")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not E.library_available():
        pytest.skip("engine not built")
    out = str(tmp_path_factory.mktemp("mirror") / "host_mirror_driver")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"), "-I", os.path.join(PKG, "host"),
           os.path.join(REPO, "tests", "host_mirror_driver.cpp"), os.path.join(PKG, "host", "pp_backend.cpp"),
           "-L", PKG, "-lppsolve", f"-Wl,-rpath,{PKG}", "-o", out]
    subprocess.run(cmd, check=True)
    return out


def _cfg(case, max_it=None):
    vals = [case.nx, case.ny, case.lx, case.ly, case.dt, int(case.periodic_x), int(case.periodic_y)]
    for t, v in zip(case.bc_type, case.bc_value):
        vals += [int(t), float(v)]
    vals += [max_it or case.max_iterations, case.tolerance, case.density_ratio, case.step_index]
    return " ".join(repr(v) if isinstance(v, float) else str(v) for v in vals)


BC = {
    "unset": dict(),
    "dirichlet_left": dict(bc_type=(1, 0, 0, 0), bc_value=(3.0, 0.0, 0.0, 0.0)),
    "mixed": dict(bc_type=(2, 1, 2, 1), bc_value=(0.5, -1.0, 2.0, 0.25)),
    "periodic_x": dict(periodic_x=True),
    "periodic_y_dirichlet": dict(periodic_y=True, bc_type=(0, 1, 0, 0), bc_value=(0.0, 1.5, 0.0, 0.0)),
    "periodic_xy": dict(periodic_x=True, periodic_y=True),
}


def test_scheme_mapping(driver):
    out = subprocess.run([driver, "schemes"], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {line.split()[0]: (int(line.split()[1]), int(line.split()[2])) for line in out if line}
    for name, code in E.SCHEME_CODES.items():
        assert got[name][0] == code
        assert got[name][1] == (1 if code in (E.PP_SPLIT_ICPCG, E.PP_SPLIT_ILDLT_PCG) else 0)
    assert got["icpcg"] == (0, 0) and got["petsc_pcg"] == (0, 0)


@pytest.mark.parametrize("name", sorted(BC))
def test_ghost_fill_matches_reference(driver, tmp_path, name):
    case = W.droplet_case(9, 7, seed=1, lx=1.3, ly=0.7, **BC[name])
    field = np.random.default_rng(4).standard_normal((case.nx + 6) * (case.ny + 6))
    path = tmp_path / "f.bin"
    field.tofile(path)
    subprocess.run([driver, "ghost", _cfg(case), str(path)], check=True)
    mine = np.fromfile(path)
    try:
        from oracle import ref
        if not ref.available():
            raise ImportError
        expect = field.copy()
        ref.lib().ref_apply_pressure_bc(ctypes.byref(ref.problem(case)), ref._dp(expect))
    except ImportError:
        from oracle import cpu
        expect = cpu.apply_pressure_bc(case, field.copy())
    assert np.array_equal(mine, expect)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,code,kw", [
    ("hip_icpcg", E.PP_ICPCG, dict()),
    ("hip_ildlt_pcg", E.PP_ILDLT_PCG, dict(bc_type=(0, 0, 0, 1), bc_value=(0, 0, 0, 2.0))),
    ("hip_liu_split_icpcg", E.PP_SPLIT_ICPCG, dict(periodic_x=True)),
    ("hip_paper_split_ildlt_pcg", E.PP_SPLIT_ILDLT_PCG, dict(periodic_y=True)),
])
def test_backend_solve_through_cpp_mirror(driver, engine, oracle_lib, tmp_path, scheme, code, kw):
    case = W.droplet_case(70, 44, seed=3, density_ratio=200.0, **kw)
    if code in (E.PP_SPLIT_ICPCG, E.PP_SPLIT_ILDLT_PCG):
        case.step_index = 2
        case.p_n = np.random.default_rng(5).standard_normal(case.n) * 10.0
        case.p_nm1 = np.random.default_rng(6).standard_normal(case.n) * 10.0
    pn = case.p_n if case.p_n is not None else np.zeros(case.n)
    pm = case.p_nm1 if case.p_nm1 is not None else np.zeros(case.n)
    inp = tmp_path / "in.bin"
    outp = tmp_path / "out.bin"
    np.concatenate([case.rho.reshape(-1), case.u.reshape(-1), case.v.reshape(-1), pn.reshape(-1),
                    pm.reshape(-1)]).tofile(inp)
    res = subprocess.run([driver, "solve", _cfg(case), scheme, str(inp), str(outp)], check=True, capture_output=True,
                         text=True)
    it, rel = res.stdout.split()
    p_ref, rep = engine.solve(case, code)
    assert int(it) == rep["iterations"] and float(rel) == rep["relative_residual"]
    full = np.fromfile(outp)
    g = 3
    interior = full.reshape(case.nx + 2 * g, case.ny + 2 * g)[g:-g, g:-g].reshape(-1)
    assert np.array_equal(interior, p_ref)
    expect = np.zeros_like(full)
    expect.reshape(case.nx + 2 * g, case.ny + 2 * g)[g:-g, g:-g] = p_ref.reshape(case.nx, case.ny)
    oracle_lib.apply_pressure_bc(case, expect)
    assert np.array_equal(full, expect)


def _vel_cfg(vbc):
    vals = []
    for types, values in ((vbc.u_type, vbc.u_value), (vbc.v_type, vbc.v_value)):
        for t, v in zip(types, values):
            vals += [str(int(t)), repr(float(v))]
    vals += [repr(float(vbc.bottom_wall_velocity_x)), repr(float(vbc.top_wall_velocity_x))]
    return " ".join(vals)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,vname,kw,step", [
    ("hip_icpcg", "lid_driven", dict(), 0),
    ("hip_ildlt_pcg", "dirichlet_values", dict(bc_type=(0, 0, 0, 1), bc_value=(0, 0, 0, 2.0)), 0),
    ("hip_liu_split_icpcg", "neumann_outflow", dict(periodic_x=True), 3),
    ("hip_paper_split_ildlt_pcg", "no_slip_defaults", dict(periodic_y=True), 0),
])
def test_velocity_correction_through_cpp_mirror(driver, oracle_lib, tmp_path, scheme, vname, kw, step):
    """PressureBackend::correct_velocity against the oracle's restatement of
    Solver::apply_pressure_velocity_correction (operators.cpp:480-534), bit for bit."""
    from tests.test_velocity_oracle import G, VELOCITY_BC, fields
    case = W.droplet_case(21, 13, seed=4, density_ratio=30.0, lx=0.8, ly=0.6, step_index=step, **kw)
    f = fields(case, seed=17 + step)
    vbc = VELOCITY_BC[vname]
    order = ("rho_mid", "u_star", "v_star", "p_corr", "u", "v", "pressure", "pressure_prev")
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([f[k] for k in order]).tofile(inp)
    res = subprocess.run([driver, "correct", _cfg(case), _vel_cfg(vbc), scheme, str(inp), str(outp)], check=True,
                         capture_output=True, text=True)
    pre, post = (float(x) for x in res.stdout.split())
    split = E.SCHEME_CODES[scheme] in (E.PP_SPLIT_ICPCG, E.PP_SPLIT_ILDLT_PCG)
    u, v, p, prev, speeds = oracle_lib.velocity_correction(case, vbc, split, G, **f)
    assert np.array_equal(np.fromfile(outp), np.concatenate([u, v, p, prev]))
    assert (pre, post) == tuple(speeds)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,code", [("hip_icpcg", E.PP_ICPCG), ("hip_split_ildlt_pcg", E.PP_SPLIT_ILDLT_PCG)])
def test_monitor_through_cpp_mirror(driver, engine, tmp_path, scheme, code):
    """PressureBackend::solve with a CGMonitor replays the engine's PCG log in order, bit for bit."""
    case = W.droplet_case(40, 36, seed=8, density_ratio=300.0, step_index=1)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    z = np.zeros(case.n)
    np.concatenate([case.rho.reshape(-1), case.u.reshape(-1), case.v.reshape(-1), z, z]).tofile(inp)
    res = subprocess.run([driver, "monitor", _cfg(case), scheme, str(inp), str(outp)], check=True,
                         capture_output=True, text=True)
    lines = res.stdout.split("\n")
    it = int(lines[0].split()[0])
    rows = [line.split() for line in lines[1:] if line]
    case.p_n, case.p_nm1 = z, z
    p, rep, rec = engine.solve_monitored(case, code)
    assert it == rep["iterations"] and len(rows) == len(rec) == it + 1
    for row, r in zip(rows, rec):
        assert int(row[0]) == r["iteration"]
        assert [float(x) for x in row[1:]] == [r["alpha"], r["beta"], r["absolute_residual"], r["relative_residual"],
                                               r["iterate_residual"], r["preconditioned_residual"]]


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,code,kw,step", [
    ("hip_icpcg", E.PP_ICPCG, dict(bc_type=(0, 1, 0, 0), bc_value=(0, 0.5, 0, 0)), 0),
    ("hip_liu_split_icpcg", E.PP_SPLIT_ICPCG, dict(periodic_y=True), 2),
])
def test_pressure_step_through_cpp_mirror(driver, engine, tmp_path, scheme, code, kw, step):
    """PressureBackend::pressure_step (solve + correction, one upload and download) equals the
    engine's pp_pressure_step bit for bit: pressure_correction with ghosts, u, v, both pressures."""
    from tests.test_velocity_oracle import G, VELOCITY_BC, fields
    case = W.droplet_case(30, 26, seed=6, density_ratio=120.0, step_index=step, **kw)
    f = fields(case, seed=51 + step)
    f["rho_mid"] = np.pad(case.rho.reshape(case.nx, case.ny), G, mode="edge").reshape(-1)
    vbc = VELOCITY_BC["lid_driven"]
    order = ("rho_mid", "u_star", "v_star", "p_corr", "u", "v", "pressure", "pressure_prev")
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([f[k] for k in order]).tofile(inp)
    res = subprocess.run([driver, "step", _cfg(case), _vel_cfg(vbc), scheme, str(inp), str(outp)], check=True,
                         capture_output=True, text=True)
    it, rel, pre, post = res.stdout.split()
    p_corr, u, v, p, prev, speeds, rep = engine.pressure_step(case, vbc, f["rho_mid"], f["u_star"], f["v_star"],
                                                              f["u"], f["v"], f["pressure"], f["pressure_prev"],
                                                              scheme=code, ghost=G)
    assert int(it) == rep["iterations"] and float(rel) == rep["relative_residual"]
    assert (float(pre), float(post)) == speeds
    assert np.array_equal(np.fromfile(outp), np.concatenate([p_corr, u, v, p, prev]))


def _bicg_input(path, s, max_it, tol, rel):
    x0 = np.zeros(s.n)
    buf = np.concatenate([[s.n, len(s.values), max_it, tol, float(rel)], s.row_ptr, s.col_idx, s.values, s.rhs, x0])
    buf.astype(np.float64).tofile(path)


@pytest.mark.gpu
@pytest.mark.parametrize("rel", [False, True])
def test_bicgstab_through_cpp_mirror(driver, oracle_lib, tmp_path, rel):
    """PressureBackend::solve_bicgstab (the reference's solve_preconditioned_bicgstab signature) on
    a momentum-shaped system: report and x bit for bit equal to the oracle in the engine's order."""
    s = W.momentum_system(48, 40, seed=3)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    _bicg_input(inp, s, 50, 1e-9, rel)
    res = subprocess.run([driver, "bicgstab", str(inp), str(outp)], check=True, capture_output=True, text=True)
    it, ab, rr, itr = res.stdout.split()
    xo, ito, abo, relo, itro = oracle_lib.bicgstab(s, max_iterations=50, tolerance=1e-9, use_relative=rel,
                                                   reduction=oracle_lib.CHUNKED)
    assert int(it) == ito and float(ab) == abo and float(rr) == relo and float(itr) == itro
    assert np.array_equal(np.fromfile(outp), xo)


@pytest.mark.gpu
def test_bicgstab_mirror_breakdown_text(driver, tmp_path):
    s = W.CsrSystem(np.array([0, 2, 4], dtype=np.int32), np.array([0, 1, 0, 1], dtype=np.int32),
                    np.array([2.0, 1.0, 1.0, 3.0]), np.array([np.nan, 1.0]))
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    _bicg_input(inp, s, 50, 1e-8, False)
    res = subprocess.run([driver, "bicgstab", str(inp), str(outp)], capture_output=True, text=True)
    assert res.returncode == 1 and res.stdout.strip() == "error bicgstab breakdown: invalid rho"
