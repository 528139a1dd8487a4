"""The C ABI library builds for gfx950, loads on the host and exports every declared entry point.

No compute call is made here (no GPU in the CPU suite)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "pp_solver.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("pp_create", "pp_destroy", "pp_last_error", "pp_solve", "pp_solve_device"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from ding_nsch_core_solver_amd import engine
    if not engine.library_available():
        pytest.skip("engine not built (python -m ding_nsch_core_solver_amd.build)")
    lib = ctypes.CDLL(engine.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} missing from {engine.LIB_PATH}"
    assert set(engine.EXPORTS) <= set(declared_functions())


def test_library_is_built_for_gfx950():
    """The offload bundle embedded by hipcc names its target triple and processor."""
    from ding_nsch_core_solver_amd import engine
    if not engine.library_available():
        pytest.skip("engine not built")
    data = open(engine.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_one_hip_runtime_when_engine_loads_before_torch():
    """Loading the engine and then importing torch must map a single libamdhip64/libhsa-runtime64:
    a second runtime in the process sees no GPU and cannot use torch's streams."""
    pytest.importorskip("torch")
    from ding_nsch_core_solver_amd import engine
    if not engine.library_available():
        pytest.skip("engine not built")
    code = ("from ding_nsch_core_solver_amd import engine as E; E.lib(); import torch; "
            "m = open('/proc/self/maps').read().splitlines(); "
            "print('\\n'.join(sorted({l.split()[-1] for l in m if 'libamdhip64' in l or 'libhsa-runtime64' in l})))")
    out = subprocess.run([sys.executable, "-c", code], cwd=REPO, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    libs = out.stdout.split()
    assert sum("libamdhip64" in p for p in libs) == 1, libs
    assert sum("libhsa-runtime64" in p for p in libs) == 1, libs


def test_report_struct_layout_matches_header():
    """The ctypes mirror of pp_report must have the header's field order and sizes."""
    from ding_nsch_core_solver_amd import engine
    names = [f[0] for f in engine.PPReport._fields_]
    assert names == ["iterations", "outer_cycles", "relative_residual", "absolute_residual", "iterate_residual",
                     "kernel_ms", "phase_ms", "status", "message"]
    text = open(HEADER).read()
    body = text[text.index("typedef struct pp_report"):text.index("} pp_report;")]
    for n in names:
        assert re.search(rf"\b{n}\b", body), n


def test_velocity_bc_struct_layout_matches_header():
    from ding_nsch_core_solver_amd import engine
    names = [f[0] for f in engine.PPVelocityBC._fields_]
    text = open(HEADER).read()
    body = text[text.index("typedef struct pp_velocity_bc"):text.index("} pp_velocity_bc;")]
    fields = re.findall(r"\b(?:int|double)\s+([a-z_]+)", body)
    assert names == fields
    assert ctypes.sizeof(engine.PPVelocityBC) == 4 * 4 + 4 * 8 + 4 * 4 + 4 * 8 + 2 * 8


def test_monitor_record_layout_matches_header():
    """MONITOR_DTYPE (engine and oracle) has pp_monitor_record's fields in order: 56 bytes, no padding
    beyond the explicit pad after the iteration index."""
    from ding_nsch_core_solver_amd import engine
    from oracle import cpu
    text = open(HEADER).read()
    body = text[text.index("typedef struct pp_monitor_record"):text.index("} pp_monitor_record;")]
    fields = re.findall(r"\b(?:int|double)\s+([a-z_]+)", body)
    assert list(engine.MONITOR_DTYPE.names) == fields
    assert engine.MONITOR_DTYPE == cpu.MONITOR_DTYPE
    assert engine.MONITOR_DTYPE.itemsize == 2 * 4 + 6 * 8
    assert engine.MONITOR_DTYPE.fields["alpha"][1] == 8


def test_field_size_check_rejects_wrong_buffers():
    """The host Field2D wrappers refuse buffers whose size does not match the grid and ghost count
    (the library would copy past their end)."""
    import numpy as np
    from ding_nsch_core_solver_amd import engine, workloads as W
    case = W.droplet_case(6, 5, seed=0)
    g = 2
    ok = dict(rho_mid=np.zeros(10 * 9), u_star=np.zeros(11 * 9), v_star=np.zeros(10 * 10), p_corr=np.zeros(10 * 9),
              u=np.zeros(11 * 9), v=np.zeros(10 * 10), pressure=np.zeros(10 * 9), pressure_prev=np.zeros(10 * 9))
    engine.Engine._check_field_sizes(case, g, **ok)
    bad = dict(ok, v=np.zeros(11 * 9))
    with pytest.raises(ValueError, match="v: 99 values, expected 100"):
        engine.Engine._check_field_sizes(case, g, **bad)


def test_status_codes_match_header():
    """engine.py's PP_* status constants are the header's pp_status values."""
    import re
    from ding_nsch_core_solver_amd import engine as E
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                             "pp_solver.h")).read()
    body = text[text.index("enum pp_status"):]
    body = body[:body.index("};")]
    codes = dict(re.findall(r"(PP_\w+)\s*=\s*(\d+)", body))
    assert codes and all(getattr(E, k) == int(v) for k, v in codes.items()), codes


@pytest.mark.parametrize("nx,ny,periodic_x,text", [
    (600, 8, False, "grid size outside nx 1..512, ny 1..65535"),
    (40, 70000, False, "grid size outside nx 1..512, ny 1..65535"),
    (40, 1100, True, "a grid periodic in x takes ny 1..1024"),
    (40, 1100, False, "the allocation-free solve takes ny 1..1024"),
])
def test_grid_limits_are_checked_before_any_device_call(nx, ny, periodic_x, text):
    """The problem checks run on the host before any HIP call, so they answer without a GPU: nx up
    to 512 (one row per lane of the 8 waves), ny up to 65535, 1024 where the grid is periodic in x
    (its wrap line sits in LDS) and in the allocation-free entry (its level table is sized so)."""
    from ding_nsch_core_solver_amd import engine as E
    from ding_nsch_core_solver_amd import workloads as W
    if not E.library_available():
        pytest.skip("engine not built (python -m ding_nsch_core_solver_amd.build)")
    L = E.lib()
    case = W.PressureCase(nx=nx, ny=ny, lx=1.0, ly=1.0, dt=1e-4, periodic_x=periodic_x, periodic_y=False,
                          tolerance=1e-8, max_iterations=10, density_ratio=1.0, rho=None, u=None, v=None)
    pb = E.make_problem(case)
    rep = E.PPReport()
    rc = L.pp_solve_device_noalloc(ctypes.byref(pb), None, None, None, None, None, None, ctypes.byref(rep), None)
    assert rc == E.PP_ERR_ARG
    assert rep.message.decode() == text
