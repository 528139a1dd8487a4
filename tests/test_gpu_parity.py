"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Bar, per case and scheme:
* bit-for-bit equality with the oracle run in the engine's reduction order (every element-wise
  operation of assembly, factorization, sweeps, SpMV and updates is the reference's own, so only the
  summation order of reductions can differ, and the oracle replays that order);
* the same iteration count as the oracle in the reference's sequential order (which is itself
  bit-identical to the compiled reference, see test_oracle.py), except where the final residual
  sits within rounding noise of the tolerance (one iteration apart, see _check_iterations);
* pressure within the case's reorder bound of the reference-order result. Changing only the
  summation order of the reference's reductions moves its own dr=1000 pressure by 1e-11 to 6e-8
  (max-abs relative), so no parallel summation can be tighter than that. Each (case, scheme) has its
  own bound, three times the spread the oracle measures between the sequential order and the
  pairwise and canonical orders (tests/golden/reorder_bounds.json, written by
  tests/golden/make_reorder_bounds.py): from 3e-15 on the small cases to 2e-7 on the split scheme
  with a pressure history.
"""
import json
import os
import re
import ctypes

import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W
from tests.cases import SCHEMES, large_cases, small_cases

pytestmark = pytest.mark.gpu

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reorder_bounds.json")) as _f:
    REORDER_BOUNDS = json.load(_f)["cases"]
CODES = {"icpcg": E.PP_ICPCG, "ildlt_pcg": E.PP_ILDLT_PCG, "split_icpcg": E.PP_SPLIT_ICPCG,
         "split_ildlt_pcg": E.PP_SPLIT_ILDLT_PCG}
SMALL = small_cases()


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


KNIFE_EDGE = 1e-3


def _check_iterations(it_a, res_a, it_b, res_b, tol):
    """Iteration counts in two summation orders must agree, unless the order that stops first does
    so with a residual within KNIFE_EDGE (relative) of the tolerance: there the last iteration's
    residual is decided by rounding noise, and the other order needs exactly one more iteration
    (seen at 256x256 dr=100 icpcg: 293 iterations at 9.99956e-9 vs 294 sequentially)."""
    if it_a == it_b:
        return
    first_res = res_a if it_a < it_b else res_b
    assert abs(it_a - it_b) == 1, (it_a, it_b)
    assert tol * (1.0 - KNIFE_EDGE) <= first_res <= tol, (it_a, res_a, it_b, res_b)


def _check(engine, oracle_lib, case, scheme, name):
    try:
        p_seq, it_seq, res_seq = oracle_lib.pressure_solve(scheme, case, oracle_lib.SEQUENTIAL)
    except RuntimeError as err:
        with pytest.raises(E.EngineError, match=re.escape(str(err))):
            engine.solve(case, CODES[scheme])
        return
    p_gpu, rep = engine.solve(case, CODES[scheme])
    p_can, it_can, res_can = oracle_lib.pressure_solve(scheme, case, oracle_lib.CANONICAL)
    assert rep["iterations"] == it_can
    _check_iterations(it_can, res_can, it_seq, res_seq, case.tolerance)
    assert rep["relative_residual"] == res_can
    assert np.array_equal(p_gpu, p_can), f"max diff {np.abs(p_gpu - p_can).max():.3e}"
    assert _rel(p_gpu, p_seq) <= REORDER_BOUNDS[f"{name}/{scheme}"]["bound"]


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("name,case", SMALL, ids=[n for n, _ in SMALL])
def test_small_cases_bitwise(engine, oracle_lib, name, case, scheme):
    _check(engine, oracle_lib, case, scheme, name)


LARGE = large_cases()


@pytest.mark.parametrize("scheme", ["icpcg", "ildlt_pcg", "split_icpcg"])
@pytest.mark.parametrize("name,case", LARGE, ids=[n for n, _ in LARGE])
def test_large_cases_bitwise(engine, oracle_lib, name, case, scheme):
    if name == "192x192_split_history":
        # pressure history from two earlier solves, so the split explicit term is non-zero
        c1 = W.droplet_case(192, 192, seed=1)
        c2 = W.droplet_case(192, 192, seed=2)
        case.p_n, _ = engine.solve(c1, E.PP_ICPCG)
        case.p_nm1, _ = engine.solve(c2, E.PP_ICPCG)
    _check(engine, oracle_lib, case, scheme, name)


def test_factor_failure_message(engine):
    case = W.droplet_case(8, 8, seed=0, density_ratio=10.0)
    case.rho = case.rho.copy()
    case.rho[:3, :3] = -5.0
    with pytest.raises(E.EngineError, match="ICC preconditioner construction failed"):
        engine.solve(case, E.PP_ICPCG)
    with pytest.raises(E.EngineError, match="ILDLT preconditioner construction failed"):
        engine.solve(case, E.PP_ILDLT_PCG)


def test_silent_nonconvergence(engine, oracle_lib):
    case = W.droplet_case(40, 40, seed=0, density_ratio=100.0, max_iterations=1)
    p, rep = engine.solve(case, E.PP_ICPCG)
    assert rep["iterations"] == 1 and rep["relative_residual"] > case.tolerance
    p_can, _, _ = oracle_lib.pressure_solve("icpcg", case, oracle_lib.CANONICAL)
    assert np.array_equal(p, p_can)


def test_argument_errors(engine):
    case = W.droplet_case(600, 8, seed=0)
    with pytest.raises(E.EngineError, match="grid size"):
        engine.solve(case, E.PP_ICPCG)
    case = W.droplet_case(3, 8, seed=0, periodic_x=True)
    with pytest.raises(E.EngineError, match="periodic direction"):
        engine.solve(case, E.PP_ICPCG)
    case = W.droplet_case(40, 1100, seed=0, periodic_x=True)
    with pytest.raises(E.EngineError, match="a grid periodic in x takes ny 1..1024"):
        engine.solve(case, E.PP_ICPCG)


TALL = [("32x1100_walls", {"nx": 32, "ny": 1100}), ("70x1030_periodic_y", {"nx": 70, "ny": 1030, "periodic_y": True}),
        ("200x1500_walls", {"nx": 200, "ny": 1500})]


@pytest.mark.parametrize("scheme", ["icpcg", "ildlt_pcg"])
@pytest.mark.parametrize("name,kw", TALL, ids=[n for n, _ in TALL])
def test_grids_taller_than_1024(engine, oracle_lib, name, kw, scheme):
    """ny above 1024, which the engine takes where the grid is not periodic in x: bitwise equal to
    the oracle's replay of the engine order, the reference's iteration count, and within three
    times the summation-order spread of the reference-order result."""
    case = W.droplet_case(seed=3, density_ratio=10.0, **kw)
    p_gpu, rep = engine.solve(case, CODES[scheme])
    p_can, it_can, res_can = oracle_lib.pressure_solve(scheme, case, oracle_lib.CANONICAL)
    p_seq, it_seq, res_seq = oracle_lib.pressure_solve(scheme, case, oracle_lib.SEQUENTIAL)
    p_pw, _, _ = oracle_lib.pressure_solve(scheme, case, oracle_lib.PAIRWISE)
    assert rep["iterations"] == it_can and rep["relative_residual"] == res_can
    assert np.array_equal(p_gpu, p_can)
    _check_iterations(it_can, res_can, it_seq, res_seq, case.tolerance)
    assert _rel(p_gpu, p_seq) <= 3.0 * max(_rel(p_pw, p_seq), _rel(p_can, p_seq), 1.0e-15)


def test_tall_grid_mixed_precision(engine, oracle_lib):
    """The mixed-precision solve on a grid taller than 1024 rows meets the tolerance in the FP64
    residual of the oracle's assembled system (the solution itself may differ from the FP64 one by
    more than the tolerance: on a 200x1500 domain the smooth modes are poorly conditioned)."""
    case = W.droplet_case(200, 1500, seed=3, density_ratio=10.0)
    pm, rep = engine.solve(case, E.PP_ICPCG, precision=E.PP_MIXED)
    assert rep["relative_residual"] < case.tolerance
    s = oracle_lib.OracleSystem(case, use_split=False)
    r = s.rhs - s.apply_matrix(pm)
    r = r - r.mean()  # Neumann walls: the consistent part
    assert np.linalg.norm(r) / np.linalg.norm(s.rhs) < 2.0 * case.tolerance


@pytest.mark.parametrize("edge", [0, 1, 2, 3])
def test_periodic_edge_error_names_the_edge(engine, edge):
    """An explicit pressure BC on a periodic axis is rejected with the name of the first edge set, in
    the reference's check order left, right, bottom, top (operators.cpp:11-34)."""
    names = ("pressure_bc_left", "pressure_bc_right", "pressure_bc_bottom", "pressure_bc_top")
    bc = [W.BC_UNSET] * 4
    bc[edge] = W.BC_NEUMANN
    case = W.droplet_case(8, 8, seed=0, periodic_x=edge < 2, periodic_y=edge >= 2, bc_type=tuple(bc))
    with pytest.raises(E.EngineError, match=re.escape(f"{names[edge]} cannot be set when the direction is periodic")):
        engine.solve(case, E.PP_ICPCG)
    if edge in (1, 3):  # both edges of the axis set: the first one in the reference's order is named
        bc[edge - 1] = W.BC_DIRICHLET
        case = W.droplet_case(8, 8, seed=0, periodic_x=edge < 2, periodic_y=edge >= 2, bc_type=tuple(bc))
        with pytest.raises(E.EngineError, match=re.escape(names[edge - 1])):
            engine.solve(case, E.PP_ICPCG)


@pytest.mark.parametrize("nx,ny,precision", [(192, 192, E.PP_FP64), (48, 40, E.PP_FP64), (48, 40, E.PP_MIXED),
                                             (192, 192, E.PP_MIXED)])
def test_hand_off_watchdog_fails_fast_and_context_recovers(engine, oracle_lib, nx, ny, precision):
    """A wait limit of one sleep makes the first sweep wait that finds its counter behind give up
    (192x192: the factorization's hand-off between sweep waves; 48x40, one sweep wave: the first
    preconditioner application, where the sweep wave waits for its loaders). The solve must come back
    with PP_ERR_SWEEP quickly, and the next solve on the same context must be exact again."""
    import time
    case = W.droplet_case(nx, ny, seed=3, density_ratio=100.0)
    p0, rep0 = engine.solve(case, E.PP_ICPCG, precision)  # warm: layouts and buffers exist
    engine.set_wait_limit(1)
    try:
        t0 = time.perf_counter()
        with pytest.raises(E.EngineError, match="wavefront hand-off did not complete") as ei:
            engine.solve(case, E.PP_ICPCG, precision)
        elapsed = time.perf_counter() - t0
    finally:
        engine.set_wait_limit(0)
    assert ei.value.status == E.PP_ERR_SWEEP
    assert elapsed < 0.5, f"failing solve took {elapsed:.3f} s"
    p1, rep1 = engine.solve(case, E.PP_ICPCG, precision)
    assert rep1["iterations"] == rep0["iterations"] and np.array_equal(p1, p0)
    if precision == E.PP_FP64:
        p_can, it_can, _ = oracle_lib.pressure_solve("icpcg", case, oracle_lib.CANONICAL)
        assert rep1["iterations"] == it_can and np.array_equal(p1, p_can)


def test_pitched_host_buffers_match_contiguous(engine):
    """Ghosted Field2D buffers (ld = ny + 6) go straight through the host entry point."""
    case = W.droplet_case(37, 29, seed=4, density_ratio=50.0)
    p_ref, rep_ref = engine.solve(case, E.PP_ICPCG)
    nx, ny, g = case.nx, case.ny, 3
    ld = ny + 2 * g

    def ghosted(a, rows, cols):
        f = np.full((rows + 2 * g, cols + 2 * g), np.nan)
        f[g:g + rows, g:g + cols] = a.reshape(rows, cols)
        return f

    rho = ghosted(case.rho, nx, ny)
    u = ghosted(case.u, nx + 1, ny)
    v = ghosted(case.v, nx, ny + 1)
    out = np.full((nx + 2 * g, ny + 2 * g), -7.0)
    dp = lambda a, off: ctypes.cast(a.ctypes.data + 8 * off, ctypes.POINTER(ctypes.c_double))  # noqa: E731
    first = g * ld + g
    rep = E.PPReport()
    pb = E.make_problem(case, E.PP_ICPCG)
    rc = engine._lib.pp_solve(engine.h, ctypes.byref(pb), dp(rho, first), ld, dp(u, first), ld, dp(v, g * (ld + 1) + g),
                              ld + 1, None, ld, None, ld, dp(out, first), ld, ctypes.byref(rep))
    assert rc == 0, rep.message
    assert rep.iterations == rep_ref["iterations"]
    assert np.array_equal(out[g:g + nx, g:g + ny].reshape(-1), p_ref)
    assert np.all(out[:g] == -7.0) and np.all(out[:, :g] == -7.0)  # ghosts untouched


def test_device_entry_point(engine):
    torch = pytest.importorskip("torch")
    case = W.droplet_case(96, 80, seed=3, density_ratio=100.0)
    p_ref, rep_ref = engine.solve(case, E.PP_ICPCG)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(getattr(case, k)).reshape(-1)).to(dev) for k in ("rho", "u", "v")}
    out = torch.zeros(case.n, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    rep = engine.solve_device(case, t["rho"].data_ptr(), t["u"].data_ptr(), t["v"].data_ptr(), None, None,
                              out.data_ptr(), scheme=E.PP_ICPCG, stream=stream.cuda_stream)
    assert rep["iterations"] == rep_ref["iterations"]
    assert rep["kernel_ms"] > 0.0
    assert np.array_equal(out.cpu().numpy(), p_ref)
